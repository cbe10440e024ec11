"""Batched ensemble of uneven cylinder cases stepping together on one GPU.

Every stage of the AB2/CN step is one kernel launch over all cases (cases index the
second grid dimension) and steady steps replay one HIP graph for the whole ensemble
(``mcf_batch_*`` in include/ghostflow.h).  Host tables of the cases are built in a
process pool.  The reference runs one case per process; this module has no
counterpart there beyond the per-case construction it reuses (solver/main.py:40-242).
"""
import ctypes
import multiprocessing
import os
import signal
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import _native as nat
from .device import prepare_case
from .flow import MAC_CylinderFlow, diagnostics_from_row
from .tables import GhostOperators

U_M = 1.5


def channel_case(m, d_over_h, delta, re, nx_range=(330, 1320), ny_range=(62, 246), courant=0.05):
    """Channel 22D long and ~D/(D/H) high on h = D/m cells (D = 1), cylinder at (2D, H/2 + delta H),
    parabolic inflow U_m = 1.5 (mean 1.0, so nu = 1/Re), dt = courant h / U_m; the grid is clamped
    to nx_range x ny_range.  courant = 0.05 is the dt U_m / h of the plain channel cases (0.06 at
    h = D/20, 0.045 at h = D/60) and keeps the step's CFL number (dt (max|u|/dx + max|v|/dy), gate
    0.5 in main.py) below 0.3.  The cases start from the zero field (inlet column at the profile,
    then projected): the impulsive start from the profile copied across the channel drives the
    centred AB2 advection unstable on the coarse, high-Re members (at a fixed physical time,
    whatever dt)."""
    D = 1.0
    h = D / m
    nx = int(np.clip(round(22 * D / h), *nx_range))
    ny = int(np.clip(round((D / d_over_h) / h), *ny_range))
    Lx, Ly = nx * h, ny * h
    u_ref = 2. / 3. * U_M
    return dict(nx=nx, ny=ny, Lx=Lx, Ly=Ly, D=D, xc=2 * D, yc=Ly / 2 + delta * Ly, nu=u_ref * D / re, rho=1.,
                inflow='parabolic', U_m=U_M, dt=courant * h / U_M, Re=re)


def ensemble_case(k, seed_base=4000, re_range=(40., 200.)):
    """Case k of the seeded ensemble: Re ~ U[40,200], D/H ~ U[0.15,0.30], delta ~ U[-0.1,0.1],
    m ~ U{15..60} (h = D/m), drawn from default_rng(seed_base + k)."""
    rng = np.random.default_rng(seed_base + k)
    re = rng.uniform(*re_range)
    d_over_h = rng.uniform(0.15, 0.30)
    delta = rng.uniform(-0.1, 0.1)
    m = int(rng.integers(15, 61))
    return channel_case(m, d_over_h, delta, re)


def _inlet(case):
    """The facade's parabolic inlet profile, evaluated in the same operation order."""
    y = (np.arange(case['ny']) + 0.5) * (case['Ly'] / case['ny'])
    return 4.0 * case['U_m'] * y * (case['Ly'] - y) / (case['Ly']**2)


def prepare(case):
    """All host-side tables of one case as recorded uploads (run in a worker process)."""
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        ops = GhostOperators(case['nx'], case['ny'], case['Lx'], case['Ly'], case['D'], case['xc'], case['yc'],
                             case['nu'], case['rho'], _inlet(case))
        return prepare_case(ops, _inlet(case), 2. / 3. * case['U_m'], case['dt'])


def solver_kwargs(case):
    return {k: case[k] for k in ('nx', 'ny', 'Lx', 'Ly', 'D', 'xc', 'yc', 'nu', 'rho', 'inflow', 'U_m')}


def _worker_pool(workers):
    """A process pool whose workers do not inherit this process's GPU runtime: they fork from a
    fresh forkserver process, which loads the package by path (``cfd_amd_loader``).  None when the
    package was not loaded through that module."""
    loader = sys.modules.get('cfd_amd_loader')
    main = sys.modules.get('__main__')
    main_file = getattr(main, '__file__', None)
    if loader is None or (getattr(main, '__spec__', None) is None and not (main_file and os.path.isfile(main_file))):
        return None   # workers re-import the main module; without a file to import, prepare serially
    if signal.getsignal(signal.SIGINT) is None or signal.getsignal(signal.SIGCHLD) is None:
        return None   # a handler installed outside Python (a profiler's): the forkserver cannot restore it
    ctx = multiprocessing.get_context('forkserver')
    ctx.set_forkserver_preload(['cfd_amd_loader'])
    return ProcessPoolExecutor(max_workers=workers, mp_context=ctx, initializer=loader.load)


def build_flows(cases, workers=None, initial='zero', failures=None):
    """One device-resident ``MAC_CylinderFlow`` per case, its Helmholtz factors built for its dt;
    the host tables are prepared in a process pool.  With a ``failures`` dict, a case whose start
    fails the reference's projection gate is recorded there (index -> message) and given None;
    without one the error propagates."""
    cases = list(cases)
    workers = workers or min(16, len(os.sched_getaffinity(0)), len(cases))
    pool = _worker_pool(workers) if workers > 1 else None
    if pool is not None:
        with pool:
            prepared = list(pool.map(prepare, cases))
    else:
        prepared = [prepare(c) for c in cases]
    flows = []
    for k, (case, prep) in enumerate(zip(cases, prepared)):
        try:
            f = MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, initial=initial,
                                 prebuilt=prep, **solver_kwargs(case))
        except RuntimeError as e:
            if failures is None:
                raise
            failures[k] = dict(step=0, message=f'start: {e}')
            flows.append(None)
            continue
        f._push()
        flows.append(f)
    return flows


def _status_message(row, step):
    st = int(row[26])
    what = ('non-finite force' if st & 16 else 'CN momentum solve' if st & 1 else 'pressure solve' if st & 2
            else 'projection' if st & 4 else 'divergence/CFL gate')
    return f'{what} failed at step {step}: div={row[10]:.3g} cfl={row[11]:.3g}'


class Ensemble:
    """Cases advanced together by ``mcf_batch_run`` (one launch per stage over all cases).

    A case that fails a gate of the reference step (main.py:414-426) is dropped, with its step and
    reason in ``failures``, and the others continue; ``drop_failed=False`` raises instead."""

    def __init__(self, cases, workers=None, initial='zero', flows=None, drop_failed=True):
        self.cases = list(cases)
        self.drop_failed = drop_failed
        self.failures = {}
        if flows is None:
            flows = build_flows(self.cases, workers, initial, failures=self.failures if drop_failed else None)
        self.flows = flows
        self.cells = np.array([c['nx'] * c['ny'] for c in self.cases])
        self.steps_done = 0
        self._lib = nat.library()
        self.b = None
        self._launch = None
        self._spec = None
        self._rebuild()

    @property
    def active(self):
        return [k for k, f in enumerate(self.flows) if f is not None and k not in self.failures]

    def _rebuild(self):
        self.close()
        idx = self.active
        if not idx:
            raise RuntimeError('no ensemble case left to run: ' + '; '.join(
                f'case {k}: {v["message"]}' for k, v in sorted(self.failures.items())))
        arr = (ctypes.c_void_p * len(idx))(*[self.flows[k].gg.device.h.h.value for k in idx])
        self.dts = np.array([self.cases[k]['dt'] for k in idx], dtype=np.float64)
        b = ctypes.c_void_p()
        rc = self._lib.mcf_batch_create(arr, nat.dptr(self.dts), len(idx), ctypes.byref(b))
        if rc:
            # a failed create returns no batch; the message is on the first case's handle
            raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(
                self._lib.mcf_last_error(self.flows[idx[0]].gg.device.h.h).decode())
        self.b, self._idx = b, idx
        if self._launch is not None:
            self._call('mcf_batch_set_launch', *self._launch)
        if self._spec is not None:
            self._call('mcf_batch_set_speculation', self._spec)

    def _call(self, name, *args):
        rc = getattr(self._lib, name)(self.b, *[nat.as_arg(a) for a in args])
        if rc:
            raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(self._lib.mcf_batch_last_error(self.b).decode())

    def set_launch(self, gemm=-1, tri=-1):
        """Pin the batch's transform tile / sweep width variants (-1, -1: autotune at the next run)."""
        self._launch = None if gemm < 0 else (int(gemm), int(tri))
        self._call('mcf_batch_set_launch', int(gemm), int(tri))

    def set_speculation(self, mode=1):
        """As GhostDevice.set_speculation, for the whole batch."""
        self._spec = int(mode)
        self._call('mcf_batch_set_speculation', self._spec)

    def get_speculation(self):
        out = np.zeros(2, dtype=np.int32)
        self._call('mcf_batch_get_speculation', out)
        return dict(mode=int(out[0]), replays=int(out[1]))

    def get_launch(self):
        out = np.zeros(5, dtype=np.int32)
        self._call('mcf_batch_get_launch', out)
        return dict(gemm_pressure=int(out[0]), gemm_velocity=int(out[1]), tri_pressure=int(out[2]),
                    tri_velocity=int(out[3]), chosen=bool(out[4]))

    def run(self, nsteps):
        """Advance every active case by nsteps at its own dt.  Returns diagnostics
        (n_cases, nsteps, 32) indexed by case; rows of inactive cases are NaN."""
        idx = self._idx
        diag = np.zeros((len(idx), nsteps, nat.DIAG_LEN))
        rc = self._lib.mcf_batch_run(self.b, int(nsteps), nat.dptr(diag))
        out = np.full((len(self.cases), nsteps, nat.DIAG_LEN), np.nan)
        out[idx] = diag
        if rc:
            msg = self._lib.mcf_batch_last_error(self.b).decode()
            bad = [i for i in range(len(idx)) if np.any(diag[i, :, 26] != 0)]
            if not self.drop_failed or not bad:
                raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(msg)
            for i in bad:
                s = int(np.flatnonzero(diag[i, :, 26] != 0)[0])
                self.failures[idx[i]] = dict(step=self.steps_done + s, message=_status_message(diag[i, s], self.steps_done + s))
                out[idx[i], s + 1:] = np.nan
            self._rebuild()
        self.steps_done += nsteps
        for k in idx:
            if k not in self.failures:
                self.flows[k].time = float(out[k, -1, 0])
                self.flows[k].last_step_diagnostics = diagnostics_from_row(out[k, -1])
        return out

    def pull(self):
        for k in self.active:
            self.flows[k]._pull()

    def close(self):
        if getattr(self, 'b', None) is not None and self.b.value:
            self._lib.mcf_batch_destroy(self.b)
        self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
