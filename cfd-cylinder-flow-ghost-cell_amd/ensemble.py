"""Batched ensemble of uneven cylinder cases stepping together on one GPU.

Every stage of the AB2/CN step is one kernel launch over all cases (cases index the
second grid dimension) and steady steps replay one HIP graph for the whole ensemble
(``mcf_batch_*`` in include/ghostflow.h).  Host tables of the cases are built in a
process pool.  The reference runs one case per process; this module has no
counterpart there beyond the per-case construction it reuses (solver/main.py:40-242).
"""
import ctypes
import os
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
    0.5 in main.py) near 0.3 from the projected start."""
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


def build_flows(cases, workers=None, initial='profile'):
    """One device-resident ``MAC_CylinderFlow`` per case, its Helmholtz factors built for its dt;
    the host tables are prepared in a process pool."""
    cases = list(cases)
    workers = workers or min(16, len(os.sched_getaffinity(0)), len(cases))
    if workers > 1:
        with ProcessPoolExecutor(max_workers=workers) as pool:
            prepared = list(pool.map(prepare, cases))
    else:
        prepared = [prepare(c) for c in cases]
    flows = []
    for case, prep in zip(cases, prepared):
        f = MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, initial=initial,
                             prebuilt=prep, **solver_kwargs(case))
        f._push()
        flows.append(f)
    return flows


class Ensemble:
    """A list of cases advanced together by ``mcf_batch_run``."""

    def __init__(self, cases, workers=None, initial='profile', flows=None):
        self.cases = list(cases)
        self.flows = flows if flows is not None else build_flows(self.cases, workers, initial)
        self.cells = np.array([c['nx'] * c['ny'] for c in self.cases])
        lib = self._lib = nat.library()
        arr = (ctypes.c_void_p * len(self.flows))(*[f.gg.device.h.h.value for f in self.flows])
        self.dts = np.array([c['dt'] for c in self.cases], dtype=np.float64)
        b = ctypes.c_void_p()
        rc = lib.mcf_batch_create(arr, nat.dptr(self.dts), len(self.flows), ctypes.byref(b))
        if rc:
            # a failed create returns no batch; the message is on the first case's handle
            raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(lib.mcf_last_error(self.flows[0].gg.device.h.h).decode())
        self.b = b

    def run(self, nsteps):
        """Advance every case by nsteps at its own dt; returns diagnostics (cases, nsteps, 32)."""
        diag = np.zeros((len(self.flows), nsteps, nat.DIAG_LEN))
        rc = self._lib.mcf_batch_run(self.b, int(nsteps), nat.dptr(diag))
        if rc:
            raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(self._lib.mcf_batch_last_error(self.b).decode())
        for f, d in zip(self.flows, diag):
            f.time = float(d[-1, 0])
            f.last_step_diagnostics = diagnostics_from_row(d[-1])
        return diag

    def pull(self):
        for f in self.flows:
            f._pull()

    def close(self):
        if getattr(self, 'b', None) is not None and self.b.value:
            self._lib.mcf_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
