"""Benchmark: cell-steps/s of the ghost-cell AB2/CN step on one MI355X.

Workload (one step = one full AB2/CN fractional step of one case): the channel
22D x 4.1D with a D=1 cylinder at (2D, 2D), h = D/60 -> 1320 x 246 pressure cells,
parabolic inflow U_m = 1.5, Re = 100 (nu = 0.01), dt = 5e-4, FP64, forces every step,
started from the inlet profile with a seeded 1e-3 Gaussian transverse v perturbation
(seed 20261203).  The state stays resident in HBM for the whole timed region.

Prints ONE JSON line.  ``--gpus N`` with N > 1 exits: the benchmark runs one GPU.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WORKLOAD = dict(nx=1320, ny=246, Lx=22., Ly=4.1, D=1., xc=2., yc=2., nu=0.01, rho=1., inflow='parabolic', U_m=1.5)
DT = 5e-4
SEED = 20261203
HBM_PEAK = 8.0e12          # B/s
FP64_PEAK = 78.6e12        # FLOP/s
STEP_BYTES_PER_CELL = 928  # algorithmic HBM bytes per cell-step of the whole AB2/CN step
REFINE_BYTES_PER_CELL = 96  # one executed refinement pass of a Woodbury solve: 12 field sweeps
# HBM bytes per forward launch of the pressure transform by tile variant (FETCH_SIZE x 2.000 +
# WRITE_SIZE x 1.000 from rocprofv3 --pmc, scaled by a 256 MiB copy; profiles/k_gemm_1320x246_traffic.md)
GEMM_TRAFFIC = {0: 10.262e6, 1: 9.697e6, 2: 10.009e6, 3: 9.688e6, 4: 9.657e6, 5: 9.754e6, 6: 9.732e6,
                8: 7.710e6, 9: 7.714e6}


BOX = dict(nx=800, ny=820, Lx=40., Ly=41., D=1., xc=10., yc=20.5, nu=0.01, rho=1., inflow='uniform', U_inf=1.)
BOX_DT = 2e-3


def box_leg(pkg, steps, warmup, launch=None):
    """The 40D x 41D box with uniform inflow U = 1 (800 x 820 cells, h = 0.05, Re = 100, dt = 2e-3, the
    reference's outlet) from the reference's default start, with the transient inflow perturbation on
    (amplitude 0.15 over 2.5 D/U, set on the device every step): warmup steps, then `steps` timed steps."""
    t0 = time.perf_counter()
    flow = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **BOX)
    dev = flow.gg.device
    dev.set_dt(BOX_DT)
    if launch:
        dev.set_launch(*launch)
    setup_s = time.perf_counter() - t0
    print(f'bench: 800x820 box ready in {setup_s:.1f} s; {warmup} + {steps} steps', file=sys.stderr, flush=True)
    flow.solve(n_steps=warmup, dt=BOX_DT, perturb_transient=True, save_artifacts=False, verbose=False)
    t0 = time.perf_counter()
    flow.solve(n_steps=steps, dt=BOX_DT, perturb_transient=True, save_artifacts=False, verbose=False)
    elapsed = time.perf_counter() - t0
    diag = flow.diagnostics_history
    cells = BOX['nx'] * BOX['ny']
    value = cells * steps / elapsed
    return dict(workload='box 40Dx41D, cylinder D=1 at (10D, 20.5D), 800x820 cells (h=0.05), uniform inflow U=1 '
                         'with the transient inflow perturbation, Re=100, dt=2e-3, FP64, reference outlet',
                value=value, unit='cell-steps/s', ms_per_step=elapsed / steps * 1e3, steps=steps, warmup=warmup,
                window=f'steps {warmup} .. {warmup + steps} from the default start', max_divergence=float(np.max(diag[:, 10])),
                step_roofline=step_roofline(dev, diag[-steps:], value),
                launch=dev.get_launch(), speculation=dev.get_speculation(), setup_seconds=setup_s)


def cpu_baseline(seconds=20.0, gmres_seconds=0.0, single_seconds=10.0):
    """The oracle on this host's cores, on a bounded sample of steps: its stencil loops and sparse
    products as C++/OpenMP loops (oracle/cpu_kernels.cpp, the reference's Numba prange loops, with
    results equal to the NumPy/SciPy ones bit for bit), scipy.fft with workers = the same thread
    count, the Schur LU through scipy.linalg.  Then the same with one thread (OpenMP and FFT), and,
    with gmres_seconds, a bounded sample on the oracle's GMRES backend.  Returns (baseline, gmres
    sample or None)."""
    from oracle.ghost_oracle import CylinderFlow
    cores = len(os.sched_getaffinity(0))
    cells = WORKLOAD['nx'] * WORKLOAD['ny']

    def sample(threads, budget, cap):
        s = CylinderFlow(initial='profile', v_perturbation=dict(seed=SEED), fft_workers=threads,
                         openmp_threads=threads, **WORKLOAD)
        s.step(DT)       # startup step builds the per-dt Helmholtz factors (setup, not timed)
        s.step(DT)
        times = []
        start = time.perf_counter()
        while time.perf_counter() - start < budget and len(times) < cap:
            t0 = time.perf_counter()
            s.step(DT)
            times.append(time.perf_counter() - t0)
        return s, float(np.median(times)), len(times), s.gg.kernels.threads

    # OpenMP threads sleep between the many short parallel regions of a step instead of spinning
    # against scipy.fft's workers (read when the kernels' OpenMP runtime starts)
    os.environ.setdefault('OMP_WAIT_POLICY', 'PASSIVE')
    s, med, n, threads = sample(cores, seconds, 400)
    runs = [dict(threads=threads, ms_per_step=med * 1e3, steps=n, value=cells / med)]
    if single_seconds:
        _, med1, n1, _ = sample(1, single_seconds, 100)
        runs.append(dict(threads=1, ms_per_step=med1 * 1e3, steps=n1, value=cells / med1))
    best = max(runs, key=lambda r: r['value'])
    base = dict(value=best['value'], unit='cell-steps/s', cores=best['threads'], kind='port',
                sample=f'median step time of the 1320x246 Re=100 case after 2 warm-up steps, the faster of '
                       f'{" and ".join(str(r["threads"]) for r in runs)} thread(s): {best["steps"]} steps at '
                       f'{best["ms_per_step"]:.1f} ms (oracle/ghost_oracle.py with the OpenMP stencil and '
                       f'sparse-product loops of oracle/cpu_kernels.cpp, scipy.fft workers = OpenMP threads)',
                runs=runs)
    if not gmres_seconds:
        return base, None
    s.gg.set_pressure_solver('iterative')
    times, arnoldi = [], []
    start = time.perf_counter()
    while time.perf_counter() - start < gmres_seconds and len(times) < 50:
        t0 = time.perf_counter()
        s.step(DT)
        times.append(time.perf_counter() - t0)
        d = s.last_step_diagnostics
        arnoldi.append(d['pressure_iterations'] / d['pressure_solve_count'])
    med = float(np.median(times))
    return base, dict(ms_per_step=med * 1e3, value=cells / med, steps=len(times), cores=threads,
                      arnoldi_steps_per_solve=float(np.mean(arnoldi)),
                      sample=f'{len(times)} GMRES-backend steps of the same case after the baseline sample, median')


def gmres_leg(dev, steps):
    """The bench case on the GMRES pressure backend (pressure_solver='iterative': restarted GMRES on
    the pressure operator, left-preconditioned by the rectangle inverse, <= 2 refinements).  It
    continues from the plain leg's state; GMRES steps run eagerly (the host reads the solver state
    every 8 Arnoldi steps), so the wall clock around the synchronised run is the measure."""
    import importlib
    nat = importlib.import_module(type(dev).__module__.rsplit('.', 1)[0] + '._native')
    dev.set_pressure_backend(True)
    try:
        dev.run(2, DT)
        t0 = time.perf_counter()
        diag = dev.run(steps, DT)
        elapsed = time.perf_counter() - t0
    finally:
        dev.set_pressure_backend(False)
    col = nat.DIAG_NAMES.index('pressure_solve_count')
    its = [diag[s, nat.DIAG_SOLVE0 + nat.SOLVE_LEN * q] for s in range(steps) for q in range(int(diag[s, col]))]
    secs = [diag[s, nat.DIAG_SOLVE0 + nat.SOLVE_LEN * q + 8] for s in range(steps) for q in range(int(diag[s, col]))]
    cells = WORKLOAD['nx'] * WORKLOAD['ny']
    return dict(value=cells * steps / elapsed, unit='cell-steps/s', ms_per_step=elapsed / steps * 1e3, steps=steps,
                arnoldi_steps_per_solve=float(np.mean(its)), solves_per_step=len(its) / steps,
                ms_per_solve=float(np.mean(secs)) * 1e3, max_divergence=float(np.max(diag[:, 10])))


def ensemble_leg(pkg, steps, launch=None):
    """The seeded 64-case ensemble (uneven 330x64 .. 1320x246 channels, Re 40..200, zero start,
    dt = 0.05 h/U_m): aggregate cell-steps/s of the batched path (one launch per stage over all
    cases, graph-captured steps) against the same cases stepped one after another on the plain
    path.  Both legs start from fresh copies of the cases and time the same steps (3 .. 3+steps);
    members that fail a gate of the reference step are dropped and listed, and only members that
    ran every timed step are counted in either leg."""
    import importlib
    ens = importlib.import_module(pkg.__name__ + '.ensemble')
    cases = [ens.ensemble_case(k) for k in range(64)]
    t0 = time.perf_counter()
    batch = ens.Ensemble(cases)
    setup_s = time.perf_counter() - t0
    if launch:
        batch.set_launch(*launch)
    plain_flows = ens.build_flows(cases, failures={})
    batch.run(3)
    shapes = batch.get_launch()
    print(f'bench: ensemble of 64 cases ready in {setup_s:.1f} s; timing {steps} batched steps', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    d = batch.run(steps)
    tb = time.perf_counter() - t0
    spec = batch.get_speculation()
    batch.close()
    ran = [k for k in range(64) if k not in batch.failures and plain_flows[k] is not None]
    print(f'bench: timing {steps} plain steps of each of {len(ran)} cases', file=sys.stderr, flush=True)
    tp = 0.
    for k in ran:
        f, c = plain_flows[k], cases[k]
        f.gg.device.run(3, c['dt'])
        t0 = time.perf_counter()
        dp = f.gg.device.run(steps, c['dt'])
        tp += time.perf_counter() - t0
        assert np.max(dp[:, 10]) < 1e-12
        f.gg.device.h.close()
    cells = np.array([cases[k]['nx'] * cases[k]['ny'] for k in ran], dtype=float)
    # the batched time also covers the dropped members' steps up to their failure: it is not rescaled
    batched = cells.sum() * steps / tb
    plain = cells.sum() * steps / tp
    return dict(cases=64, cases_run=len(ran), failures={str(k): v['message'] for k, v in batch.failures.items()},
                cells=int(cells.sum()), grids='330x64 .. 1320x246, seeds 4000..4063', steps=steps,
                window='steps 3 .. 3+steps of both legs, fresh copies', batched_value=batched, plain_value=plain,
                ratio=batched / plain, max_divergence=float(np.nanmax(d[ran][:, :, 10])), setup_seconds=setup_s,
                unit='cell-steps/s', batched_speculation=spec, batched_launch=shapes)


def ensemble_long(pkg, total, chunk=500):
    """The 64-case ensemble for `total` steps in chunks, one progress line per chunk (stderr) and
    one JSON summary line (stdout)."""
    import importlib
    ens = importlib.import_module(pkg.__name__ + '.ensemble')
    cases = [ens.ensemble_case(k) for k in range(64)]
    t0 = time.perf_counter()
    b = ens.Ensemble(cases)
    setup_s = time.perf_counter() - t0
    done, cell_steps, wall, chunks = 0, 0., 0., []
    while done < total:
        act = b.active
        t = time.perf_counter()
        d = b.run(min(chunk, total - done))
        dt = time.perf_counter() - t
        n = d.shape[1]
        done += n
        survivors = b.active
        cells = sum(cases[k]['nx'] * cases[k]['ny'] for k in survivors)
        cell_steps += cells * n
        wall += dt
        row = dict(steps=done, active=len(survivors), rate=cells * n / dt, max_div=float(np.nanmax(d[survivors][:, :, 10])),
                   max_cfl=float(np.nanmax(d[survivors][:, :, 11])), dropped=sorted(set(act) - set(survivors)))
        chunks.append(row)
        print('bench: ' + json.dumps(row), file=sys.stderr, flush=True)
    out = dict(workload='64-case seeded ensemble, zero start, dt = 0.05 h/U_m', steps=total, setup_seconds=setup_s,
               wall_seconds=wall, cell_steps_per_s=cell_steps / wall, cases_completed=len(b.active),
               failures={str(k): v for k, v in sorted(b.failures.items())}, chunks=chunks)
    print(json.dumps(out))


COPY_BYTES = 256 << 20


def pin(dev, launch=None, classes=None):
    if launch:
        dev.set_launch(*launch)
    if classes:
        dev.set_launch_classes(dict(gemm_pressure=classes[0], gemm_velocity=classes[1], tri_pressure=classes[2],
                                    tri_velocity=classes[3]))


def traffic_probe(pkg, steps, launch=None, classes=None):
    """Counter calibration and the measured kernels in one process: a plain device copy of
    COPY_BYTES (reads and writes COPY_BYTES each), then `steps` steps of the 1320x246 workload
    (with the transform tile / sweep width pinned to `launch` when given)."""
    import torch
    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    flow = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, initial='profile',
                                v_perturbation=dict(seed=SEED), **WORKLOAD)
    pin(flow.gg.device, launch, classes)
    flow._push()
    flow.gg.device.run(steps, DT)
    print(f'bench: traffic probe ran a {COPY_BYTES} B copy and {steps} steps', file=sys.stderr, flush=True)


# Algorithmic HBM bytes of each kernel of the lean 1320x246 step, in launch order: every input read
# once and every output written once (N = 1320 x 246 cells; "vel" = the u and v arrays, 651,006
# doubles; "p" = one pressure-sized array, 324,720 doubles; V = the y eigenbasis, yy^2 doubles per
# kind; LU = the x tridiagonal factors, two yy x xx arrays per kind; r = the capacitance rank, 340
# for the pressure and 336 for each velocity problem; npts = r).
_VEL, _P = 651006 * 8, 324720 * 8
_VV = (246 ** 2 + 245 ** 2) * 8                      # V of the u and v problems
_LUV = 2 * (246 + 245) * 1320 * 8                    # LU factors of the u and v problems
_LUP = 2 * 246 * 1320 * 8
_PTS_P, _PTS_V = 2 * 246 * 340 * 8, 2 * (246 + 245) * 336 * 8     # V rows at the points + F gathered there
_CAP_P, _CAP_V = 340 * 340 * 8, 2 * 336 * 336 * 8                 # S^-1 C
STEP_KERNEL_BYTES = [
    ('k_adv_rhs', 'advection + AB2 + CN right-hand side: reads u, v, the half pressure and the previous '
                  'advection, writes the advection and the rhs', 4 * _VEL + _P),
    ('velocity transform, forward', 'reads rhs, V_u, V_v; writes F', 2 * _VEL + _VV),
    ('velocity sweep, mode 0', 'reads F and LU, writes F', 2 * _VEL + _LUV),
    ('velocity points', 'V rows and F at the capacitance points', _PTS_V),
    ('velocity capacitance', 'S^-1 C product', _CAP_V),
    ('velocity sweep, mode 1', 'reads F and LU, writes F', 2 * _VEL + _LUV),
    ('velocity transform, inverse', 'reads F, V^T; writes u*, v*', 2 * _VEL + _VV),
    ('Helmholtz statistics + divergence', 'reads u*, v*, rhs; writes div', 2 * _VEL + _P),
    ('pressure transform, forward', 'reads div, V; writes F', 2 * _P + 246 ** 2 * 8),
    ('pressure sweep, mode 0', 'reads F and LU, writes F', 2 * _P + _LUP),
    ('pressure points', '', _PTS_P),
    ('pressure capacitance', '', _CAP_P),
    ('pressure sweep, mode 1', 'reads F and LU, writes F', 2 * _P + _LUP),
    ('pressure transform, inverse', 'reads F, V^T; writes phi', 2 * _P + 246 ** 2 * 8),
    ('solve statistics + projection update', 'reads phi, div, u*, v*; writes u, v', 2 * _P + 2 * _VEL),
    ('gate 1 + momentum defect', 'reads u, v, rhs; writes defect, projected field', 4 * _VEL),
    ('divergence', 'reads the projected field, writes div', _VEL + _P),
    ('pressure transform, forward', '', 2 * _P + 246 ** 2 * 8),
    ('pressure sweep, mode 0', '', 2 * _P + _LUP),
    ('pressure points', '', _PTS_P),
    ('pressure capacitance', '', _CAP_P),
    ('pressure sweep, mode 1', '', 2 * _P + _LUP),
    ('pressure transform, inverse', '', 2 * _P + 246 ** 2 * 8),
    ('solve statistics + projection update', '', 2 * _P + 2 * _VEL),
    ('gate 2 + pressures + ghost fill', 'reads u, v, half pressure, phi; writes half and output pressure',
     _VEL + 4 * _P),
    ('k_forces', 'surface stencils (756 points); one block gathers and clears the reduction slots',
     756 * 40 * 12 + 2 * 64 * 64 * 8),
    ('k_finalize', 'force partial sums into the diagnostics row', 189 * 4 * 8),
]


def step_table(trace_dir, fetch_dir, write_dir, last=40):
    """Markdown table of the lean step's kernels from three rocprofv3 CSV runs of the same
    `--traffic-probe` command (kernel trace, --pmc FETCH_SIZE, --pmc WRITE_SIZE): per launch position
    the mean duration and scaled traffic over the last `last` steps, against STEP_KERNEL_BYTES."""
    import csv
    import re

    def rows(path, value):
        out = list(csv.DictReader(open(path)))
        out.sort(key=lambda r: int(r['Dispatch_Id']))
        name = lambda r: re.sub(r'\(anonymous namespace\)::', '', r['Kernel_Name']).split('(')[0].replace('void ', '')
        return [(name(r), value(r)) for r in out]

    trace = rows(Path(trace_dir) / 'run_kernel_trace.csv',
                 lambda r: (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) / 1e3)
    fetch = rows(Path(fetch_dir) / 'run_counter_collection.csv', lambda r: float(r['Counter_Value']))
    write = rows(Path(write_dir) / 'run_counter_collection.csv', lambda r: float(r['Counter_Value']))
    copy_kib = COPY_BYTES / 1024
    sf = copy_kib / next(v for n, v in fetch if 'copyBuffer' in n)
    sw = copy_kib / next(v for n, v in write if 'copyBuffer' in n)

    def per_position(seq):
        starts = [i for i, (n, _) in enumerate(seq) if n == 'k_adv_rhs']
        spans = list(zip(starts, starts[1:]))[-last:]
        width = spans[0][1] - spans[0][0]
        if any(b - a != width for a, b in spans) or width != len(STEP_KERNEL_BYTES):
            raise RuntimeError('the timed steps are not all lean steps of the expected shape')
        names = [n for n, _ in seq[spans[0][0]:spans[0][1]]]
        vals = np.array([[v for _, v in seq[a:b]] for a, b in spans])
        return names, vals.mean(axis=0)

    names, dur = per_position(trace)
    names_f, fet = per_position(fetch)
    names_w, wri = per_position(write)
    if not names == names_f == names_w:
        raise RuntimeError('the three runs launched different kernels (pin --launch-classes)')
    lines = [f'Counter scale factors from the {COPY_BYTES >> 20} MiB copy: FETCH_SIZE x {sf:.3f}, WRITE_SIZE x {sw:.3f}.',
             '', '| # | kernel | stage | us | fetched MB | written MB | traffic GB/s | algorithmic MB | algorithmic GB/s | % of 8 TB/s |',
             '|---|---|---|---|---|---|---|---|---|---|']
    tot = [0.0] * 4
    for k, (n, (stage, _, alg)) in enumerate(zip(names, STEP_KERNEL_BYTES)):
        if n.startswith('k_gemm_par'):   # the parity-split transform reads half of V
            alg -= (_VV if 'velocity' in stage else 246 ** 2 * 8) / 2
        f, w = fet[k] * sf * 1024 / 1e6, wri[k] * sw * 1024 / 1e6
        t = dur[k]
        lines.append(f'| {k} | `{n}` | {stage} | {t:.2f} | {f:.2f} | {w:.2f} | {(f + w) / t * 1e3:.0f} | '
                     f'{alg / 1e6:.2f} | {alg / t / 1e3:.0f} | {alg / t / 1e3 / 80:.1f} |')
        tot = [tot[0] + t, tot[1] + f, tot[2] + w, tot[3] + alg / 1e6]
    lines.append(f'| | **step** | sum of the launches | {tot[0]:.1f} | {tot[1]:.1f} | {tot[2]:.1f} | '
                 f'{(tot[1] + tot[2]) / tot[0] * 1e3:.0f} | {tot[3]:.1f} | {tot[3] / tot[0] * 1e3:.0f} | '
                 f'{tot[3] / tot[0] * 1e3 / 80:.1f} |')
    return '\n'.join(lines)


def step_roofline(dev, diag, rate):
    """Whole-step roofline of a timed window: the 928 B/cell-step of a steady step, and beside it the
    bytes the window executed, each refinement pass of a Woodbury solve adding its two rectangle
    inverses (12 field sweeps, 96 B per cell; a velocity component counts as one solve)."""
    import importlib
    nat = importlib.import_module(type(dev).__module__.rsplit('.', 1)[0] + '._native')
    passes = (diag[:, nat.DIAG_NAMES.index('pressure_refinement_iterations')]
              + diag[:, nat.DIAG_NAMES.index('helmholtz_refinement_iterations')])
    per_step = float(np.mean(passes)) if len(passes) else 0.0
    executed = STEP_BYTES_PER_CELL + REFINE_BYTES_PER_CELL * per_step
    return dict(bound='hbm', bytes_per_cell_step=STEP_BYTES_PER_CELL, achieved=rate * STEP_BYTES_PER_CELL / 1e9,
                peak=HBM_PEAK / 1e9, unit='GB/s', frac=rate * STEP_BYTES_PER_CELL / HBM_PEAK,
                refinement_passes_per_step=per_step, executed_bytes_per_cell_step=executed,
                executed_frac=rate * executed / HBM_PEAK)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2000)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-seconds', type=float, default=20.0)
    ap.add_argument('--no-ensemble', action='store_true')
    ap.add_argument('--box-steps', type=int, default=100,
                    help='timed steps of the 800x820 uniform-inflow box leg (0: skip)')
    ap.add_argument('--ensemble-steps', type=int, default=200)
    ap.add_argument('--no-warm-start', action='store_true')
    ap.add_argument('--no-kernel-timing', action='store_true',
                    help='skip the HIP-event timing of the dominant kernels (roofline becomes null)')
    ap.add_argument('--traffic-probe', action='store_true',
                    help='for rocprofv3 --pmc runs: one 256 MiB device copy (the counter calibration), then '
                         '--steps steps of the workload; prints nothing else')
    ap.add_argument('--launch', type=int, nargs=2, default=None, metavar=('TILE', 'WIDTH'),
                    help='pin the transform tile and sweep width variants (default: autotuned)')
    ap.add_argument('--step-table', nargs=3, default=None, metavar=('TRACE_DIR', 'FETCH_DIR', 'WRITE_DIR'),
                    help='print the per-kernel table of the lean step from three rocprofv3 CSV runs of '
                         '--traffic-probe (kernel trace, --pmc FETCH_SIZE, --pmc WRITE_SIZE) and exit')
    ap.add_argument('--launch-classes', type=int, nargs=4, default=None, metavar=('GP', 'GV', 'TP', 'TV'),
                    help='pin the transform tile and sweep width per problem class (pressure, velocity)')
    ap.add_argument('--ensemble-launch', type=int, nargs=2, default=None, metavar=('TILE', 'WIDTH'),
                    help="pin the ensemble batch's transform tile and sweep width (default: autotuned)")
    ap.add_argument('--gmres-steps', type=int, default=20,
                    help='steps of the GMRES pressure backend leg (0: skip)')
    ap.add_argument('--ensemble-long', type=int, default=0,
                    help='only run the 64-case ensemble for this many steps and print its summary')
    ap.add_argument('--box-probe', type=int, default=0,
                    help='for rocprofv3 runs: only this many steps of the 800x820 box (after 20 warm-up steps)')
    ap.add_argument('--warm-start-workload', action='store_true',
                    help='only run the learned warm start on its benchmark workload (about 200k training pairs '
                         'from the three channel / box cases, the 64 ensemble geometries at Re ~ U[200, 250]) '
                         'and print its summary')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported', file=sys.stderr)
        sys.exit(2)
    if args.step_table:
        print(step_table(*args.step_table))
        return
    import cfd_amd_loader
    pkg = cfd_amd_loader.load()
    if args.ensemble_long:
        ensemble_long(pkg, args.ensemble_long)
        return
    if args.warm_start_workload:
        import importlib
        ws = importlib.import_module(pkg.__name__ + '.warmstart')
        res = ws.config5_study(log=lambda m: print('bench: ' + m, file=sys.stderr, flush=True))
        print(json.dumps(res))
        return
    if args.traffic_probe:
        traffic_probe(pkg, args.steps, args.launch, args.launch_classes)
        return
    if args.box_probe:
        print(json.dumps(box_leg(pkg, args.box_probe, 20, args.launch)))
        return
    t0 = time.perf_counter()
    flow = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, initial='profile',
                                v_perturbation=dict(seed=SEED), **WORKLOAD)
    dev = flow.gg.device
    flow._push()
    dev.set_dt(DT)
    pin(dev, args.launch, args.launch_classes)
    setup_s = time.perf_counter() - t0
    print(f'bench: setup {setup_s:.2f} s; warmup {args.warmup} steps', file=sys.stderr, flush=True)
    if args.warmup:
        dev.run(args.warmup, DT)
    print(f'bench: timing {args.steps} steps', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    diag = dev.run(args.steps, DT)     # synchronises before returning
    elapsed = time.perf_counter() - t0
    cells = WORKLOAD['nx'] * WORKLOAD['ny']
    value = rate = cells * args.steps / elapsed
    max_div = float(np.max(diag[:, 10]))
    if not max_div < 1e-12:
        # the metric is quoted at div < 1e-12: no value is reported for a run that misses it
        print(f'bench: max divergence {max_div:.3e} >= 1e-12 in the timed steps; value withheld', file=sys.stderr)
        value = None
    # dominant kernels timed with HIP events on the library's stream
    launch = dev.get_launch()
    tile = launch['gemm_pressure']
    lean_misses = None
    if args.no_kernel_timing:
        gemm_ms = pcr_ms = step_ms = gemm_isolated_ms = None
    else:
        prof = dev.profile(20, DT)
        gemm_ms, pcr_ms, step_ms = prof[0], prof[1], prof[2]
        gemm_isolated_ms = prof[3]
        lean_misses = int(prof[4])
    yy, xx = WORKLOAD['ny'], WORKLOAD['nx']
    # the parity-split transform (tiles 8, 9) computes two half-size products: yy^2 xx flops and half of
    # V read, against 2 yy^2 xx and all of V for the dense tiles
    par = tile in (8, 9)
    gemm_flops = (1.0 if par else 2.0) * yy * yy * xx
    gemm_achieved = gemm_flops / (gemm_ms * 1e-3) if gemm_ms else None
    out = dict(metric='cell-steps/s on 1 MI355X (plain vs batched+NN warm-start) at div<1e-12', value=value,
               unit='cell-steps/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=elapsed / args.steps * 1e3, higher_is_better=True, scaling='weak', vs_baseline=None,
               dtype='f64', data='synthetic',
               config=dict(workload='channel 22Dx4.1D, cylinder D=1 at (2D,2D), 1320x246 cells (h=D/60), '
                                    'parabolic Um=1.5, Re=100, dt=5e-4, FP64, plain ghost AB2/CN step, forces every step',
                           path='plain', cells=cells, max_divergence=max_div),
               roofline=None if gemm_achieved is None else dict(
                   bound='mfma', achieved=gemm_achieved / 1e12, peak=FP64_PEAK / 1e12, unit='TFLOP/s',
                   frac=gemm_achieved / FP64_PEAK, traffic=GEMM_TRAFFIC.get(tile),
                   traffic_unit='bytes per launch', algorithmic_bytes=8.0 * ((0.5 if par else 1.0) * yy * yy + 2 * yy * xx),
                   tile=tile, kernel=('k_gemm_par (parity-split FP64 y-eigenbasis transform, two 123x123 by 123x1320 products)'
                                      if par else ('k_gemm_sk' if tile >= 3 else 'k_gemm')
                                      + ' (dense FP64 y-eigenbasis transform, 246x246 by 246x1320)'),
                   launch_ms=gemm_ms),
               step_roofline=step_roofline(dev, diag, rate),
               kernel_ms=dict(gemm=gemm_ms, gemm_isolated=gemm_isolated_ms, pcr=pcr_ms, step=step_ms,
                              lean_misses=lean_misses), launch=launch,
               speculation=dev.get_speculation(), setup_seconds=setup_s)
    if args.box_steps:
        out['box'] = box_leg(pkg, args.box_steps, 20, args.launch)
    if args.gmres_steps:
        print(f'bench: GMRES pressure backend, {args.gmres_steps} steps', file=sys.stderr, flush=True)
        out['pressure_iterative'] = gmres_leg(dev, args.gmres_steps)
    out['ensemble'] = None if args.no_ensemble else ensemble_leg(pkg, args.ensemble_steps, args.ensemble_launch)
    if not args.no_warm_start:
        import importlib
        ws = importlib.import_module(pkg.__name__ + '.warmstart')
        print('bench: warm-start study (pair dump, device training, GMRES on held-out channels)', file=sys.stderr, flush=True)
        res = ws.study(log=lambda m: print('bench: ' + m, file=sys.stderr, flush=True))
        conv_flops = 2.0 * 64 * 340 * 64 * 64 * 5
        out['warm_start'] = dict(
            training_pairs=res['training_pairs'], training_steps=res['training_steps'],
            training_seconds=res['training_seconds'], loss_first=res['loss_first'], loss_last=res['loss_last'],
            gmres_rtol=1e-12, heldout_cases=[dict(group=h['group'], case=h['case'], rank=h['rank']) for h in res['heldout']],
            heldout_failed=res['failed'],
            guesses={grp: {g: dict(mean_arnoldi_steps=v['mean_iterations'], median_initial_residual=v['median_initial_residual'],
                                   max_relres=v['max_relres'], max_error_vs_direct=v['max_error_vs_direct'], systems=v['systems'])
                           for g, v in gs.items()} for grp, gs in res['summary'].items()},
            conv_roofline=dict(bound='mfma', kernel='k_conv (BF16 16x16x16 MFMA, 64 -> 64 channels, 5 taps, 21760 positions)',
                               launch_ms=res['conv_ms_per_launch'],
                               achieved=conv_flops / (res['conv_ms_per_launch'] * 1e-3) / 1e12, peak=2500.0,
                               unit='TFLOP/s', frac=conv_flops / (res['conv_ms_per_launch'] * 1e-3) / 2.5e15),
            seconds=res['seconds'])
    if not args.no_cpu_baseline:
        print('bench: CPU baseline (oracle)', file=sys.stderr, flush=True)
        out['cpu_baseline'], cpu_gmres = cpu_baseline(args.cpu_seconds, 15.0 if args.gmres_steps else 0.0)
        if cpu_gmres is not None:
            out['pressure_iterative']['cpu'] = cpu_gmres
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out))


if __name__ == '__main__':
    main()
