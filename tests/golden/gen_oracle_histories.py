"""Write the oracle's long trajectories used by tests/test_gpu_long_runs.py (CPU, a few minutes).

Cases (the workloads of the benchmark configurations):
  * 440x82 Re=20 channel, zero start, dt=2e-3, all 2000 steps;
  * 1320x246 Re=100 channel, inlet-profile start with the seeded 1e-3 Gaussian v perturbation
    (seed 20261203), dt=5e-4, 500 steps.
Stored per case: time, C_D, C_L and max divergence of every step; the spread of C_D and C_L of every
step, |F - F'| with F' from a second oracle run whose Woodbury solves take six refinements past the
reference's stopping rule (``extra_refinements=6``: the residual stays at its round-off floor while
the solution moves along the operator's weakly determined directions, which measures how far the
rule determines each step's forces); of the final u, v, p four seeded
random projections (seed 70, 71, 72) with their absolute sums, for checks of the form
|w.x_gpu - w.x_oracle| <= tol * |w|.|x|; and the final u, v, p themselves on a block of every row and
column within 0.6 D of the cylinder centre and on a strided lattice of about 4000 points (rows and
columns stored as index arrays), with the field's max |.| over the whole grid, for max-norm checks
relative to it.

Usage:  python tests/golden/gen_oracle_histories.py
"""
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.ghost_oracle import CylinderFlow  # noqa: E402

OUT = Path(__file__).resolve().parent / 'oracle_histories.npz'
CASES = {
    'channel440x82': (dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5,
                           initial='zero'), 2e-3, 2000),
    'channel1320x246': (dict(nx=1320, ny=246, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.01, inflow='parabolic', U_m=1.5,
                             initial='profile', v_perturbation=dict(seed=20261203)), 5e-4, 500),
}


def projections(x, seed):
    w = np.random.default_rng(seed).normal(size=(4, x.size))
    return w @ x.ravel(), abs(w) @ abs(x.ravel())


def sample_blocks(x, h_row, h_col, centre_row, centre_col):
    """(rows, cols) index arrays of the window block and of the lattice block of field x."""
    near = lambda n, h, c: np.flatnonzero(abs((np.arange(n) + 0.5) * h - c) <= 0.6)
    stride = max(1, int(np.sqrt(x.size / 4000)))
    return {'win': (near(x.shape[0], h_row, centre_row), near(x.shape[1], h_col, centre_col)),
            'lat': (np.arange(0, x.shape[0], stride), np.arange(0, x.shape[1], stride))}


def main():
    out = {}
    workers = len(os.sched_getaffinity(0))
    for name, (kw, dt, steps) in CASES.items():
        t0 = time.perf_counter()
        s = CylinderFlow(fft_workers=workers, **kw)
        hist = np.empty((steps, 4))
        for k in range(steps):
            s.step(dt)
            d = s.last_step_diagnostics
            hist[k] = d['time'], d['cd'], d['cl'], d['max_divergence']
            if k % 100 == 0:
                print(f'{name}: step {k} cd {d["cd"]:.10f} cl {d["cl"]:.3e}', flush=True)
        out[f'{name}/history'] = hist
        t1 = time.perf_counter()
        alt = CylinderFlow(fft_workers=workers, extra_refinements=6, **kw)
        spread = np.empty((steps, 2))
        for k in range(steps):
            _, cd, cl = alt.step(dt)
            spread[k] = abs(cd - hist[k, 1]), abs(cl - hist[k, 2])
        out[f'{name}/spread'] = spread
        print(f'{name}: spread run {time.perf_counter() - t1:.0f} s; largest C_D / C_L spread '
              f'{spread[:, 0].max():.2e} / {spread[:, 1].max():.2e} (steps {spread.argmax(axis=0)}); from step 10 '
              f'{spread[10:].max(axis=0)}; final field moves u {np.max(abs(alt.u - s.u)):.2e} '
              f'v {np.max(abs(alt.v - s.v)):.2e} p {np.max(abs(alt.p - s.p)):.2e}', flush=True)
        out[f'{name}/dt'] = np.array([dt])
        for key, seed in (('u', 70), ('v', 71), ('p', 72)):
            pr, sc = projections(getattr(s, key), seed)
            out[f'{name}/{key}_proj'] = pr
            out[f'{name}/{key}_scale'] = sc
            x = getattr(s, key)
            for block, (rows, cols) in sample_blocks(x, s.dy, s.dx, s.yc, s.xc).items():
                out[f'{name}/{key}_{block}_rows'], out[f'{name}/{key}_{block}_cols'] = rows, cols
                out[f'{name}/{key}_{block}'] = x[np.ix_(rows, cols)]
            out[f'{name}/{key}_max'] = np.array([np.max(abs(x))])
        print(f'{name}: {steps} steps in {time.perf_counter() - t0:.0f} s', flush=True)
    np.savez_compressed(OUT, **out)
    print(f'wrote {OUT} ({OUT.stat().st_size / 1e3:.0f} kB)')


if __name__ == '__main__':
    main()
