"""GPU parity at the production sizes, through the C ABI: the 1320x246 Re=100 channel of the bench
(inlet-profile start with the seeded 1e-3 Gaussian transverse perturbation, dt = 5e-4) and the
800x820 uniform-inflow box (40D x 41D, cylinder at (10D, 20.5D), h = 0.05, Re = 100, dt = 2e-3),
each against the oracle after 10 steps: fields within 1e-11 relative, C_D / C_L within 1e-10 of
the step's force magnitude max(|C_D|, |C_L|) at every step, divergence below 1e-12 at every step.
A checkpoint restart at the bench size continues bit for bit.
"""
import os

import numpy as np
import pytest

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    '1320x246': (dict(nx=1320, ny=246, Lx=22., Ly=4.1, D=1., xc=2., yc=2., nu=0.01, inflow='parabolic', U_m=1.5,
                      initial='profile', v_perturbation=dict(seed=20261203)), 5e-4),
    '800x820': (dict(nx=800, ny=820, Lx=40., Ly=41., D=1., xc=10., yc=20.5, nu=0.01, inflow='uniform', U_inf=1.0,
                     initial='profile'), 2e-3),
}
STEPS = 10


def flow(pkg, **kw):
    return pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)


def rel(a, b):
    return float(np.max(abs(np.asarray(a) - np.asarray(b))) / max(np.max(abs(np.asarray(b))), 1e-300))


@pytest.mark.parametrize('name', list(CASES))
def test_full_size_steps_match_oracle(pkg, name):
    kw, dt = CASES[name]
    print('full size', name, flush=True)
    ref = O.CylinderFlow(fft_workers=min(16, len(os.sched_getaffinity(0))), **kw)
    gpu = flow(pkg, **kw)
    assert rel(gpu.u, ref.u) < 1e-12 and rel(gpu.v, ref.v) < 1e-12     # initial projection
    rows = []
    for _ in range(STEPS):
        ref.step(dt)
        rows.append(dict(ref.last_step_diagnostics))
    t, cd, cl = gpu.solve(n_steps=STEPS, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    errs = dict(u=rel(gpu.u, ref.u), v=rel(gpu.v, ref.v), p=rel(gpu.p, ref.p))
    print(name, 'field errors after', STEPS, 'steps', errs, flush=True)
    assert max(errs.values()) < 1e-11, errs
    for k, r in enumerate(rows):
        scale = max(abs(r['cd']), abs(r['cl']))
        assert abs(cd[k] - r['cd']) <= 1e-10 * scale and abs(cl[k] - r['cl']) <= 1e-10 * scale, (k, cd[k], r['cd'])
        assert abs(t[k] - r['time']) <= 1e-12 * max(1.0, r['time'])
    div = gpu.diagnostics_history[:, 10]
    assert np.all(div < 1e-12), div.max()


def test_full_size_restart_is_exact(pkg, tmp_path):
    kw, dt = CASES['1320x246']
    a = flow(pkg, **kw)
    a.solve(n_steps=5, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    path = tmp_path / 'checkpoint.npz'
    a.save_checkpoint(path)
    b = pkg.MAC_CylinderFlow.from_checkpoint(path)
    for key in ('u', 'v', 'p'):
        assert np.array_equal(getattr(a, key), getattr(b, key)), key
    for f in (a, b):
        f.gg.device.set_launch(3, 1)     # the same launch shapes for both (sweep widths differ in round-off)
        f.solve(n_steps=4, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    for key in ('u', 'v', 'p'):
        assert np.array_equal(getattr(a, key), getattr(b, key)), key
    assert np.array_equal(a.diagnostics_history[:, 2], b.diagnostics_history[:, 2])
