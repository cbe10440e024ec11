"""GPU: the x-FFT form of the rectangle inverse (DCT-IV along x through a half-length complex FFT,
tridiagonal solves along y per x mode; rectangular_inverse.py:48-53 factorises the same operator with
DCTs in both directions) against the oracle: the separable inverse at the five reference grids, the
capacitance-corrected pressure solve (pressure_schur.py:103-124) and whole steps."""
import os

import numpy as np
import pytest

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

GRIDS = {
    '88x16': dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.),
    '40x30': dict(nx=40, ny=30, Lx=4., Ly=3., xc=1.5, yc=1.5),
    '440x82': dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5),
    '1320x246': dict(nx=1320, ny=246, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.01, inflow='parabolic', U_m=1.5),
    '800x820': dict(nx=800, ny=820, Lx=40., Ly=41., xc=10., yc=20.5, nu=0.01),
}
XFFT = 7


def gpu_flow(pkg, **kw):
    return pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)


def rel(a, b):
    return float(np.max(abs(np.asarray(a) - np.asarray(b))) / max(np.max(abs(np.asarray(b))), 1e-300))


@pytest.mark.parametrize('name', list(GRIDS))
def test_xfft_rectangle_inverse(pkg, name):
    """All three kinds: pressure (DCT-IV), u (mirrored-outlet sine basis), v (DST-IV)."""
    print('x-FFT rectangle inverse', name, flush=True)
    f = gpu_flow(pkg, **GRIDS[name])
    dev = f.gg.device
    dt = 1e-3
    dev.set_dt(dt)
    dev.set_launch(XFFT, 1)
    launch = dev.get_launch()
    assert launch['gemm_pressure'] == XFFT and launch['tri_pressure'] == 3
    assert launch['gemm_velocity'] == XFFT and launch['tri_velocity'] == 3
    factors = [dev.pressure_factor] + [hf[0] for hf in dev.helmholtz_factors]
    for code, kind in enumerate(('pressure', 'u', 'v')):
        b = np.random.default_rng(3 + code).normal(size=factors[code].shape)
        x = dev.rect_inverse(code, b)
        _, inv, _ = O.rectangle(f.nx, f.ny, f.dx, f.dy, kind, f.nu, dt, 1.)
        err = rel(x, inv(b.copy()))
        print(name, kind, 'x-FFT inverse error', err, flush=True)
        assert err < 1e-12, (kind, err)
        assert np.max(abs(factors[code].matrix @ x.ravel() - b.ravel())) < 1e-10 * max(1., np.max(abs(x)))


def test_xfft_helmholtz_solve(pkg):
    """The capacitance-corrected u/v Helmholtz solve (optimized_ghost_cell.py:243-255) on the x-FFT path."""
    kw = GRIDS['440x82']
    ref = O.CylinderFlow(**kw)
    gpu = gpu_flow(pkg, **kw)
    dt = 2e-3
    g = ref.gg
    g.build_helmholtz(dt)
    dev = gpu.gg.device
    dev.set_dt(dt)
    dev.set_launch(XFFT, 1)
    q = np.random.default_rng(8).normal(size=g.qsize)
    xr = g._helmholtz(q)
    x, info = dev.schur_solve(1, dev.q_to_vel(q))
    assert rel(dev.vel_to_q(x), xr) < 1e-11


def test_xfft_falls_back_for_odd_widths(pkg):
    f = gpu_flow(pkg, nx=87, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)
    f.gg.device.set_launch(XFFT, 1)
    assert f.gg.device.get_launch()['gemm_pressure'] == 4


@pytest.mark.parametrize('name', ['88x16', '440x82'])
def test_xfft_pressure_schur_solve(pkg, name):
    ref = O.CylinderFlow(**GRIDS[name])
    gpu = gpu_flow(pkg, **GRIDS[name])
    gpu.gg.device.set_launch(XFFT, 1)
    g = ref.gg
    vals = np.random.default_rng(4).normal(size=len(g.pi))
    xr = g.pressure_solve(vals)
    rhs = np.zeros(gpu.nx * gpu.ny)
    rhs[g.pi] = vals
    x, info = gpu.gg.device.schur_solve(0, rhs)
    assert rel(x[g.pi], xr) < 1e-11
    assert np.max(abs(g.A @ x[g.pi] - vals)) <= info[3]


@pytest.mark.parametrize('name,steps,dt,initial', [('440x82', 20, 2e-3, 'zero'), ('1320x246', 10, 5e-4, 'profile')])
def test_xfft_steps_match_oracle(pkg, name, steps, dt, initial):
    kw = dict(GRIDS[name], initial=initial)
    if name == '1320x246':
        kw['v_perturbation'] = dict(seed=20261203)
    print('x-FFT steps', name, flush=True)
    ref = O.CylinderFlow(fft_workers=min(16, len(os.sched_getaffinity(0))), **kw)
    gpu = gpu_flow(pkg, **kw)
    gpu.gg.device.set_launch(XFFT, 1)
    rows = []
    for _ in range(steps):
        ref.step(dt)
        rows.append(dict(ref.last_step_diagnostics))
    t, cd, cl = gpu.solve(n_steps=steps, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    errs = dict(u=rel(gpu.u, ref.u), v=rel(gpu.v, ref.v), p=rel(gpu.p, ref.p))
    print(name, 'x-FFT field errors', errs, flush=True)
    assert max(errs.values()) < 1e-11, errs
    for k, r in enumerate(rows):
        scale = max(abs(r['cd']), abs(r['cl']))
        assert abs(cd[k] - r['cd']) <= 1e-10 * scale and abs(cl[k] - r['cl']) <= 1e-10 * scale, k
    assert np.all(gpu.diagnostics_history[:, 10] < 1e-12)
