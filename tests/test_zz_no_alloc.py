"""mcf_rect_inverse_noalloc (the allocation-free rectangle inverse) against mcf_rect_inverse, through
the C ABI: equal bit for bit on two calls in a row whose second input is smaller, for every kind,
and a clear error beyond the scratch limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {'440x82': dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5),
         '88x16': dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)}
LIMIT = 786432


def device(pkg, name, dt=1e-3):
    print(f'no-alloc: building the {name} case', flush=True)
    flow = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **CASES[name])
    dev = flow.gg.device
    dev.set_dt(dt)        # uploads the u and v Helmholtz rectangles
    return flow, dev


def inverses(dev, kind, b):
    b = np.ascontiguousarray(b, dtype=np.float64)
    ref, mine = np.empty_like(b), np.full_like(b, np.nan)
    dev.h.call('mcf_rect_inverse', kind, b, ref)
    dev.h.call('mcf_rect_inverse_noalloc', kind, b, mine)
    return ref, mine


def test_noalloc_equals_allocating_entry_point(pkg):
    big_flow, big = device(pkg, '440x82')
    small_flow, small = device(pkg, '88x16')
    rng = np.random.default_rng(7)
    shapes = {0: lambda f: (f.ny, f.nx), 1: lambda f: (f.ny, f.nx), 2: lambda f: (f.ny - 1, f.nx)}
    for kind in (0, 1, 2):
        for flow, dev in ((big_flow, big), (small_flow, small)):      # second call: smaller input
            print(f'no-alloc: kind {kind} on {flow.nx}x{flow.ny}', flush=True)
            b = rng.normal(size=shapes[kind](flow))
            ref, mine = inverses(dev, kind, b)
            assert np.isfinite(ref).all()
            assert np.array_equal(mine, ref), (kind, flow.nx, float(np.max(abs(mine - ref))))
    # the same handle again after the smaller call: the shared scratch holds nothing stale
    b = rng.normal(size=(big_flow.ny, big_flow.nx))
    ref, mine = inverses(big, 0, b)
    assert np.array_equal(mine, ref)


def test_noalloc_rejects_rectangles_beyond_the_limit(pkg):
    import importlib
    nat = importlib.import_module(pkg.__name__ + '._native')
    nx, ny = 1000, 800                       # 800,000 pressure values > 786,432
    assert nx * ny > LIMIT
    f = pkg.RectangleFactor('pressure', nx, ny, 40. / nx, 41. / ny)
    h = nat.Handle(nx, ny, 40. / nx, 41. / ny, 0.01, 1.)
    try:
        h.call('mcf_set_rectangle', 0, ny, nx, float(f.a), float(f.b), nat.f64(f.V), nat.f64(f.lam),
               nat.f64(f.lower), nat.f64(f.diag), nat.f64(f.upper))
        b = np.zeros((ny, nx))
        out = np.zeros_like(b)
        with pytest.raises(ValueError, match='allocation-free limit'):
            h.call('mcf_rect_inverse_noalloc', 0, b, out)
    finally:
        h.close()
