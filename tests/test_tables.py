"""The product's host tables (vectorised builder) against the oracle's loop restatement:
integer index tables bit-exact, operator values bit-exact (same IEEE expressions), and the
eigenbasis rectangle inverse / capacitance matrix within round-off of the SciPy DCT path."""
import numpy as np
import pytest

from oracle import ghost_oracle as O

CASES = [dict(nx=88, ny=16, Lx=22., Ly=4.1, D=1., xc=2., yc=2., nu=0.01),
         dict(nx=40, ny=30, Lx=4., Ly=3., D=1., xc=1.5, yc=1.5, nu=0.01),
         dict(nx=440, ny=82, Lx=22., Ly=4.1, D=1., xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5)]


def build(pkg, case, interpolation='bilinear', pressure='neumann'):
    ref = O.CylinderFlow(ghost_velocity_interpolation=interpolation, ghost_pressure_extension=pressure, **case)
    ops = pkg.GhostOperators(case['nx'], case['ny'], case['Lx'], case['Ly'], case['D'], case['xc'], case['yc'],
                             case['nu'], 1., ref.u_inlet_base, interpolation=interpolation, pressure_extension=pressure)
    return ref, ops


def same_csr(a, b):
    a = a.tocsr(); b = b.tocsr()
    a.sort_indices(); b.sort_indices()
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


@pytest.mark.parametrize('case', CASES, ids=['88x16', '40x30', '440x82'])
def test_operator_tables_bit_exact(pkg, case):
    ref, ops = build(pkg, case)
    g = ref.gg
    for name in ('ui', 'vi', 'pi', 'outlet_u'):
        assert np.array_equal(getattr(ops, name), getattr(g, name)), name
    pairs = dict(E=g.E, P=g.P, Dop=g.D, G=g.G, A=g.A, L=g.L, pressure_surface=g.pressure_surface,
                 du_dn=g.du_dn, dv_dn=g.dv_dn)
    for name, other in pairs.items():
        assert same_csr(getattr(ops, name), other), name
    assert np.array_equal(ops.div_offset, g.div_offset) and np.array_equal(ops.diff_offset, g.diff_offset)
    assert np.array_equal(ops.normal_weights, g.normal_weights)


def test_ghost_probe_records_match(pkg):
    """Ghost nodes, probe growth s and tries of every ghost node equal the oracle's loop."""
    case = CASES[2]
    ref, ops = build(pkg, case)
    gc = ref.gc
    active_u = gc.dist_u > 1e-13
    active_u[:, 0] = False
    _, _, probes = O.extension(gc.xu, gc.yu, gc.sdf, active_u, return_probes=True)
    info = ops.ghost_u
    assert np.array_equal(info['ghost_nodes'], [p['node'] for p in probes])
    assert np.array_equal(info['s'], [p['s'] for p in probes])
    assert np.array_equal(info['tries'], [p['tries'] for p in probes])


def test_biquadratic_tables_bit_exact(pkg):
    ref, ops = build(pkg, CASES[1], interpolation='biquadratic')
    assert same_csr(ops.E, ref.gg.E) and same_csr(ops.L, ref.gg.L)


@pytest.mark.parametrize('case', CASES[1:], ids=['40x30', '440x82'])
def test_quadratic_pressure_extension_tables(pkg, case):
    """ghost_pressure_extension='quadratic' (ghost_reconstruction.py:7-45): the product's P and the
    operators built on it equal the oracle's bit for bit; P reproduces any quadratic polynomial at
    the ghost cells and its rows sum to one."""
    ref, ops = build(pkg, case, pressure='quadratic')
    g = ref.gg
    for name, a, b in (('P', ops.P, g.P), ('G', ops.G, g.G), ('A', ops.A, g.A)):
        assert same_csr(a, b), name
    assert np.array_equal(ops.pi, g.pi)
    gc = ops.gc
    X, Y = np.meshgrid(gc.xp, gc.yp)
    f = 1.3 - 0.7 * X + 0.2 * Y + 0.05 * X * X - 0.11 * X * Y + 0.3 * Y * Y
    ext = ops.P @ f.ravel()[ops.pi]
    ghost = ops.ghost_p['ghost_nodes']
    assert len(ghost) > 0 and np.max(abs(ext[ghost] - f.ravel()[ghost])) < 1e-9 * np.max(abs(f))
    sums = np.asarray(ops.P.sum(axis=1)).ravel()
    assert np.max(abs(sums[ghost] - 1.0)) < 1e-12
    assert ops.ghost_p['polynomial_reproduction_linf'] <= 1e-10 and ops.ghost_p['neighbor_count_max'] >= 12


@pytest.mark.parametrize('kind', ['pressure', 'u', 'v'])
def test_eigenbasis_rectangle_inverse(pkg, kind):
    nx, ny, dx, dy = 440, 82, 22 / 440, 4.1 / 82
    f = pkg.RectangleFactor(kind, nx, ny, dx, dy, nu=.05, dt=.002)
    rect, inv, shape = O.rectangle(nx, ny, dx, dy, kind, .05, .002)
    assert abs(rect - f.matrix).max() == 0
    b = np.random.default_rng(4).normal(size=shape)
    x = f.solve(b)
    assert np.max(abs(x - inv(b))) < 1e-13 * np.max(abs(x))
    V = f.V
    assert np.max(abs(V.T @ V - np.eye(V.shape[0]))) < 1e-13


def test_capacitance_rows_and_matrix(pkg):
    ref, ops = build(pkg, CASES[2])
    factor = pkg.RectangleFactor('pressure', ops.nx, ops.ny, ops.dx, ops.dy)
    schur = pkg.SchurFactor(ops.A, ops.pi, factor)
    info = ref.gg.pressure_inverse_info
    assert schur.rank == info['rank'] and np.array_equal(schur.rows, info['rows'])
    assert np.max(abs(schur.S - info['schur'])) < 1e-12 * np.max(abs(info['schur']))


def test_force_floor_of_exact_inverses(pkg, monkeypatch):
    """Two exact FP64 inverses of the same rectangle (SciPy DCT vs eigenbasis + banded) change the
    impulsive-start C_L of the 440x82 case by ~1e-10 absolute while C_D ~ 36: the parity floor for
    forces is relative to the force magnitude, not to |C_L|."""
    cfg = dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5)
    a = O.CylinderFlow(initial='zero', **cfg)
    orig = O.rectangle

    def eigen_rectangle(nx, ny, dx, dy, kind, nu=0., dt=0., rho=1., workers=None):
        m, _, shape = orig(nx, ny, dx, dy, kind, nu, dt, rho, workers)
        return m, pkg.RectangleFactor(kind, nx, ny, dx, dy, nu, dt, rho).solve, shape
    monkeypatch.setattr(O, 'rectangle', eigen_rectangle)
    b = O.CylinderFlow(initial='zero', **cfg)
    monkeypatch.setattr(O, 'rectangle', orig)
    a.step(2e-3); b.step(2e-3)
    da, db = a.last_step_diagnostics, b.last_step_diagnostics
    dcl = abs(da['cl'] - db['cl'])
    assert 1e-12 < dcl < 1e-10 * abs(da['cd'])
    assert np.max(abs(a.u - b.u)) < 1e-12 * np.max(abs(a.u))
