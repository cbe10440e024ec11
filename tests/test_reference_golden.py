"""The oracle and the product's host tables against vectors the reference itself produced.

``tests/golden/reference_tables.npz`` was written by ``tests/golden/gen_reference_tables.py`` from the
reference's own functions (solver/optimized_ghost_cell.py:14-58, 61-172; solver/cut_cell_solver.py:25-77;
solver/pressure_schur.py:75-125; solver/rectangular_inverse.py:8-58; solver/ghost_reconstruction.py:7-45)
on the 88x16, 40x30, 440x82, 1320x246 and 800x820 grids.  Integer decisions -- the active and ghost
node sets, each ghost node's probe cells after stencil growth, the kNN-with-ties sets of the surface
pressure and wall-gradient stencils, the Schur rank -- must match exactly; weights are compared bit
for bit where the reference's evaluation order is followed and to 1e-12 where a library's internal
order differs (the quadratic pressure continuation's single-point kd-tree query); transform outputs
to 1e-12.
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import cKDTree

from oracle import ghost_oracle as O

GOLDEN = np.load(Path(__file__).resolve().parent / 'golden' / 'reference_tables.npz')
TAGS = sorted({k.split('/')[0] for k in GOLDEN.files})
BASE = [t for t in TAGS if str(GOLDEN[f'{t}/variant']) == 'bilinear/neumann']
_cache = {}


def g(tag, key):
    return GOLDEN[f'{tag}/{key}']


def case_kwargs(tag):
    nx, ny, Lx, Ly, D, xc, yc, nu, parabolic, dt = g(tag, 'params')
    velocity, pressure = str(g(tag, 'variant')).split('/')
    kw = dict(nx=int(nx), ny=int(ny), Lx=Lx, Ly=Ly, D=D, xc=xc, yc=yc, nu=nu,
              ghost_velocity_interpolation=velocity, ghost_pressure_extension=pressure)
    if parabolic:
        kw.update(inflow='parabolic', U_m=1.5)
    return kw, dt


def oracle(tag):
    if ('oracle', tag) not in _cache:
        kw, _ = case_kwargs(tag)
        _cache['oracle', tag] = O.CylinderFlow(**kw)
    return _cache['oracle', tag]


def product(pkg, tag):
    if ('product', tag) not in _cache:
        kw, _ = case_kwargs(tag)
        s = oracle(tag)
        _cache['product', tag] = pkg.GhostOperators(kw['nx'], kw['ny'], kw['Lx'], kw['Ly'], kw['D'], kw['xc'], kw['yc'],
                                                    kw['nu'], 1., s.u_inlet_base,
                                                    interpolation=kw['ghost_velocity_interpolation'],
                                                    pressure_extension=kw['ghost_pressure_extension'])
    return _cache['product', tag]


def canonical(matrix):
    m = sp.csr_matrix(matrix, copy=True)
    m.sum_duplicates()
    m.sort_indices()
    return m


def check_ghost_rows(tag, field, matrix, ids, size, exact_weights=True):
    inactive = np.setdiff1d(np.arange(size), ids)
    assert np.array_equal(inactive, g(tag, f'{field}_inactive')), f'{tag} {field}: active node set'
    m = canonical(matrix)
    active = np.zeros(m.shape[0], bool)
    active[ids] = True
    rows = np.flatnonzero((np.diff(m.indptr) > 0) & ~active)
    assert np.array_equal(rows, g(tag, f'{field}_ghost_rows')), f'{tag} {field}: ghost nodes'
    sub = m[rows]
    assert np.array_equal(sub.indptr, g(tag, f'{field}_ghost_ptr')), f'{tag} {field}: stencil sizes'
    assert np.array_equal(sub.indices, g(tag, f'{field}_ghost_col')), f'{tag} {field}: probe cells'
    w = g(tag, f'{field}_ghost_w')
    if exact_weights:
        assert np.array_equal(sub.data, w), f'{tag} {field}: weights'
    else:
        assert np.max(abs(sub.data - w)) <= 1e-12 * np.max(abs(w)), f'{tag} {field}: weights'


def check_surface(tag, ops):
    assert np.array_equal(ops.surface_points, g(tag, 'surface_points'))
    assert np.array_equal(ops.normal_weights, g(tag, 'normal_weights'))
    for key in ('pressure_surface', 'du_dn', 'dv_dn'):
        m = canonical(getattr(ops, key))
        assert np.array_equal(m.indptr, g(tag, f'{key}_ptr')), f'{tag} {key}: stencil sizes'
        assert np.array_equal(m.indices, g(tag, f'{key}_col')), f'{tag} {key}: kNN-with-ties sets'
        # both stencils query all points at once, so the neighbours arrive in ascending index
        # order and the min-norm least-squares weights are the reference's bit for bit
        assert np.array_equal(m.data, g(tag, f'{key}_w')), f'{tag} {key}: weights'


# ------------------------------------------------------------------------------ oracle
@pytest.mark.parametrize('tag', TAGS)
def test_oracle_extension_tables(tag):
    """Active sets, ghost nodes, probe cells and weights of E_u, E_v and P (optimized_ghost_cell.py:14-58,
    ghost_reconstruction.py:7-45)."""
    s = oracle(tag)
    gg = s.gg
    quadratic = str(g(tag, 'variant')).endswith('/quadratic')
    check_ghost_rows(tag, 'u', gg.Eu, gg.ui, s.u.size)
    check_ghost_rows(tag, 'v', gg.Ev, gg.vi, s.v.size)
    # the quadratic continuation's neighbours come from a single-point kd-tree query, whose result
    # order (not set) is the tree's: its least-squares weights agree to round-off
    check_ghost_rows(tag, 'p', gg.P, gg.pi, s.p.size, exact_weights=not quadratic)
    assert gg.pressure_inverse_info['rank'] == int(g(tag, 'pressure_rank')[0])


@pytest.mark.parametrize('tag', BASE)
def test_oracle_surface_stencils(tag):
    """Surface points, kNN-with-ties sets and LSQ weights of the pressure quadrature and the wall
    gradients (cut_cell_solver.py:25-77, optimized_ghost_cell.py:152-172)."""
    check_surface(tag, oracle(tag).gg)


def solution_matches(tag, key, x, seed):
    x = np.asarray(x).ravel()
    if f'{tag}/{key}_x' in GOLDEN.files:
        ref = g(tag, f'{key}_x').ravel()
        assert np.max(abs(x - ref)) <= 1e-12 * np.max(abs(ref)), key
    else:
        w = np.random.default_rng(seed).normal(size=(4, x.size))
        scale = abs(w) @ abs(x)
        assert np.all(abs(w @ x - g(tag, f'{key}_proj')) <= 1e-12 * scale), key


@pytest.mark.parametrize('tag', BASE)
def test_oracle_rectangle_and_woodbury_values(tag):
    """Rectangle inverses (rectangular_inverse.py:48-53), the monitored Woodbury pressure solve
    (pressure_schur.py:16-124) and the per-dt Helmholtz boundary-row inverses
    (optimized_ghost_cell.py:243-255) on the reference's seeded right-hand sides."""
    s = oracle(tag)
    kw, dt = case_kwargs(tag)
    nx, ny = kw['nx'], kw['ny']
    for k, kind in enumerate(('pressure', 'u', 'v')):
        _, inverse, shape = O.rectangle(nx, ny, s.dx, s.dy, kind, kw['nu'], dt, 1.)
        b = np.random.default_rng(40 + k).normal(size=shape)
        solution_matches(tag, f'rect_{kind}', inverse(b.copy()), 50 + k)
    b = np.random.default_rng(60).normal(size=len(s.gg.pi))
    solution_matches(tag, 'woodbury', s.gg.pressure_solve(b), 61)
    assert s.gg.pressure_inverse_info['last_solve']['refinement_iterations'] == int(g(tag, 'woodbury_refinements')[0])
    s.gg.build_helmholtz(dt)
    nu_act = len(s.gg.ui)
    bu = np.random.default_rng(31).normal(size=nu_act)
    bv = np.random.default_rng(31).normal(size=len(s.gg.vi))
    x = s.gg._helmholtz(np.concatenate((bu, bv)))
    solution_matches(tag, 'helmholtz_u', x[:nu_act], 32)
    solution_matches(tag, 'helmholtz_v', x[nu_act:], 32)
    for info, kind in zip(s.gg.helmholtz_inverse_info, 'uv'):
        assert info['rank'] == int(g(tag, f'helmholtz_{kind}_rank')[0]), kind


# ------------------------------------------------------------------------------ product host tables
@pytest.mark.parametrize('tag', TAGS)
def test_product_extension_tables(pkg, tag):
    ops = product(pkg, tag)
    variant = str(g(tag, 'variant'))
    check_ghost_rows(tag, 'u', ops.Eu, ops.ui, ops.usize)
    check_ghost_rows(tag, 'v', ops.Ev, ops.vi, ops.vsize)
    check_ghost_rows(tag, 'p', ops.P, ops.pi, ops.nx * ops.ny, exact_weights=not variant.endswith('/quadratic'))


@pytest.mark.parametrize('tag', BASE)
def test_product_surface_and_schur_rows(pkg, tag):
    ops = product(pkg, tag)
    check_surface(tag, ops)
    _, dt = case_kwargs(tag)
    factor = pkg.RectangleFactor('pressure', ops.nx, ops.ny, ops.dx, ops.dy)
    assert pkg.SchurFactor(ops.A, ops.pi, factor).rank == int(g(tag, 'pressure_rank')[0])
    for kind, selected, actual in ops.helmholtz_actual(dt):
        rows, _, _ = pkg.tables.boundary_rows(actual, selected, pkg.rectangle_matrix(kind, ops.nx, ops.ny, ops.dx, ops.dy,
                                                                                     ops.nu, dt, 1.))
        assert len(rows) == int(g(tag, f'helmholtz_{kind}_rank')[0]), kind


@pytest.mark.parametrize('tag', BASE)
def test_product_rectangle_values(pkg, tag):
    """The eigenbasis + tridiagonal factorisation the device uploads, on the reference's seeded
    right-hand sides, against the reference's DCT/DST inverse."""
    kw, dt = case_kwargs(tag)
    ops = product(pkg, tag)
    for k, kind in enumerate(('pressure', 'u', 'v')):
        f = pkg.RectangleFactor(kind, ops.nx, ops.ny, ops.dx, ops.dy, kw['nu'], dt, 1.)
        b = np.random.default_rng(40 + k).normal(size=f.shape)
        solution_matches(tag, f'rect_{kind}', f.solve(b), 50 + k)


# ------------------------------------------------------------------------------ library routines
@pytest.mark.parametrize('seed', [0, 1, 2])
def test_knn_with_ties_equals_kdtree_queries(pkg, seed):
    """The brute-force kNN-with-ties of the oracle and the product equal the reference's calls
    ``tree.query(points, k)`` then ``tree.query_ball_point(points, d_k (1 + 1e-12))``
    (cut_cell_solver.py:51-55, optimized_ghost_cell.py:157-160), on lattice centres, where ties are
    the rule, and on a jittered cloud."""
    rng = np.random.default_rng(seed)
    h = 0.05
    jj, ii = np.meshgrid(np.arange(40), np.arange(50), indexing='ij')
    centers = np.column_stack(((jj.ravel() + .5) * h, (ii.ravel() + .5) * h))
    if seed == 2:
        centers = centers + rng.uniform(-.2, .2, size=centers.shape) * h
    theta = rng.uniform(0, 2 * np.pi, 200)
    points = np.column_stack((1. + .4 * np.sin(theta), 1.25 + .4 * np.cos(theta)))
    points[:20] = centers[rng.integers(0, len(centers), 20)] + h / 2   # points on the lattice's symmetry lines
    tree = cKDTree(centers)
    for k in (12, 14):
        dist, _ = tree.query(points, k=k)
        balls = tree.query_ball_point(points, dist[:, -1] * (1 + 1e-12))
        mine = pkg.tables.nearest_with_ties(centers, points, k)
        for q, ball in enumerate(balls):
            ref = sorted(ball)
            assert np.array_equal(O.knn_with_ties(centers, points[q], k), ref)
            assert np.array_equal(np.sort(mine[q]), ref)
