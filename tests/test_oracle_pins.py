"""Pins of the CPU oracle (oracle/ghost_oracle.py) against closed forms, brute force and the
reference's own known-answer tests.  Each test names the reference test or equation it follows."""
import importlib.util
import tempfile
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import ghost_oracle as O

ROOT = Path(__file__).resolve().parents[1]


def channel(nx=88, ny=16, **kw):
    """The reference's small ghost fixture (tests/test_numba_ghost_cell.py:12-14)."""
    return O.CylinderFlow(nx=nx, ny=ny, Lx=22., Ly=4.1, xc=2., yc=2., **kw)


def box(nx=40, ny=30, **kw):
    """The 40x30 fixture of tests/test_gemini_followup.py:24-26."""
    return O.CylinderFlow(nx=nx, ny=ny, Lx=4., Ly=3., xc=1.5, yc=1.5, **kw)


# ----------------------------------------------------------------------------- geometry
def test_circle_sdf_exact_values():
    """tests/test_ghost_cell_sharp.py:9-27."""
    sdf = O.CircleSDF(xc=2.5, yc=1.5, radius=0.5)
    assert abs(sdf.distance(3.0, 1.5)) < 1e-15
    assert abs(sdf.distance(2.5, 2.0)) < 1e-15
    assert abs(sdf.distance(2.0, 1.5)) < 1e-15
    assert sdf.distance(2.5, 1.5) == -0.5
    assert sdf.distance(3.5, 1.5) == 0.5
    nx, ny = sdf.normal(3.0, 1.5)
    assert abs(nx - 1.0) < 1e-15 and abs(ny) < 1e-15
    nx, ny = sdf.normal(2.5, 2.0)
    assert abs(nx) < 1e-15 and abs(ny - 1.0) < 1e-15


class FlatWall:
    """Wall y = 0 with fluid above (the Patch SDF of tests/test_gemini_followup.py:32-34)."""

    def __init__(self, h):
        self.h = h

    def distance(self, xx, yy):
        return np.where(abs(xx) < self.h / 4, yy, 1.)

    def normal(self, xx, yy):
        return 0., 1.


def flat_setup(h=.1):
    x, y = np.arange(-2, 3) * h, np.arange(-2.5, 5) * h
    sdf = FlatWall(h)
    xx, yy = np.meshgrid(x, y)
    return x, y, xx, yy, sdf, sdf.distance(xx, yy) > 0


def test_flat_wall_quadratic_laplacian():
    """Biquadratic probes reproduce the Laplacian of y^2, bilinear ones do not (test_gemini_followup.py:29-44)."""
    h = .1
    x, y, xx, yy, sdf, active = flat_setup(h)
    errors = []
    for mode in ('bilinear', 'biquadratic'):
        matrix, ids = O.extension(x, y, sdf, active, interpolation=mode)
        field = (matrix @ (yy * yy).ravel()[ids]).reshape(yy.shape)
        j = int(np.argmin(abs(y + .5 * h))); i = 2
        lap = (field[j, i] - 2 * field[j + 1, i] + field[j + 2, i]) / h**2
        errors.append(abs(lap - 2))
    assert errors[0] > .05 and errors[1] < 1e-12, errors


def test_flat_wall_linear_dirichlet_and_neumann():
    """Dirichlet extension reproduces y (zero on the wall) exactly; Neumann reproduces fields constant in y."""
    h = .1
    x, y, xx, yy, sdf, active = flat_setup(h)
    E, ids = O.extension(x, y, sdf, active)
    ghost = (sdf.distance(xx, yy) <= 1e-13) & (sdf.distance(xx, yy) >= -2 * h)
    assert ghost.sum() > 0
    lin = (E @ yy.ravel()[ids]).reshape(yy.shape)
    assert np.max(abs(lin[ghost] - yy[ghost])) < 1e-13
    P, pid = O.extension(x, y, sdf, active, neumann=True)
    f = 3. + 0. * yy + xx
    ext = (P @ f.ravel()[pid]).reshape(yy.shape)
    assert np.max(abs(ext[ghost] - f[ghost])) < 1e-13
    # y^2 has zero normal derivative on the wall: biquadratic probes + the Neumann quadratic are exact
    Pq, pq = O.extension(x, y, sdf, active, neumann=True, interpolation='biquadratic')
    g = (Pq @ (yy * yy).ravel()[pq]).reshape(yy.shape)
    assert np.max(abs(g[ghost] - yy[ghost]**2)) < 1e-13


# ----------------------------------------------------------------------------- rectangle
def brute_rectangle(kind, nx, ny, dx, dy, nu=0., dt=0., rho=1.):
    """Cylinder-free operator assembled node by node from the boundary conditions
    (rectangular_inverse.py:9-39): pressure Neumann inlet/walls + Dirichlet outlet face;
    u faces 1..nx with known inlet, mirrored outlet, no-slip walls; v rows 1..ny-1, v=0 inlet,
    Neumann outlet, v=0 walls."""
    yy, xx = (ny - 1, nx) if kind == 'v' else (ny, nx)
    n = yy * xx
    lap = np.zeros((n, n))
    for j in range(yy):
        for i in range(xx):
            r = j * xx + i
            # x direction
            if kind == 'pressure':
                if i > 0: lap[r, r - 1] += 1 / dx**2; lap[r, r] -= 1 / dx**2
                if i < xx - 1: lap[r, r + 1] += 1 / dx**2; lap[r, r] -= 1 / dx**2
                else: lap[r, r] -= 2 / dx**2          # phi = -phi_ghost at the outlet face
            elif kind == 'u':
                lap[r, r] -= 2 / dx**2
                if i > 0: lap[r, r - 1] += (2. if i == xx - 1 else 1.) / dx**2
                if i < xx - 1: lap[r, r + 1] += 1 / dx**2
            else:
                lap[r, r] -= 2 / dx**2
                if i > 0: lap[r, r - 1] += 1 / dx**2
                else: lap[r, r] -= 1 / dx**2           # v = 0 at the inlet wall by mirror
                if i < xx - 1: lap[r, r + 1] += 1 / dx**2
                else: lap[r, r] += 1 / dx**2           # Neumann outlet
            # y direction
            if kind == 'pressure':
                if j > 0: lap[r, r - xx] += 1 / dy**2; lap[r, r] -= 1 / dy**2
                if j < yy - 1: lap[r, r + xx] += 1 / dy**2; lap[r, r] -= 1 / dy**2
            elif kind == 'u':
                lap[r, r] -= 2 / dy**2
                if j > 0: lap[r, r - xx] += 1 / dy**2
                else: lap[r, r] -= 1 / dy**2           # no-slip wall between faces: ghost = -u
                if j < yy - 1: lap[r, r + xx] += 1 / dy**2
                else: lap[r, r] -= 1 / dy**2
            else:
                lap[r, r] -= 2 / dy**2
                if j > 0: lap[r, r - xx] += 1 / dy**2
                if j < yy - 1: lap[r, r + xx] += 1 / dy**2
    if kind == 'pressure':
        return lap / rho
    return np.eye(n) - .5 * dt * nu * lap


@pytest.mark.parametrize('kind', ['pressure', 'u', 'v'])
def test_rectangle_matches_brute_force_and_dense_solve(kind):
    nx, ny, dx, dy, nu, dt = 12, 7, .3, .2, .05, .01
    matrix, solve, shape = O.rectangle(nx, ny, dx, dy, kind, nu, dt, 1.3 if kind == 'pressure' else 1.)
    dense = brute_rectangle(kind, nx, ny, dx, dy, nu, dt, 1.3 if kind == 'pressure' else 1.)
    assert np.max(abs(matrix.toarray() - dense)) < 1e-12 * np.max(abs(dense))
    b = np.random.default_rng(5).normal(size=shape)
    x = solve(b.copy())
    assert np.max(abs(x.ravel() - np.linalg.solve(dense, b.ravel()))) < 1e-10 * max(1, np.max(abs(x)))


def test_rectangle_self_check_residual():
    """rectangular_inverse.py:54-57 at the 440x82 shape."""
    for kind in ('pressure', 'u', 'v'):
        O.rectangle(440, 82, 22 / 440, 4.1 / 82, kind, .05, .002)


# ----------------------------------------------------------------------------- Woodbury
def test_boundary_row_inverse_equals_dense_solve():
    s = box()
    g = s.gg
    rng = np.random.default_rng(11)
    b = rng.normal(size=len(g.pi))
    x = g.pressure_solve(b)
    exact = spla.spsolve(g.A.tocsc(), b)
    assert np.max(abs(x - exact)) < 1e-9 * np.max(abs(exact))
    assert np.max(abs(g.A @ x - b)) < 1e-9
    g.build_helmholtz(.003)
    q = rng.normal(size=g.qsize)
    star = g._helmholtz(q)
    H = sp.eye(g.qsize) - .5 * .003 * g.L
    assert np.max(abs(star - spla.spsolve(H.tocsc(), q))) < 1e-11


# ----------------------------------------------------------------------------- operators
def test_diffusion_matches_flux_form():
    """L q + offset equals the flux-form diffusion of the extended field (test_numba_ghost_cell.py:22-29)."""
    s = channel()
    g = s.gg
    q = np.random.default_rng(6).normal(size=g.qsize)
    u, v = [a.copy() for a in g.unpack(q)]
    ru, rv = O.flux_rhs(u, v, s.dx, s.dy, s.nu, False)
    diffusion = g.pack(ru, rv) / (s.dx * s.dy)
    diffusion[:len(g.ui)][np.unravel_index(g.ui, s.u.shape)[1] == s.nx] *= 2
    assert np.max(abs(diffusion - g.diffusion(q))) < 1e-11


def test_projection_divergence_and_poisson():
    """Projected divergence < 1e-10 and A phi = div/dt (test_numba_ghost_cell.py:30-34)."""
    s = channel()
    g = s.gg
    q = np.random.default_rng(6).normal(size=g.qsize)
    projected, phi = g.project(q, .003)
    assert np.max(abs(g.divergence(projected))) < 1e-10
    assert np.max(abs(g.A @ phi - g.divergence(q) / .003)) < 1e-9
    u, v = g.unpack(projected)
    assert np.max(abs(g.pack(u, v) - projected)) < 1e-12


def test_advection_of_uniform_flow_vanishes_in_channel_interior():
    """-N(u) is zero for a uniform u field away from the cylinder and the outlet (centred fluxes)."""
    s = channel()
    g = s.gg
    u = np.ones_like(s.u); v = np.zeros_like(s.v)
    q = g.pack(u, v)
    adv = g.advection(q)
    ju, iu = np.unravel_index(g.ui, s.u.shape)
    far = (iu > 30) & (iu < s.nx - 2)
    assert np.max(abs(adv[:len(g.ui)][far])) < 1e-12


def test_advection_of_linear_fields_closed_form():
    """For u = a + b x + c y and v = d + e x + f y the centred flux form (numba_kernels.py:17-142) is
    exact: interior u faces give -d(u^2)/dx - d(uv)/dy = -(2 b u + c v + f u), interior v faces
    -d(uv)/dx - d(v^2)/dy = -(b v + e u + 2 f v).  The outlet u row (numba_kernels.py:65-74, doubled
    at optimized_ghost_cell.py:197) replaces the missing cell by the outlet face itself, so
    (u_c^2 - u_nx^2)/dx, doubled, is -d(u^2)/dx a quarter cell upstream, and its corner fluxes carry
    the last v column: -2 b u(x_nx - dx/4) - f u(x_nx) - c v(x_nx - dx/2).  The wall rows carry no
    advective flux through the wall (numba_kernels.py:61-63)."""
    s = channel()
    g = s.gg
    a, b, c, d, e, f = 0.7, -0.3, 0.45, -0.2, 0.25, 0.6
    U = lambda x, y: a + b * x + c * y
    V = lambda x, y: d + e * x + f * y
    gc = s.gc
    xu, yu = np.meshgrid(gc.xu, gc.yu)
    xv, yv = np.meshgrid(gc.xv, gc.yv)
    adv_u, adv_v = g.unpack(g.advection(g.pack(U(xu, yu), V(xv, yv))))
    adv_u, adv_v = adv_u.copy(), adv_v.copy()
    h, dx, dy, nx, ny = max(s.dx, s.dy), s.dx, s.dy, s.nx, s.ny
    # faces whose stencils reach neither the inlet column nor a ghost node near the cylinder
    far_u = (np.hypot(xu - s.xc, yu - s.yc) > s.D / 2 + 4 * h) & (xu > 1.5 * dx)
    far_v = (np.hypot(xv - s.xc, yv - s.yc) > s.D / 2 + 4 * h) & (xv > 1.5 * dx)
    uu, vu = U(xu, yu), V(xu, yu)
    interior_u = far_u & (xu < nx * dx - dx / 2) & (yu > dy) & (yu < (ny - 1) * dy)
    exact_u = -(2 * b * uu + c * vu + f * uu)
    assert interior_u.sum() > 100
    assert np.max(abs(adv_u - exact_u)[interior_u]) < 1e-12
    outlet = far_u & (xu > nx * dx - dx / 2) & (yu > dy) & (yu < (ny - 1) * dy)
    xo = nx * dx
    exact_out = -2 * b * U(xo - dx / 4, yu) - f * U(xo, yu) - c * V(xo - dx / 2, yu)
    assert outlet.sum() == ny - 2
    assert np.max(abs(adv_u - exact_out)[outlet]) < 1e-12
    # bottom wall row: the wall flux is zero, the first interior corner carries u v
    bottom = far_u & (yu < dy) & (xu < nx * dx - dx / 2)
    exact_bottom = -2 * b * uu - U(xu, dy) * V(xu, dy) / dy
    assert np.max(abs(adv_u - exact_bottom)[bottom]) < 1e-11
    uv, vv = U(xv, yv), V(xv, yv)
    # v on the wall rows is the no-slip zero, so the faces next to a wall are left out
    interior_v = far_v & (xv < (nx - 1) * dx) & (yv > 1.5 * dy) & (yv < (ny - 1.5) * dy)
    exact_v = -(b * vv + e * uv + 2 * f * vv)
    assert interior_v.sum() > 100
    assert np.max(abs(adv_v - exact_v)[interior_v]) < 1e-12
    # wall rows of v are not unknowns: the extension keeps them at zero
    assert np.all(adv_v[[0, -1]] == 0)


# ----------------------------------------------------------------------------- stepper
def test_noncommuting_pressure_recovery_order():
    """Closed-form constrained ODE (tests/test_cut_cell_time_restart.py:65-86)."""
    L = np.array([[-2., -1.], [-1., -3.]])

    def project(q, dt):
        return np.array([0., q[1]]), np.array([q[0] / dt])
    errors = []
    for dt in (.02, .01, .005):
        q, p = np.array([0., 1.]), np.array([-1.])
        history = previous_dt = None
        for _ in range(round(.2 / dt)):
            q, p, output, history, _ = O.ab2_cn_step(
                q, p, dt, advection=lambda q: np.zeros_like(q), diffusion=lambda q: L @ q,
                gradient=lambda p: np.array([p[0], 0.]), project=project, nu=1.,
                previous_advection=history, previous_dt=previous_dt)
            previous_dt = dt
        errors.append(abs(output[0] + np.exp(-.6)))
        assert q[0] == 0.
    orders = np.log2(np.array(errors[:-1]) / errors[1:])
    assert np.all((orders > 1.8) & (orders < 2.2)), orders


def test_ghost_time_order():
    """Coupled velocity/pressure time order in (1.8, 2.2) (tests/test_numba_ghost_cell.py:37-46)."""
    states = []
    for dt in (.004, .002, .001, .00025):
        s = channel()
        s.solve(n_steps=round(.064 / dt), dt=dt, perturb_transient=False)
        states.append((s.gg.pack(s.u, s.v), s.p.ravel()[s.gg.pi]))
    errors = np.array([[np.sqrt(np.mean((st[k] - states[-1][k])**2)) for k in (0, 1)] for st in states[:-1]])
    orders = np.log2(errors[:-1] / errors[1:])
    assert np.all((orders > 1.8) & (orders < 2.2)), orders


# ----------------------------------------------------------------------------- forces
def test_physical_surface_traction():
    """p = x gives F_x,p = -pi R^2; torque of u_theta = r - R converges to 2 pi nu R^2
    (tests/test_numba_ghost_cell.py:49-69)."""
    errors = []
    for nx, ny in ((40, 30), (80, 60), (120, 90)):
        s = box(nx, ny)
        gc, g = s.gc, s.gg
        xx, yy = np.meshgrid(gc.xu, gc.yu); rr = np.hypot(xx - s.xc, yy - s.yc)
        u = -(rr - s.D / 2) * (yy - s.yc) / rr
        xx, yy = np.meshgrid(gc.xv, gc.yv); rr = np.hypot(xx - s.xc, yy - s.yc)
        v = (rr - s.D / 2) * (xx - s.xc) / rr
        p = np.broadcast_to(gc.xp, s.p.shape)
        forces = g.surface_forces(u, v, p, s.nu * s.rho)
        assert abs(forces[2] + np.pi * (s.D / 2)**2) < 1e-10
        assert abs(forces[3]) < 1e-10
        dn = np.column_stack((g.du_dn @ u.ravel(), g.dv_dn @ v.ravel()))
        tangential = -g.normals[:, 1] * dn[:, 0] + g.normals[:, 0] * dn[:, 1]
        torque = s.nu * (s.D / 2) * np.sum(tangential) * (np.pi * s.D / len(tangential))
        exact = 2 * np.pi * s.nu * (s.D / 2)**2
        errors.append(abs(torque - exact) / exact)
    assert errors[2] < errors[1] < errors[0] and errors[-1] < .01, errors


def test_uniform_pressure_gives_no_force():
    s = box()
    f = s.gg.surface_forces(np.zeros_like(s.u), np.zeros_like(s.v), np.full(s.p.shape, 3.7), 1.)
    assert max(abs(np.array(f))) < 1e-12


# ----------------------------------------------------------------------------- divergence reduction
def _reference_engine():
    for path in (ROOT / 'oracle' / '_ref').glob('mac_engine*.so'):
        spec = importlib.util.spec_from_file_location('mac_engine', path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        return mod
    return None


def test_divergence_stats_against_reference_engine():
    """oracle.divergence_stats == calculate_mac_divergence_cpp built from solver/src/bindings.cpp."""
    eng = _reference_engine()
    if eng is None:
        pytest.skip('oracle/_ref not built (make -C oracle needs /root/reference)')
    s = channel()
    rng = np.random.default_rng(2)
    u = rng.normal(size=s.u.shape); v = rng.normal(size=s.v.shape)
    ref = eng.calculate_mac_divergence_cpp(u, v, s.mask, s.dx, s.dy)
    mine = O.divergence_stats(u, v, s.mask, s.dx, s.dy)
    assert ref[2] == mine[2] and ref[4] == mine[4]
    assert abs(ref[0] - mine[0]) <= 1e-15 * ref[0]
    assert abs(ref[1] - mine[1]) <= 1e-12 * ref[1] and abs(ref[3] - mine[3]) <= 1e-12 * ref[3]


# ----------------------------------------------------------------------------- restart
def test_complete_restart():
    """7 + 13 steps through a checkpoint equal 20 steps (tests/test_numba_ghost_cell.py:87-106)."""
    continuous = channel(); divided = channel()
    continuous.solve(n_steps=20, dt=.001, perturb_transient=False)
    divided.solve(n_steps=7, dt=.001, perturb_transient=False)
    with tempfile.TemporaryDirectory() as directory:
        path = Path(directory) / 'checkpoint.npz'
        divided.save_checkpoint(path)
        resumed = O.CylinderFlow.from_checkpoint(path)
        resumed.solve(n_steps=13, dt=.001, perturb_transient=False)
        for key in ('u', 'v', 'p'):
            assert np.max(abs(getattr(continuous, key) - getattr(resumed, key))) < 1e-11, key
        with np.load(path, allow_pickle=False) as data:
            arrays = {k: data[k].copy() for k in data.files if k != 'half_pressure'}
        np.savez(path, **arrays)
        with pytest.raises(ValueError):
            O.CylinderFlow.from_checkpoint(path)


def test_committed_long_run_history_is_the_oracles():
    """The first 15 steps of the committed 440x82 trajectory (tests/golden/oracle_histories.npz, the
    reference for the whole-run GPU tests) are what the oracle computes now, to the last bit."""
    hist = np.load(ROOT / 'tests' / 'golden' / 'oracle_histories.npz')['channel440x82/history']
    s = O.CylinderFlow(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5,
                       initial='zero')
    for k in range(15):
        s.step(2e-3)
        d = s.last_step_diagnostics
        assert np.array_equal([d['time'], d['cd'], d['cl'], d['max_divergence']], hist[k]), k


@pytest.fixture(scope='module')
def openmp_kernels():
    """The OpenMP loops of the CPU baseline (oracle/cpu_kernels.cpp), built if missing."""
    if not (ROOT / 'oracle' / 'libghost_cpu.so').exists():
        import subprocess
        subprocess.run(['make', '-C', str(ROOT / 'oracle'), 'cpu'], check=True, capture_output=True)
    return O.OpenMPKernels(3)


def test_openmp_fluxes_and_products_equal_numpy(openmp_kernels):
    """The C++/OpenMP flux loops and CSR products equal the NumPy/SciPy ones bit for bit, also on odd
    sizes and for CSC input (summed in csc_matvec's column order)."""
    rng = np.random.default_rng(11)
    for ny, nx in ((16, 88), (5, 7), (4, 4)):
        u, v = rng.normal(size=(ny, nx + 1)), rng.normal(size=(ny + 1, nx))
        for adv in (True, False):
            a = O.flux_rhs(u, v, 0.1, 0.2, 0.03, adv, 1.3)
            b = openmp_kernels.flux_rhs(u, v, 0.1, 0.2, 0.03, adv, 1.3)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (ny, nx, adv)
    m = sp.random(300, 200, density=0.05, random_state=3, format='csr')
    x = rng.normal(size=200)
    assert np.array_equal(openmp_kernels.matvec(m, x), m @ x)
    c = m.tocsc()
    assert np.array_equal(openmp_kernels.matvec(c, x), c @ x)


def test_openmp_steps_equal_numpy_steps(openmp_kernels):
    """Whole AB2/CN steps with the OpenMP loops (the CPU baseline) are the oracle's steps bit for bit."""
    a = channel(initial='zero', nu=0.05, inflow='parabolic', U_m=1.5)
    b = channel(initial='zero', nu=0.05, inflow='parabolic', U_m=1.5, openmp_threads=3)
    for k in range(5):
        a.step(2e-3)
        b.step(2e-3)
        for key in ('u', 'v', 'p'):
            assert np.array_equal(getattr(a, key), getattr(b, key)), (k, key)
        assert a.last_step_diagnostics['cd'] == b.last_step_diagnostics['cd']
        assert a.last_step_diagnostics['cl'] == b.last_step_diagnostics['cl']


def test_extra_refinements_keep_the_residual_at_its_floor():
    """The study option behind the long-run force spread: refinements past the reference's rule keep
    the pressure residual at its round-off floor (no growth) while the forces move by at most the
    impulsive-start spread."""
    kw = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2., initial='zero')
    a, b = O.CylinderFlow(**kw), O.CylinderFlow(extra_refinements=6, **kw)
    for _ in range(3):
        a.step(1e-3)
        b.step(1e-3)
        ra = a.last_step_diagnostics['pressure_residual_linf']
        rb = b.last_step_diagnostics['pressure_residual_linf']
        assert rb <= 4 * max(ra, 1e-14), (ra, rb)
        assert abs(a.last_step_diagnostics['cd'] - b.last_step_diagnostics['cd']) < 1e-9
