"""GPU parity: libghostflow.so (through the C ABI) against the CPU oracle on the same inputs."""
import numpy as np
import pytest

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

CH = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)
CFG1 = dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5)


def rel(a, b):
    return float(np.max(abs(np.asarray(a) - np.asarray(b))) / max(np.max(abs(np.asarray(b))), 1e-300))


@pytest.fixture(scope='module')
def pair88(pkg):
    print('case 88x16 channel', flush=True)
    ref = O.CylinderFlow(**CH)
    gpu = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **CH)
    return ref, gpu


def test_initial_projection(pair88):
    ref, gpu = pair88
    assert rel(gpu.u, ref.u) < 1e-12 and rel(gpu.v, ref.v) < 1e-12


@pytest.mark.parametrize('kind', ['pressure', 'u', 'v'])
def test_rectangle_inverse(pair88, kind):
    ref, gpu = pair88
    dev = gpu.gg.device
    dev.set_dt(1e-3)
    code = {'pressure': 0, 'u': 1, 'v': 2}[kind]
    factor = dev.pressure_factor if kind == 'pressure' else dev.helmholtz_factors[code - 1][0]
    b = np.random.default_rng(3).normal(size=factor.shape)
    x = dev.rect_inverse(code, b)
    _, inv, _ = O.rectangle(gpu.nx, gpu.ny, gpu.dx, gpu.dy, kind, gpu.nu, 1e-3, gpu.rho)
    assert rel(x, inv(b)) < 1e-12
    # residual self-check of rectangular_inverse.py:54-57
    assert np.max(abs(factor.matrix @ x.ravel() - b.ravel())) < 1e-10


def test_pressure_schur_solve(pair88):
    ref, gpu = pair88
    g = ref.gg
    vals = np.random.default_rng(4).normal(size=len(g.pi))
    xr = g.pressure_solve(vals)
    rhs = np.zeros(gpu.nx * gpu.ny); rhs[g.pi] = vals
    x, info = gpu.gg.device.schur_solve(0, rhs)
    assert rel(x[g.pi], xr) < 1e-11
    assert np.max(abs(g.A @ x[g.pi] - vals)) <= info[3]
    covered = np.ones(gpu.nx * gpu.ny, bool); covered[g.pi] = False
    assert np.all(x[covered] == 0.)


def test_operators(pair88):
    ref, gpu = pair88
    g, d = ref.gg, gpu.gg
    rng = np.random.default_rng(6)
    q = rng.normal(size=g.qsize)
    assert rel(d.advection(q), g.advection(q)) < 1e-12
    assert rel(d.diffusion(q), g.diffusion(q)) < 1e-12
    assert rel(d.divergence(q), g.divergence(q)) < 1e-12
    assert rel(d.H_matvec(q, 1e-3), g.H_matvec(q, 1e-3)) < 1e-12
    ph = rng.normal(size=len(g.pi))
    assert rel(d.gradient(ph), g.gradient(ph)) < 1e-12
    qp, phi = d.project(q, .003)
    qr, phr = g.project(q, .003)
    assert rel(qp, qr) < 1e-11 and rel(phi, phr) < 1e-10
    assert np.max(abs(g.divergence(qp))) < 1e-10


def test_steps_88x16(pair88):
    """12 steps vs the oracle (the reference's compiled/pure parity uses 1e-10, test_numba_ghost_cell.py:72-84)."""
    ref, gpu = pair88
    ref = O.CylinderFlow(**CH)
    t, cd, cl = gpu.solve(n_steps=12, dt=1e-3, perturb_transient=False, save_artifacts=False, verbose=False)
    for _ in range(12):
        ref.step(1e-3)
    for key in ('u', 'v', 'p'):
        assert rel(getattr(gpu, key), getattr(ref, key)) < 1e-11, key
    dr = ref.last_step_diagnostics
    assert abs(cd[-1] - dr['cd']) <= 1e-10 * abs(dr['cd'])
    assert abs(cl[-1] - dr['cl']) <= 1e-10 * abs(dr['cl'])
    assert gpu.last_step_diagnostics['max_divergence'] < 1e-12


def test_surface_forces_closed_form(pkg):
    """p = x gives F_x,p = -pi R^2 on the device too (tests/test_numba_ghost_cell.py:49-62)."""
    s = pkg.MAC_CylinderFlow(nx=80, ny=60, Lx=4., Ly=3., xc=1.5, yc=1.5, boundary='ghost_cell',
                             time_scheme='ab2_cn', verbose=False)
    gc = s.gc
    p = np.broadcast_to(gc.xp, s.p.shape).copy()
    f = s.gg.surface_forces(np.zeros_like(s.u), np.zeros_like(s.v), p, s.nu * s.rho)
    assert abs(f[2] + np.pi * (s.D / 2)**2) < 1e-10 and abs(f[3]) < 1e-10


def test_config1_steps(pkg):
    """440x82 Re=20 channel from the zero field: 20 steps, fields 1e-11 and forces 1e-10 relative."""
    print('case 440x82 Re=20, 20 steps', flush=True)
    ref = O.CylinderFlow(initial='zero', **CFG1)
    gpu = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, initial='zero', **CFG1)
    t, cd, cl = gpu.solve(n_steps=20, dt=2e-3, perturb_transient=False, save_artifacts=False, verbose=False)
    cdr, clr = [], []
    for _ in range(20):
        ref.step(2e-3)
        cdr.append(ref.last_step_diagnostics['cd']); clr.append(ref.last_step_diagnostics['cl'])
    for key in ('u', 'v', 'p'):
        assert rel(getattr(gpu, key), getattr(ref, key)) < 1e-11, key
    # C_L is a small difference of pressure integrals of size C_D; two exact inverses of the same
    # operator move it by ~2e-10 at the impulsive first step (test_force_floor_of_exact_inverses), so the
    # force tolerance is 1e-10 relative to the force magnitude max(|C_D|, |C_L|) of each step.
    scale = np.maximum(abs(np.array(cdr)), abs(np.array(clr)))
    assert np.max(abs(cd - cdr) / scale) < 1e-10
    assert np.max(abs(cl - clr) / scale) < 1e-10
    assert gpu.last_step_diagnostics['max_divergence'] < 1e-12


def test_divergence_stats(pkg, pair88):
    ref, gpu = pair88
    rng = np.random.default_rng(8)
    u = rng.normal(size=gpu.u.shape); v = rng.normal(size=gpu.v.shape)
    mine = gpu.gg.device.divergence_stats(u, v, gpu.mask)
    want = O.divergence_stats(u, v, gpu.mask, gpu.dx, gpu.dy)
    assert mine[2] == want[2] and mine[4] == want[4]
    assert abs(mine[0] - want[0]) <= 1e-15 * want[0]
    assert abs(mine[1] - want[1]) <= 1e-12 * want[1]


@pytest.mark.parametrize('perturbed', [False, True], ids=['plain', 'perturbed inflow'])
def test_graph_replays_match_eager(pkg, monkeypatch, perturbed):
    """Steps replayed from the captured HIP graphs equal eagerly launched steps bit for bit,
    across several solve() calls with host reads (synchronisations) between the replays; with the
    inflow perturbation (its window ends in the second call) the replayed steps set each inlet from
    the clock while the previous step's clock advance is still pending."""
    kw = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)
    pk = dict(perturb_transient=True, perturb_duration=0.005, perturb_amp=0.15) if perturbed else \
        dict(perturb_transient=False)
    monkeypatch.setenv('GHOSTFLOW_NO_GRAPH', '1')
    eager = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)
    monkeypatch.setenv('GHOSTFLOW_NO_GRAPH', '0')
    graph = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)
    for f in (eager, graph):
        f.gg.device.set_launch(0, 1)   # the same launch shapes for both (autotuned shapes can differ in round-off)
    for chunk in (3, 5, 4):
        te, cde, cle = eager.solve(n_steps=chunk, dt=1e-3, save_artifacts=False, **pk)
        tg, cdg, clg = graph.solve(n_steps=chunk, dt=1e-3, save_artifacts=False, **pk)
        for key in ('u', 'v', 'p'):
            assert np.array_equal(getattr(eager, key), getattr(graph, key)), key
        assert np.array_equal(cde, cdg) and np.array_equal(cle, clg) and np.array_equal(te, tg)


def test_deferred_step_tails_give_the_eager_rows(pkg, monkeypatch):
    """A replayed step leaves its tail (surface forces, slot gather into its diagnostics row, clock) to
    the next step's launches or to the end of the call: the rows of calls of 1, 4, 1 and 3 steps equal
    those of eagerly launched steps bit for bit, every row holds its forces, and the clock advances
    by dt per row."""
    kw = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)
    monkeypatch.setenv('GHOSTFLOW_NO_GRAPH', '1')
    eager = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)
    monkeypatch.setenv('GHOSTFLOW_NO_GRAPH', '0')
    graph = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)
    for f in (eager, graph):
        f.gg.device.set_launch(0, 1)
    names = list(nat_names(pkg))
    dt = 1e-3
    rows_e, rows_g = [], []
    for chunk in (1, 4, 1, 3):
        rows_e.append(eager.gg.device.run(chunk, dt))
        rows_g.append(graph.gg.device.run(chunk, dt))
    re, rg = np.concatenate(rows_e), np.concatenate(rows_g)
    assert np.array_equal(re, rg)
    t = rg[:, names.index('time')]
    assert np.allclose(np.diff(t), dt, rtol=0, atol=1e-15)
    assert np.all(rg[:, names.index('cd')] != 0.0) and np.all(rg[:, names.index('status')] == 0)


def nat_names(pkg):
    import importlib
    return importlib.import_module(pkg.__name__ + '._native').DIAG_NAMES


def test_launch_shapes_agree(pkg):
    """Every staged transform tile gives bit-identical steps, and so do the two split-K tile heights
    of one block size (8 or 4 waves); the split-K transforms, the parity-split transforms (8 and 4
    waves; the 81-row v rectangle has a middle row) and every sweep width agree with the staged tiles
    to round-off."""
    runs = {}
    for g, t in [(0, 1), (1, 1), (2, 1), (0, 0), (0, 2), (3, 1), (4, 1), (5, 1), (6, 1), (8, 1), (9, 1)]:
        f = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **CFG1)
        f.gg.device.set_launch(g, t)
        f.solve(n_steps=6, dt=2e-3, perturb_transient=False, save_artifacts=False, verbose=False)
        runs[(g, t)] = (f.u.copy(), f.v.copy(), f.p.copy())
        assert f.gg.device.get_launch()['tri_pressure'] == t
    ref = runs[(0, 1)]
    for key in [(1, 1), (2, 1)]:
        for a, b in zip(runs[key], ref):
            assert np.array_equal(a, b), key
    for key in [(0, 0), (0, 2), (3, 1), (5, 1), (8, 1), (9, 1)]:
        for a, b in zip(runs[key], ref):
            assert np.max(abs(a - b)) <= 1e-12 * np.max(abs(b)), key
    for a, b in zip(runs[(4, 1)], runs[(3, 1)]):
        assert np.array_equal(a, b)
    for a, b in zip(runs[(6, 1)], runs[(5, 1)]):
        assert np.array_equal(a, b)
    f = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **CFG1)
    f.solve(n_steps=2, dt=2e-3, perturb_transient=False, save_artifacts=False, verbose=False)
    ln = f.gg.device.get_launch()
    print('autotuned launch shapes (440x82):', ln)
    assert ln['chosen'] and all(ln[k] in (0, 1, 2, 3, 4, 5, 6, 8, 9) for k in ('gemm_pressure', 'gemm_velocity'))
    assert all(0 <= ln[k] < 3 for k in ('tri_pressure', 'tri_velocity'))


@pytest.mark.parametrize('tile', [3, 4, 5, 6, 8, 9], ids=['32-row', '16-row', '32-row-4-waves', '16-row-4-waves',
                                                   'parity-split', 'parity-split-4-waves'])
@pytest.mark.parametrize('kind', ['pressure', 'u', 'v'])
def test_split_k_transform_inverse(pair88, kind, tile):
    """The split-K and parity-split transforms inside the rectangle inverse against the oracle's
    DCT/DST inverse (the 88x16 grid's v rectangle has 15 rows: a middle row of its own)."""
    ref, gpu = pair88
    dev = gpu.gg.device
    dev.set_dt(1e-3)
    before = dev.get_launch()
    dev.set_launch(tile, 1)
    code = {'pressure': 0, 'u': 1, 'v': 2}[kind]
    factor = dev.pressure_factor if kind == 'pressure' else dev.helmholtz_factors[code - 1][0]
    b = np.random.default_rng(31).normal(size=factor.shape)
    x = dev.rect_inverse(code, b)
    _, inv, _ = O.rectangle(gpu.nx, gpu.ny, gpu.dx, gpu.dy, kind, gpu.nu, 1e-3, gpu.rho)
    assert rel(x, inv(b)) < 1e-12
    dev.set_launch(before['gemm_pressure'], before['tri_pressure'])


def test_parity_split_needs_the_reflection_parity(pkg):
    """The parity-split transform runs only for a class whose bases have V[yy-1-j][m] = (-1)^m V[j][m]:
    with the pressure basis's first two columns rotated into each other (still orthonormal, no longer
    of one parity each), a pinned parity-split transform runs as the 16-row split-K transform for
    pressure only."""
    import importlib
    nat = importlib.import_module(pkg.__name__ + '._native')
    f = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **CH)
    dev = f.gg.device
    dev.set_launch(8, 1)
    f.solve(n_steps=1, dt=1e-3, perturb_transient=False, save_artifacts=False, verbose=False)
    ln = dev.get_launch()
    assert ln['gemm_pressure'] == 8 and ln['gemm_velocity'] == 8
    fac = dev.pressure_factor
    V = np.array(fac.V, dtype=float)
    c0, c1 = V[:, 0].copy(), V[:, 1].copy()
    V[:, 0], V[:, 1] = (c0 + c1) / np.sqrt(2), (c0 - c1) / np.sqrt(2)
    yy, xx = fac.shape
    dev.h.call('mcf_set_rectangle', 0, yy, xx, float(fac.a), float(fac.b), nat.f64(V), nat.f64(fac.lam),
               nat.f64(fac.lower), nat.f64(fac.diag), nat.f64(fac.upper))
    ln = dev.get_launch()
    assert ln['gemm_pressure'] == 4 and ln['gemm_velocity'] == 8, ln
    dev.set_launch_classes(dict(gemm_pressure=9, gemm_velocity=9, tri_pressure=1, tri_velocity=1))
    ln = dev.get_launch()
    assert ln['gemm_pressure'] == 4 and ln['gemm_velocity'] == 9, ln


def test_speculative_steps_match_full(pkg):
    """Lean steady steps (refinement launches left out, replayed in full on a miss) give the same
    fields and diagnostics as always launching the refinements; forcing a miss at every lean step
    replays every chunk and still gives the same result."""
    runs = {}
    for mode in (0, 1, 2):
        f = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **CFG1)
        dev = f.gg.device
        dev.set_launch(0, 1)          # same sweep width in every run (widths differ in round-off)
        dev.set_speculation(mode)
        f._push()
        d = np.concatenate([dev.run(7, 2e-3), dev.run(5, 2e-3)])
        f._pull()
        runs[mode] = (f.u.copy(), f.v.copy(), f.p.copy(), d, dev.get_speculation())
    for mode in (1, 2):
        for a, b in zip(runs[mode][:3], runs[0][:3]):
            assert np.array_equal(a, b), mode
        for col in (0, 2, 3, 10, 11, 12, 14, 18, 25, 26):
            assert np.array_equal(runs[mode][3][:, col], runs[0][3][:, col]), (mode, col)
    print('replays with speculation on:', runs[1][4]['replays'], '(a chunk replays when one of its steps refines)')
    # after a miss the next chunk runs in full until a chunk's last steps refine nothing, so the
    # start-up refinements cost at most a couple of replays; mode 2 replays every lean chunk
    assert runs[1][4]['replays'] <= 2 and runs[2][4]['replays'] >= 1


@pytest.mark.parametrize('kw', [CH, dict(CFG1, initial='zero')])
def test_biquadratic_velocity_extension_steps(pkg, kw):
    """ghost_velocity_interpolation='biquadratic' (9-point probes, optimized_ghost_cell.py:34-44):
    the wider ghost rows and their override rows through the same kernels match the oracle."""
    n, dt = 10, 1e-3
    gpu = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False,
                               ghost_velocity_interpolation='biquadratic', **kw)
    t, cd, cl = gpu.solve(n_steps=n, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    ref = O.CylinderFlow(ghost_velocity_interpolation='biquadratic', **kw)
    for _ in range(n):
        ref.step(dt)
    for key in ('u', 'v', 'p'):
        assert rel(getattr(gpu, key), getattr(ref, key)) < 1e-11, key
    d = ref.last_step_diagnostics
    scale = max(abs(d['cd']), abs(d['cl']))
    assert abs(cd[-1] - d['cd']) <= 1e-10 * scale and abs(cl[-1] - d['cl']) <= 1e-10 * scale
    assert gpu.last_step_diagnostics['max_divergence'] < 1e-12


@pytest.mark.parametrize('kw', [CH, dict(CFG1, initial='zero')])
def test_quadratic_pressure_extension_steps(pkg, kw):
    """ghost_pressure_extension='quadratic' (ghost_reconstruction.py:7-45): wider pressure ghost rows
    and the operators built on them through the same kernels match the oracle."""
    n, dt = 10, 1e-3
    gpu = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False,
                               ghost_pressure_extension='quadratic', **kw)
    t, cd, cl = gpu.solve(n_steps=n, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    ref = O.CylinderFlow(ghost_pressure_extension='quadratic', **kw)
    for _ in range(n):
        ref.step(dt)
    for key in ('u', 'v', 'p'):
        assert rel(getattr(gpu, key), getattr(ref, key)) < 1e-11, key
    d = ref.last_step_diagnostics
    scale = max(abs(d['cd']), abs(d['cl']))
    assert abs(cd[-1] - d['cd']) <= 1e-10 * scale and abs(cl[-1] - d['cl']) <= 1e-10 * scale
    assert gpu.last_step_diagnostics['max_divergence'] < 1e-12
    assert gpu.gg.reconstruction_info['pressure'] == 'quadratic'
    assert gpu.gg.reconstruction_info['polynomial_reproduction_linf'] <= 1e-10
