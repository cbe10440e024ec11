"""GPU parity of the pressure backends through the C ABI: the GMRES backend (restarted GMRES on the
pressure operator, left-preconditioned by the rectangle inverse, pressure_schur.py:16-72) against the
oracle's SciPy GMRES, a sparse LU solve and the direct backend; the reference's direct/GMRES step
comparison (tests/test_gemini_followup.py:47-64); a non-convergence forced from the host.

Tolerances: a GMRES solution stops at ||b - A x||_2 <= 1e-12 ||b||_2, so two such solutions (or one
and the exact solution) differ by up to cond(A) times that; the bounds below are the reference's
direct/GMRES field bounds (1e-9 for u, v and 1e-8 for p) and 1e-8 relative for single solves.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

REF = dict(nx=40, ny=30, Lx=4., Ly=3., xc=1.5, yc=1.5)   # test_gemini_followup.py:24-26
CFG1 = dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5)


def flow(pkg, **kw):
    return pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)


def rel(a, b):
    return float(np.max(abs(np.asarray(a) - np.asarray(b))) / max(np.max(abs(np.asarray(b))), 1e-300))


@pytest.mark.parametrize('kw', [REF, CFG1], ids=['40x30', '440x82'])
@pytest.mark.parametrize('backend', ['direct', 'iterative'])
def test_monitored_solve_matches_oracle(pkg, kw, backend):
    print('pressure solve', backend, kw['nx'], kw['ny'], flush=True)
    ref = O.CylinderFlow(pressure_solver=backend, **kw).gg
    gpu = flow(pkg, pressure_solver=backend, **kw).gg
    vals = np.random.default_rng(11).normal(size=len(ref.pi))
    xr = ref.pressure_solve(vals)
    rr = dict(ref.pressure_inverse_info['last_solve'])
    xd = gpu.pressure_solve(vals)
    rd = gpu.pressure_inverse_info['last_solve']
    exact = spla.splu(ref.A.tocsc()).solve(vals)
    print('record', rd, 'oracle', rr, 'error vs LU', rel(xd, exact), flush=True)
    assert gpu.pressure_inverse_info['backend'] == ref.pressure_inverse_info['backend']
    assert rd['status'] == 0 and rd['residual_linf'] <= rd['residual_target_linf']
    assert np.max(abs(ref.A @ xd - vals)) <= rd['residual_target_linf']
    assert rel(xd, exact) < 1e-8 and rel(xd, xr) < 1e-8
    assert rd['refinement_iterations'] == rr['refinement_iterations']
    if backend == 'iterative':
        assert rd['iterations'] > 0 and abs(rd['iterations'] - rr['iterations']) <= 2
        assert rd['seconds'] > 0
    else:
        assert rd['iterations'] == 0 and rel(xd, xr) < 1e-11


def test_direct_and_gmres_steps_agree(pkg):
    outputs, last = [], {}
    for backend in ('direct', 'iterative'):
        print('12 steps', backend, flush=True)
        s = flow(pkg, pressure_solver=backend, **REF)
        saved = []
        s.solve(n_steps=12, dt=.001, perturb_transient=False, save_artifacts=False, verbose=False,
                step_callback=lambda _, d: saved.append(d))
        assert all(len(d['pressure_solves']) == d['pressure_solve_count'] for d in saved)
        assert len({id(d['pressure_solves']) for d in saved}) == len(saved)
        assert saved[0]['pressure_solve_count'] == 3 and saved[-1]['pressure_solve_count'] == 2
        d = s.last_step_diagnostics
        assert d['pressure_residual_linf'] < 1e-9 and d['pressure_solve_count'] >= 2, d
        assert d['pressure_iterations'] >= 0 and d['pressure_seconds'] >= 0, d
        assert d['pressure_backend'] == {'direct': 'fft_boundary_schur', 'iterative': 'gmres_rectangle'}[backend]
        outputs.append((s.u.copy(), s.v.copy(), s.p.copy()))
        last[backend] = d
    errors = [float(np.max(abs(a - b))) for a, b in zip(*outputs)]
    print('direct/GMRES field differences', errors, flush=True)
    assert max(errors[:2]) < 1e-9 and errors[2] < 1e-8, errors
    assert last['iterative']['pressure_iterations'] > 0 and last['direct']['pressure_iterations'] == 0


def test_gmres_steps_match_oracle(pkg):
    print('12 GMRES steps vs the oracle', flush=True)
    ref = O.CylinderFlow(pressure_solver='iterative', **REF)
    gpu = flow(pkg, pressure_solver='iterative', **REF)
    rows = []
    ref.solve(n_steps=12, dt=.001, perturb_transient=False, step_callback=lambda _, d: rows.append(d))
    gpu.solve(n_steps=12, dt=.001, perturb_transient=False, save_artifacts=False, verbose=False)
    print('u', rel(gpu.u, ref.u), 'v', rel(gpu.v, ref.v), 'p', rel(gpu.p, ref.p), flush=True)
    assert rel(gpu.u, ref.u) < 1e-9 and rel(gpu.v, ref.v) < 1e-9 and rel(gpu.p, ref.p) < 1e-8
    d, r = gpu.last_step_diagnostics, rows[-1]
    assert abs(d['pressure_iterations'] - r['pressure_iterations']) <= 2 * d['pressure_solve_count']
    assert d['max_divergence'] < 1e-10 and abs(d['cd'] - r['cd']) <= 1e-8 * max(abs(r['cd']), abs(r['cl']))


def test_gmres_nonconvergence_is_reported(pkg):
    """Two Arnoldi steps in one cycle cannot reach 1e-15: status 1, no refinement, RuntimeError."""
    s = flow(pkg, pressure_solver='iterative', **REF)
    dev = s.gg.device
    dev.set_pressure_backend(True, restart=2, max_cycles=1, rtol=1e-15, atol=1e-15)
    vals = np.random.default_rng(12).normal(size=len(s.gg.pi))
    with pytest.raises(RuntimeError):
        s.gg.pressure_solve(vals)
    rec = s.gg.pressure_inverse_info['last_solve']
    assert rec['status'] == 1 and rec['iterations'] == 2 and rec['refinement_iterations'] == 0, rec
    with pytest.raises(ValueError):
        s.gg.pressure_solve(np.full(len(s.gg.pi), np.nan))
    # back to the reference's settings: the same handle solves again
    dev.set_pressure_backend(True)
    x = s.gg.pressure_solve(vals)
    assert s.gg.pressure_inverse_info['last_solve']['status'] == 0
    assert np.max(abs(s.gg.A @ x - vals)) <= s.gg.pressure_inverse_info['last_solve']['residual_target_linf']
    assert dev.pressure_backend == 1
    dev.set_pressure_backend(False)
    assert dev.pressure_backend == 0 and s.gg.pressure_inverse_info['backend'] == 'fft_boundary_schur'


def test_gmres_graph_groups_match_eager(pkg, monkeypatch):
    """Arnoldi groups replayed from the captured graph equal eagerly launched groups bit for bit,
    over several solves and steps with host reads (synchronisations) between the replays."""
    monkeypatch.setenv('GHOSTFLOW_NO_GRAPH', '1')
    eager = flow(pkg, pressure_solver='iterative', **REF)
    monkeypatch.setenv('GHOSTFLOW_NO_GRAPH', '0')
    graph = flow(pkg, pressure_solver='iterative', **REF)
    for f in (eager, graph):
        f.gg.device.set_launch(0, 1)      # same sweep width for both (widths differ in round-off)
    rng = np.random.default_rng(13)
    for _ in range(3):
        vals = rng.normal(size=len(eager.gg.pi))
        assert np.array_equal(eager.gg.pressure_solve(vals), graph.gg.pressure_solve(vals))
        assert (eager.gg.pressure_inverse_info['last_solve']['iterations']
                == graph.gg.pressure_inverse_info['last_solve']['iterations'])
    for chunk in (2, 3):
        for f in (eager, graph):
            f.solve(n_steps=chunk, dt=1e-3, perturb_transient=False, save_artifacts=False)
        for key in ('u', 'v', 'p'):
            assert np.array_equal(getattr(eager, key), getattr(graph, key)), key
