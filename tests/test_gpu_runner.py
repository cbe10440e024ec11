"""GPU: wake samples of a live solver and the stage-JSON runner end to end (with a restart)."""
import importlib
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mods(pkg):
    return (importlib.import_module(pkg.__name__ + '.observables'), importlib.import_module(pkg.__name__ + '.runner'))


def test_cylinder_samples_of_linear_fields(pkg, mods, tmp_path):
    """Linear-in-x fields (tests/test_gemini_followup.py:88-98): the quadratic reconstruction makes
    p_front - p_back = -D exactly and the centre-line velocity crosses zero 0.1 D behind the rear."""
    ob, _ = mods
    s = pkg.MAC_CylinderFlow(nx=40, ny=30, Lx=4., Ly=3., xc=1.5, yc=1.5, boundary='ghost_cell',
                             time_scheme='ab2_cn', verbose=False)
    path = tmp_path / 'samples.bin'
    samples = ob.CylinderSamples(s, path, stride=1)
    for t in (0., .2, .7, 1.1, 1.5):
        s.time = t
        s.u[:] = np.arange(s.nx + 1) * s.dx - (s.xc + s.D / 2) - .4 + t
        s.p[:] = np.broadcast_to(s.gc.xp, s.p.shape) + 2 * t
        samples.record()
    samples.close()
    r = ob.mean_observables(path, .1, .5)
    assert abs(r['delta_p_mean'] + s.D) < 1e-10, r
    assert abs(r['recirculation']['length_D'] - .1) < 1e-10, r


def test_runner_end_to_end_with_restart(pkg, mods, tmp_path):
    _, rn = mods
    solver = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2., boundary='ghost_cell', time_scheme='ab2_cn')
    case = dict(dt=1e-3, end_time=0.03, sample_start=0.01, kind='steady', profile_stride=5, solver=solver)
    cfg = tmp_path / 'case.json'
    cfg.write_text(json.dumps(case))
    rn.run(cfg, tmp_path / 'a')
    rows = [json.loads(x) for x in (tmp_path / 'a' / 'history.jsonl').read_text().splitlines()]
    assert len(rows) == 30 and rows[-1]['step'] == 30 and abs(rows[-1]['time'] - 0.03) < 1e-12
    assert all(r['max_divergence'] < 1e-12 for r in rows) and 'delta_p' in rows[0]
    summary = json.loads((tmp_path / 'a' / 'summary.json').read_text())
    assert summary['completed'] and summary['steps'] == 30 and summary['sample_window'][0] == 0.01
    st = summary['statistics']
    cd = np.array([r['cd'] for r in rows])
    assert min(cd) - 1e-12 <= st['mean'][0] <= max(cd) + 1e-12
    assert json.loads((tmp_path / 'a' / 'mean_observables.json').read_text())['window'][1] == pytest.approx(0.03)
    manifest = json.loads((tmp_path / 'a' / 'manifest.json').read_text())
    assert manifest['case'] == case and manifest['metadata']['n_steps'] == 30
    # delta_p rows come from the device probes: equal to the host reconstruction of the final state
    ob, _ = mods
    end = pkg.MAC_CylinderFlow.from_checkpoint(tmp_path / 'a' / 'checkpoint.npz')
    p = ob.stagnation_operator(end) @ end.p.ravel()
    assert abs(rows[-1]['delta_p'] - (p[0] - p[1])) <= 1e-12 * np.max(abs(p))
    # samples as the reference's stride counter takes them: the initial one, steps 4, 9, .., 29, the final one
    meta = json.loads((tmp_path / 'a' / 'centerline.json').read_text())
    assert (tmp_path / 'a' / 'centerline.bin').stat().st_size == 8 * 8 * meta['columns']
    # restart to 0.05 equals a direct run of 50 steps
    case2 = dict(case, end_time=0.05, sample_start=0.035)   # after the restart time (30 accumulated steps)
    cfg2 = tmp_path / 'case2.json'
    cfg2.write_text(json.dumps(case2))
    rn.run(cfg2, tmp_path / 'b', restart=tmp_path / 'a' / 'checkpoint.npz')
    rows2 = [json.loads(x) for x in (tmp_path / 'b' / 'history.jsonl').read_text().splitlines()]
    assert len(rows2) == 20 and abs(rows2[-1]['time'] - 0.05) < 1e-12
    direct = pkg.MAC_CylinderFlow(verbose=False, **solver)
    t, cd, cl = direct.solve(n_steps=50, dt=1e-3, perturb_transient=False, save_artifacts=False, verbose=False,
                             step_callback=lambda state, diag: None)
    assert abs(cd[-1] - rows2[-1]['cd']) <= 1e-12 * abs(cd[-1])
    with pytest.raises(FileExistsError):
        rn.run(cfg2, tmp_path / 'b', restart=tmp_path / 'a' / 'checkpoint.npz')


def test_runner_history_matches_oracle(pkg, mods, tmp_path):
    """The runner's rows and end state against the oracle's uninterrupted steps of the same case
    (validation/run_ghost_validation.py:19-102 drives main.py's step loop): C_D and C_L of every
    history row within 1e-10 of the step's force magnitude, the divergence under 1e-12, the final
    checkpointed fields within 1e-10 of the oracle's, and delta_p equal to the oracle state's."""
    from oracle.ghost_oracle import CylinderFlow
    ob, rn = mods
    kw = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)
    case = dict(dt=1e-3, end_time=0.03, sample_start=0.01, kind='steady', profile_stride=5,
                solver=dict(kw, boundary='ghost_cell', time_scheme='ab2_cn'))
    cfg = tmp_path / 'case.json'
    cfg.write_text(json.dumps(case))
    print('runner: 88x16 channel, 30 steps at dt=1e-3, against the oracle', flush=True)
    rn.run(cfg, tmp_path / 'out')
    rows = [json.loads(x) for x in (tmp_path / 'out' / 'history.jsonl').read_text().splitlines()]
    ref = CylinderFlow(**kw)
    worst = 0.0
    for row in rows:
        ref.step(1e-3)
        d = ref.last_step_diagnostics
        scale = max(abs(d['cd']), abs(d['cl']))
        err = max(abs(row['cd'] - d['cd']), abs(row['cl'] - d['cl'])) / scale
        worst = max(worst, err)
        assert err <= 1e-10, (row['step'], row['cd'], d['cd'], row['cl'], d['cl'])
        assert row['max_divergence'] < 1e-12
    print(f'worst force difference {worst:.2e} of the force magnitude over {len(rows)} rows', flush=True)
    assert len(rows) == 30
    end = pkg.MAC_CylinderFlow.from_checkpoint(tmp_path / 'out' / 'checkpoint.npz')
    for key in ('u', 'v', 'p'):
        a, b = getattr(end, key), getattr(ref, key)
        err = float(np.max(abs(a - b)) / np.max(abs(b)))
        print(f'final {key}: relative difference {err:.2e}', flush=True)
        assert err < 1e-10, (key, err)
    p = ob.stagnation_operator(end) @ ref.p.ravel()
    assert abs(rows[-1]['delta_p'] - (p[0] - p[1])) <= 1e-10 * np.max(abs(p))
