"""GPU: whole benchmark-length runs against the oracle.

* the 440x82 Re 20 channel from a zero field, all 2000 steps at dt = 2e-3;
* the 1320x246 Re 100 channel from the profile with the seeded v perturbation, 500 steps at dt = 5e-4;
* the seeded 64-case ensemble, 200 steps of every case;
* a run restarted from a checkpoint against the oracle's uninterrupted run.

The oracle trajectories of the two channels are committed (tests/golden/oracle_histories.npz,
written by tests/golden/gen_oracle_histories.py).  Bars, for every step k:

* |C_D - C_D'| <= 1e-10 |C_D'| + 2 s_k and |C_L - C_L'| <= 1e-10 max(|C_L'|, CL_FLOOR) + 2 s_k (primes:
  the oracle).  s_k is the spread of that force at step k: how far it moves in a second oracle run
  whose Woodbury solves take six refinements past the reference's stopping rule, their residual
  staying at its round-off floor (``extra_refinements=6``).  It is 1e-10 .. 2e-10 in C_L at the
  impulsive start, where a pressure of ~100 is determined by the rule only to ~1e-9 along the
  operator's weakly determined directions, and below 1e-12 from the third step on.  CL_FLOOR = 0.1
  makes the C_L bar absolute (1e-11) where the lift is small: the lift integral's terms have the
  size of C_D (5 .. 40 here), so its round-off does not shrink with the lift.
* max |div u| < 1e-12.
* Final fields: max-norm error within 1e-11 of the field's max |.| on the stored samples (a block
  within 0.6 D of the cylinder centre and a lattice over the channel), and four seeded random
  projections within 1e-11 of |w|.|x|.
"""
import importlib
import os
from pathlib import Path

import numpy as np
import pytest

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = np.load(Path(__file__).resolve().parent / 'golden' / 'oracle_histories.npz')
CL_FLOOR = 0.1
CHANNELS = {
    'channel440x82': (dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic', U_m=1.5,
                           initial='zero'), 2000),
    'channel1320x246': (dict(nx=1320, ny=246, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.01, inflow='parabolic', U_m=1.5,
                             initial='profile', v_perturbation=dict(seed=20261203)), 500),
}


def gpu_flow(pkg, **kw):
    return pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)


def workers():
    return min(16, len(os.sched_getaffinity(0)))


def force_errors(cd, cl, cd_ref, cl_ref, spread):
    """Per-step errors in units of the bars above (<= 1 passes); spread is (steps, 2): s_k of C_D, C_L."""
    e_cd = abs(np.asarray(cd) - cd_ref) / (1e-10 * abs(cd_ref) + 2 * spread[:, 0])
    e_cl = abs(np.asarray(cl) - cl_ref) / (1e-10 * np.maximum(abs(cl_ref), CL_FLOOR) + 2 * spread[:, 1])
    return e_cd, e_cl


def check_final_fields(flow, name):
    for key, seed in (('u', 70), ('v', 71), ('p', 72)):
        x = getattr(flow, key)
        err = 0.0
        for block in ('win', 'lat'):
            rows, cols = GOLDEN[f'{name}/{key}_{block}_rows'], GOLDEN[f'{name}/{key}_{block}_cols']
            err = max(err, np.max(abs(x[np.ix_(rows, cols)] - GOLDEN[f'{name}/{key}_{block}'])))
        err /= GOLDEN[f'{name}/{key}_max'][0]
        w = np.random.default_rng(seed).normal(size=(4, x.size))
        perr = np.max(abs(w @ x.ravel() - GOLDEN[f'{name}/{key}_proj']) / GOLDEN[f'{name}/{key}_scale'])
        print(name, key, 'sampled max-norm error', err, 'projection error', perr, flush=True)
        assert err < 1e-11 and perr < 1e-11, (key, err, perr)


@pytest.mark.parametrize('name', list(CHANNELS))
def test_channel_whole_run(pkg, name):
    kw, steps = CHANNELS[name]
    hist = GOLDEN[f'{name}/history']
    dt = float(GOLDEN[f'{name}/dt'][0])
    assert len(hist) == steps
    print('long run', name, steps, 'steps', flush=True)
    gpu = gpu_flow(pkg, **kw)
    t, cd, cl = gpu.solve(n_steps=steps, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    print(name, 'launch', gpu.gg.device.get_launch(), 'speculation', gpu.gg.device.get_speculation(), flush=True)
    assert np.max(abs(np.asarray(t) - hist[:, 0])) < 1e-12 * hist[-1, 0]
    e_cd, e_cl = force_errors(cd, cl, hist[:, 1], hist[:, 2], GOLDEN[f'{name}/spread'])
    print(name, 'worst C_D error / bar', e_cd.max(), 'at step', int(e_cd.argmax()),
          '; worst C_L error / bar', e_cl.max(), 'at step', int(e_cl.argmax()), flush=True)
    assert e_cd.max() <= 1 and e_cl.max() <= 1
    div = gpu.diagnostics_history[:, 10]
    assert len(div) == steps and np.all(div < 1e-12), div.max()
    check_final_fields(gpu, name)


def test_ensemble_all_cases_200_steps(pkg):
    """Every case of the seeded 64-case ensemble for 200 steps.  The members that run through keep
    max|div| < 1e-12, finite forces and a clean status word on every step, and end at 200 dt.  The
    members stopped by the reference's CFL gate (main.py:414-415) are checked against the oracle: it
    raises the gate at the same step, with the same CFL number to 1e-8.  One member that runs through
    is also compared with the oracle over all 200 steps."""
    ens = importlib.import_module(pkg.__name__ + '.ensemble')
    cases = [ens.ensemble_case(k) for k in range(64)]
    steps = 200
    print('ensemble: 64 cases,', steps, 'steps', flush=True)
    batch = ens.Ensemble(cases)
    d = batch.run(steps)
    batch.pull()
    failed = dict(batch.failures)
    print('ensemble failures', failed, flush=True)
    assert len(failed) <= 8
    for c, case in enumerate(cases):
        if c in failed:
            s = failed[c]['step']
            assert 'CFL' in failed[c]['message'] and d[c, s, 11] >= 0.5 and d[c, s, 10] < 1e-12, c
            assert np.all(d[c, :s, 10] < 1e-12) and np.all(d[c, :s, 11] < 0.5) and np.all(np.isnan(d[c, s + 1:, 0]))
            continue
        assert np.all(d[c, :, 10] < 1e-12), (c, d[c, :, 10].max())
        assert np.all(np.isfinite(d[c, :, 2])) and np.all(np.isfinite(d[c, :, 3])), c
        assert np.all(d[c, :, 26] == 0), c
        assert d[c, -1, 0] == pytest.approx(steps * case['dt'], rel=1e-13), c
    for c in sorted(failed):
        s = failed[c]['step']
        ref = O.CylinderFlow(initial='zero', fft_workers=workers(), **ens.solver_kwargs(cases[c]))
        print('oracle: case', c, (cases[c]['nx'], cases[c]['ny']), 'to its gate at step', s, flush=True)
        for k in range(s):
            ref.step(cases[c]['dt'])
        with pytest.raises(RuntimeError, match='CFL'):
            ref.step(cases[c]['dt'])
        cfl = ref.last_step_diagnostics['cfl']
        assert abs(d[c, s, 11] - cfl) <= 1e-8 * cfl, (c, d[c, s, 11], cfl)
    # the smallest member that ran through, against the oracle for all 200 steps
    c = min((c for c in range(64) if c not in failed), key=lambda c: cases[c]['nx'] * cases[c]['ny'])
    print('oracle: case', c, (cases[c]['nx'], cases[c]['ny']), steps, 'steps', flush=True)
    ref = O.CylinderFlow(initial='zero', fft_workers=workers(), **ens.solver_kwargs(cases[c]))
    alt = O.CylinderFlow(initial='zero', fft_workers=workers(), extra_refinements=6, **ens.solver_kwargs(cases[c]))
    cd_ref, cl_ref, spread = np.empty(steps), np.empty(steps), np.empty((steps, 2))
    for k in range(steps):
        _, cd_ref[k], cl_ref[k] = ref.step(cases[c]['dt'])
        _, cd_alt, cl_alt = alt.step(cases[c]['dt'])
        spread[k] = abs(cd_alt - cd_ref[k]), abs(cl_alt - cl_ref[k])
    e_cd, e_cl = force_errors(d[c, :, 2], d[c, :, 3], cd_ref, cl_ref, spread)
    print('case', c, 'worst C_D error / bar', e_cd.max(), '; worst C_L error / bar', e_cl.max(), flush=True)
    assert e_cd.max() <= 1 and e_cl.max() <= 1
    f = batch.flows[c]
    for key in ('u', 'v', 'p'):
        a, r = getattr(f, key), getattr(ref, key)
        assert np.max(abs(a - r)) <= 1e-11 * np.max(abs(r)), key
    batch.close()


def test_restart_matches_uninterrupted_oracle(pkg, tmp_path):
    """440x82 channel: 7 steps, checkpoint, a new solver from the checkpoint, 13 more steps.  The
    restarted run matches the oracle's uninterrupted 20 steps, and, with the launch shapes the
    checkpoint restores, equals an uninterrupted GPU run bit for bit."""
    kw, _ = CHANNELS['channel440x82']
    dt = 2e-3
    hist = GOLDEN['channel440x82/history']
    first = gpu_flow(pkg, **kw)
    first.solve(n_steps=7, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    launch = first.gg.device.get_launch()
    assert launch['chosen']
    first.save_checkpoint(tmp_path / 'step7.npz')
    resumed = pkg.MAC_CylinderFlow.from_checkpoint(tmp_path / 'step7.npz')
    assert resumed.gg.device.get_launch() == launch
    t, cd, cl = resumed.solve(n_steps=13, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    assert abs(resumed.time - 20 * dt) < 1e-15
    e_cd, e_cl = force_errors(cd, cl, hist[7:20, 1], hist[7:20, 2], GOLDEN['channel440x82/spread'][7:20])
    assert e_cd.max() <= 1 and e_cl.max() <= 1
    ref = O.CylinderFlow(**kw)
    for _ in range(20):
        ref.step(dt)
    for key in ('u', 'v', 'p'):
        a, r = getattr(resumed, key), getattr(ref, key)
        assert np.max(abs(a - r)) <= 1e-11 * np.max(abs(r)), key
    whole = gpu_flow(pkg, **kw)
    whole.gg.device.set_launch_classes(launch)
    _, cd_w, cl_w = whole.solve(n_steps=20, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
    assert np.array_equal(cd_w[7:], cd) and np.array_equal(cl_w[7:], cl)
    for key in ('u', 'v', 'p'):
        assert np.array_equal(getattr(resumed, key), getattr(whole, key)), key
