"""GPU: the batched ensemble (mcf_batch_*) against the same cases stepped one by one, and the oracle.

Batched and single-case runs execute the same kernels on the same per-case data, so their fields
must agree bit for bit; one case is also checked against the CPU oracle.
"""
import importlib

import numpy as np
import pytest

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

# three uneven small channels: (m, D/H, offset, Re)
SMALL = [(4, 0.25, 0.0, 100.), (5, 0.2, 0.06, 60.), (6, 0.28, -0.08, 150.)]


@pytest.fixture(scope='module')
def ens(pkg):
    return importlib.import_module(pkg.__name__ + '.ensemble')


def small_cases(ens):
    return [ens.channel_case(m, dh, dl, re, nx_range=(40, 400), ny_range=(12, 80)) for m, dh, dl, re in SMALL]


@pytest.mark.parametrize('launch', [(0, 1), (7, 1)], ids=['dense y transform', 'x-FFT'])
def test_batched_equals_single(ens, launch):
    cases = small_cases(ens)
    print('ensemble: shapes', [(c['nx'], c['ny']) for c in cases], launch, flush=True)
    batch = ens.Ensemble(cases, workers=1)
    singles = ens.build_flows(cases, workers=1)
    # same launch shapes on both sides (the autotuner may pick other sweep widths for the batch,
    # which changes the round-off)
    batch.set_launch(*launch)
    for f in singles:
        f.gg.device.set_launch(*launch)
    assert batch.get_launch()['gemm_pressure'] == launch[0]
    nsteps = 7                        # eager startup step, then graph replays of both parities
    d_b = batch.run(3)
    d_b = np.concatenate([d_b, batch.run(nsteps - 3)], axis=1)
    batch.pull()
    for c, (case, f) in enumerate(zip(cases, singles)):
        d_s = f.gg.device.run(nsteps, case['dt'])
        f._pull()
        b = batch.flows[c]
        for key in ('u', 'v', 'p'):
            assert np.array_equal(getattr(b, key), getattr(f, key)), (c, key)
        assert np.array_equal(b.gg.half_pressure, f.gg.half_pressure), c
        for col in (0, 2, 3, 10):     # time, C_D, C_L, max divergence
            assert np.array_equal(d_b[c, :, col], d_s[:, col]), (c, col)
        assert np.all(d_b[c, :, 10] < 1e-12)
        assert b.time == pytest.approx(nsteps * case['dt'], rel=1e-14)
    batch.close()


def test_batched_case_matches_oracle(ens):
    case = small_cases(ens)[1]
    batch = ens.Ensemble([case], workers=1)
    nsteps = 5
    d = batch.run(nsteps)
    batch.pull()
    ref = O.CylinderFlow(initial='zero', **ens.solver_kwargs(case))
    for _ in range(nsteps):
        ref.step(case['dt'])
    f = batch.flows[0]
    for key in ('u', 'v', 'p'):
        a, r = getattr(f, key), getattr(ref, key)
        assert np.max(abs(a - r)) <= 1e-11 * np.max(abs(r)), key
    dr = ref.last_step_diagnostics
    scale = max(abs(dr['cd']), abs(dr['cl']))
    assert abs(d[0, -1, 2] - dr['cd']) <= 1e-10 * scale
    assert abs(d[0, -1, 3] - dr['cl']) <= 1e-10 * scale
    batch.close()


def test_seeded_ensemble_divergence(ens):
    """The first eight cases of the seeded 64-case ensemble at their real sizes: every step of every
    case keeps max|div| < 1e-12 and finite forces."""
    cases = [ens.ensemble_case(k) for k in range(8)]
    print('ensemble: shapes', [(c['nx'], c['ny']) for c in cases], flush=True)
    batch = ens.Ensemble(cases)
    d = batch.run(12)
    assert not batch.failures, batch.failures
    assert np.all(d[:, :, 10] < 1e-12), d[:, :, 10].max(axis=1)
    assert np.all(np.isfinite(d[:, :, 2])) and np.all(np.isfinite(d[:, :, 3]))
    assert np.all(d[:, :, 26] == 0)   # status word
    batch.close()


def test_failed_case_is_dropped(ens):
    """A member that fails the reference's CFL gate (dt five times too large) is dropped at its first
    step with the reason recorded; the other members continue bit-identical to single-case runs."""
    good = small_cases(ens)
    bad = ens.channel_case(5, 0.2, 0.06, 60., nx_range=(40, 400), ny_range=(12, 80), courant=0.25)
    cases = [good[0], bad, good[2]]
    batch = ens.Ensemble(cases, workers=1)
    batch.set_launch(0, 1)
    d = batch.run(3)
    assert list(batch.failures) == [1] and batch.failures[1]['step'] == 0
    # the other members ran every step of the call
    assert np.allclose(d[[0, 2], -1, 0], [3 * cases[0]['dt'], 3 * cases[2]['dt']], rtol=1e-12)
    assert 'CFL' in batch.failures[1]['message'] and d[1, 0, 11] >= 0.5
    assert np.all(np.isnan(d[1, 1:, 0]))
    d2 = batch.run(3)
    assert np.all(np.isnan(d2[1]))
    batch.pull()
    singles = ens.build_flows([good[0], good[2]], workers=1)
    for k, f in zip((0, 2), singles):
        f.gg.device.set_launch(0, 1)
        f.gg.device.run(6, cases[k]['dt'])
        f._pull()
        for key in ('u', 'v', 'p'):
            assert np.array_equal(getattr(batch.flows[k], key), getattr(f, key)), (k, key)
    batch.close()


def test_autotuned_batch_matches_single_to_roundoff(ens):
    """Without pinned shapes the batch and the single cases may run other sweep widths: the fields
    then agree to round-off, and the chosen shapes are valid variants."""
    cases = small_cases(ens)
    batch = ens.Ensemble(cases, workers=1)
    singles = ens.build_flows(cases, workers=1)
    batch.run(5)
    batch.pull()
    ln = batch.get_launch()
    # transform variants: 0-6 dense, 7 the x-FFT form (with its y sweep, 3), 8-9 the parity split
    assert ln['chosen'] and ln['gemm_pressure'] in range(10) and 0 <= ln['tri_velocity'] < 4
    for c, (case, f) in enumerate(zip(cases, singles)):
        f.gg.device.run(5, case['dt'])
        f._pull()
        for key in ('u', 'v', 'p'):
            a, b = getattr(batch.flows[c], key), getattr(f, key)
            assert np.max(abs(a - b)) <= 1e-12 * np.max(abs(b)), (c, key)
    batch.close()


def test_batch_speculation_matches_full(ens):
    cases = small_cases(ens)
    out = {}
    for mode in (0, 1, 2):
        batch = ens.Ensemble(cases, workers=1)
        batch.set_launch(0, 1)
        batch.set_speculation(mode)
        batch.run(4)
        batch.run(4)
        batch.pull()
        out[mode] = ([(f.u.copy(), f.v.copy(), f.p.copy()) for f in batch.flows], batch.get_speculation())
        batch.close()
    for mode in (1, 2):
        for fa, fb in zip(out[mode][0], out[0][0]):
            for a, b in zip(fa, fb):
                assert np.array_equal(a, b), mode
    assert out[2][1]['replays'] >= 1


def test_member_failing_mid_run_does_not_stop_the_others(ens):
    """A member that fails after several speculative chunks (dt too large for its grid) is dropped
    at its step; the others complete every step of the same call."""
    good = small_cases(ens)
    bad = ens.channel_case(5, 0.2, 0.06, 60., nx_range=(40, 400), ny_range=(12, 80), courant=0.12)
    batch = ens.Ensemble([good[0], bad], workers=1)
    d = batch.run(60)
    print('failures', batch.failures)
    assert np.allclose(d[0, -1, 0], 60 * good[0]['dt'], rtol=1e-12)
    assert np.all(d[0, :, 26] == 0)
    if 1 in batch.failures:
        s = batch.failures[1]['step']
        assert np.all(np.isnan(d[1, s + 1:, 0]))
    batch.close()
