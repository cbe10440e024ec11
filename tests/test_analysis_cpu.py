"""Force-history statistics and wake-sample means against closed forms: the known answers of the
reference's own tests (tests/test_periodic_window_audit.py:16-80, tests/test_gemini_followup.py:88-107)
plus exact integrals of polynomials and sinusoids."""
import importlib
import json

import numpy as np
import pytest


@pytest.fixture(scope='module')
def an(pkg):
    return importlib.import_module(pkg.__name__ + '.analysis')


@pytest.fixture(scope='module')
def ob(pkg):
    return importlib.import_module(pkg.__name__ + '.observables')


def test_linear_window_on_nonuniform_grid(an):
    t = np.array([0., .1, .4, 1.])
    values = np.column_stack((1e6 + t, 3 - 2 * t))
    r = an.window_statistics(t, values, .2, .8)
    assert np.allclose(r['mean'], [1e6 + .5, 2.], rtol=0, atol=2e-10)
    assert np.allclose(r['rms'], [.6 / np.sqrt(12), 1.2 / np.sqrt(12)], rtol=0, atol=1e-10)
    assert np.allclose(r['mid_extrema'], [1e6 + .5, 2.], rtol=0, atol=2e-10)
    assert np.allclose(r['half_amplitude'], [.3, .6], rtol=0, atol=2e-10)
    assert np.allclose(r['amplitude'], [.6, 1.2], rtol=0, atol=2e-10)
    for a, b in ((-.1, .8), (.2, 1.1), (.8, .2), (0., float('nan'))):
        with pytest.raises(ValueError):
            an.window_statistics(t, values, a, b)


def test_quadratic_on_unequal_intervals(an):
    dt = 0.001 * (1.0 + 0.6 * np.sin(np.linspace(0, 20 * np.pi, 8000)))
    t = np.r_[0.0, np.cumsum(dt)]
    t = t[t <= 8.0]
    r = an.window_statistics(t, (t ** 2)[:, None], 1.0, 7.0)
    assert abs(r['mean'][0] - 19.0) < 1e-3          # (7^3 - 1^3) / 3 / 6


def test_piecewise_linear_integrals_are_exact(an):
    # a piecewise-linear history integrates exactly: mean and variance from the closed forms of
    # each linear piece
    t = np.array([0., 1., 3., 4.])
    v = np.array([0., 2., -2., 1.])
    r = an.window_statistics(t, v[:, None], 0., 4.)
    mean = (1 * (0 + 2) / 2 + 2 * (2 - 2) / 2 + 1 * (-2 + 1) / 2) / 4
    assert abs(r['mean'][0] - mean) < 1e-15
    # variance: integral of (v - mean)^2 over each piece = h (a^2 + ab + b^2) / 3
    c = v - mean
    var = sum(h * (a * a + a * b + b * b) / 3 for h, a, b in zip(np.diff(t), c[:-1], c[1:])) / 4
    assert abs(r['rms'][0] - np.sqrt(var)) < 1e-15


def test_periodic_metrics_of_a_sinusoid(an):
    t = np.linspace(0, 10.3, 10301)
    cd = 3.0 + 0.1 * np.cos(4 * np.pi * t)
    cl = 0.2 + 1.0 * np.sin(2 * np.pi * t)
    r = an.periodic_metrics(t, cd, cl, 1.0, D=1.0, U_ref=1.0)
    assert r['cycles'] >= 8 and r['stationary']
    assert abs(r['strouhal'] - 1.0) < 1e-5 and abs(r['period'] - 1.0) < 1e-5
    assert abs(r['cd_mean'] - 3.0) < 1e-6 and abs(r['cl_mean'] - 0.2) < 1e-6
    assert abs(r['cl_rms'] - 1 / np.sqrt(2)) < 5e-6 and abs(r['cd_rms'] - 0.1 / np.sqrt(2)) < 5e-6
    assert abs(r['cl_half_amplitude'] - 1.0) < 1e-6 and r['cl_amp'] == r['cl_half_amplitude']
    assert r['amplitude_drift'] < 1e-6 and r['drag_drift'] < 1e-6
    # a constant lift has no resolved period
    flat = an.periodic_metrics(t, cd, np.ones_like(t), 1.0)
    assert flat['cycles'] == 0 and np.isnan(flat['strouhal']) and not flat['stationary']
    with pytest.raises(ValueError):
        an.periodic_metrics(t[::-1], cd, cl, 1.0)


def test_recirculation_length(ob):
    r = ob.recirculation_length([1., 2., 3., 4.], [-1., -.5, .5, 1.], 1., 1.)
    assert r['status'] == 'resolved' and abs(r['length_D'] - 1.5) < 1e-15
    assert ob.recirculation_length([1., 2., 3.], [-1., -.5, -.1], 1., 1.)['length_D'] is None
    assert ob.recirculation_length([1., 2., 3.], [1., 1., 1.], 1., 1.)['length_D'] is None
    assert ob.recirculation_length([1., 2., 3., 4.], [1., 1., -1., 1.], 1., 1.)['status'] == 'detached_negative_region'


def test_mean_observables_of_written_samples(ob, tmp_path):
    """Rows written in the sample format: linear-in-time columns have exact window means; a torn
    file and an uncovered window are rejected."""
    x = [2.5, 3.0, 3.5]
    meta = dict(schema=1, columns=6, dtype='<f8', stride=1, x=x, rear=2.5, diameter=1.0,
                quantities=['time', 'p_front', 'p_back', 'u_centerline'])
    path = tmp_path / 'samples.bin'
    path.with_suffix('.json').write_text(json.dumps(meta))
    times = [0., .2, .7, 1.1, 1.5]
    rows = [[t, 1 + t, 2 + 3 * t, -.4 + t, .1 + t, .6 + t] for t in times]
    path.write_bytes(np.array(rows, dtype='<f8').tobytes())
    r = ob.mean_observables(path, .1, .5)
    assert abs(r['p_front_mean'] - 1.3) < 1e-14 and abs(r['p_back_mean'] - 2.9) < 1e-14
    assert abs(r['delta_p_mean'] + 1.6) < 1e-14
    assert np.allclose(r['u_centerline_mean'], [-.1, .4, .9], atol=1e-14)
    assert r['recirculation']['status'] == 'resolved' and abs(r['recirculation']['length_D'] - .1) < 1e-12
    with pytest.raises(ValueError):
        ob.mean_observables(path, -.1, 1.)
    with path.open('ab') as f:
        f.write(b'x')
    with pytest.raises(ValueError):
        ob.mean_observables(path, .1, .5)


def test_runner_rejects_invalid_cases(pkg, tmp_path):
    rn = importlib.import_module(pkg.__name__ + '.runner')
    good = dict(dt=1e-3, end_time=0.03, sample_start=0.01, kind='steady',
                solver=dict(nx=88, ny=16, boundary='ghost_cell', time_scheme='ab2_cn'))
    for bad in (dict(good, kind='unsteady'), dict(good, dt=-1.), dict(good, sample_start=0.05),
                dict(good, solver=dict(good['solver'], boundary='immersed'))):
        cfg = tmp_path / 'bad.json'
        cfg.write_text(json.dumps(bad))
        with pytest.raises(ValueError):
            rn.run(cfg, tmp_path / 'out')
        assert not (tmp_path / 'out').exists()
