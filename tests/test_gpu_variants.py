"""GPU checks of the ghost variants the reference tests on its own 40x30 case
(tests/test_gemini_followup.py:24-26): ghost_pressure_extension='quadratic',
ghost_velocity_interpolation='biquadratic' and both together.

* Temporal order (test_gemini_followup.py:141-159): 0.064 time units at dt = 0.004 ... 0.000125
  against the finest run; the order of the two finest pairs lies in (1.8, 2.2), the reference's
  acceptance test for these candidates.
* Gates and restart (test_gemini_followup.py:65-85): 20 steps at dt = 0.001 keep the divergence
  below 1e-10 and the CFL number below 0.5; a checkpoint restored into a new solver continues
  three steps within 1e-10 of the original.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REF = dict(nx=40, ny=30, Lx=4., Ly=3., xc=1.5, yc=1.5)
CHOICES = [dict(ghost_pressure_extension='quadratic'), dict(ghost_velocity_interpolation='biquadratic'),
           dict(ghost_pressure_extension='quadratic', ghost_velocity_interpolation='biquadratic')]
IDS = ['quadratic-pressure', 'biquadratic-velocity', 'both']


def solver(pkg, **options):
    f = pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **REF, **options)
    f.gg.device.set_launch(0, 1)   # one set of launch shapes for every run compared
    return f


@pytest.mark.parametrize('choice', CHOICES, ids=IDS)
def test_candidate_time_order(pkg, choice):
    print('time order', choice, flush=True)
    states = []
    timesteps = (.004, .002, .001, .0005, .000125)
    for dt in timesteps:
        s = solver(pkg, **choice)
        s.solve(n_steps=round(.064 / dt), dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
        states.append((s.gg.pack(s.u, s.v).copy(), s.p.ravel()[s.gg.pi].copy()))
    errors = np.array([[np.sqrt(np.mean((a[k] - states[-1][k]) ** 2)) for k in (0, 1)] for a in states[:-1]])
    orders = np.log2(errors[:-1] / errors[1:])
    print(choice, 'orders', orders.tolist(), flush=True)
    assert np.all((orders[-2:] > 1.8) & (orders[-2:] < 2.2)), orders


@pytest.mark.parametrize('choice', CHOICES, ids=IDS)
def test_variant_gates_and_restart(pkg, choice, tmp_path):
    s = solver(pkg, **choice)
    s.solve(n_steps=20, dt=.001, perturb_transient=False, save_artifacts=False, verbose=False)
    assert s.last_step_diagnostics['max_divergence'] < 1e-10
    assert s.last_step_diagnostics['cfl'] < .5
    path = tmp_path / 'checkpoint.npz'
    s.save_checkpoint(path)
    restored = pkg.MAC_CylinderFlow.from_checkpoint(path)
    restored.gg.device.set_launch(0, 1)
    for key in ('u', 'v', 'p'):
        assert np.array_equal(getattr(s, key), getattr(restored, key)), key
    for f in (s, restored):
        f.solve(n_steps=3, dt=.001, perturb_transient=False, save_artifacts=False, verbose=False)
    for key in ('u', 'v', 'p'):
        assert np.max(abs(getattr(s, key) - getattr(restored, key))) < 1e-10, key
