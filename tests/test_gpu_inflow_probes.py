"""GPU: the transient inflow perturbation set by the device from each step's time (solver/main.py:280-286)
against the oracle, segment plans of ``advance`` against one-shot runs, and the device's stagnation
pressure probes (postprocessing/cylinder_observables.py:37-38) against the host reconstruction."""
import importlib

import numpy as np
import pytest

from oracle import ghost_oracle as O

pytestmark = pytest.mark.gpu

CH = dict(nx=88, ny=16, Lx=22., Ly=4.1, xc=2., yc=2.)     # uniform inflow U = 1
DT = 1e-3


def gpu_flow(pkg, **kw):
    return pkg.MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **dict(CH, **kw))


def rel(a, b):
    return float(np.max(abs(np.asarray(a) - np.asarray(b))) / max(np.max(abs(np.asarray(b))), 1e-300))


@pytest.mark.parametrize('calls', [(20,), (5, 15)], ids=['one call', 'two calls'])
def test_perturbed_inflow_matches_oracle(pkg, calls):
    """12 steps inside a 0.012 window, then 8 with the base profile; the window may span two calls."""
    print('perturbed inflow', calls, flush=True)
    ref, gpu = O.CylinderFlow(**CH), gpu_flow(pkg)
    kw = dict(perturb_transient=True, perturb_duration=0.012, perturb_amp=0.15)
    cd_ref, cl_ref, inlets = [], [], []
    for _ in range(sum(calls)):
        _, cd, cl = ref.step(DT, **kw)
        cd_ref.append(cd); cl_ref.append(cl)
        inlets.append(ref.u_inlet.copy())
    cd_gpu, cl_gpu = [], []
    for n in calls:
        _, cd, cl = gpu.solve(n_steps=n, dt=DT, save_artifacts=False, verbose=False, **kw)
        cd_gpu.extend(cd); cl_gpu.extend(cl)
        assert rel(gpu.u_inlet, inlets[len(cd_gpu) - 1]) < 1e-15
    errs = dict(u=rel(gpu.u, ref.u), v=rel(gpu.v, ref.v), p=rel(gpu.p, ref.p))
    print('perturbed inflow field errors', errs, flush=True)
    assert max(errs.values()) < 1e-11, errs
    for k in range(len(cd_ref)):
        scale = max(abs(cd_ref[k]), abs(cl_ref[k]))
        assert abs(cd_gpu[k] - cd_ref[k]) <= 1e-10 * scale and abs(cl_gpu[k] - cl_ref[k]) <= 1e-10 * scale, k
    # the perturbation really acted: a run without it differs
    plain = gpu_flow(pkg)
    plain.solve(n_steps=sum(calls), dt=DT, perturb_transient=False, save_artifacts=False, verbose=False)
    assert rel(plain.u, gpu.u) > 1e-6


def test_segment_plan_equals_one_shot(pkg):
    a, b = gpu_flow(pkg), gpu_flow(pkg)
    a.gg.device.set_launch(1, 1)
    b.gg.device.set_launch(1, 1)
    kw = dict(perturb_transient=True, perturb_duration=0.007)
    _, cd_a, _ = a.solve(n_steps=16, dt=DT, save_artifacts=False, verbose=False, **kw)
    rows = list(b.advance(16, DT, segment=[3, 1, 7, 5], **kw))
    assert [len(r) for r in rows] == [3, 1, 7, 5]
    cd_b = np.concatenate(rows)[:, 2]
    assert np.array_equal(cd_a, cd_b)
    for key in ('u', 'v', 'p', 'u_inlet'):
        assert np.array_equal(getattr(a, key), getattr(b, key)), key
    assert a.time == b.time
    with pytest.raises(ValueError):
        next(b.advance(4, DT, segment=[3, 2]))


def test_step_is_one_iteration(pkg):
    a, b = gpu_flow(pkg), gpu_flow(pkg)
    a.gg.device.set_launch(1, 1)
    b.gg.device.set_launch(1, 1)
    a.solve(n_steps=3, dt=DT, perturb_transient=False, save_artifacts=False, verbose=False)
    for _ in range(3):
        d = b.step(DT, perturb_transient=False)
    assert np.array_equal(a.u, b.u) and np.array_equal(a.p, b.p)
    assert d['cd'] == a.last_step_diagnostics['cd'] and d['time'] == a.time


def test_device_probes_equal_host_reconstruction(pkg):
    ob = importlib.import_module(pkg.__name__ + '.observables')
    nat = importlib.import_module(pkg.__name__ + '._native')
    s = gpu_flow(pkg)
    P = ob.stagnation_operator(s)
    s.gg.device.set_probes(P)
    c0, c1 = nat.DIAG_NAMES.index('probe0'), nat.DIAG_NAMES.index('probe1')
    for rows in s.advance(12, DT, segment=4, perturb_transient=False):
        host = P @ s.p.ravel()
        dev = rows[-1, [c0, c1]]
        assert np.max(abs(dev - host)) <= 1e-12 * np.max(abs(host)), (dev, host)
    s.gg.device.set_probes(None)
    rows = s.gg.device.run(1, DT)
    assert rows[0, c0] == 0.0 and rows[0, c1] == 0.0
