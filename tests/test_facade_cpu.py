"""Host-side behaviour of the facade that needs no GPU: argument validation mirrors solver/main.py:45-80."""
import pytest


@pytest.mark.parametrize('kwargs, exc', [
    (dict(boundary='nope'), ValueError),
    (dict(time_scheme='rk4'), ValueError),
    (dict(time_scheme='ab2_cn', boundary='ghost_cell', transport='upwind'), ValueError),
    (dict(boundary='ghost_cell', time_scheme='ab2_cn', nx=3), ValueError),
    (dict(boundary='ghost_cell', time_scheme='ab2_cn', xc=0.1), ValueError),
    (dict(boundary='ghost_cell', time_scheme='ab2_cn', inflow='shear'), ValueError),
    (dict(boundary='ghost_cell', time_scheme='ab2_cn', ghost_pressure_extension='unknown'), ValueError),
    (dict(boundary='ghost_cell', time_scheme='ab2_cn', slip_walls=True), ValueError),
    (dict(boundary='immersed'), NotImplementedError),
])
def test_constructor_validation(pkg, kwargs, exc):
    with pytest.raises(exc):
        pkg.MAC_CylinderFlow(verbose=False, **kwargs)


def test_prepared_case_records_uploads(pkg):
    """A case prepared without a device records every table upload as plain arrays (picklable for
    the ensemble's process pool) and keeps the facade's inlet profile bit for bit."""
    import importlib
    import pickle
    import numpy as np
    ens = importlib.import_module(pkg.__name__ + '.ensemble')
    case = ens.channel_case(5, 0.2, 0.06, 60., nx_range=(40, 400), ny_range=(12, 80))
    prep = pickle.loads(pickle.dumps(ens.prepare(case)))
    names = [name for name, _ in prep.calls]
    for name, count in (('mcf_set_inlet', 1), ('mcf_set_solid_intervals', 3), ('mcf_set_ghost', 3),
                        ('mcf_set_override', 6), ('mcf_set_rectangle', 3), ('mcf_set_schur', 3),
                        ('mcf_set_surface', 1)):
        assert names.count(name) == count, name
    for _, args in prep.calls:
        for a in args:
            assert not isinstance(a, np.ndarray) or a.dtype in (np.float64, np.int32)
    ny, Ly = case['ny'], case['Ly']
    y = (np.arange(ny) + 0.5) * (Ly / ny)
    assert np.array_equal(prep.u_inlet, 4.0 * case['U_m'] * y * (Ly - y) / (Ly**2))
    assert prep.layout.L is None and prep.layout.ui is not None and prep.dt == case['dt']


def test_ensemble_cases_seeded(pkg):
    import importlib
    ens = importlib.import_module(pkg.__name__ + '.ensemble')
    cases = [ens.ensemble_case(k) for k in range(64)]
    assert cases[5] == ens.ensemble_case(5)
    assert min(c['nx'] for c in cases) >= 330 and max(c['nx'] for c in cases) <= 1320
    assert min(c['ny'] for c in cases) >= 62 and max(c['ny'] for c in cases) <= 246
    assert all(40 <= c['Re'] <= 200 for c in cases)
    assert all(abs(c['nu'] * c['Re'] - 1.0) < 1e-12 for c in cases)
