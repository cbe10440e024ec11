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
