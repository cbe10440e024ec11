import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs libghostflow.so kernels)')


@pytest.fixture(scope='session')
def pkg():
    import cfd_amd_loader
    return cfd_amd_loader.load()
