"""Import helper: the package directory name contains hyphens, so it is loaded by path as ``cfd_amd``."""
import importlib.util
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent / 'cfd-cylinder-flow-ghost-cell_amd'


def load():
    if 'cfd_amd' in sys.modules:
        return sys.modules['cfd_amd']
    spec = importlib.util.spec_from_file_location('cfd_amd', PACKAGE_DIR / '__init__.py',
                                                  submodule_search_locations=[str(PACKAGE_DIR)])
    module = importlib.util.module_from_spec(spec)
    sys.modules['cfd_amd'] = module
    spec.loader.exec_module(module)
    return module
