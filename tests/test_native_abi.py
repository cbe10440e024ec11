"""The C-ABI library loads and exports every symbol include/ghostflow.h declares (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / 'include' / 'ghostflow.h').read_text()
    return sorted(set(re.findall(r'^\s*(?:int|void|const char\*)\s+(mcf_\w+)\s*\(', text, re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for required in ('mcf_create', 'mcf_step', 'mcf_run', 'mcf_project', 'mcf_forces', 'mcf_rect_inverse',
                     'mcf_schur_solve', 'mcf_set_state', 'mcf_get_state', 'mcf_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol(pkg):
    import importlib
    nat = importlib.import_module('cfd_amd._native')
    if not nat.LIB_PATH.exists():
        pytest.fail('libghostflow.so is not built; run __graft_entry__.build()')
    lib = ctypes.CDLL(str(nat.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(declared_symbols()) == set(nat.exported_symbols()) | {'mcf_destroy', 'mcf_last_error',
                                                                     'mcf_device_name'} | set(nat.exported_symbols())
    assert set(nat.exported_symbols()) >= set(declared_symbols())


def test_binding_fails_loudly_without_library(pkg, tmp_path, monkeypatch):
    import importlib
    nat = importlib.import_module('cfd_amd._native')
    monkeypatch.setattr(nat, 'LIB_PATH', tmp_path / 'missing.so')
    monkeypatch.setattr(nat, '_lib', None)
    with pytest.raises(ImportError):
        nat.library()
