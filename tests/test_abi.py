"""The C-ABI library loads and exports every symbol include/hiprmsd.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from mdtraj_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "hiprmsd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "hr_rmsd_one_vs_all" in names and "hr_superpose" in names
    assert set(names) == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.hr_version() >= 100
    assert lib.hr_rmsd_workspace_bytes(1500) >= 1500 * 12
    assert lib.hr_superpose_workspace_bytes(1000, 100) >= 1000 * (24 + 36)
