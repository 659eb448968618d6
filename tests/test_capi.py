"""The C-ABI library loads and exports every symbol include/atomistica_amd.h
declares (no GPU calls), and host-side argument checks raise before any
device work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "atomistica_amd.h")
LIB = os.path.join(ROOT, "atomistica_amd", "libatomistica_amd.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(agpu_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "agpu_potential_energy_and_forces" in names
    assert "agpu_neighbors_update" in names
    assert len(names) >= 35


@pytest.mark.skipif(not os.path.exists(LIB), reason="native library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="native library not built")
def test_python_prototypes_match_header():
    from atomistica_amd import _lib
    assert set(_lib.PROTOTYPES) == set(declared())


@pytest.mark.skipif(not os.path.exists(LIB), reason="native library not built")
def test_error_reporting_without_device():
    from atomistica_amd import _lib
    lib = _lib.load()
    assert lib.agpu_version().startswith(b"atomistica_amd")
    # NULL out-pointer is rejected on the host with a message
    rc = lib.agpu_particles_create(None)
    assert rc == _lib.ERR_ARG
    assert "NULL" in _lib.last_error()
    rc = lib.agpu_ewald_create(None, 10.0, 0.0, 0.0)
    assert rc == _lib.ERR_ARG
