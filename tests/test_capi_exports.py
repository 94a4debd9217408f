"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute call: runs without a GPU)."""

import ctypes
import glob
import os
import re

import pytest

from fluidsim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"^int\s+(fsx_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "fsx_plan_create" in names and "fsx_rk_stage" in names
    # the ctypes binding covers exactly what the header declares
    assert set(names) == set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_library_exports_all_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    lib.fsx_version.restype = ctypes.c_int
    assert lib.fsx_version() == 1


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_argument_errors_without_gpu():
    """Argument validation happens before any HIP call."""
    lib = _lib.get_lib()
    h = ctypes.c_void_p()
    # any size >= 2 per axis is taken (mixed-radix passes); 1 and lengths
    # beyond the LDS tile are refused
    rc = lib.fsx_plan_create(ctypes.byref(h), 2, 48, 1, 1, 1.0, 1.0, 1.0, 8, None, 4)
    assert rc == _lib.FSX_ERR_ARG
    assert ">= 2" in _lib.last_error()
    rc = lib.fsx_plan_create(ctypes.byref(h), 3, 48, 48, 3000, 1.0, 1.0, 1.0, 8, None, 7)
    assert rc == _lib.FSX_ERR_ARG
    assert "too large" in _lib.last_error()
    rc = lib.fsx_plan_create(ctypes.byref(h), 4, 64, 64, 1, 1.0, 1.0, 1.0, 8, None, 4)
    assert rc == _lib.FSX_ERR_ARG
    rc = lib.fsx_plan_create(ctypes.byref(h), 2, 64, 64, 1, 1.0, 1.0, 1.0, 2, None, 4)
    assert rc == _lib.FSX_ERR_ARG
    with pytest.raises(ValueError):
        _lib.check(lib.fsx_rk_stage(None, 4, 0, 1, None, None, None, None, 0, 0, None, None))
