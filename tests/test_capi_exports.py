"""CPU-side checks of the C ABI library: it loads without a GPU and exports
every function include/freeflow_lbm.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from freeflow_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "freeflow_lbm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ff_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "ff_step" in names and "ff_step_proxy" in names and len(names) >= 20
    assert sorted(capi.EXPORTS) == names


def test_library_loads_and_exports_all_symbols():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_version_and_error_string_without_gpu():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    L = capi.lib()
    assert b"gfx950" in L.ff_version()
    # an invalid descriptor fails before touching the device
    d = capi.DomainDesc(4, 1, 1, 1, 0.01, 1.0, 1.0, 1.0, 0, 0, 0)
    h = ctypes.c_void_p()
    rc = L.ff_domain_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == 1 and b"dim" in L.ff_last_error()
