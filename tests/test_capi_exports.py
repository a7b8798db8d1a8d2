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


def test_footprint_without_gpu():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    from freeflow_amd import workloads as wl
    nx, ny = 1200, 400
    plane = (nx * ny + 255) // 256 * 256
    base32 = capi.footprint(2, nx, ny, storage="f32")
    base16 = capi.footprint(2, nx, ny, storage="f16")
    assert base32 - base16 == 2 * plane * 6 * 2  # two ping-pong buffers, 6 planes, 2 bytes saved each
    assert base32 == 2 * plane * 6 * 4 + nx * ny + 8192
    sw = wl.swimmer2d_swimmer()
    with_b = capi.footprint(2, nx, ny, storage="f32", verts=sw.rest, elems=sw.elems, dx=0.005)
    assert with_b > base32
    p3 = (96 ** 3 + 255) // 256 * 256
    assert capi.footprint(3, 96, 96, 96, storage="bf16") == 2 * p3 * 10 * 2 + 96 ** 3 + 8192
