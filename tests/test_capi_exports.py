"""CPU-side checks of the C ABI library: it loads without a GPU and exports
every function include/freeflow_lbm.h declares (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
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


def _al(count, size):
    """Bytes of one array in a carved allocation: 256-byte aligned."""
    return (max(count, 1) * size + 255) // 256 * 256


def _round2m(b):
    return (b + (2 << 20) - 1) // (2 << 20) * (2 << 20)


def _core_bytes(dim, n, sbytes, split_velocity=False):
    """Moments (two ping-pong buffers of nm planes, dim more for the velocity
    low halves of the split-velocity storage; 256-element plane stride),
    flags (one byte a cell), 1024 doubles of reduction scratch."""
    nm = (6 if dim == 2 else 10) + (dim if split_velocity else 0)
    plane = (n + 255) // 256 * 256
    return _round2m(2 * _al(plane * nm * sbytes, 1) + _al(n, 1) + _al(1024, 8))


def _boundary_bytes(dim, nv, ne, nb, cap, nx=1, ny=1):
    """Every array of the ff_set_boundary allocation, in its order."""
    D, K, nm = dim, (8 if dim == 2 else 26), (6 if dim == 2 else 10)
    vc = max(nv, 1) * D
    b = 4 * _al(vc, 4)                                      # positions and velocities, two generations
    b += _al(ne * D, 4) + _al(ne, 4) + _al(nv, 4)           # elements, element -> body, vertex -> body
    b += _al(nb + 1, 4) + _al(ne, 4)                        # CSR body -> elements
    b += _al(nv + 1, 4) + _al(ne * D, 4)                    # CSR vertex -> (element, end)
    b += _al(nb * D, 8) + _al(1, 16)                        # body reference points, step state
    b += _al(vc, 4) + _al(nb * (3 if dim == 2 else 6), 8)   # vertex forces, body wrench
    b += _al(nb, 8)                                         # body power
    b += _al(2 * cap, 8) + _al(2 * cap * K, 8)              # link tables: cells and keys, two by tag parity
    b += _al(ne, 8 * (8 if dim == 2 else 16))               # per-element FP64 sums (with power)
    b += 3 * _al(cap * K * D, 4) + _al(cap * nm, 4)         # link force, barycentrics, hit point; cell rows
    b += _al(ny * ((nx + 31) // 32) if dim == 2 else 1, 4)  # fill-solid row marks
    b += _al(vc, 4) + _al(nv, 4) + _al(nb * 8, 8)           # wave inputs (4 FP64 + 8 FP32 per body)
    return _round2m(b)


def test_footprint_without_gpu():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    from freeflow_amd import workloads as wl
    nx, ny = 1200, 400
    assert capi.footprint(2, nx, ny, storage="f32") == _core_bytes(2, nx * ny, 4)
    assert capi.footprint(2, nx, ny, storage="f16") == _core_bytes(2, nx * ny, 2)
    assert capi.footprint(3, 96, 96, 96, storage="bf16") == _core_bytes(3, 96 ** 3, 2)
    assert capi.footprint(3, 96, 96, 96, storage="f16u") == _core_bytes(3, 96 ** 3, 2, True)
    assert capi.footprint(2, nx, ny, storage="f16u") == _core_bytes(2, nx * ny, 2, True)
    sw = wl.swimmer2d_swimmer()
    nv, ne = len(sw.rest), len(sw.elems)
    with_b = capi.footprint(2, nx, ny, storage="f32", verts=sw.rest, elems=sw.elems, dx=0.005)
    extra = with_b - _core_bytes(2, nx * ny, 4)
    # the link table's capacity is a power of two >= 1024; ne bodies bound the per-body arrays
    caps = [1 << k for k in range(10, 24) if _boundary_bytes(2, nv, ne, ne, 1 << k, nx, ny) == extra]
    assert caps, f"boundary bytes {extra} match no table capacity"


def _distinct_cells_near(verts, elems, dx):
    """Independent count of the link-table estimate: distinct cells of the
    element boxes widened by two cells (floor(min) - 2 .. floor(max) + 2)."""
    cells = set()
    for e in elems:
        p = verts[e] / dx
        lo = np.floor(p.min(0)).astype(int) - 2
        hi = np.floor(p.max(0)).astype(int) + 2
        for x in range(lo[0], hi[0] + 1):
            for y in range(lo[1], hi[1] + 1):
                cells.add((x, y))
    return len(cells)


def test_link_table_sized_from_distinct_cells_near_the_mesh():
    """Two 2D polylines with the same vertex and element counts, one packed
    into a few cells and one spread over the grid: their footprints differ by
    their link tables, each sized next_pow2(2 * distinct cells + 1024) slots
    (360 B per 2D slot: two tables of cells and 8 keys, force/barycentric/hit
    per link, 6 moments)."""
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    n, nx, ny = 400, 2000, 200
    e = np.stack([np.arange(n), np.arange(1, n + 1)], 1).astype(np.uint32)
    packed = np.stack([100.0 + 0.01 * np.arange(n + 1), np.full(n + 1, 50.3)], 1).astype(np.float32)
    t = np.linspace(0.0, 1.0, n + 1)
    spread = np.stack([10.0 + 1900.0 * t, 100.0 + 60.0 * np.sin(12.0 * t)], 1).astype(np.float32)

    def cap(v):
        c = 2 * _distinct_cells_near(v.astype(np.float64), e, 1.0) + 1024
        return 1 << (c - 1).bit_length()

    fa = capi.footprint(2, nx, ny, storage="f32", verts=packed, elems=e, dx=1.0)
    fb = capi.footprint(2, nx, ny, storage="f32", verts=spread, elems=e, dx=1.0)
    assert cap(spread) > cap(packed)
    core = _core_bytes(2, nx * ny, 4)
    assert fa - core == _boundary_bytes(2, n + 1, n, n, cap(packed), nx, ny)
    assert fb - core == _boundary_bytes(2, n + 1, n, n, cap(spread), nx, ny)


def test_invalid_arguments_rejected_without_gpu():
    # descriptor and footprint checks run on the host before any device call
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    L = capi.lib()
    bad = [
        (dict(nx=0), b"grid sizes"),
        (dict(ny=-4), b"grid sizes"),
        (dict(nx=1 << 16, ny=1 << 16), b"2^31"),
        (dict(storage=4), b"storage"),
        (dict(dx=0.0), b"dx"),
        (dict(dx=float("nan")), b"dx"),
        (dict(viscosity=0.0), b"viscosity"),
        (dict(dt=-1.0), b"dt"),
        (dict(options=8), b"option"),
    ]
    base = dict(dim=2, nx=8, ny=8, nz=1, viscosity=0.01, dx=1.0, dt=1.0, fluid_density=1.0, storage=0,
                options=0, device=0)
    for change, msg in bad:
        d = capi.DomainDesc(**{**base, **change})
        h = ctypes.c_void_p()
        assert L.ff_domain_create(ctypes.byref(d), ctypes.byref(h)) == 1, change
        assert msg in L.ff_last_error(), (change, L.ff_last_error())
        assert not h.value
    # footprint: same grid checks, plus element indices and vertex coordinates
    with pytest.raises(capi.FreeFlowError, match="grid sizes"):
        capi.footprint(2, 0, 8)
    with pytest.raises(capi.FreeFlowError, match="dx"):
        capi.footprint(2, 8, 8, dx=-1.0)
    v = np.array([[1.0, 1.0], [2.0, 1.0]], np.float32)
    with pytest.raises(capi.FreeFlowError, match="out of range"):
        capi.footprint(2, 8, 8, verts=v, elems=np.array([[0, 2]], np.uint32))
    for x in (np.nan, np.inf, 3e9):
        w = v.copy()
        w[1, 0] = x
        with pytest.raises(capi.FreeFlowError, match="not finite or beyond"):
            capi.footprint(2, 8, 8, verts=w, elems=np.array([[0, 1]], np.uint32))
    # a long thin segment far from the mesh box: the element-box product saturates
    w = np.array([[0.0, 0.0, 0.0], [1e9, 1e9, 1e9], [0.0, 1e9, 0.0]], np.float32)
    assert capi.footprint(3, 8, 8, 8, verts=w, elems=np.array([[0, 1, 2]], np.uint32)) > 0
