"""GPU parity of the D3Q27 path against the CPU oracle, through the C ABI.

Bar: rel-L2 <= 1e-5 on the moments after the stated steps (FP32), with the
number of bit-identical values reported; the 3D kernels are compiled without
FMA contraction in the reference's operation order, so they are expected to
be bit-exact as well.  Forces: rel-L2 <= 1e-5.
"""
import sys

import numpy as np
import pytest

import oracle_lib as ol
from freeflow_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


icosphere = wl.icosphere


def pair3(nx, ny, nz, flags, m, periodic=False, storage="f32", nu=0.01, dx=1.0, dt=1.0, rho=1.0):
    g = capi.Domain(3, nx, ny, nz, viscosity=nu, dx=dx, dt=dt, fluid_density=rho, storage=storage,
                    periodic=periodic)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, periodic=periodic, nu=nu, dx=dx, dt=dt,
                        cf=capi.cf_of(3, rho, dx, dt), storage={"f32": 0, "f16": 1, "bf16": 2, "f16u": 3}[storage])
    g.set_flags(flags)
    o.set_flags(flags)
    mf = np.asarray(m, np.float32)
    g.set_moments(mf)
    o.set_moments(mf.astype(np.float64))
    return g, o


def compare(g, o, which=0, tol=1e-5):
    a = g.moments(which).astype(np.float64)
    b = o.moments(which)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    same = np.mean(a == b)
    note(f"  rel-L2 {rel:.3e}, bit-identical fraction {same:.6f}")
    assert rel <= tol
    return rel, same


def test_taylor_green_periodic():
    nx, ny, nz = 32, 24, 16
    note(f"case 3d taylor-green periodic {nx}x{ny}x{nz}")
    m = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=5.0, seed=11)
    flags = np.ones(nx * ny * nz, np.uint8)
    g, o = pair3(nx, ny, nz, flags, m, periodic=True)
    g.step(50)
    o.step(50)
    compare(g, o)


def test_channel_walls_inlet_outlet_clamped():
    nx, ny, nz = 40, 20, 16
    note(f"case 3d channel {nx}x{ny}x{nz}")
    f = np.ones((nz, ny, nx), np.uint8)
    f[:, :, 0] = 13  # InletLeft
    f[:, :, nx - 1] = 20  # OutletRight
    f[:, 0, :] = 5
    f[:, ny - 1, :] = 5
    f[0, :, :] = 5
    f[nz - 1, :, :] = 5
    flags = f.ravel()
    g = capi.Domain(3, nx, ny, nz, viscosity=0.01)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, nu=0.01)
    g.set_flags(flags)
    o.set_flags(flags)
    g.init_state([0.05, 0, 0, 0, 0, 0])
    o.init_reference([0.05, 0, 0, 0, 0, 0])
    compare(g, o, 0, tol=0)
    g.step(40)
    o.step(40)
    compare(g, o, 0)
    compare(g, o, 1)


@pytest.mark.parametrize("spin,storage", [(0.0, "f32"), (0.02, "f32"), (0.02, "f16"), (0.02, "f16u")])
def test_sphere_moving_wall_links(spin, storage):
    nx, ny, nz = 40, 32, 32
    note(f"case 3d sphere spin={spin} storage={storage}")
    flags = np.ones(nx * ny * nz, np.uint8)
    rng = np.random.default_rng(5)
    u = np.zeros((nx * ny * nz, 3))
    u[:, 0] = 0.03
    u += rng.uniform(-1e-3, 1e-3, u.shape)
    m = wl.equilibrium_moments(u)
    g, o = pair3(nx, ny, nz, flags, m, periodic=True, storage=storage)
    c = np.array([19.3, 15.6, 16.2])
    v, tri = icosphere(c, 6.1, subdiv=2)
    vel = spin * np.cross([0.0, 0.0, 1.0], v - c)
    g.set_boundary(v, vel, tri)
    o.set_boundary(v, vel, tri)
    g.step(20)
    o.step(20)
    compare(g, o)
    fg = g.vertex_forces(reset=True).astype(np.float64)
    fo = o.vertex_forces()
    assert np.abs(fo).max() > 0
    assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5
    g.step(1)
    f1 = g.vertex_forces().astype(np.float64)
    w = g.body_wrench()[0]
    scale = np.abs(f1).sum()
    assert np.allclose(w[:3], f1.sum(0), rtol=0, atol=1e-5 * scale)
    r = v - v.mean(0)
    T = np.cross(r, f1).sum(0)
    assert np.allclose(w[3:], T, rtol=0, atol=1e-4 * np.abs(r).max() * scale)


def test_step_proxy_with_device_arrays_3d():
    # ff_step_proxy with the caller's device arrays (SolidGeometryProxy_Device<3>
    # layout): a sphere translated each step; moments equal the oracle's,
    # forces added into the caller's array equal the library's own
    import ctypes as C
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]

    def dev(a):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    nx, ny, nz = 40, 32, 32
    note(f"case 3d step_proxy {nx}x{ny}x{nz}")
    rng = np.random.default_rng(9)
    u = np.zeros((nx * ny * nz, 3))
    u[:, 0] = 0.02
    u += rng.uniform(-1e-3, 1e-3, u.shape)
    g, o = pair3(nx, ny, nz, np.ones(nx * ny * nz, np.uint8), wl.equilibrium_moments(u), periodic=True)
    v0, tri = icosphere([18.2, 15.7, 16.4], 5.5, subdiv=2)
    vel = np.tile(np.float32([0.03, -0.01, 0.02]), (len(v0), 1))
    tri = np.ascontiguousarray(tri, np.uint32)
    v32 = np.ascontiguousarray(v0, np.float32)
    g.set_boundary(v32, vel, tri)
    o.set_boundary(v32, vel, tri)
    dv, du, de = dev(v32), dev(vel), dev(tri)
    dF = dev(np.zeros_like(v32))
    L = capi.lib()
    try:
        for s_ in range(15):
            v32 = np.ascontiguousarray(v0 + s_ * vel, np.float32)
            assert hip.hipMemcpy(dv, v32.ctypes.data, v32.nbytes, 1) == 0
            assert L.ff_step_proxy(g.h, dv, du, de, dF, len(v32), len(tri)) == 0, L.ff_last_error()
            g.sync()  # the proxy arrays are rewritten next iteration
            o.update_boundary(v32, vel)
            o.step(1)
        _, same = compare(g, o)
        assert same == 1.0
        F = np.empty_like(v32)
        assert hip.hipMemcpy(F.ctypes.data, dF, F.nbytes, 2) == 0
    finally:
        for p in (dv, du, de, dF):
            hip.hipFree(p)
    fo = o.vertex_forces()
    assert np.abs(fo).max() > 0
    assert np.linalg.norm(F.astype(np.float64) - fo) / np.linalg.norm(fo) < 1e-5
    assert np.array_equal(F, g.vertex_forces())


@pytest.mark.parametrize("periodic", [False, True])
def test_spheres_across_the_grid_edge(periodic):
    # one sphere through the x = 0 face, one through a corner: links whose
    # upstream cell leaves the grid are not moving-wall links
    nx, ny, nz = 40, 32, 24
    note(f"case 3d spheres across the edge periodic={periodic}")
    rng = np.random.default_rng(8)
    u = np.zeros((nx * ny * nz, 3))
    u[:, 0] = 0.02
    u += rng.uniform(-1e-3, 1e-3, u.shape)
    g, o = pair3(nx, ny, nz, np.ones(nx * ny * nz, np.uint8), wl.equilibrium_moments(u), periodic=periodic)
    v1, t1 = icosphere([1.7, 16.3, 12.2], 5.5, subdiv=2)
    v2, t2 = icosphere([38.9, 31.2, 0.6], 4.5, subdiv=2)
    v = np.concatenate([v1, v2])
    tri = np.concatenate([t1, t2 + len(v1)])
    vel = np.tile([0.01, 0.0, -0.01], (len(v), 1))
    g.set_boundary(v, vel, tri)
    o.set_boundary(v, vel, tri)
    g.step(15)
    o.step(15)
    _, same = compare(g, o)
    assert same == 1.0
    fg = g.vertex_forces().astype(np.float64)
    fo = o.vertex_forces()
    assert np.abs(fo[len(v1):]).max() > 0
    assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5


@pytest.mark.parametrize("storage", ["f16", "bf16", "f16u"])
def test_16bit_storage_3d(storage):
    nx, ny, nz = 24, 16, 16
    note(f"case 3d {storage} storage")
    m = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=4.0, seed=12)
    flags = np.ones(nx * ny * nz, np.uint8)
    g, o = pair3(nx, ny, nz, flags, m, periodic=True, storage=storage)
    g.step(30)
    o.step(30)
    # the storage rounding is the oracle's, so the 16-bit state matches bit for bit
    _, same = compare(g, o, tol=1e-5)
    assert same == 1.0


@pytest.mark.parametrize("tile", [0, 2, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15])
def test_grid_smaller_than_tile(tile):
    """6x5x3 periodic cells: every box or column is partial and the halo
    wraps around the whole grid."""
    nx, ny, nz = 6, 5, 3
    note(f"case 3d tiny tile {tile}")
    m = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=1.5, seed=17)
    g, o = pair3(nx, ny, nz, np.ones(nx * ny * nz, np.uint8), m, periodic=True)
    g.set_tile(tile)
    g.step(10)
    o.step(10)
    rel, same = compare(g, o, 0)
    assert same == 1.0


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15])
def test_tile_variants_match_oracle(tile):
    # walls on every side (3D flag bounce-back), inlet/outlet, nz = 40 so the
    # 32-plane columns have a full and a partial chunk
    nx, ny, nz = 36, 20, 40
    note(f"case 3d tile {tile}")
    f = np.ones((nz, ny, nx), np.uint8)
    f[:, :, 0] = 13
    f[:, :, nx - 1] = 20
    f[:, 0, :] = 5
    f[:, ny - 1, :] = 5
    f[0, :, :] = 5
    f[nz - 1, :, :] = 5
    flags = f.ravel()
    g = capi.Domain(3, nx, ny, nz, viscosity=0.01)
    g.set_tile(tile)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, nu=0.01)
    g.set_flags(flags)
    o.set_flags(flags)
    g.init_state([0.05, 0, 0, 0, 0, 0])
    o.init_reference([0.05, 0, 0, 0, 0, 0])
    g.step(15)
    o.step(15)
    rel, same = compare(g, o, 0)
    assert same == 1.0
    # periodic Taylor-Green with the same tile
    m = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=5.0, seed=13)
    g2, o2 = pair3(nx, ny, nz, np.ones(nx * ny * nz, np.uint8), m, periodic=True)
    g2.set_tile(tile)
    g2.step(15)
    o2.step(15)
    rel, same = compare(g2, o2, 0)
    assert same == 1.0


def _channel_flags(nx, ny, nz):
    f = np.ones((nz, ny, nx), np.uint8)
    f[:, :, 0] = 13
    f[:, :, nx - 1] = 20
    f[:, 0, :] = 5
    f[:, ny - 1, :] = 5
    f[0, :, :] = 5
    f[nz - 1, :, :] = 5
    return f.ravel()


@pytest.mark.parametrize("tile", list(range(16)))
def test_fast_math_tiles_within_tolerance(tile):
    """FF_OPT_FAST_MATH (pair-decomposed reconstruction and moments, FMA):
    moments within 1e-5 relative L2 of the reference-order oracle after 100
    steps, for a periodic Taylor-Green field and a walled channel (flag
    bounce-back on all sides, inlet and outlet); nz = 40 leaves partial
    column chunks; the velocity error against the FP64 oracle stays at the
    reference-order kernel's level; and every tile shape gives the same bits
    as the 64x4x16 column (the fast form does not depend on the tile)."""
    nx, ny, nz = 36, 30, 40
    steps = 100
    note(f"case 3d fast math tile {tile}")
    m = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=6.0, seed=21)
    g = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True, fast_math=True)
    g.set_tile(tile)
    ex = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True)
    ex.set_tile(tile)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, periodic=True, nu=0.01)
    o64 = ol.OracleDomain(3, nx, ny, nz, fp64=True, stage_fp16=False, periodic=True, nu=0.01)
    for d in (g, ex):
        d.set_moments(np.asarray(m, np.float32))
    o.set_moments(np.asarray(m, np.float32).astype(np.float64))
    o64.set_moments(np.asarray(m, np.float32).astype(np.float64))
    for d in (g, ex, o, o64):
        d.step(steps)
    a, b, r64 = g.moments(0).astype(np.float64), o.moments(0), o64.moments(0)
    t9 = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True, fast_math=True)
    t9.set_tile(9)
    t9.set_moments(np.asarray(m, np.float32))
    t9.step(steps)
    assert np.array_equal(g.moments(0), t9.moments(0)), "fast form differs between tile shapes"
    t9.close()
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    e_fast = np.linalg.norm(a[:, 1:4] - r64[:, 1:4]) / np.linalg.norm(r64[:, 1:4])
    e_ref = np.linalg.norm(ex.moments(0)[:, 1:4].astype(np.float64) - r64[:, 1:4]) / np.linalg.norm(r64[:, 1:4])
    note(f"  fast vs FP32 oracle rel-L2 {rel:.3e}; velocity vs FP64: fast {e_fast:.3e}, reference order {e_ref:.3e}")
    assert rel <= 1e-5
    assert e_fast <= max(2.0 * e_ref, 1e-6)
    # walled channel: flag bounce-back through the re-reconstructed centre classes
    flags = _channel_flags(nx, ny, nz)
    g = capi.Domain(3, nx, ny, nz, viscosity=0.01, fast_math=True)
    g.set_tile(tile)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, nu=0.01)
    g.set_flags(flags)
    o.set_flags(flags)
    g.init_state([0.05, 0, 0, 0, 0, 0])
    o.init_reference([0.05, 0, 0, 0, 0, 0])
    g.step(steps)
    o.step(steps)
    for which in (0, 1):
        a, b = g.moments(which).astype(np.float64), o.moments(which)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        note(f"  channel buffer {which}: fast vs FP32 oracle rel-L2 {rel:.3e}")
        assert rel <= 1e-5


def test_step_timed_equals_step():
    """ff_step_timed advances the state exactly as ff_step does, and its
    per-launch stream time fits inside the region's wall time."""
    nx, ny, nz = 48, 40, 32
    note("case step_timed vs step")
    m = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=6.0, seed=3)
    a = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True)
    b = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True)
    for d in (a, b):
        d.set_moments(np.asarray(m, np.float32))
    a.step(12)
    total, stream = b.step_timed(12)
    assert 0.0 < stream <= total
    assert np.array_equal(a.moments(0), b.moments(0))
    assert np.array_equal(a.moments(1), b.moments(1))
    a.close()
    b.close()
