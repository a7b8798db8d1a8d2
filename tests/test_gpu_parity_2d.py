"""GPU parity of the D2Q9 path against the CPU oracle, through the C ABI.

The 2D kernels are written in the reference's operation order and compiled
without FMA contraction, so the bar is bit-exact moments (FP32 and the 16-bit
storage modes, with and without the reference's FP16 tile staging).  Forces
are summed in a different (deterministic) order than the oracle, so they are
compared to a relative tolerance of 1e-5.
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


def pair(nx, ny, flags, m, stage=True, periodic=False, storage="f32", nu=0.005, dx=1.0, dt=1.0, rho=1.0):
    g = capi.Domain(2, nx, ny, viscosity=nu, dx=dx, dt=dt, fluid_density=rho, storage=storage,
                    stage_fp16=stage, periodic=periodic)
    cf = capi.cf_of(2, rho, dx, dt)
    o = ol.OracleDomain(2, nx, ny, fp64=False, stage_fp16=stage, periodic=periodic, nu=nu, dx=dx, dt=dt, cf=cf,
                        storage={"f32": 0, "f16": 1, "bf16": 2, "f16u": 3}[storage])
    g.set_flags(flags)
    o.set_flags(flags)
    mf = np.asarray(m, np.float32)
    g.set_moments(mf)
    o.set_moments(mf.astype(np.float64))
    return g, o


def assert_same(g, o, which=0):
    a = g.moments(which).astype(np.float64)
    b = o.moments(which)
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        raise AssertionError(f"{len(bad)} differing values, first {bad[:5].tolist()}, rel-L2 {rel:.3e}")


@pytest.mark.parametrize("stage,periodic", [(True, False), (False, False), (True, True), (False, True)])
def test_vortex_street_bitexact(stage, periodic):
    nx, ny = 96, 48
    note(f"case vortex_street nx={nx} ny={ny} stage={stage} periodic={periodic}")
    flags = wl.vortex_street_flags(nx, ny)
    m = wl.noisy_inflow_moments(flags, nx, ny, 0.05, 1e-3, seed=7)
    g, o = pair(nx, ny, flags, m, stage=stage, periodic=periodic)
    for _ in range(4):
        g.step(25)
        o.step(25)
        assert_same(g, o, 0)
    assert_same(g, o, 1)


@pytest.mark.parametrize("tile,periodic", [(0, True), (4, True), (4, False)])
def test_grid_smaller_than_tile_bitexact(tile, periodic):
    """7x3 cells: every tile is partial and the periodic halo wraps."""
    nx, ny = 7, 3
    note(f"case 2d tiny tile {tile} periodic {periodic}")
    rng = np.random.default_rng(3)
    m = wl.equilibrium_moments(0.02 * rng.standard_normal((nx * ny, 2)))
    g, o = pair(nx, ny, np.ones(nx * ny, np.uint8), m, periodic=periodic)
    g.set_tile(tile)
    g.step(10)
    o.step(10)
    assert_same(g, o)


def test_reference_initializer_bitexact():
    nx, ny = 64, 40
    note("case init_state")
    flags = wl.vortex_street_flags(nx, ny)
    g = capi.Domain(2, nx, ny, viscosity=0.005)
    o = ol.OracleDomain(2, nx, ny, fp64=False)
    g.set_flags(flags)
    o.set_flags(flags)
    g.init_state([0.05, 0.0, 0.0, 0.0])
    o.init_reference([0.05, 0.0, 0.0, 0.0])
    assert_same(g, o, 0)
    assert_same(g, o, 1)


@pytest.mark.parametrize("storage", ["f16", "bf16", "f16u"])
def test_16bit_storage_bitexact(storage):
    nx, ny = 80, 40
    note(f"case storage={storage}")
    flags = wl.vortex_street_flags(nx, ny)
    m = wl.noisy_inflow_moments(flags, nx, ny, 0.05, 1e-3, seed=8)
    g, o = pair(nx, ny, flags, m, stage=False, storage=storage)
    assert_same(g, o, 0)
    g.step(60)
    o.step(60)
    assert_same(g, o, 0)
    assert_same(g, o, 1)


def test_config1_cylinder_links_bitexact():
    # 400x200 vortex street, static 128-gon cylinder r=20 at (100,100), through
    # the moving-wall link path with zero wall velocity (the cylinder case).
    c = wl.CONFIG1
    nx, ny = c["nx"], c["ny"]
    note(f"case config1 cylinder {nx}x{ny}")
    flags = wl.vortex_street_flags(nx, ny)
    m = wl.noisy_inflow_moments(flags, nx, ny, c["u_in"], 1e-3, seed=c["seed"])
    g, o = pair(nx, ny, flags, m, nu=c["viscosity"])
    v, e = wl.cylinder(c["n_poly"], c["cx"], c["cy"], c["r"])
    g.set_boundary(v, np.zeros_like(v), e)
    o.set_boundary(v, np.zeros_like(v), e)
    g.step(100)
    o.step(100)
    assert_same(g, o, 0)
    fg = g.vertex_forces(reset=True).astype(np.float64)
    fo = o.vertex_forces()
    assert np.abs(fo).max() > 1e-3
    assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5
    # one more step: the FP64 body wrench equals that step's vertex forces
    # summed, and its torque equals sum (x_v - ref) x f_v (barycentric split)
    g.step(1)
    f1 = g.vertex_forces().astype(np.float64)
    w = g.body_wrench()[0]
    scale = np.abs(f1).sum()
    assert np.allclose(w[:2], f1.sum(0), rtol=0, atol=1e-5 * scale)
    r = v.astype(np.float64) - v.astype(np.float64).mean(0)
    tz = np.sum(r[:, 0] * f1[:, 1] - r[:, 1] * f1[:, 0])
    assert abs(w[2] - tz) < 1e-4 * np.sum(np.abs(r).max() * np.abs(f1))


@pytest.mark.parametrize("periodic", [False, True])
def test_bodies_across_the_grid_edge_bitexact(periodic):
    # polygons partly outside the grid (left edge, top edge, a corner): links
    # whose upstream cell leaves the grid are not moving-wall links, and the
    # candidate window is clipped to the grid, not wrapped
    nx, ny = 120, 80
    note(f"case bodies across the edge periodic={periodic}")
    rng = np.random.default_rng(33)
    u = np.zeros((nx * ny, 2))
    u[:, 0] = 0.02
    u += rng.uniform(-1e-3, 1e-3, u.shape)
    g, o = pair(nx, ny, np.ones(nx * ny, np.uint8), wl.equilibrium_moments(u), periodic=periodic, nu=0.01)
    vs, es = [], []
    for k, (cx, cy, r) in enumerate([(2.3, 40.2, 9.0), (60.4, 78.1, 7.5), (118.6, 1.2, 6.0)]):
        v, e = wl.cylinder(96, cx, cy, r)
        vs.append(v)
        es.append(e + 96 * k)
    v, e = np.concatenate(vs), np.concatenate(es)
    vel = np.tile([0.01, -0.015], (len(v), 1))
    g.set_boundary(v, vel, e)
    o.set_boundary(v, vel, e)
    g.step(25)
    o.step(25)
    assert_same(g, o, 0)
    assert_same(g, o, 1)
    fg = g.vertex_forces().astype(np.float64)
    fo = o.vertex_forces()
    assert np.abs(fo).max() > 0
    assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5


def test_max_elements_65536_bitexact():
    # the largest boundary: 256 moving 256-gons, 65536 segments, so element
    # ids fill the 16-bit field of the link key; one body per polygon.
    nx = ny = 512
    note(f"case 65536 segments, 256 bodies, {nx}x{ny} periodic")
    rng = np.random.default_rng(21)
    u = np.zeros((nx * ny, 2))
    u[:, 0] = 0.03
    u += rng.uniform(-1e-3, 1e-3, u.shape)
    g, o = pair(nx, ny, np.ones(nx * ny, np.uint8), wl.equilibrium_moments(u), periodic=True, nu=0.01)
    vs, es, vels = [], [], []
    for k in range(256):
        cx, cy = 16.0 + 32.0 * (k % 16) + rng.uniform(-2, 2), 16.0 + 32.0 * (k // 16) + rng.uniform(-2, 2)
        v, e = wl.cylinder(256, cx, cy, 7.0 + rng.uniform(0, 2))
        es.append(e + 256 * k)
        vs.append(v)
        vels.append(np.tile(rng.uniform(-0.02, 0.02, 2), (256, 1)))
    v, e, vel = np.concatenate(vs), np.concatenate(es), np.concatenate(vels)
    assert len(e) == 65536
    body = np.repeat(np.arange(256), 256)
    g.set_boundary(v, vel, e, elem_body=body, n_bodies=256)
    o.set_boundary(v, vel, e)
    g.step(10)
    o.step(10)
    assert_same(g, o, 0)
    fg = g.vertex_forces(reset=True).astype(np.float64)
    fo = o.vertex_forces()
    assert np.abs(fo[-256:]).max() > 0  # the last body (element ids near 65535) is hit
    assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5
    g.step(1)
    f1 = g.vertex_forces().astype(np.float64).reshape(256, 256, 2).sum(1)
    w = g.body_wrench()
    assert np.allclose(w[:, :2], f1, rtol=0, atol=1e-5 * np.abs(f1).max())


@pytest.mark.parametrize("storage,stage", [("f32", True), ("f16", False), ("bf16", False), ("f16u", False)])
def test_moving_swimmer_explicit_boundary(storage, stage):
    # explicit per-step vertex positions/velocities (the reference proxy
    # contract), an elliptical swimmer with a travelling wave, outlet box;
    # also in the 16-bit storage form the batched and served paths use.
    nx, ny, dx, dt = 240, 120, 0.005, 0.0025
    note(f"case swimmer explicit {nx}x{ny} storage={storage} stage={stage}")
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.zeros((nx * ny, 2)))
    g, o = pair(nx, ny, flags, m, stage=stage, storage=storage, nu=0.005, dx=dx, dt=dt, rho=10.0)
    sw = wl.Swimmer2D(160, 0.6, 0.3, 0.2, 0.05, A=0.02, f=1.5, phi=0.3)
    v, vel = sw.state(0.0)
    g.set_boundary(v, vel, sw.elems)
    o.set_boundary(v, vel, sw.elems)
    for s in range(40):
        v, vel = sw.state(s * dt)
        v32, vel32 = v.astype(np.float32), vel.astype(np.float32)
        g.update_boundary(v32, vel32)
        o.update_boundary(v32, vel32)
        g.step(1)
        o.step(1)
    assert_same(g, o, 0)
    assert_same(g, o, 1)
    fg = g.vertex_forces().astype(np.float64)
    fo = o.vertex_forces()
    assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5


def test_step_proxy_with_device_arrays():
    # ff_step_proxy: the caller's own device arrays in the proxy layout
    # (SolidGeometryProxy_Device<2>), new positions each step, forces added
    # into the caller's force array; moments equal the oracle fed the same
    # positions, forces equal those of the host-array path
    import ctypes as C
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    h2d, d2h = 1, 2

    def dev(a):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, h2d) == 0
        return p

    nx, ny, dx, dt = 240, 120, 0.005, 0.0025
    note(f"case step_proxy {nx}x{ny}")
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.zeros((nx * ny, 2)))
    g, o = pair(nx, ny, flags, m, nu=0.005, dx=dx, dt=dt, rho=10.0)
    hpath, _ = pair(nx, ny, flags, m, nu=0.005, dx=dx, dt=dt, rho=10.0)
    sw = wl.Swimmer2D(160, 0.6, 0.3, 0.2, 0.05, A=0.02, f=1.5, phi=0.3)
    v, vel = sw.state(0.0)
    v32, vel32 = v.astype(np.float32), vel.astype(np.float32)
    elems = np.ascontiguousarray(sw.elems, np.uint32)
    nv, ne = len(v32), len(elems)
    for x in (g, o, hpath):
        x.set_boundary(v32, vel32, elems)
    dv, du, de = dev(v32), dev(vel32), dev(elems)
    dF = dev(np.zeros((nv, 2), np.float32))
    L = capi.lib()
    try:
        for s_ in range(30):
            v, vel = sw.state(s_ * dt)
            v32 = np.ascontiguousarray(v, np.float32)
            vel32 = np.ascontiguousarray(vel, np.float32)
            assert hip.hipMemcpy(dv, v32.ctypes.data, v32.nbytes, h2d) == 0
            assert hip.hipMemcpy(du, vel32.ctypes.data, vel32.nbytes, h2d) == 0
            assert L.ff_step_proxy(g.h, dv, du, de, dF, nv, ne) == 0, L.ff_last_error()
            g.sync()  # the proxy arrays are rewritten next iteration
            o.update_boundary(v32, vel32)
            o.step(1)
            hpath.update_boundary(v32, vel32)
            hpath.step(1)
        assert_same(g, o, 0)
        assert_same(g, o, 1)
        F = np.empty((nv, 2), np.float32)
        assert hip.hipMemcpy(F.ctypes.data, dF, F.nbytes, d2h) == 0
    finally:
        for p in (dv, du, de, dF):
            hip.hipFree(p)
    fh = hpath.vertex_forces()
    assert np.abs(fh).max() > 0
    assert np.array_equal(F, fh)
    assert np.array_equal(g.vertex_forces(), fh)


def test_step_vertex_forces_and_body_ref():
    # the per-step forces equal one step's accumulation from zero; the torque
    # is taken about the reference point given by ff_set_body_ref
    nx, ny = 160, 80
    note(f"case step vertex forces and body ref {nx}x{ny}")
    d = capi.Domain(2, nx, ny, viscosity=0.01)
    d.set_flags(wl.outlet_box_flags(nx, ny))
    d.set_moments(wl.equilibrium_moments(np.tile([0.02, 0.0], (nx * ny, 1))))
    v, e = wl.cylinder(64, 60.3, 40.2, 10.0)
    vel = np.zeros_like(v)
    vel[:, 1] = 0.01
    d.set_boundary(v, vel, e)
    assert not np.any(d.step_vertex_forces())  # no step yet
    ref = np.array([[55.0, 37.5]])
    d.set_body_ref(ref)
    d.step(5)
    d.vertex_forces(reset=True)
    d.step(1)
    fs = d.step_vertex_forces()
    assert np.abs(fs).max() > 0
    assert np.array_equal(fs, d.vertex_forces())
    f = fs.astype(np.float64)
    w = d.body_wrench()[0]
    assert np.allclose(w[:2], f.sum(0), rtol=0, atol=1e-5 * np.abs(f).sum())

    def torque(about):
        r = v.astype(np.float64) - about
        return np.sum(r[:, 0] * f[:, 1] - r[:, 1] * f[:, 0]), 1e-4 * np.abs(r).max() * np.abs(f).sum()

    tz, tol = torque(ref[0])
    tc, _ = torque(v.astype(np.float64).mean(0))
    assert abs(w[2] - tz) < tol
    assert abs(tc - tz) > 10 * tol  # the default centroid would not pass


def test_wave_kinematics_on_device():
    nx, ny, dx, dt = 240, 120, 0.005, 0.0025
    note("case wave kinematics")
    flags = wl.outlet_box_flags(nx, ny)
    g = capi.Domain(2, nx, ny, viscosity=0.005, dx=dx, dt=dt, fluid_density=10.0)
    g.set_flags(flags)
    g.set_moments(wl.equilibrium_moments(np.zeros((nx * ny, 2))))
    sw = wl.Swimmer2D(160, 0.6, 0.3, 0.2, 0.05, A=0.02, f=1.5, phi=0.3)
    v0, vel0 = sw.state(0.0)
    g.set_boundary(v0, vel0, sw.elems)
    g.set_wave(sw.rest, sw.s, sw.params[None, :])
    g.step(7)  # kinematics of step k are evaluated at t = k dt before the step
    v, vel = g.boundary()
    ve, vele = sw.state(6 * dt)
    assert np.abs(v - ve).max() < 1e-6
    assert np.abs(vel - vele).max() < 1e-5


def test_wave_mode_matches_oracle_fed_with_device_kinematics():
    nx, ny, dx, dt = 240, 120, 0.005, 0.0025
    note("case wave mode parity")
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.zeros((nx * ny, 2)))
    g, o = pair(nx, ny, flags, m, nu=0.005, dx=dx, dt=dt, rho=10.0)
    sw = wl.Swimmer2D(160, 0.6, 0.3, 0.2, 0.05, A=0.03, f=2.0, phi=1.0)
    v0, vel0 = sw.state(0.0)
    g.set_boundary(v0, vel0, sw.elems)
    o.set_boundary(v0, vel0, sw.elems)
    g.set_wave(sw.rest, sw.s, sw.params[None, :])
    g.set_wrench_history(64)
    for _ in range(30):
        g.step(1)
        v, vel = g.boundary()
        o.update_boundary(v, vel)
        o.step(1)
    assert_same(g, o, 0)
    hist = g.wrench_history(64)
    assert hist.shape == (30, 1, 3)
    assert np.array_equal(hist[-1], g.body_wrench())


def _fill_case(kind, rng):
    """(nx, ny, dx, flags, verts, elems) of one fill-solid case."""
    if kind == "ellipse":
        nx, ny, dx = 160, 80, 0.01
        v, e = wl.polygon(90, 0.8, 0.4, 0.3, 0.15)
        return nx, ny, dx, wl.vortex_street_flags(nx, ny), v, e
    if kind == "on_centres":  # axis-aligned edges through cell centres and corners on them (ties)
        nx, ny, dx = 64, 48, 1.0
        v = np.array([[10.5, 7.5], [40.5, 7.5], [40.5, 30.5], [22.0, 30.5], [22.0, 41.0], [10.5, 41.0]], np.float32)
        e = np.array([[i, (i + 1) % len(v)] for i in range(len(v))], np.uint32)
        return nx, ny, dx, np.ones(nx * ny, np.uint8), v, e
    if kind == "outside":  # a star partly beyond every edge of the grid, and a second body
        nx, ny, dx = 96, 64, 0.0173
        n = 57
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(0.3, 1.0, n) * 70 * dx
        c = np.array([0.5 * nx, 0.5 * ny]) * dx
        v1 = c + np.stack([r * np.cos(ang), r * np.sin(ang)], 1)
        v2 = 0.3 * (v1 - c) + c + np.array([0.2, 0.1]) * nx * dx
        v = np.vstack([v1, v2]).astype(np.float32)
        e1 = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
        e = np.vstack([e1, e1 + n]).astype(np.uint32)
        return nx, ny, dx, wl.vortex_street_flags(nx, ny), v, e
    if kind == "wide":  # rows longer than one 8192-cell chunk of the row scan
        nx, ny, dx = 9000, 24, 1.0
        v = np.array([[3.2, 2.7], [8990.4, 5.5], [8100.0, 21.3], [40.7, 19.9], [4000.5, 11.5]], np.float32)
        e = np.array([[i, (i + 1) % len(v)] for i in range(len(v))], np.uint32)
        return nx, ny, dx, np.ones(nx * ny, np.uint8), v, e
    if kind == "swimmer":  # the 1200x400 workload's ~600-segment swimmer
        c = wl.SWIMMER2D
        sw = wl.swimmer2d_swimmer()
        v, _ = sw.state(0.37)
        return c["nx"], c["ny"], c["dx"], wl.outlet_box_flags(c["nx"], c["ny"]), v.astype(np.float32), sw.elems
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["ellipse", "on_centres", "outside", "wide", "swimmer"])
def test_fill_solid_matches_oracle(kind):
    """ff_fill_solid (row scan) against the oracle's per-cell ray cast: the
    same flags, byte for byte; called twice, with the boundary moved between
    the calls (SolidDynamic cells left behind return to Fluid)."""
    rng = np.random.default_rng(11)
    nx, ny, dx, flags, v, e = _fill_case(kind, rng)
    note(f"case fill_solid {kind} {nx}x{ny}, {len(e)} segments")
    g = capi.Domain(2, nx, ny, viscosity=0.005, dx=dx)
    g.set_flags(flags)
    g.set_boundary(v, np.zeros_like(v), e)
    cur = flags.copy()
    for shift in (0.0, 3.3 * dx):
        vv = (v + np.float32(shift)).astype(np.float32)
        if shift:
            g.update_boundary(vv, np.zeros_like(vv))
        g.fill_solid()
        ref = cur.copy()
        ol.lib().orc_fill_solid2d(nx, ny, dx, len(vv), np.ascontiguousarray(vv, np.float32).ravel(), len(e),
                                  np.ascontiguousarray(e, np.uint32).ravel(), ref)
        got = g.flags()
        assert np.array_equal(got, ref), f"{int(np.sum(got != ref))} cells differ"
        assert (ref == 4).sum() > 10
        cur = ref
    g.close()


def test_fill_solid_empty_boundary():
    nx, ny = 40, 30
    note("case fill_solid empty boundary")
    flags = wl.vortex_street_flags(nx, ny)
    flags[5 * nx + 7] = 4
    g = capi.Domain(2, nx, ny, viscosity=0.005)
    g.set_flags(flags)
    g.set_boundary(np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32), np.zeros((0, 2), np.uint32))
    g.fill_solid()
    ref = flags.copy()
    ref[ref == 4] = 1
    assert np.array_equal(g.flags(), ref)
    g.close()


def test_mass_fp64():
    nx, ny = 128, 64
    note("case mass")
    flags = wl.vortex_street_flags(nx, ny)
    m = wl.noisy_inflow_moments(flags, nx, ny, 0.05, 1e-3, seed=3)
    m[:, 0] += np.random.default_rng(4).uniform(-1e-3, 1e-3, nx * ny)
    g = capi.Domain(2, nx, ny, viscosity=0.005)
    g.set_flags(flags)
    g.set_moments(m)
    upd = (flags == 1) | (flags == 4)
    ref = np.sum(m[upd, 0].astype(np.float32).astype(np.float64))
    assert abs(g.mass() - ref) < 1e-10 * ref


def test_swimmer2d_1200x400_parity_and_determinism():
    # the 1200x400 swimmer workload at full size (~600-segment swimmer, wave
    # kinematics on device): 100 steps against the oracle (bit-exact moments
    # after 20 and after 100), then a repeat run must be bit-identical
    # (moments and per-step wrench).
    c = wl.SWIMMER2D
    nx, ny = c["nx"], c["ny"]
    note(f"case swimmer2d full size {nx}x{ny}")
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.zeros((nx * ny, 2)))
    sw = wl.swimmer2d_swimmer()
    runs = []
    for rep in range(2):
        g = capi.Domain(**wl.swimmer2d_domain_args())
        g.set_flags(flags)
        g.set_moments(m)
        v0, vel0 = sw.state(0.0)
        g.set_boundary(v0, vel0, sw.elems)
        g.set_wave(sw.rest, sw.s, sw.params[None, :])
        g.set_wrench_history(32)
        if rep == 0:
            o = ol.OracleDomain(2, nx, ny, fp64=False, stage_fp16=True, nu=c["viscosity"], dx=c["dx"], dt=c["dt"],
                                cf=capi.cf_of(2, c["rho"], c["dx"], c["dt"]))
            o.set_flags(flags)
            o.set_moments(m)
            o.set_boundary(v0, vel0, sw.elems)
            for s in range(100):
                g.step(1)
                v, vel = g.boundary()
                o.update_boundary(v, vel)
                o.step(1)
                if s + 1 in (20, 100):
                    assert_same(g, o, 0)
            assert_same(g, o, 1)
            fo = o.vertex_forces()
            fg = g.vertex_forces().astype(np.float64)
            assert np.linalg.norm(fg - fo) / np.linalg.norm(fo) < 1e-5
        else:
            g.step(100)
        runs.append((g.moments(), g.wrench_history(32)))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
