"""Per-step readback for RL loops and frame output, through the C ABI and the drop-in.

Power: the reference's RL loop (rl/env.py:167-178) copies every vertex force
and velocity to the host each simulation step and reduces
power = sum(-force * velocity).  ff_get_body_power returns that sum per body,
reduced on the device in FP64; it is checked against the oracle's vertex
forces (FP64) and the velocities the step used, per body, every step, in 2D
(two bodies) and 3D, and the power ring against the per-step values.

Frames: save_frame_data writes legacy-VTK files (the fluid velocity on the
grid, the solid's vertices and cells); the files are parsed and compared with
get_fluid_moments and getVertices.
"""
import sys

import numpy as np
import pytest

import meshes
import oracle_lib as ol
from freeflow_amd import capi, workloads as wl
import freeflow_amd

pytestmark = pytest.mark.gpu
fs = freeflow_amd.enable_dropin()


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_power(o, vel, body_of_vertex, nb):
    f = o.vertex_forces()  # FP64, this step only (zeroed before it)
    pv = -(f * vel.astype(np.float64)).sum(1)
    return np.array([pv[body_of_vertex == b].sum() for b in range(nb)]), \
        np.array([np.abs(f[body_of_vertex == b] * vel[body_of_vertex == b]).sum() for b in range(nb)])


def test_body_power_2d_two_bodies():
    nx, ny, dx, dt = 240, 120, 0.005, 0.0025
    note(f"case body power 2D two bodies {nx}x{ny}")
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.tile([0.01, 0.0], (nx * ny, 1)))
    g = capi.Domain(2, nx, ny, viscosity=0.005, dx=dx, dt=dt, fluid_density=10.0)
    o = ol.OracleDomain(2, nx, ny, fp64=False, stage_fp16=True, nu=0.005, dx=dx, dt=dt,
                        cf=capi.cf_of(2, 10.0, dx, dt))
    g.set_flags(flags)
    o.set_flags(flags)
    g.set_moments(m)
    o.set_moments(m.astype(np.float32).astype(np.float64))
    sw1 = wl.Swimmer2D(120, 0.35, 0.3, 0.15, 0.04, A=0.03, f=2.0, phi=0.4)
    sw2 = wl.Swimmer2D(80, 0.8, 0.28, 0.1, 0.03, A=0.02, f=3.0, phi=1.9)
    e = np.vstack([sw1.elems, sw2.elems + len(sw1.rest)])
    eb = np.r_[np.zeros(len(sw1.elems), np.int32), np.ones(len(sw2.elems), np.int32)]
    vb = np.r_[np.zeros(len(sw1.rest), int), np.ones(len(sw2.rest), int)]

    def state(t):
        v1, u1 = sw1.state(t)
        v2, u2 = sw2.state(t)
        return np.vstack([v1, v2]).astype(np.float32), np.vstack([u1, u2]).astype(np.float32)

    v, u = state(0.0)
    g.set_boundary(v, u, e, elem_body=eb, n_bodies=2)
    o.set_boundary(v, u, e)
    g.set_wrench_history(16)
    got = []
    for s in range(12):
        v, u = state(s * dt)
        g.update_boundary(v, u)
        o.update_boundary(v, u)
        o.zero_vertex_forces()
        g.step(1)
        o.step(1)
        p = g.body_power()
        ref, scale = _oracle_power(o, u, vb, 2)
        assert np.all(scale > 0)
        assert np.all(np.abs(p - ref) <= 1e-5 * scale), f"step {s}: {p} vs {ref}"
        got.append(p.copy())
    hist = g.power_history(12)
    assert hist.shape == (12, 2)
    assert np.array_equal(hist, np.array(got))
    note(f"  last step power per body {got[-1]}")


def test_body_power_3d_moving_sphere():
    nx, ny, nz = 40, 32, 32
    note(f"case body power 3D sphere {nx}x{ny}x{nz}")
    rng = np.random.default_rng(4)
    u0 = np.zeros((nx * ny * nz, 3))
    u0[:, 0] = 0.02
    u0 += rng.uniform(-1e-3, 1e-3, u0.shape)
    m = np.asarray(wl.equilibrium_moments(u0), np.float32)
    g = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, periodic=True, nu=0.01)
    g.set_moments(m)
    o.set_moments(m.astype(np.float64))
    v0, tri = wl.icosphere([18.2, 15.7, 16.4], 5.5, subdiv=2)
    vel = np.tile(np.float32([0.03, -0.01, 0.02]), (len(v0), 1))
    for s in range(8):
        v = np.asarray(v0 + s * vel, np.float32)
        if s == 0:
            g.set_boundary(v, vel, tri)
            o.set_boundary(v, vel, tri)
        else:
            g.update_boundary(v, vel)
            o.update_boundary(v, vel)
        o.zero_vertex_forces()
        g.step(1)
        o.step(1)
        p = g.body_power()
        ref, scale = _oracle_power(o, vel, np.zeros(len(v0), int), 1)
        assert scale[0] > 0
        assert abs(p[0] - ref[0]) <= 1e-5 * scale[0], f"step {s}: {p[0]} vs {ref[0]}"


def _read_vtk_points(path):
    with open(path) as f:
        lines = f.read().split("\n")
    return lines


def test_save_frame_data_2d_and_3d(tmp_path):
    out = tmp_path / "frames"
    out.mkdir()
    v, t = meshes.disk(24)
    mesh = tmp_path / "disk.mesh"
    meshes.write_mesh(mesh, 2, v, tris=t)
    nx, ny = 64, 40
    path = meshes.write_config(
        tmp_path / "f2.json", dimension=2, fluid_viscosity=0.01, fluid_density=1000.0, fluid_nx=nx, fluid_ny=ny,
        fluid_dx=0.01, dt=0.005, output_frequency=1, output_path=str(out),
        boundary_velocities=[0.03, 0.0, 0.0, 0.0], boundaries=meshes.channel_boundaries(nx, ny),
        solids=[{"mesh_path": str(mesh), "translate": [0.2, 0.2], "scale": [0.06, 0.06]}])
    cfg = fs.Config2D()
    assert cfg.load(path)
    sim = fs.Simulator2D(cfg.get_params())
    sim.initialize()
    for _ in range(5):
        sim.step()
    sim.save_frame_data(5)
    lines = _read_vtk_points(out / "fluid_frame_5.vtk")
    assert lines[0].startswith("# vtk DataFile") and "STRUCTURED_POINTS" in lines[3]
    assert lines[4].split() == ["DIMENSIONS", str(nx), str(ny), "1"]
    assert lines[6].split()[0] == "SPACING" and float(lines[6].split()[1]) == pytest.approx(0.01)
    assert lines[7].split() == ["POINT_DATA", str(nx * ny)]
    vec = np.array([[float(x) for x in l.split()] for l in lines[9:9 + nx * ny]])
    m = sim.get_fluid_moments().reshape(-1, 6)
    assert vec.shape == (nx * ny, 3)
    np.testing.assert_allclose(vec[:, :2], m[:, 1:3], rtol=1e-5, atol=1e-9)
    assert np.all(vec[:, 2] == 0.0)
    assert np.abs(vec[:, 0]).max() > 0
    sl = _read_vtk_points(out / "solid_frame_5.vtk")
    assert "UNSTRUCTURED_GRID" in sl[3]
    npts = int(sl[4].split()[1])
    pts = np.array([[float(x) for x in l.split()] for l in sl[5:5 + npts]])
    np.testing.assert_allclose(pts[:, :2], sim.getVertices(), rtol=1e-5, atol=1e-9)
    ncell = int(sl[5 + npts].split()[1])
    assert ncell == len(t)
    # 3D
    corners, tets = meshes.kuhn_cube()
    mesh3 = tmp_path / "cube.mesh"
    meshes.write_mesh(mesh3, 3, corners, tets=tets)
    nx, ny, nz = 24, 16, 16
    path = meshes.write_config(
        tmp_path / "f3.json", dimension=3, fluid_viscosity=0.02, fluid_density=1000.0, fluid_nx=nx, fluid_ny=ny,
        fluid_nz=nz, fluid_dx=0.01, dt=0.005, output_frequency=1, output_path=str(out),
        boundary_velocities=[0.0] * 6,
        solids=[{"mesh_path": str(mesh3), "translate": [0.08, 0.05, 0.05], "scale": [0.05, 0.05, 0.05]}])
    cfg = fs.Config3D()
    assert cfg.load(path)
    sim = fs.Simulator3D(cfg.get_params())
    sim.initialize()
    sim.apply_kinematic_control(0, np.array([[0.15, 0.08, 0.08], [0.08, 0.05, 0.05]]), 0.2)
    for _ in range(4):
        sim.step()
    sim.save_frame_data(4)
    lines = _read_vtk_points(out / "fluid_frame_4.vtk")
    assert lines[4].split() == ["DIMENSIONS", str(nx), str(ny), str(nz)]
    vec = np.array([[float(x) for x in l.split()] for l in lines[9:9 + nx * ny * nz]])
    m = sim.get_fluid_moments().reshape(-1, 10)
    np.testing.assert_allclose(vec, m[:, 1:4], rtol=1e-5, atol=1e-9)
    assert np.abs(vec).max() > 0
    sl = _read_vtk_points(out / "solid_frame_4.vtk")
    npts = int(sl[4].split()[1])
    pts = np.array([[float(x) for x in l.split()] for l in sl[5:5 + npts]])
    np.testing.assert_allclose(pts, sim.getVertices(), rtol=1e-5, atol=1e-9)
    assert int(sl[5 + npts].split()[1]) == len(tets)


def test_dropin_body_power_and_fast_math(tmp_path):
    """The drop-in's get_body_power equals the host sum the reference's RL
    loop computes from getForce and getVelocity (rl/env.py:175-178), in the
    reference form and with the `fast_math` device option, whose flow agrees
    with the reference form to 1e-5."""
    v, t = meshes.disk(32)
    mesh = tmp_path / "disk.mesh"
    meshes.write_mesh(mesh, 2, v, tris=t)
    nx, ny = 96, 64
    sims = {}
    for fast in (False, True):
        extra = {"freeflow_amd": {"fast_math": True, "stage_fp16": False}} if fast else \
            {"freeflow_amd": {"stage_fp16": False}}
        path = meshes.write_config(
            tmp_path / f"p{int(fast)}.json", dimension=2, fluid_viscosity=0.01, fluid_density=1000.0, fluid_nx=nx,
            fluid_ny=ny, fluid_dx=0.01, dt=0.005, solid_solver_type="static", output_frequency=0,
            boundary_velocities=[0.03, 0.0, 0.0, 0.0],
            boundaries=meshes.channel_boundaries(nx, ny),
            solids=[{"mesh_path": str(mesh), "translate": [0.3, 0.3], "scale": [0.08, 0.08]}], **extra)
        cfg = fs.Config2D()
        assert cfg.load(path)
        params = cfg.get_params()
        assert params.device_options.fast_math == fast
        sim = fs.Simulator2D(params)
        sim.initialize()
        sim.apply_kinematic_control(0, np.array([[0.45, 0.33], [0.3, 0.3]]), 0.05)  # needs two targets
        nonzero = 0
        for _ in range(12):
            U = np.asarray(sim.getVelocity(), np.float64)  # the velocities this step's boundary carries
            sim.step()
            p = sim.get_body_power()
            F = np.asarray(sim.getForce(), np.float64)
            ref = -(F * U).sum()
            scale = np.abs(F * U).sum()
            assert abs(p[0] - ref) <= 1e-4 * scale + 1e-30, f"{p[0]} vs {ref}"
            nonzero += scale > 0
        assert nonzero >= 8
        sims[fast] = (sim, cfg, params)
    a = sims[False][0].get_fluid_moments().astype(np.float64)
    b = sims[True][0].get_fluid_moments().astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    note(f"  drop-in fast math vs reference form after 12 steps: rel-L2 {rel:.3e}")
    assert 0.0 < rel <= 1e-5
