"""The pybind11 drop-in `fsi_simulator` (src/python/bindings.cpp surface).

CPU tests: configuration loading (Config.hpp / Config.cpp semantics), the
Medit mesh reader, the initial transforms and the boundary extraction, pinned
by closed forms (a fan-triangulated disk has its ring as boundary; a
tetrahedralised cube has a closed, outward-oriented 12-triangle surface of
known volume; the 3D rotation equals scipy's intrinsic Y-X-Z Euler matrix).

GPU tests drive Simulator2D/3D the way the reference's scripts do
(scripts/vortex_street_2d.py: load, initialize, apply_kinematic_control,
step) and compare every step with the oracle fed the same geometry and the
kinematic mover restated here in Python.
"""
import math
import sys

import numpy as np
import pytest

import meshes
import oracle_lib as ol
from freeflow_amd import capi, workloads as wl
import freeflow_amd

fs = freeflow_amd.enable_dropin()


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ config
def test_json_config_2d_fields(tmp_path):
    mesh = tmp_path / "d.mesh"
    v, t = meshes.disk(12)
    meshes.write_mesh(mesh, 2, v, tris=t)
    path = meshes.write_config(
        tmp_path / "c.json", dimension=2, fluid_viscosity=0.02, fluid_density=998.0, fluid_nx=33, fluid_ny=17,
        fluid_dx=0.01, solid_solver_type="static", total_time=2.5, dt=0.004, output_frequency=7,
        output_path=str(tmp_path / "out"), log_level="warn", log_file="x.log", use_newton=False,
        boundary_velocities=[0.1, 0.0, 0.0, 0.0],
        global_fem_options={"iterations": 3, "ls_beta": 0.25, "linear_solver_type": "direct_ldlt"},
        solids=[{"mesh_path": str(mesh), "translate": [0.1, 0.2], "scale": [0.5, 0.25], "rotate": 15.0,
                 "density": 500.0, "initial_velocity": [0.01, -0.02]}],
        boundaries=[{"type": "Wall", "pos": [0, 0]}, {"type": "InletLeft", "pos": [0, 5]},
                    {"type": "OutletRight", "pos": [32, 5]}, {"type": "Bogus", "pos": [1, 1]}],
        freeflow_amd={"storage": "bf16", "stage_fp16": False, "periodic": True})
    c = fs.Config2D()
    assert c.load(path)
    p = c.get_params()
    assert (p.fluid_nx, p.fluid_ny, p.fluid_viscosity, p.fluid_density, p.fluid_dx) == (33, 17, 0.02, 998.0, 0.01)
    assert (p.total_time, p.dt, p.output_frequency, p.log_level, p.log_file, p.use_newton) == \
        (2.5, 0.004, 7, "warn", "x.log", False)
    assert np.allclose(p.boundary_velocities, [0.1, 0, 0, 0])
    assert [b.flag for b in p.boundaries] == [5, 13, 20, 0]
    assert [list(b.position) for b in p.boundaries] == [[0, 0], [0, 5], [32, 5], [1, 1]]
    s = p.solids[0]
    assert (list(s.translate), list(s.scale), s.rotate, s.density) == ([0.1, 0.2], [0.5, 0.25], 15.0, 500.0)
    assert list(s.initial_velocity) == [0.01, -0.02]
    assert s.youngs_modulus == 1.0e6 and s.poisson_ratio == 0.45
    g = p.global_fem_options
    assert (g.iterations, g.ls_beta, g.linear_solver_type, g.substeps) == (3, 0.25, "direct_ldlt", 3)
    d = p.device_options
    assert (d.storage, d.stage_fp16, d.periodic) == ("bf16", False, True)
    # getCf (Config.hpp:96-100): float cm = rho dx^2, cm dx / dt^2
    cm = np.float32(998.0 * 0.01 * 0.01)
    assert p.getCf() == np.float32(float(cm) * 0.01 / (0.004 * 0.004))


def test_config_defaults(tmp_path):
    path = meshes.write_config(tmp_path / "e.json")
    c2 = fs.Config2D()
    assert c2.load(path)
    p = c2.get_params()
    assert (p.fluid_viscosity, p.fluid_density, p.fluid_nx, p.fluid_ny, p.fluid_dx) == (0.001, 1000.0, 400, 400, 1.0)
    assert (p.solid_solver_type, p.total_time, p.dt, p.output_frequency, p.output_path) == \
        ("static", 10.0, 0.001, 100, "./output")
    assert list(p.boundary_velocities) == [0.0] * 4 and len(p.boundaries) == 0 and len(p.solids) == 0
    c3 = fs.Config3D()
    path3 = meshes.write_config(tmp_path / "e3.json", output_frequency=5)
    assert c3.load(path3)
    q = c3.get_params()
    assert (q.fluid_nx, q.fluid_ny, q.fluid_nz) == (128, 128, 128)
    assert list(q.boundary_velocities) == []
    assert q.output_frequency == 100  # the 3D reader does not read it (Config.hpp:161-219)


def test_config_errors(tmp_path, capfd):
    c = fs.Config2D()
    assert not c.load(str(tmp_path / "missing.json"))
    bad = tmp_path / "bad.json"
    bad.write_text('{"fluid_nx": 10,, }')
    assert not c.load(str(bad))
    nopos = meshes.write_config(tmp_path / "nopos.json", boundaries=[{"type": "Wall"}])
    assert not c.load(nopos)
    notype = meshes.write_config(tmp_path / "notype.json", boundaries=[{"pos": [1, 2]}])
    assert not c.load(notype)
    capfd.readouterr()


def test_flag_names():
    for name, val in capi.FLAG.items():
        expect = val if (name == "Wall" or name.startswith("Inlet") or name.startswith("Outlet")) else 0
        assert fs.flag_from_name(name) == expect
    assert fs.flag_from_name("wall") == 0


# ------------------------------------------------------------------ meshes
def _params2d(tmp_path, solids, nx=64, ny=32):
    path = meshes.write_config(tmp_path / "s.json", fluid_nx=nx, fluid_ny=ny, fluid_dx=0.01, dt=0.005,
                               solids=solids, output_frequency=0)
    c = fs.Config2D()
    assert c.load(path)
    return c, c.get_params()


def _transform2(v, scale, rot, trans):
    """Mesh2D::applyInitialTransform in double: float-rounded scale/translate,
    angle converted in float, then (cos x + sin y, -sin x + cos y) + t."""
    sx, sy = float(np.float32(scale[0])), float(np.float32(scale[1]))
    ang = float(np.float32(np.float32(rot) * np.float32(0.017453292519943295)))
    cs, sn = math.cos(ang), math.sin(ang)
    tx, ty = float(np.float32(trans[0])), float(np.float32(trans[1]))
    out = []
    for x, y in v:
        x, y = sx * x, sy * y
        out.append((cs * x + sn * y + tx, -sn * x + cs * y + ty))
    return np.array(out)


def test_disk_boundary_and_transform(tmp_path):
    n = 24
    v, t = meshes.disk(n)
    # a co-dimensional tail: one extra vertex joined to ring vertex 1 by an Edge
    v = np.vstack([v, [0.9, 0.0]])
    mesh = tmp_path / "disk.mesh"
    meshes.write_mesh(mesh, 2, v, edges=[(n + 1, 1)], tris=t)
    c, p = _params2d(tmp_path, [{"mesh_path": str(mesh), "translate": [0.3, 0.2], "scale": [0.1, 0.05],
                                 "rotate": 30.0}])
    sim = fs.Simulator2D(p)
    sim.load_solids()
    got = sim.getVertices()
    exp = _transform2(v, (0.1, 0.05), 30.0, (0.3, 0.2))
    assert got.shape == (n + 2, 2)
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-15)
    # boundary = ring edges + the tail, as sorted pairs in lexicographic order,
    # indexed by surface vertex (volume vertices 1..n+1 -> 0..n)
    ring = {tuple(sorted((1 + i, 1 + (i + 1) % n))) for i in range(n)}
    edges = sorted(ring | {(1, n + 1)})
    assert list(sim.getBoundaryPointIdx()) == list(range(1, n + 2))
    be = np.array(sim.getBoundaryElements()).reshape(-1, 2)
    assert be.tolist() == [[a - 1, b - 1] for a, b in edges]
    bv = sim.getBoundaryVertices()
    np.testing.assert_array_equal(bv, exp[1:].astype(np.float32).astype(np.float64))
    assert list(sim.getTriangles()) == [k for tri in t for k in tri]
    # the ring polygon's area equals the triangles' area (shoelace)
    ring_xy = exp[1:n + 1]
    shoelace = 0.5 * abs(np.sum(ring_xy[:, 0] * np.roll(ring_xy[:, 1], -1) - np.roll(ring_xy[:, 0], -1) * ring_xy[:, 1]))
    def cross2(p, q):
        return p[0] * q[1] - p[1] * q[0]

    tri_area = sum(0.5 * abs(cross2(exp[b] - exp[a], exp[cc] - exp[a])) for a, b, cc in t)
    assert abs(shoelace - tri_area) < 1e-12


def test_two_bodies_are_concatenated(tmp_path):
    v, t = meshes.disk(10)
    mesh = tmp_path / "d.mesh"
    meshes.write_mesh(mesh, 2, v, tris=t)
    c, p = _params2d(tmp_path, [{"mesh_path": str(mesh), "translate": [0.2, 0.15], "scale": [0.1, 0.1]},
                                {"mesh_path": str(mesh), "translate": [0.45, 0.15], "scale": [0.05, 0.05]}])
    sim = fs.Simulator2D(p)
    sim.load_solids()
    assert sim.getVertices().shape == (22, 2)
    idx = list(sim.getBoundaryPointIdx())
    assert idx == list(range(1, 11)) + list(range(12, 22))
    be = np.array(sim.getBoundaryElements()).reshape(-1, 2)
    assert len(be) == 20 and be[:10].max() == 9 and be[10:].min() == 10
    assert list(sim.getTriangles())[len(t) * 3:] == [k + 11 for tri in t for k in tri]


def _signed_volume(v, tris):
    return sum(np.dot(v[a], np.cross(v[b], v[c])) for a, b, c in tris) / 6.0


def test_cube_surface_and_transform3d(tmp_path):
    from scipy.spatial.transform import Rotation

    corners, tets = meshes.kuhn_cube()
    mesh = tmp_path / "cube.mesh"
    meshes.write_mesh(mesh, 3, corners, tets=tets)
    path = meshes.write_config(tmp_path / "c3.json", fluid_nx=32, fluid_ny=32, fluid_nz=32,
                               solids=[{"mesh_path": str(mesh), "translate": [1.0, 2.0, 3.0],
                                        "scale": [2.0, 1.0, 0.5], "rotate": [10.0, 20.0, 30.0]}])
    c = fs.Config3D()
    assert c.load(path)
    sim = fs.Simulator3D(c.get_params())
    sim.load_solids()
    got = sim.getVertices()
    # eulerAngleYXZ(yaw=rotate[1], pitch=rotate[0], roll=rotate[2]) = Ry Rx Rz
    R = Rotation.from_euler("YXZ", [20.0, 10.0, 30.0], degrees=True).as_matrix()
    exp = (R @ (corners * [2.0, 1.0, 0.5]).T).T + [1.0, 2.0, 3.0]
    np.testing.assert_allclose(got, exp, rtol=0, atol=2e-6)
    tris = np.array(sim.getBoundaryElements()).reshape(-1, 3)
    assert len(tris) == 12 and list(sim.getBoundaryPointIdx()) == list(range(8))
    directed = {}
    for a, b, cc in tris:
        for e in ((a, b), (b, cc), (cc, a)):
            directed[e] = directed.get(e, 0) + 1
    assert all(k == 1 for k in directed.values())
    assert all((b, a) in directed for a, b in directed)  # closed, consistently oriented
    assert abs(_signed_volume(corners, tris) - 1.0) < 1e-12  # outward normals
    assert list(sim.getTetrahedra()) == [k for tt in tets for k in tt]


def test_unsupported_solver_paths(tmp_path):
    c, p = _params2d(tmp_path, [])
    sim = fs.Simulator2D(p)
    with pytest.raises(NotImplementedError):
        sim.apply_lbs_control(0, [], [])
    with pytest.raises(NotImplementedError):
        sim.apply_active_strain(0, 0, np.zeros(2), 1.0)
    assert list(sim.getControlPointIdx()) == []


# ------------------------------------------------------------------ GPU
Mover = meshes.Mover


@pytest.mark.gpu
def test_dropin_vortex_street_with_kinematic_disk(tmp_path):
    nx, ny, dx, dt, nu, rho = 160, 64, 0.005, 0.0025, 0.01, 1000.0
    note(f"case dropin 2D vortex street {nx}x{ny}")
    v, t = meshes.disk(48)
    mesh = tmp_path / "disk.mesh"
    meshes.write_mesh(mesh, 2, v, tris=t)
    path = meshes.write_config(
        tmp_path / "vs.json", dimension=2, fluid_viscosity=nu, fluid_density=rho, fluid_nx=nx, fluid_ny=ny,
        fluid_dx=dx, dt=dt, solid_solver_type="static", output_frequency=0,
        boundary_velocities=[0.05, 0.0, 0.0, 0.0], boundaries=meshes.channel_boundaries(nx, ny),
        solids=[{"mesh_path": str(mesh), "translate": [0.2, 0.16], "scale": [0.05, 0.05]}])
    cfg = fs.Config2D()
    assert cfg.load(path)
    sim = fs.Simulator2D(cfg.get_params())
    sim.initialize()
    targets = [[0.26, 0.18], [0.26, 0.13], [0.2, 0.16]]
    sim.apply_kinematic_control(0, np.array(targets), 0.1)

    flags = wl.vortex_street_flags(nx, ny)
    o = ol.OracleDomain(2, nx, ny, fp64=False, stage_fp16=True, nu=nu, dx=dx, dt=dt, cf=capi.cf_of(2, rho, dx, dt))
    o.set_flags(flags)
    o.init_reference([0.05, 0.0, 0.0, 0.0])
    mover = Mover(sim.getVertices(), targets, 0.1, dt)
    np.testing.assert_array_equal(mover.pos, _transform2(v, (0.05, 0.05), 0.0, (0.2, 0.16)))
    surf = np.array(sim.getBoundaryPointIdx())
    elems = np.array(sim.getBoundaryElements()).reshape(-1, 2)
    o.set_boundary(mover.pos[surf].astype(np.float32), np.zeros((len(surf), 2), np.float32), elems)
    for s in range(60):
        if s == 30:
            cells = np.array(sim.fillSolid())
            ref = flags.copy()
            ol.lib().orc_fill_solid2d(nx, ny, dx, len(surf), np.ascontiguousarray(mover.pos[surf], np.float32).ravel(),
                                      len(elems), np.ascontiguousarray(elems, np.uint32).ravel(), ref)
            flags = ref
            o.set_flags(flags)
            assert len(cells) == int((ref == 4).sum()) > 10
            assert np.all(ref.reshape(ny, nx)[cells[:, 1], cells[:, 0]] == 4)
        if s == 59:
            o.zero_vertex_forces()
        sim.step()
        o.step(1)
        mover.step()
        o.update_boundary(mover.pos[surf].astype(np.float32), mover.vel[surf].astype(np.float32))
        np.testing.assert_array_equal(sim.getVertices(), mover.pos)
    m = sim.get_fluid_moments()
    assert m.shape == (ny, nx, 6)
    a, b = m.reshape(-1, 6).astype(np.float64), o.moments(0)
    assert np.array_equal(a, b), f"rel-L2 {np.linalg.norm(a - b) / np.linalg.norm(b):.3e}"
    np.testing.assert_array_equal(sim.getVelocity(), mover.vel)
    f = sim.getForce()
    fo = o.vertex_forces()
    assert f.shape == (49, 2) and np.all(f[0] == 0)
    assert np.abs(fo).max() > 0
    assert np.linalg.norm(f[surf] - fo) / np.linalg.norm(fo) < 1e-5
    w = sim.get_body_wrench()
    assert w.shape == (1, 3)
    assert np.allclose(w[0, :2], f.sum(0), rtol=0, atol=1e-5 * np.abs(f).sum())


@pytest.mark.gpu
def test_dropin_3d_moving_cube(tmp_path):
    nx, ny, nz, dx, dt, nu, rho = 40, 24, 24, 0.01, 0.005, 0.02, 1000.0
    note(f"case dropin 3D cube {nx}x{ny}x{nz}")
    corners, tets = meshes.kuhn_cube()
    mesh = tmp_path / "cube.mesh"
    meshes.write_mesh(mesh, 3, corners, tets=tets)
    path = meshes.write_config(
        tmp_path / "c3.json", dimension=3, fluid_viscosity=nu, fluid_density=rho, fluid_nx=nx, fluid_ny=ny,
        fluid_nz=nz, fluid_dx=dx, dt=dt, solid_solver_type="static", boundary_velocities=[0.0] * 6,
        solids=[{"mesh_path": str(mesh), "translate": [0.12, 0.08, 0.08], "scale": [0.08, 0.07, 0.06],
                 "rotate": [0.0, 15.0, 5.0]}])
    cfg = fs.Config3D()
    assert cfg.load(path)
    sim = fs.Simulator3D(cfg.get_params())
    sim.initialize()
    targets = [[0.3, 0.1, 0.1], [0.12, 0.08, 0.08]]
    sim.apply_kinematic_control(0, np.array(targets), 0.2)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, stage_fp16=False, nu=nu, dx=dx, dt=dt,
                        cf=capi.cf_of(3, rho, dx, dt))
    o.set_flags(np.full(nx * ny * nz, 1, np.uint8))
    o.init_reference([0.0] * 6)
    mover = Mover(sim.getVertices(), targets, 0.2, dt)
    surf = np.array(sim.getBoundaryPointIdx())
    elems = np.array(sim.getBoundaryElements()).reshape(-1, 3)
    o.set_boundary(mover.pos[surf].astype(np.float32), np.zeros((len(surf), 3), np.float32), elems)
    for s in range(12):
        if s == 11:
            o.zero_vertex_forces()
        sim.step()
        o.step(1)
        mover.step()
        o.update_boundary(mover.pos[surf].astype(np.float32), mover.vel[surf].astype(np.float32))
    np.testing.assert_array_equal(sim.getVertices(), mover.pos)
    m = sim.get_fluid_moments()
    assert m.shape == (nz, ny, nx, 10)
    a, b = m.reshape(-1, 10).astype(np.float64), o.moments(0)
    assert np.abs(b[:, 1]).max() > 1e-4
    assert np.array_equal(a, b), f"rel-L2 {np.linalg.norm(a - b) / np.linalg.norm(b):.3e}"
    f = sim.getForce()
    fo = o.vertex_forces()
    assert np.linalg.norm(f[surf] - fo) / np.linalg.norm(fo) < 1e-5
