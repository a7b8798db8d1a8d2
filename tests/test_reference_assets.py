"""The reference repository's own inputs through the drop-in (tests/golden/assets,
copied unchanged by tests/golden/make_assets.sh; SHA256SUMS pins them).

  * assets/configs/vortexstreet2d.json: the configuration scripts/vortex_street_2d.py
    writes and runs (801x201, dx 0.005, dt 0.0025, 1602 Wall, 201 InletLeft,
    201 OutletRight cells, a static sphere2d.mesh body scaled 0.15 at (0.5, 0.5));
  * assets/mesh/sphere2d.mesh (a 100-gon fan), fish2d.mesh, sphere_coarse.mesh (3D).

CPU: the configuration loads with the reference's values, the boundary list
gives the vortex-street flag field, and each mesh's extracted boundary is a
closed, consistently oriented curve or surface enclosing the mesh's area or
volume.  GPU: the vortex street is driven as the script drives it
(initialize, apply_kinematic_control with the script's targets, step) and
compared every step with the oracle fed the same geometry; the fish and the
3D sphere run the same way in smaller domains.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import meshes
import oracle_lib as ol
from freeflow_amd import capi, workloads as wl
import freeflow_amd

fs = freeflow_amd.enable_dropin()
ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def asset(name):
    return os.path.join(ASSETS, name)


def vortexstreet_config(tmp_path):
    """The reference configuration with its author's absolute paths pointed
    at the fixtures (nothing else changed)."""
    with open(asset("vortexstreet2d.json")) as f:
        c = json.load(f)
    c["solids"][0]["mesh_path"] = asset("sphere2d.mesh")
    c["output_path"] = str(tmp_path / "out")
    c["output_frequency"] = 0
    p = tmp_path / "vortexstreet2d.json"
    with open(p, "w") as f:
        json.dump(c, f)
    return str(p), c


def test_fixtures_are_the_reference_files():
    with open(asset("SHA256SUMS")) as f:
        for line in f:
            digest, name = line.split()
            with open(asset(name), "rb") as g:
                assert hashlib.sha256(g.read()).hexdigest() == digest, name


def test_vortexstreet2d_config(tmp_path):
    path, raw = vortexstreet_config(tmp_path)
    cfg = fs.Config2D()
    assert cfg.load(path)
    p = cfg.get_params()
    assert (p.fluid_nx, p.fluid_ny, p.fluid_dx, p.dt, p.fluid_viscosity, p.fluid_density) == \
        (801, 201, 0.005, 0.0025, 0.01, 1000.0)
    assert list(p.boundary_velocities) == [np.float32(0.2), 0.0, 0.0, 0.0]  # float fields, as Config.hpp
    names = {5: "Wall", 13: "InletLeft", 20: "OutletRight"}
    assert [names[b.flag] for b in p.boundaries] == [b["type"] for b in raw["boundaries"]]
    # later entries overwrite earlier ones (LbmInitializer2D.cu:106-122): the
    # vortex-street field with walls on the bottom and top rows and the inlet
    # and outlet columns winning the corners
    f = np.full((p.fluid_ny, p.fluid_nx), 1, np.uint8)
    for b in p.boundaries:
        x, y = b.position
        f[y, x] = b.flag
    assert np.array_equal(f.ravel(), wl.vortex_street_flags(801, 201))
    s = p.solids[0]
    assert (list(s.translate), list(s.scale)) == ([0.5, 0.5], [0.15, 0.15])


def _loop_order(edges):
    """Vertices of a closed polyline given as undirected edges, in walk order."""
    adj = {}
    for a, b in edges:
        adj.setdefault(a, []).append(b)
        adj.setdefault(b, []).append(a)
    assert all(len(v) == 2 for v in adj.values()), "every boundary vertex has two boundary edges"
    start = edges[0][0]
    order, prev, cur = [start], None, start
    while True:
        nxt = adj[cur][0] if adj[cur][0] != prev else adj[cur][1]
        if nxt == start:
            break
        order.append(nxt)
        prev, cur = cur, nxt
    return order


@pytest.mark.parametrize("name", ["sphere2d.mesh", "fish2d.mesh"])
def test_2d_mesh_boundary_encloses_its_triangles(tmp_path, name):
    path = meshes.write_config(tmp_path / "m.json", fluid_nx=400, fluid_ny=200, fluid_dx=0.005, dt=0.0025,
                               output_frequency=0,
                               solids=[{"mesh_path": asset(name), "translate": [0.5, 0.5], "scale": [0.15, 0.15]}])
    cfg = fs.Config2D()
    assert cfg.load(path)
    sim = fs.Simulator2D(cfg.get_params())
    sim.load_solids()
    v = sim.getVertices()
    surf = np.array(sim.getBoundaryPointIdx())
    edges = np.array(sim.getBoundaryElements()).reshape(-1, 2)
    order = _loop_order([tuple(e) for e in edges])
    assert len(order) == len(surf) == len(edges)  # one closed loop through every boundary vertex
    ring = v[surf[order]]
    shoelace = 0.5 * abs(np.sum(ring[:, 0] * np.roll(ring[:, 1], -1) - np.roll(ring[:, 0], -1) * ring[:, 1]))
    tris = np.array(sim.getTriangles()).reshape(-1, 3)
    a, b, c = v[tris[:, 0]], v[tris[:, 1]], v[tris[:, 2]]
    area = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])).sum()
    note(f"{name}: {len(v)} vertices, {len(tris)} triangles, {len(edges)} boundary edges, area {area:.6f}")
    assert abs(shoelace - area) < 1e-9 * max(area, 1e-12)
    if name == "sphere2d.mesh":
        assert len(v) == 101 and len(edges) == 100


def test_3d_sphere_surface_encloses_its_tetrahedra(tmp_path):
    path = meshes.write_config(tmp_path / "s3.json", dimension=3, fluid_nx=48, fluid_ny=48, fluid_nz=48,
                               fluid_dx=0.01, dt=0.005, output_frequency=0,
                               solids=[{"mesh_path": asset("sphere_coarse.mesh"), "translate": [0.24, 0.24, 0.24],
                                        "scale": [0.1, 0.1, 0.1]}])
    cfg = fs.Config3D()
    assert cfg.load(path)
    sim = fs.Simulator3D(cfg.get_params())
    sim.load_solids()
    v = sim.getVertices()
    surf = np.array(sim.getBoundaryPointIdx())
    tris = np.array(sim.getBoundaryElements()).reshape(-1, 3)
    sv = v[surf]
    directed = {}
    for a, b, c in tris:
        for e in ((a, b), (b, c), (c, a)):
            directed[e] = directed.get(e, 0) + 1
    assert all(k == 1 for k in directed.values())
    assert all((b, a) in directed for a, b in directed)  # closed and consistently oriented
    vol_surf = sum(np.dot(sv[a], np.cross(sv[b], sv[c])) for a, b, c in tris) / 6.0
    tets = np.array(sim.getTetrahedra()).reshape(-1, 4)
    p0, p1, p2, p3 = (v[tets[:, k]] for k in range(4))
    signed = np.einsum("ij,ij->i", p1 - p0, np.cross(p2 - p0, p3 - p0)) / 6.0
    note(f"sphere_coarse.mesh: {len(v)} vertices, {len(tets)} tetrahedra, {len(tris)} surface triangles")
    # the file's tetrahedra all have one orientation (negative here), and the
    # extracted faces keep it: the surface's signed volume is the tetrahedra's
    assert np.all(np.sign(signed) == np.sign(signed[0]))
    assert abs(vol_surf - signed.sum()) < 1e-9 * abs(signed.sum())


def _drive_2d(sim, cfg_params, targets, vel, steps, flags, bc):
    p = cfg_params
    nx, ny = p.fluid_nx, p.fluid_ny
    o = ol.OracleDomain(2, nx, ny, fp64=False, stage_fp16=True, nu=p.fluid_viscosity, dx=p.fluid_dx, dt=p.dt,
                        cf=capi.cf_of(2, p.fluid_density, p.fluid_dx, p.dt))
    o.set_flags(flags)
    o.init_reference(bc)
    mover = meshes.Mover(sim.getVertices(), targets, vel, p.dt)
    surf = np.array(sim.getBoundaryPointIdx())
    elems = np.array(sim.getBoundaryElements()).reshape(-1, 2)
    o.set_boundary(mover.pos[surf].astype(np.float32), np.zeros((len(surf), 2), np.float32), elems)
    for s in range(steps):
        if s == steps - 1:
            o.zero_vertex_forces()
        sim.step()
        o.step(1)
        mover.step()
        o.update_boundary(mover.pos[surf].astype(np.float32), mover.vel[surf].astype(np.float32))
    np.testing.assert_array_equal(sim.getVertices(), mover.pos)
    m = sim.get_fluid_moments()
    a, b = m.reshape(-1, 6).astype(np.float64), o.moments(0)
    assert np.array_equal(a, b), f"rel-L2 {np.linalg.norm(a - b) / np.linalg.norm(b):.3e}"
    f = sim.getForce()
    fo = o.vertex_forces()
    assert np.abs(fo).max() > 0
    assert np.linalg.norm(f[surf] - fo) / np.linalg.norm(fo) < 1e-5


@pytest.mark.gpu
def test_vortexstreet2d_as_the_script_runs_it(tmp_path):
    path, raw = vortexstreet_config(tmp_path)
    cfg = fs.Config2D()
    assert cfg.load(path)
    p = cfg.get_params()
    note(f"case reference vortexstreet2d.json {p.fluid_nx}x{p.fluid_ny}")
    sim = fs.Simulator2D(p)
    sim.initialize()
    targets = [[0.75, 0.75], [1.0, 0.5], [0.75, 0.25], [0.5, 0.5]]  # scripts/vortex_street_2d.py:149-150
    sim.apply_kinematic_control(0, np.array(targets), 0.1)
    _drive_2d(sim, p, targets, 0.1, 40, wl.vortex_street_flags(p.fluid_nx, p.fluid_ny), [0.2, 0.0, 0.0, 0.0])


@pytest.mark.gpu
def test_fish2d_mesh_kinematic_swim(tmp_path):
    nx, ny = 400, 160
    note(f"case reference fish2d.mesh {nx}x{ny}")
    path = meshes.write_config(
        tmp_path / "fish.json", dimension=2, fluid_viscosity=0.005, fluid_density=1000.0, fluid_nx=nx, fluid_ny=ny,
        fluid_dx=0.005, dt=0.0025, solid_solver_type="static", output_frequency=0,
        boundary_velocities=[0.05, 0.0, 0.0, 0.0], boundaries=meshes.channel_boundaries(nx, ny),
        solids=[{"mesh_path": asset("fish2d.mesh"), "translate": [0.8, 0.4], "scale": [0.3, 0.3], "rotate": 10.0}])
    cfg = fs.Config2D()
    assert cfg.load(path)
    p = cfg.get_params()
    sim = fs.Simulator2D(p)
    sim.initialize()
    targets = [[0.6, 0.45], [0.8, 0.4]]
    sim.apply_kinematic_control(0, np.array(targets), 0.15)
    _drive_2d(sim, p, targets, 0.15, 40, wl.vortex_street_flags(nx, ny), [0.05, 0.0, 0.0, 0.0])


@pytest.mark.gpu
def test_sphere_coarse_mesh_3d(tmp_path):
    nx, ny, nz, dx, dt, nu, rho = 48, 32, 32, 0.01, 0.005, 0.02, 1000.0
    note(f"case reference sphere_coarse.mesh {nx}x{ny}x{nz}")
    path = meshes.write_config(
        tmp_path / "s3.json", dimension=3, fluid_viscosity=nu, fluid_density=rho, fluid_nx=nx, fluid_ny=ny,
        fluid_nz=nz, fluid_dx=dx, dt=dt, solid_solver_type="static", boundary_velocities=[0.0] * 6,
        output_frequency=0,
        solids=[{"mesh_path": asset("sphere_coarse.mesh"), "translate": [0.16, 0.16, 0.16], "scale": [0.07] * 3}])
    cfg = fs.Config3D()
    assert cfg.load(path)
    sim = fs.Simulator3D(cfg.get_params())
    sim.initialize()
    targets = [[0.3, 0.17, 0.15], [0.16, 0.16, 0.16]]
    sim.apply_kinematic_control(0, np.array(targets), 0.2)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, stage_fp16=False, nu=nu, dx=dx, dt=dt,
                        cf=capi.cf_of(3, rho, dx, dt))
    o.set_flags(np.full(nx * ny * nz, 1, np.uint8))
    o.init_reference([0.0] * 6)
    mover = meshes.Mover(sim.getVertices(), targets, 0.2, dt)
    surf = np.array(sim.getBoundaryPointIdx())
    elems = np.array(sim.getBoundaryElements()).reshape(-1, 3)
    o.set_boundary(mover.pos[surf].astype(np.float32), np.zeros((len(surf), 3), np.float32), elems)
    for s in range(15):
        if s == 14:
            o.zero_vertex_forces()
        sim.step()
        o.step(1)
        mover.step()
        o.update_boundary(mover.pos[surf].astype(np.float32), mover.vel[surf].astype(np.float32))
    a, b = sim.get_fluid_moments().reshape(-1, 10).astype(np.float64), o.moments(0)
    assert np.abs(b[:, 1]).max() > 1e-4
    assert np.array_equal(a, b), f"rel-L2 {np.linalg.norm(a - b) / np.linalg.norm(b):.3e}"
    f = sim.getForce()
    fo = o.vertex_forces()
    assert np.linalg.norm(f[surf] - fo) / np.linalg.norm(fo) < 1e-5
