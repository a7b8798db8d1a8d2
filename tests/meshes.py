"""Small Medit .mesh files and JSON configs for the drop-in tests.

The meshes are built here (a fan-triangulated disk, a Kuhn-triangulated
cube) so that their boundaries are known in closed form.
"""
import json
import math

import numpy as np


def write_mesh(path, dim, verts, edges=(), tris=(), tets=()):
    with open(path, "w") as f:
        f.write(f"MeshVersionFormatted 1\nDimension {dim}\n\nVertices\n{len(verts)}\n")
        for v in verts:
            f.write(" ".join(repr(float(x)) for x in v) + " 0\n")  # exact round trip
        for name, block in (("Edges", edges), ("Triangles", tris), ("Tetrahedra", tets)):
            if len(block):
                f.write(f"\n{name}\n{len(block)}\n")
                for e in block:
                    f.write(" ".join(str(int(k) + 1) for k in e) + " 0\n")
        f.write("\nEnd\n")


def disk(n, r=0.5):
    """Centre vertex 0 plus n ring vertices, fan triangles (0, i, i+1)."""
    a = 2 * np.pi * np.arange(n) / n
    verts = np.vstack([[0.0, 0.0], np.stack([r * np.cos(a), r * np.sin(a)], 1)])
    tris = [(0, 1 + i, 1 + (i + 1) % n) for i in range(n)]
    return verts, tris


def kuhn_cube():
    """Unit cube as 6 positively oriented tetrahedra (one per axis order)."""
    corners = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], float)
    idx = {tuple(c.astype(int)): i for i, c in enumerate(corners)}
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        p = np.zeros(3, int)
        t = [idx[tuple(p)]]
        for ax in perm:
            p = p.copy()
            p[ax] = 1
            t.append(idx[tuple(p)])
        v = corners[t]
        if np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])) < 0:
            t[1], t[2] = t[2], t[1]
        tets.append(t)
    return corners, tets


def channel_boundaries(nx, ny):
    """Walls on the bottom and top rows, then an inlet on the left column and an
    outlet on the right one, listed in that order (later entries win)."""
    b = []
    for x in range(nx):
        b.append({"type": "Wall", "pos": [x, 0]})
        b.append({"type": "Wall", "pos": [x, ny - 1]})
    for y in range(ny):
        b.append({"type": "InletLeft", "pos": [0, y]})
        b.append({"type": "OutletRight", "pos": [nx - 1, y]})
    return b


def write_config(path, **kw):
    with open(path, "w") as f:
        json.dump(kw, f, indent=1)
    return str(path)


class Mover:
    """StaticBodySolver2D/3D::advanceTimeStep (StaticBodySolver.cpp:39-68,
    122-151) restated: marker = volume vertex 0; switch target when closer
    than dt*vel; all vertices get vel*normalize(target-marker), glm's double
    arithmetic (length = sqrt(dot), normalize = d * (1/sqrt(dot)))."""

    def __init__(self, pos, targets, vel, dt):
        self.pos = np.array(pos, np.float64)
        self.vel = np.zeros_like(self.pos)
        self.targets = [np.array(t, np.float64) for t in targets]
        self.v, self.dt, self.idx = vel, dt, -1

    def step(self):
        m = self.pos[0].copy()
        if self.idx < 0:
            self.idx = 0

        def dot(d):
            s = d[0] * d[0] + d[1] * d[1]
            return s + d[2] * d[2] if len(d) == 3 else s

        if math.sqrt(dot(self.targets[self.idx] - m)) < self.dt * self.v:
            self.idx = (self.idx + 1) % len(self.targets)
        d = self.targets[self.idx] - m
        inv = 1.0 / math.sqrt(dot(d))
        vel = np.array([self.v * (d[k] * inv) for k in range(len(d))])
        shift = vel * self.dt
        self.vel[:] = vel
        self.pos += shift
