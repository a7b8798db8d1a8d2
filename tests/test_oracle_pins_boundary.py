"""Pins for the oracle's boundary paths (CPU-only, no GPU needed).

The reference ships no tests or golden vectors, so each boundary rule of the
oracle is pinned here by something fixed independently of its source text:
a brute-force search over every triangle for the 3D link key, the analytic
decay of a shear mode between no-slip walls and a rest state that ignores
arbitrary wall moments (3D wall bounce-back), Galilean invariance and a
closed-form first-step force (moving walls, 2D and 3D), the physical sense of
every outlet face (outflow copied, inflow kept, the face's own inward
neighbour) and the equilibrium closed form of the 3D initial state.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as ol
from freeflow_amd import workloads as wl

L = ol.lib()

EX2 = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
EY2 = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
EX3 = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1], float)
EY3 = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1], float)
EZ3 = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1], float)

FLUID, SOLID, WALL = 1, 3, 5
WALL_DOWN, WALL_UP = 10, 11
INLET = {"Left": 13, "Right": 14, "Up": 15, "Down": 16, "Front": 17, "Back": 18}
OUTLET = {"Left": 19, "Right": 20, "Up": 21, "Down": 22, "Front": 23, "Back": 24}


def lattice(dim):
    if dim == 2:
        c = np.stack([EX2, EY2], 1)
    else:
        c = np.stack([EX3, EY3, EZ3], 1)
    # D2Q9 / D3Q27 weights are products of the D1Q3 weights 2/3 and 1/6
    w = np.prod(np.where(c == 0, 2.0 / 3.0, 1.0 / 6.0), axis=1)
    inv = np.array([int(np.where((c == -ci).all(1))[0][0]) for ci in c])
    return c, w, inv


def feq_second_order(rho, u, c, w):
    """Second-order equilibrium populations w (rho + 3 c.u + 9/2 (c.u)^2 - 3/2 u.u)
    written from the lattice alone (no reconstruction code).  The moment-encoded
    reconstruction of an equilibrium state (Pi = u u) adds third-order Hermite
    terms, so the tests below use this only where those cancel."""
    cu = c @ u
    return w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u @ u)


def hermite_feq(rho, u, c, w):
    """Populations of the equilibrium moment state (rho, u, Pi = u u) as the
    tensor-product Hermite synthesis of its coefficients up to the orders the
    lattice carries (third-order terms a_{aab} = rho u_a u_a u_b, a_{xyz} =
    rho u_x u_y u_z: the recursive form with Pi = u u)."""
    D = c.shape[1]
    cs2 = 1.0 / 3.0
    N = [1.0, cs2, 2.0 / 9.0]

    def h1(x, k):
        return [np.ones_like(x), x, x * x - cs2][k]

    f = np.zeros(len(c))
    for orders in itertools.product(range(3), repeat=D):
        # a_{k1..kD} = rho prod_d u_d^{k_d} for an equilibrium; the scheme
        # carries total order <= 3 with at most one axis at order 2
        if sum(1 for k in orders if k == 2) >= 2 or sum(orders) > 3:
            continue
        coef = rho * np.prod([u[d] ** k for d, k in enumerate(orders)])
        term = coef * np.ones(len(c))
        for d, k in enumerate(orders):
            term = term * h1(c[:, d], k) / N[k]
        f += term
    return w * f


def eq_moments(u):
    n, D = u.shape
    rho = np.ones((n, 1))
    if D == 2:
        pi = np.stack([u[:, 0] ** 2, u[:, 1] ** 2, u[:, 0] * u[:, 1]], 1)
    else:
        pi = np.stack([u[:, 0] ** 2, u[:, 1] ** 2, u[:, 2] ** 2, u[:, 0] * u[:, 1], u[:, 1] * u[:, 2],
                       u[:, 0] * u[:, 2]], 1)
    return np.column_stack([rho, u, pi])


def test_hermite_feq_helper_matches_oracle_reconstruction():
    # the helper is the independent statement used below; its agreement with the
    # oracle's reconstruction of an equilibrium state is itself a pin (the third
    # order terms of an equilibrium reduce to rho u u u)
    rng = np.random.default_rng(40)
    for dim in (2, 3):
        c, w, _ = lattice(dim)
        for _ in range(10):
            u = rng.uniform(-0.08, 0.08, dim)
            m = eq_moments(u[None, :])[0]
            f = np.empty(len(c))
            if dim == 2:
                L.orc_recon2d_all_d(np.ascontiguousarray(m), f)
            else:
                L.orc_recon3d_all_d(np.ascontiguousarray(m), f)
            assert np.max(np.abs(f - hermite_feq(1.0, u, c, w))) < 1e-15


# ---------------------------------------------------------------- 3D link key
def test_hit_table3d_equals_bruteforce_min_over_all_triangles():
    # FsiCoupling3D.cu:31-81: for every cell and direction the key is the
    # minimum of depthIndexCompact(depth, tri) over the triangles the link hits.
    # Brute force over every cell, direction and triangle; a conservative numpy
    # box test (independent of the oracle's bounding-box enumeration) skips
    # only pairs that cannot touch.
    nx, ny, nz, dx = 11, 10, 9, 0.1
    verts, tris = wl.icosphere((0.52, 0.47, 0.41), 0.27, subdiv=1)
    verts = verts.astype(np.float32)
    dom = ol.OracleDomain(3, nx, ny, nz, dx=dx)
    dom.set_boundary(verts, np.zeros_like(verts), tris)
    table = dom.hit_table()
    c, _, _ = lattice(3)
    cells = np.array([(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)], np.float32)
    p1 = cells * np.float32(dx)
    tri_v = verts[tris]  # (nt, 3, 3)
    tmin = tri_v.min(1) - 1e-3
    tmax = tri_v.max(1) + 1e-3
    out = np.zeros(7, np.float32)
    nhit = 0
    expect = np.full((len(cells), 27), 0xFFFFFFFF, np.uint64)
    for l in range(1, 27):
        p2 = (cells - c[l].astype(np.float32)) * np.float32(dx)
        lo = np.minimum(p1, p2)
        hi = np.maximum(p1, p2)
        cand = np.all((lo[:, None, :] <= tmax[None]) & (hi[:, None, :] >= tmin[None]), axis=2)
        for ci, ti in zip(*np.nonzero(cand)):
            r = L.orc_intersect3(np.ascontiguousarray(p1[ci]), np.ascontiguousarray(p2[ci]),
                                 np.ascontiguousarray(tri_v[ti].ravel()), out)
            if r == 0:
                key = L.orc_depth_index_compact(float(out[0]), int(ti))
                expect[ci, l] = min(expect[ci, l], key)
    nhit = int(np.sum(expect != 0xFFFFFFFF))
    assert nhit > 200
    assert np.array_equal(table.astype(np.uint64), expect)


# ---------------------------------------------------------------- 3D walls
WALLS3 = {"Solid": 3, "Wall": 5, "WallLeft": 6, "WallRight": 7, "WallFront": 8, "WallBack": 9, "WallDown": 10,
          "WallUp": 11}


@pytest.mark.parametrize("kind", list(WALLS3))
def test_rest_between_walls_3d_ignores_wall_moments(kind):
    # FsiCoupling3D.cu:208-222: a link whose upstream cell carries one of the
    # eight wall flags takes the cell's own population in the opposite
    # direction, never the wall's.  Rest fluid enclosed by walls of that kind,
    # whose moments are arbitrary, stays exactly at rest.
    nx, ny, nz = 7, 8, 6
    dom = ol.OracleDomain(3, nx, ny, nz, fp64=True, periodic=True, nu=0.03)
    flags = np.full((nz, ny, nx), FLUID, np.uint8)
    w = WALLS3[kind]
    flags[:, 0, :] = w
    flags[:, -1, :] = w
    flags[0, :, :] = w
    flags[-1, :, :] = w
    flags[2:4, 3:5, 2:4] = w
    dom.set_flags(flags.ravel())
    rng = np.random.default_rng(41)
    m = eq_moments(np.zeros((dom.n, 3)))
    wall = flags.ravel() != FLUID
    m[wall] = np.column_stack([1 + rng.uniform(0, 3, wall.sum()), rng.uniform(-0.3, 0.3, (wall.sum(), 9))])
    dom.set_moments(m)
    dom.step(8)
    r = dom.moments()
    fl = ~wall
    assert np.max(np.abs(r[fl] - m[fl])) < 1e-14
    assert np.array_equal(r[wall], m[wall])  # wall cells are never written


def test_wall_corner_is_pulled_from_in_3d():
    # WallCorner (12) is not among the bounce-back flags of FsiCoupling3D.cu:209-216:
    # its (frozen) moments are pulled like a fluid neighbour's, so a moving
    # corner state leaks into the rest fluid next to it.
    nx, ny, nz = 6, 6, 6
    dom = ol.OracleDomain(3, nx, ny, nz, fp64=True, periodic=True, nu=0.03)
    flags = np.full(dom.n, FLUID, np.uint8)
    flags[0] = 12
    dom.set_flags(flags)
    m = eq_moments(np.zeros((dom.n, 3)))
    m[0] = eq_moments(np.array([[0.05, 0.0, 0.0]]))[0]
    dom.set_moments(m)
    dom.step(1)
    r = dom.moments()
    assert r[1, 1] > 1e-4  # +x neighbour of the corner cell picks up its momentum
    assert np.array_equal(r[0], m[0])


def test_channel_shear_mode_decays_at_viscous_rate_3d():
    # No-slip walls at y = 0 and y = ny-1 (halfway bounce-back: the wall sits
    # half a cell outside the first fluid row).  u_x(y) = A sin(pi (y - 1/2) / H)
    # is the slowest shear mode of the channel of width H and decays as
    # exp(-nu (pi/H)^2 t).  A bounce-back of the wrong population, a wall that
    # is pulled from or a wall at the wrong place changes the rate by O(1).
    H = 20
    nx, ny, nz = 4, H + 2, 3
    nu = 0.04
    dom = ol.OracleDomain(3, nx, ny, nz, fp64=True, periodic=True, nu=nu)
    flags = np.full((nz, ny, nx), FLUID, np.uint8)
    flags[:, 0, :] = WALL
    flags[:, -1, :] = WALL
    dom.set_flags(flags.ravel())
    ys = np.tile(np.repeat(np.arange(ny), nx), nz)
    A = 1e-4
    k = np.pi / H
    u = np.zeros((dom.n, 3))
    fluid = flags.ravel() == FLUID
    u[fluid, 0] = A * np.sin(k * (ys[fluid] - 0.5))
    m = eq_moments(u)
    m[~fluid] = eq_moments(np.tile([0.2, -0.1, 0.05], ((~fluid).sum(), 1)))  # wall moments are never used
    dom.set_moments(m)
    basis = np.where(fluid, np.sin(k * (ys - 0.5)), 0.0)
    dom.step(100)  # non-equilibrium stress settles onto the mode
    a1 = np.sum(dom.moments()[:, 1] * basis)
    T = 600
    dom.step(T)
    a2 = np.sum(dom.moments()[:, 1] * basis)
    omega = 1 / (3 * nu + 0.5)
    nu_eff = (1 / omega - 0.5) / 3
    expect = np.exp(-nu_eff * k * k * T)
    assert abs(a2 / a1 / expect - 1) < 1e-2, (a2 / a1, expect)
    # the same mode with periodic y (no walls) decays with (2 pi / ny)^2: walls matter
    assert abs(np.exp(-nu_eff * (2 * np.pi / ny) ** 2 * T) / expect - 1) > 0.3


# ---------------------------------------------------------------- moving walls
def _links(table):
    return [(cell, l) for cell, l in zip(*np.nonzero(table != 0xFFFFFFFF))]


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_flow_with_comoving_body_is_fixed_point(dim):
    # Galilean consistency of the moving-wall rule (FsiCoupling2D.cu:187-214,
    # FsiCoupling3D.cu:226-275): a body moving with a uniform flow disturbs
    # nothing.  u_w = barycentric vertex velocity * ct / cl = u0, Pi* = Pi + u0 u0
    # - u0 u0 = Pi, so every moving-wall population equals the pulled one.  A
    # wrong velocity scale (cl/ct), a wrong barycentric weight or Pi* sign moves
    # the field by O(u0).
    n, dxp, dtp = (40, 0.01, 0.002) if dim == 2 else (16, 0.01, 0.002)
    u0 = np.array([0.04, -0.015, 0.02][:dim])
    vel_phys = (u0 * dxp / dtp).astype(np.float32)
    if dim == 2:
        th = np.pi / 50 + 2 * np.pi * np.arange(50) / 50
        v = np.stack([0.2 + 0.09 * np.cos(th), 0.2 + 0.06 * np.sin(th)], 1).astype(np.float32)
        e = np.stack([np.arange(50), (np.arange(50) + 1) % 50], 1).astype(np.uint32)
        dom = ol.OracleDomain(2, n, n, fp64=True, stage_fp16=False, periodic=True, nu=0.01, dx=dxp, dt=dtp, cf=1.0)
    else:
        v, e = wl.icosphere((0.081, 0.077, 0.083), 0.043, subdiv=1)
        v = v.astype(np.float32)
        dom = ol.OracleDomain(3, n, n, n, fp64=True, periodic=True, nu=0.01, dx=dxp, dt=dtp, cf=1.0)
    m = eq_moments(np.tile(u0, (dom.n, 1)))
    dom.set_moments(m)
    dom.set_boundary(v, np.tile(vel_phys, (len(v), 1)), e)
    assert len(_links(dom.hit_table())) > 50
    dom.step(3)
    assert np.max(np.abs(dom.moments() - m)) < 2e-7


@pytest.mark.parametrize("dim", [2, 3])
def test_first_step_force_closed_form(dim):
    # One step from a uniform equilibrium flow u0 with the body moving at u_w:
    # every link (cell, i) of the hit table carries
    #   F = [f_inv(i)(-c_i - u_w) - f_i^w (c_i - u_w)] cf
    # where f_inv(i) = f^eq_inv(i)(u0) is the cell's own population and
    # f_i^w = f_i(rho, u_w, Pi + u_w u_w - u0 u0) the moving-wall population
    # (an equilibrium at u_w).  Their sum equals the sum of the per-vertex
    # forces (the barycentric weights of a link sum to one).
    n = 36 if dim == 2 else 14
    c, w, inv = lattice(dim)
    u0 = np.array([0.05, 0.01, -0.02][:dim])
    uw = np.array([-0.02, 0.03, 0.01][:dim])
    if dim == 2:
        th = np.pi / 64 + 2 * np.pi * np.arange(64) / 64
        v = np.stack([17.3 + 7.3 * np.cos(th), 18.1 + 5.1 * np.sin(th)], 1).astype(np.float32)
        e = np.stack([np.arange(64), (np.arange(64) + 1) % 64], 1).astype(np.uint32)
        dom = ol.OracleDomain(2, n, n, fp64=True, stage_fp16=False, periodic=True, dx=1.0, dt=1.0, cf=1.0)
    else:
        v, e = wl.icosphere((6.91, 7.13, 6.77), 3.6, subdiv=1)
        v = v.astype(np.float32)
        dom = ol.OracleDomain(3, n, n, n, fp64=True, periodic=True, dx=1.0, dt=1.0, cf=1.0)
    m = eq_moments(np.tile(u0, (dom.n, 1)))
    dom.set_moments(m)
    dom.set_boundary(v, np.tile(uw.astype(np.float32), (len(v), 1)), e)
    links = _links(dom.hit_table())
    assert len(links) > 50
    dom.step(1)
    F = dom.vertex_forces().sum(0)
    f0 = hermite_feq(1.0, u0, c, w)
    # moving-wall population: reconstruction at (rho, u_w, Pi*), Pi* = u0 u0 + u_w u_w - u0 u0
    fw = hermite_feq(1.0, uw.astype(np.float32).astype(float), c, w)
    uwl = uw.astype(np.float32).astype(float)
    expect = np.zeros(dim)
    for _, l in links:
        expect += f0[inv[l]] * (-c[l] - uwl) - fw[l] * (c[l] - uwl)
    assert np.abs(expect).max() > 1e-3
    assert np.allclose(F, expect, rtol=2e-6, atol=1e-8), (F, expect)


# ---------------------------------------------------------------- outlets
# face -> (axis, inward neighbour offset along the axis, outflow sign), from the
# physical sense of an outlet: a face at the high end of an axis lets fluid
# out when it moves towards +axis, and reads the neighbour just inside.
FACES2 = {"Left": (0, +1, -1), "Right": (0, -1, +1), "Down": (1, +1, -1), "Up": (1, -1, +1)}
FACES3 = {"Left": (0, +1, -1), "Right": (0, -1, +1), "Front": (1, +1, -1), "Back": (1, -1, +1),
          "Down": (2, +1, -1), "Up": (2, -1, +1)}


def _outlet_case(dim, face, outflow, nb_flag=FLUID):
    axis, nb_off, out_sign = (FACES2 if dim == 2 else FACES3)[face]
    shape = [3] * dim  # x, y(, z)
    shape[axis] = 5
    rng = np.random.default_rng(100 * dim + 2 * list(FACES3).index(face) + int(outflow))
    flags = np.full(shape[::-1], FLUID, np.uint8)  # indexed [z,] y, x
    pos = [1] * dim
    pos[axis] = 0 if nb_off > 0 else shape[axis] - 1
    opos = tuple(pos[::-1])
    flags[opos] = OUTLET[face]
    npos = list(pos)
    npos[axis] += nb_off
    npos = tuple(npos[::-1])
    flags[npos] = nb_flag
    if dim == 2:
        dom = ol.OracleDomain(2, shape[0], shape[1], fp64=True, stage_fp16=False, periodic=True, nu=0.02)
    else:
        dom = ol.OracleDomain(3, shape[0], shape[1], shape[2], fp64=True, periodic=True, nu=0.02)
    dom.set_flags(flags.ravel())
    # a strong flow along the axis (out of or into the face) plus distinct noise
    # everywhere, so a wrong neighbour or component shows
    u = rng.uniform(-0.01, 0.01, (dom.n, dim))
    u[:, axis] += 0.06 * (out_sign if outflow else -out_sign)
    m = eq_moments(u)
    m[:, 0] += rng.uniform(-1e-3, 1e-3, dom.n)
    m[:, dim + 1:] += rng.uniform(-1e-4, 1e-4, (dom.n, m.shape[1] - dim - 1))
    dom.set_moments(m)
    dom.step(1)
    r = dom.moments()
    flat = lambda p: int(np.ravel_multi_index(p, flags.shape))
    return r, m, flat(opos), flat(npos), 1 + axis, 1 + dim + axis, out_sign


@pytest.mark.parametrize("dim,face", [(2, f) for f in FACES2] + [(3, f) for f in FACES3])
@pytest.mark.parametrize("outflow", [True, False])
def test_outlet_face(dim, face, outflow):
    # FsiCoupling2D.cu:268-314, FsiCoupling3D.cu:348-418, in place on the new
    # buffer: the face's velocity component is copied from the inward fluid
    # neighbour only when that flows out of the face; the matching diagonal Pi
    # becomes the square of the neighbour's component in both cases; nothing
    # else of the outlet cell changes.
    r, m, o, nb, comp, pcomp, out_sign = _outlet_case(dim, face, outflow)
    un = r[nb, comp]
    assert np.sign(un) == (out_sign if outflow else -out_sign)
    if outflow:
        assert r[o, comp] == un
    else:
        assert r[o, comp] == m[o, comp]
    assert r[o, pcomp] == un * un
    rest = [k for k in range(m.shape[1]) if k not in (comp, pcomp)]
    assert np.array_equal(r[o, rest], m[o, rest])


@pytest.mark.parametrize("dim,face", [(2, "Up"), (3, "Front"), (3, "Down")])
def test_outlet_needs_a_fluid_neighbour(dim, face):
    # the rule acts only when the inward neighbour is Fluid
    r, m, o, nb, comp, pcomp, _ = _outlet_case(dim, face, True, nb_flag=WALL)
    assert np.array_equal(r[o], m[o])


# ---------------------------------------------------------------- 3D initial state
def test_init_reference_3d_closed_form():
    # LbmInitializer3D.cu:33-113: rho = 1; an inlet cell flows into the domain
    # through its face (Left: +x from bc[0], Right: -x from bc[2], Down: +z from
    # bc[1], Up: -z from bc[3], Front: +y from bc[4], Back: -y from bc[5]); Pi is
    # the second moment of the second-order equilibrium, which is u u in closed
    # form; both ping-pong buffers hold it.
    nx, ny, nz = 5, 4, 3
    dom = ol.OracleDomain(3, nx, ny, nz, fp64=False)
    names = ["Left", "Right", "Up", "Down", "Front", "Back"]
    flags = np.full(dom.n, FLUID, np.uint8)
    for k, nme in enumerate(names):
        flags[3 * k:3 * k + 2] = INLET[nme]
    flags[-3:] = WALL
    flags[-5] = OUTLET["Up"]
    dom.set_flags(flags)
    bc = np.array([0.011, 0.023, 0.037, 0.041, 0.053, 0.067], np.float32)
    dom.init_reference(bc)
    m = dom.moments()
    into = {"Left": (0, +1, 0), "Right": (0, -1, 2), "Down": (2, +1, 1), "Up": (2, -1, 3),
            "Front": (1, +1, 4), "Back": (1, -1, 5)}
    u = np.zeros((dom.n, 3))
    for nme, (axis, sgn, slot) in into.items():
        u[flags == INLET[nme], axis] = sgn * float(bc[slot])
    assert np.all(m[:, 0] == 1.0)
    assert np.array_equal(m[:, 1:4], u.astype(np.float32).astype(float))
    pi = eq_moments(u)[:, 4:]
    assert np.allclose(m[:, 4:], pi, atol=1e-7, rtol=0)  # FP32 rounding of f_eq sums (1 ulp of 1/3)
    assert np.array_equal(dom.moments(1), m)
    # the closed form Pi = u u holds for the second-order equilibrium on D3Q27
    c, w, _ = lattice(3)
    for uu in u[:18:3]:
        f = feq_second_order(1.0, uu, c, w)
        P = (f[:, None, None] * c[:, :, None] * c[:, None, :]).sum(0) - np.eye(3) / 3
        assert np.allclose(P, np.outer(uu, uu), atol=1e-15)
    # a wrong count of boundary velocities means zeros, as in the reference
    dom2 = ol.OracleDomain(3, nx, ny, nz, fp64=False)
    dom2.set_flags(flags)
    dom2.init_reference(bc[:4])
    assert np.all(dom2.moments()[:, 1:4] == 0)
