"""Synthetic inputs for the fluid-step workloads (flags, bodies, initial fields).

Layouts follow the reference's own driver scripts:
  * vortex street: walls on the bottom/top rows, InletLeft column, OutletRight
    column, later entries overwriting earlier ones (scripts/vortex_street_2d.py:105-119,
    applied in list order by LbmInitializer2D.cu:106-122);
  * swimmer box: Outlet* on all four sides (scripts/fish2d_lbs.py:85-91);
  * physical parameters of the 2D swimmer (scripts/fish2d_lbs.py:30-37).
All randomness comes from numpy's PCG64 seeded generator.
"""
import numpy as np

FLUID, WALL = 1, 5
INLET_LEFT, OUTLET_LEFT, OUTLET_RIGHT, OUTLET_UP, OUTLET_DOWN = 13, 19, 20, 21, 22


def vortex_street_flags(nx, ny):
    f = np.full((ny, nx), FLUID, np.uint8)
    f[0, :] = WALL
    f[ny - 1, :] = WALL
    f[:, 0] = INLET_LEFT
    f[:, nx - 1] = OUTLET_RIGHT
    return f.ravel()


def outlet_box_flags(nx, ny):
    f = np.full((ny, nx), FLUID, np.uint8)
    f[0, :] = OUTLET_DOWN
    f[ny - 1, :] = OUTLET_UP
    f[:, 0] = OUTLET_LEFT
    f[:, nx - 1] = OUTLET_RIGHT
    return f.ravel()


def equilibrium_moments(u, rho=None):
    """Stored-order moments of the equilibrium with velocity u (n, D):
    rho, u, Pi = u u (the normalised second moment of f_eq)."""
    n, D = u.shape
    r = np.ones(n) if rho is None else rho
    if D == 2:
        pi = np.stack([u[:, 0] ** 2, u[:, 1] ** 2, u[:, 0] * u[:, 1]], 1)
    else:
        pi = np.stack([u[:, 0] ** 2, u[:, 1] ** 2, u[:, 2] ** 2,
                       u[:, 0] * u[:, 1], u[:, 1] * u[:, 2], u[:, 0] * u[:, 2]], 1)
    return np.column_stack([r, u, pi])


def noisy_inflow_moments(flags, nx, ny, u_in, amp, seed):
    """Equilibrium at rho=1, u = inlet velocity on InletLeft cells, plus
    uniform +-amp noise per component (seeded), Pi = u u."""
    rng = np.random.default_rng(seed)
    u = np.zeros((nx * ny, 2))
    u[flags == INLET_LEFT, 0] = u_in
    u += rng.uniform(-amp, amp, u.shape)
    return equilibrium_moments(u)


def polygon(n, cx, cy, a, b, phase=0.1234567):
    """Closed polygon on the ellipse (cx + a cos t, cy + b sin t); the phase
    offset keeps segments away from exactly vertical (the reference's
    segment test ignores exactly vertical segments)."""
    t = phase + 2 * np.pi * np.arange(n) / n
    v = np.stack([cx + a * np.cos(t), cy + b * np.sin(t)], 1)
    e = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.uint32)
    return v, e


def cylinder(n, cx, cy, r):
    return polygon(n, cx, cy, r, r)


class Swimmer2D:
    """Elliptical swimmer boundary with a prescribed lateral travelling wave
    y = y_rest + A s sin(2 pi (s L/lambda - f t) + phi), s in [0,1] from head
    to tail along the body axis (x)."""

    def __init__(self, n_seg, cx, cy, a, b, A, f, phi, l_over_lambda=1.0):
        self.rest, self.elems = polygon(n_seg, cx, cy, a, b)
        self.s = (self.rest[:, 0] - (cx - a)) / (2 * a)
        self.params = np.array([A, f, phi, l_over_lambda], np.float64)

    def state(self, t):
        th = 2 * np.pi * (self.s * self.params[3] - self.params[1] * t) + self.params[2]
        v = self.rest.copy()
        v[:, 1] = self.rest[:, 1] + self.params[0] * self.s * np.sin(th)
        vel = np.zeros_like(v)
        vel[:, 1] = -2 * np.pi * self.params[1] * self.params[0] * self.s * np.cos(th)
        return v, vel


def random_swimmer(rng, n_seg, cx, cy, a=0.3, b=0.08):
    A = rng.uniform(0.02, 0.05)
    f = rng.uniform(1.0, 3.0)
    phi = rng.uniform(0, 2 * np.pi)
    return Swimmer2D(n_seg, cx, cy, a, b, A, f, phi)


def icosphere(center, r, subdiv=1):
    """Triangulated sphere (icosahedron refined `subdiv` times), outward faces."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    v = list(v / np.linalg.norm(v, axis=1)[:, None])
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = (v[a] + v[b]) / 2
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(v) * r + np.asarray(center), np.array(f, np.uint32)


class SwimmerScene2D:
    """One 2D swimmer domain: outlet box, equilibrium at rest, elliptical
    swimmer with travelling-wave kinematics (the layout of the 1200x400 swimmer workload scaled to
    the domain)."""

    def __init__(self, nx, ny, n_seg, rng, dx=0.005, dt=0.0025, viscosity=0.005, rho=10.0):
        self.nx, self.ny, self.dx, self.dt, self.viscosity, self.rho = nx, ny, dx, dt, viscosity, rho
        lx, ly = nx * dx, ny * dx
        a = min(0.3, 0.22 * lx)
        b = min(0.27 * a, 0.12 * ly)
        self.swimmer = random_swimmer(rng, n_seg, 0.45 * lx, 0.5 * ly, a, b)
        self.swimmer.params[0] = min(self.swimmer.params[0], 0.2 * ly)
        self.flags = outlet_box_flags(nx, ny)

    def domain_args(self):
        return dict(dim=2, nx=self.nx, ny=self.ny, viscosity=self.viscosity, dx=self.dx, dt=self.dt,
                    fluid_density=self.rho)

    def setup(self, dom, history=0):
        """Flags, rest state, boundary and device-side wave kinematics on `dom`."""
        dom.set_flags(self.flags)
        dom.init_state([0.0, 0.0, 0.0, 0.0])  # rest fluid, computed on the device
        v0, vel0 = self.swimmer.state(0.0)
        dom.set_boundary(v0, vel0, self.swimmer.elems)
        dom.set_wave(self.swimmer.rest, self.swimmer.s, self.swimmer.params[None, :])
        if history:
            dom.set_wrench_history(history)


def batched_scenes(n, seed, nx_range=(256, 1200), ny_range=(128, 400), seg_range=(200, 800)):
    """`n` uneven swimmer domains, sizes and swimmers from one seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        ns = int(rng.integers(seg_range[0], seg_range[1] + 1))
        out.append(SwimmerScene2D(nx, ny, ns, rng))
    return out


# ---------------------------------------------------------------- configurations
CONFIG1 = dict(nx=400, ny=200, viscosity=0.005, u_in=0.05, r=20.0, cx=100.0, cy=100.0, n_poly=128, seed=1234)
SWIMMER2D = dict(nx=1200, ny=400, dx=0.005, dt=0.0025, viscosity=0.005, rho=10.0, n_seg=600,
               cx=1.0, cy=1.0, a=0.3, b=0.08, seed=1234)
TG3D = dict(nx=256, ny=128, nz=128, viscosity=0.01, amp=0.04, ring_r=20.0, seed=1234)
BATCH64 = dict(n_domains=64, seed=1234, steps=1000, storage="f16")


def swimmer2d_domain_args():
    c = SWIMMER2D
    return dict(dim=2, nx=c["nx"], ny=c["ny"], viscosity=c["viscosity"], dx=c["dx"], dt=c["dt"], fluid_density=c["rho"])


def swimmer2d_swimmer(seed=None):
    c = SWIMMER2D
    rng = np.random.default_rng(c["seed"] if seed is None else seed)
    return random_swimmer(rng, c["n_seg"], c["cx"], c["cy"], c["a"], c["b"])


def taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=20.0, seed=1234):
    """Seeded Taylor-Green field of amplitude amp (2 pi scaled axes, random
    phase shift from the seed) plus a vortex ring of radius ring_r cells
    (Gaussian core sigma=3 cells) centred in the domain, scaled so its peak
    speed is amp.  Returns stored-order equilibrium moments (n, 10)."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 3)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = 2 * np.pi * x / nx + ph[0]
    Y = 2 * np.pi * y / ny + ph[1]
    Z = 2 * np.pi * z / nz + ph[2]
    u = amp * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -amp * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    # vortex ring in the plane z = cz, axis along z: azimuthal vorticity ->
    # swirl around the core circle; velocity from a Gaussian core
    cx, cy, cz = nx / 2, ny / 2, nz / 2
    rr = np.hypot(x - cx, y - cy)
    dr = rr - ring_r
    dz = z - cz
    sig = 3.0
    g = np.exp(-(dr ** 2 + dz ** 2) / (2 * sig ** 2))
    # circulation around the core: tangential direction in the (r, z) plane
    ur = -dz * g
    uz = dr * g
    with np.errstate(invalid="ignore", divide="ignore"):
        ex = np.where(rr > 0, (x - cx) / rr, 0.0)
        ey = np.where(rr > 0, (y - cy) / rr, 0.0)
    ring = np.stack([ur * ex, ur * ey, uz], -1)
    peak = np.max(np.linalg.norm(ring, axis=-1))
    ring *= amp / peak
    U = np.stack([u, v, w], -1) + ring
    return equilibrium_moments(U.reshape(-1, 3))
