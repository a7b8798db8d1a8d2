"""ctypes binding of the C ABI in include/freeflow_lbm.h.

This is the binding a Python caller of the C library would write; it has no
fallback: if libfreeflow_lbm.so is missing or no gfx950 device is present the
calls fail with FreeFlowError.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# FREEFLOW_LBM_LIB selects another build of the same library (e.g. an
# installed copy); default: the in-tree build.
LIB_PATH = os.environ.get("FREEFLOW_LBM_LIB") or os.path.join(PKG, "lib", "libfreeflow_lbm.so")

FF_OK = 0
STORAGE = {"f32": 0, "f16": 1, "bf16": 2, "f16u": 3}
OPT_STAGE_FP16 = 0x1
OPT_PERIODIC = 0x2
OPT_FAST_MATH = 0x4

# fsi::lbm::LbmNodeFlag ordinals (LbmDataTypes.cuh:17-56)
FLAG = {
    "Invalid": 0, "Fluid": 1, "FluidRest": 2, "Solid": 3, "SolidDynamic": 4, "Wall": 5,
    "WallLeft": 6, "WallRight": 7, "WallFront": 8, "WallBack": 9, "WallDown": 10, "WallUp": 11, "WallCorner": 12,
    "InletLeft": 13, "InletRight": 14, "InletUp": 15, "InletDown": 16, "InletFront": 17, "InletBack": 18,
    "OutletLeft": 19, "OutletRight": 20, "OutletUp": 21, "OutletDown": 22, "OutletFront": 23, "OutletBack": 24,
}

EXPORTS = [
    "ff_domain_create", "ff_domain_destroy", "ff_set_flags", "ff_get_flags", "ff_init_state", "ff_set_moments",
    "ff_get_moments", "ff_set_boundary", "ff_update_boundary", "ff_set_body_ref", "ff_set_wave", "ff_set_wave_params", "ff_step",
    "ff_step_proxy", "ff_get_vertex_forces", "ff_get_step_vertex_forces", "ff_get_step_vertex_forces_noalloc", "ff_get_body_wrench", "ff_set_wrench_history",
    "ff_get_wrench_history", "ff_get_body_power", "ff_get_power_history", "ff_get_boundary", "ff_mass", "ff_fill_solid", "ff_sync", "ff_time_bulk_kernel",
    "ff_time_steps", "ff_step_timed", "ff_domain_footprint", "ff_set_tile", "ff_autotune", "ff_batch_create", "ff_batch_destroy", "ff_batch_add",
    "ff_batch_remove", "ff_batch_size", "ff_batch_step", "ff_batch_sync", "ff_batch_set_tile", "ff_batch_get_tile",
    "ff_batch_time_steps", "ff_batch_time_bulk", "ff_last_error", "ff_device_count", "ff_version",
]


class FreeFlowError(RuntimeError):
    pass


class DomainDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("viscosity", C.c_double), ("dx", C.c_double), ("dt", C.c_double), ("fluid_density", C.c_double),
        ("storage", C.c_int), ("options", C.c_uint), ("device", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FreeFlowError(f"{LIB_PATH} is not built; run python -m freeflow_amd.build")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ff_last_error.restype = C.c_char_p
        L.ff_version.restype = C.c_char_p
        L.ff_domain_create.argtypes = [C.POINTER(DomainDesc), C.POINTER(vp)]
        L.ff_domain_destroy.argtypes = [vp]
        L.ff_set_flags.argtypes = [vp, vp]
        L.ff_get_flags.argtypes = [vp, vp]
        L.ff_init_state.argtypes = [vp, vp, C.c_int]
        L.ff_set_moments.argtypes = [vp, vp, C.c_int]
        L.ff_get_moments.argtypes = [vp, vp, C.c_int]
        L.ff_set_boundary.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, vp, C.c_int]
        L.ff_update_boundary.argtypes = [vp, vp, vp]
        L.ff_set_body_ref.argtypes = [vp, vp]
        L.ff_set_wave.argtypes = [vp, vp, vp, vp]
        L.ff_set_wave_params.argtypes = [vp, vp]
        L.ff_step.argtypes = [vp, C.c_int]
        L.ff_step_proxy.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int]
        L.ff_get_vertex_forces.argtypes = [vp, vp, C.c_int]
        L.ff_get_step_vertex_forces.argtypes = [vp, vp]
        L.ff_get_step_vertex_forces_noalloc.argtypes = [vp, vp]
        L.ff_get_body_wrench.argtypes = [vp, vp]
        L.ff_set_wrench_history.argtypes = [vp, C.c_int]
        L.ff_get_wrench_history.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
        L.ff_get_body_power.argtypes = [vp, vp]
        L.ff_get_power_history.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
        L.ff_get_boundary.argtypes = [vp, vp, vp]
        L.ff_mass.argtypes = [vp, vp]
        L.ff_fill_solid.argtypes = [vp]
        L.ff_sync.argtypes = [vp]
        L.ff_time_bulk_kernel.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.ff_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.ff_step_timed.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ff_device_count.argtypes = [C.POINTER(C.c_int)]
        L.ff_domain_footprint.argtypes = [C.POINTER(DomainDesc), C.c_int, C.c_int, vp, vp, C.POINTER(C.c_ulonglong)]
        L.ff_set_tile.argtypes = [vp, C.c_int]
        L.ff_autotune.argtypes = [vp, C.POINTER(C.c_int)]
        L.ff_batch_create.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_int, C.POINTER(vp)]
        L.ff_batch_destroy.argtypes = [vp]
        L.ff_batch_add.argtypes = [vp, vp]
        L.ff_batch_remove.argtypes = [vp, vp]
        L.ff_batch_size.argtypes = [vp, C.POINTER(C.c_int)]
        L.ff_batch_step.argtypes = [vp, C.c_int]
        L.ff_batch_sync.argtypes = [vp]
        L.ff_batch_set_tile.argtypes = [vp, C.c_int]
        L.ff_batch_get_tile.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ff_batch_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.ff_batch_time_bulk.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _check(rc):
    if rc != FF_OK:
        raise FreeFlowError(f"freeflow_lbm error {rc}: {lib().ff_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def device_count():
    n = C.c_int(0)
    rc = lib().ff_device_count(C.byref(n))
    return n.value if rc == FF_OK else 0


def footprint(dim, nx, ny, nz=1, storage="f32", verts=None, elems=None, dx=1.0):
    """Device bytes a domain with this grid and boundary allocates (no GPU needed)."""
    desc = DomainDesc(dim, nx, ny, nz if dim == 3 else 1, 0.01, dx, 1.0, 1.0, STORAGE[storage], 0, 0)
    out = C.c_ulonglong(0)
    if verts is None:
        _check(lib().ff_domain_footprint(C.byref(desc), 0, 0, None, None, C.byref(out)))
    else:
        v = np.ascontiguousarray(verts, np.float32).ravel()
        e = np.ascontiguousarray(elems, np.uint32).ravel()
        _check(lib().ff_domain_footprint(C.byref(desc), v.size // dim, e.size // dim, _ptr(v), _ptr(e), C.byref(out)))
    return out.value


def cf_of(dim, rho, dx, dt):
    """SimulationParameters2D/3D::getCf (Config.hpp:96-100, 154-158) with its float cm."""
    cm = np.float32(rho * dx * dx * (dx if dim == 3 else 1.0))
    return float(np.float32(float(cm) * dx / (dt * dt)))


class Domain:
    """One fluid domain on the GPU (the role of LbmFlowField2D/3D + the step)."""

    def __init__(self, dim, nx, ny, nz=1, viscosity=0.005, dx=1.0, dt=1.0, fluid_density=1.0,
                 storage="f32", stage_fp16=None, periodic=False, fast_math=False, device=0):
        L = lib()
        if stage_fp16 is None:  # the reference's FP16 tile staging is its 2D operation order
            stage_fp16 = dim == 2 and not fast_math
        opts = ((OPT_STAGE_FP16 if stage_fp16 else 0) | (OPT_PERIODIC if periodic else 0)
                | (OPT_FAST_MATH if fast_math else 0))
        self.desc = DomainDesc(dim, nx, ny, nz if dim == 3 else 1, viscosity, dx, dt, fluid_density,
                               STORAGE[storage], opts, device)
        self.h = C.c_void_p()
        _check(L.ff_domain_create(C.byref(self.desc), C.byref(self.h)))
        self.dim, self.nx, self.ny, self.nz = dim, nx, ny, (nz if dim == 3 else 1)
        self.nm = 6 if dim == 2 else 10
        self.n = nx * ny * self.nz
        self.nv = 0
        self.n_bodies = 0

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().ff_domain_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_flags(self, flags):
        a = np.ascontiguousarray(flags, np.uint8).ravel()
        assert a.size == self.n
        _check(lib().ff_set_flags(self.h, _ptr(a)))

    def flags(self):
        out = np.empty(self.n, np.uint8)
        _check(lib().ff_get_flags(self.h, _ptr(out)))
        return out

    def init_state(self, bc):
        b = np.ascontiguousarray(bc, np.float32)
        _check(lib().ff_init_state(self.h, _ptr(b), b.size))

    def set_moments(self, m, which=-1):
        a = np.ascontiguousarray(m, np.float32).reshape(-1)
        assert a.size == self.n * self.nm
        _check(lib().ff_set_moments(self.h, _ptr(a), which))

    def moments(self, which=0):
        out = np.empty(self.n * self.nm, np.float32)
        _check(lib().ff_get_moments(self.h, _ptr(out), which))
        return out.reshape(self.n, self.nm)

    def set_boundary(self, verts, vels, elems, elem_body=None, n_bodies=1):
        v = np.ascontiguousarray(verts, np.float32).ravel()
        u = np.ascontiguousarray(vels, np.float32).ravel()
        e = np.ascontiguousarray(elems, np.uint32).ravel()
        b = None if elem_body is None else np.ascontiguousarray(elem_body, np.int32).ravel()
        nv = v.size // self.dim
        _check(lib().ff_set_boundary(self.h, nv, _ptr(v), _ptr(u), e.size // self.dim, _ptr(e),
                                     None if b is None else _ptr(b), n_bodies))
        self.nv, self.n_bodies = nv, n_bodies  # a rejected boundary leaves the previous one

    def update_boundary(self, verts, vels):
        v = np.ascontiguousarray(verts, np.float32).ravel()
        u = np.ascontiguousarray(vels, np.float32).ravel()
        _check(lib().ff_update_boundary(self.h, _ptr(v), _ptr(u)))

    def set_body_ref(self, ref):
        r = np.ascontiguousarray(ref, np.float64).ravel()
        _check(lib().ff_set_body_ref(self.h, _ptr(r)))

    def set_wave(self, rest, s, params):
        r = np.ascontiguousarray(rest, np.float32).ravel()
        s = np.ascontiguousarray(s, np.float32).ravel()
        p = np.ascontiguousarray(params, np.float64).ravel()
        _check(lib().ff_set_wave(self.h, _ptr(r), _ptr(s), _ptr(p)))

    def set_wave_params(self, params):
        """New parameters for the wave set by set_wave (no wait for the device)."""
        p = np.ascontiguousarray(params, np.float64).ravel()
        _check(lib().ff_set_wave_params(self.h, _ptr(p)))

    def step(self, n=1):
        _check(lib().ff_step(self.h, n))

    def sync(self):
        _check(lib().ff_sync(self.h))

    def vertex_forces(self, reset=False):
        out = np.empty(self.nv * self.dim, np.float32)
        _check(lib().ff_get_vertex_forces(self.h, _ptr(out), int(reset)))
        return out.reshape(self.nv, self.dim)

    def step_vertex_forces(self):
        out = np.empty(self.nv * self.dim, np.float32)
        _check(lib().ff_get_step_vertex_forces(self.h, _ptr(out)))
        return out.reshape(self.nv, self.dim)

    def body_wrench(self):
        w = 3 if self.dim == 2 else 6
        out = np.empty(self.n_bodies * w, np.float64)
        _check(lib().ff_get_body_wrench(self.h, _ptr(out)))
        return out.reshape(self.n_bodies, w)

    def set_wrench_history(self, cap):
        _check(lib().ff_set_wrench_history(self.h, cap))

    def wrench_history(self, count):
        w = 3 if self.dim == 2 else 6
        out = np.empty(max(count, 1) * self.n_bodies * w, np.float64)
        n = C.c_int(0)
        _check(lib().ff_get_wrench_history(self.h, _ptr(out), count, C.byref(n)))
        return out[: n.value * self.n_bodies * w].reshape(n.value, self.n_bodies, w)

    def body_power(self):
        """Per-body power of the latest step: sum over the body's vertices of
        -F_v . u_v (FP64, computed on the device)."""
        out = np.empty(max(self.n_bodies, 1), np.float64)
        _check(lib().ff_get_body_power(self.h, _ptr(out)))
        return out[: self.n_bodies]

    def power_history(self, count):
        out = np.empty(max(count, 1) * max(self.n_bodies, 1), np.float64)
        n = C.c_int(0)
        _check(lib().ff_get_power_history(self.h, _ptr(out), count, C.byref(n)))
        return out[: n.value * self.n_bodies].reshape(n.value, self.n_bodies)

    def boundary(self):
        v = np.empty(self.nv * self.dim, np.float32)
        u = np.empty(self.nv * self.dim, np.float32)
        _check(lib().ff_get_boundary(self.h, _ptr(v), _ptr(u)))
        return v.reshape(self.nv, self.dim), u.reshape(self.nv, self.dim)

    def mass(self):
        out = np.zeros(1, np.float64)
        _check(lib().ff_mass(self.h, _ptr(out)))
        return float(out[0])

    def fill_solid(self):
        _check(lib().ff_fill_solid(self.h))

    def time_bulk_kernel(self, iters):
        ms = C.c_double(0)
        _check(lib().ff_time_bulk_kernel(self.h, iters, C.byref(ms)))
        return ms.value

    def time_steps(self, n):
        ms = C.c_double(0)
        _check(lib().ff_time_steps(self.h, n, C.byref(ms)))
        return ms.value

    def step_timed(self, n):
        """n steps; returns (wall ms, summed ms of the steps' stream launches)."""
        tot, stream = C.c_double(0), C.c_double(0)
        _check(lib().ff_step_timed(self.h, n, C.byref(tot), C.byref(stream)))
        return tot.value, stream.value

    def set_tile(self, tile):
        _check(lib().ff_set_tile(self.h, tile))

    def autotune(self):
        t = C.c_int(0)
        _check(lib().ff_autotune(self.h, C.byref(t)))
        return t.value


class Batch:
    """Domains of one kind stepped together, one graph replay per step."""

    def __init__(self, dim, storage="f32", stage_fp16=None, periodic=False, fast_math=False, device=0):
        if stage_fp16 is None:
            stage_fp16 = dim == 2 and not fast_math
        opts = ((OPT_STAGE_FP16 if (stage_fp16 and dim == 2) else 0) | (OPT_PERIODIC if periodic else 0)
                | (OPT_FAST_MATH if fast_math else 0))
        self.h = C.c_void_p()
        _check(lib().ff_batch_create(dim, STORAGE[storage], opts, device, C.byref(self.h)))
        self.members = []

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().ff_batch_destroy(self.h)
            self.h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, dom):
        _check(lib().ff_batch_add(self.h, dom.h))
        self.members.append(dom)

    def remove(self, dom):
        _check(lib().ff_batch_remove(self.h, dom.h))
        self.members.remove(dom)

    def __len__(self):
        n = C.c_int(0)
        _check(lib().ff_batch_size(self.h, C.byref(n)))
        return n.value

    def step(self, n=1):
        _check(lib().ff_batch_step(self.h, n))

    def sync(self):
        _check(lib().ff_batch_sync(self.h))

    def set_tile(self, tile):
        _check(lib().ff_batch_set_tile(self.h, tile))

    def tile(self):
        t, x, y = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib().ff_batch_get_tile(self.h, C.byref(t), C.byref(x), C.byref(y)))
        return t.value, x.value, y.value

    def time_steps(self, n):
        ms = C.c_double(0)
        _check(lib().ff_batch_time_steps(self.h, n, C.byref(ms)))
        return ms.value

    def time_bulk_kernel(self, iters):
        ms = C.c_double(0)
        _check(lib().ff_batch_time_bulk(self.h, iters, C.byref(ms)))
        return ms.value
