"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this.  The oracle is the checker, never the thing
measured as the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    L = C.CDLL(ORACLE_SO)
    vp = C.c_void_p
    L.orc_create.restype = vp
    L.orc_create.argtypes = [C.c_int] * 7
    L.orc_destroy.argtypes = [vp]
    L.orc_set_params.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_set_storage.argtypes = [vp, C.c_int]
    L.orc_set_flags.argtypes = [vp, U8P]
    L.orc_set_moments.argtypes = [vp, F64P, C.c_int]
    L.orc_get_moments.argtypes = [vp, F64P, C.c_int]
    L.orc_init_reference.argtypes = [vp, F32P, C.c_int]
    L.orc_set_boundary.argtypes = [vp, C.c_int, F32P, F32P, C.c_int, U32P]
    L.orc_update_boundary.argtypes = [vp, F32P, F32P]
    L.orc_get_vertex_forces.argtypes = [vp, F64P]
    L.orc_zero_vertex_forces.argtypes = [vp]
    L.orc_step.argtypes = [vp, C.c_int]
    L.orc_get_hit_table.argtypes = [vp, U32P]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = C.c_int
    L.orc_recon2d_all_f.argtypes = [F32P, F32P]
    L.orc_recon2d_all_d.argtypes = [F64P, F64P]
    L.orc_recon2d_at_f.argtypes = [F32P, C.c_int]
    L.orc_recon2d_at_f.restype = C.c_float
    L.orc_recon2d_at_d.argtypes = [F64P, C.c_int]
    L.orc_recon2d_at_d.restype = C.c_double
    L.orc_collide2d_f.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, F32P]
    L.orc_collide2d_d.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, F64P]
    L.orc_recon3d_all_f.argtypes = [F32P, F32P]
    L.orc_recon3d_all_d.argtypes = [F64P, F64P]
    L.orc_recon3d_at_f.argtypes = [F32P, C.c_int]
    L.orc_recon3d_at_f.restype = C.c_float
    L.orc_recon3d_at_d.argtypes = [F64P, C.c_int]
    L.orc_recon3d_at_d.restype = C.c_double
    L.orc_collide3d_f.argtypes = [C.c_float] * 5 + [F32P]
    L.orc_collide3d_d.argtypes = [C.c_double] * 5 + [F64P]
    L.orc_line_intersect.argtypes = [F32P, F32P]
    L.orc_line_intersect.restype = C.c_float
    L.orc_intersect3.argtypes = [F32P, F32P, F32P, F32P]
    L.orc_intersect3.restype = C.c_int
    L.orc_depth_index_compact.argtypes = [C.c_float, C.c_int]
    L.orc_depth_index_compact.restype = C.c_uint32
    L.orc_f2h.argtypes = [C.c_float]
    L.orc_f2h.restype = C.c_uint16
    L.orc_h2f.argtypes = [C.c_uint16]
    L.orc_h2f.restype = C.c_float
    L.orc_f2bf.argtypes = [C.c_float]
    L.orc_f2bf.restype = C.c_uint16
    L.orc_bf2f.argtypes = [C.c_uint16]
    L.orc_bf2f.restype = C.c_float
    L.orc_quantize.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
    L.orc_quantize.restype = C.c_double
    L.orc_fill_solid2d.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, F32P, C.c_int, U32P, U8P]
    _lib = L
    return L


class OracleDomain:
    """One oracle domain in the reference's AoS moment order."""

    def __init__(self, dim, nx, ny, nz=1, fp64=False, stage_fp16=True, periodic=False,
                 nu=0.005, dx=1.0, dt=1.0, cf=1.0, storage=0):
        L = lib()
        self.L = L
        self.dim, self.nx, self.ny, self.nz = dim, nx, ny, (nz if dim == 3 else 1)
        self.nm = 6 if dim == 2 else 10
        self.n = nx * ny * self.nz
        self.h = L.orc_create(dim, nx, ny, self.nz, int(fp64), int(stage_fp16), int(periodic))
        if not self.h:
            raise ValueError("oracle domain creation failed")
        L.orc_set_params(self.h, nu, dx, dt, cf)
        if storage:
            L.orc_set_storage(self.h, storage)
        self.nv = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_flags(self, flags):
        self.L.orc_set_flags(self.h, np.ascontiguousarray(flags, np.uint8).ravel())

    def set_moments(self, m, which=-1):
        a = np.ascontiguousarray(m, np.float64).reshape(-1)
        assert a.size == self.n * self.nm
        self.L.orc_set_moments(self.h, a, which)

    def moments(self, which=0):
        out = np.empty(self.n * self.nm, np.float64)
        self.L.orc_get_moments(self.h, out, which)
        return out.reshape(self.n, self.nm)

    def init_reference(self, bc):
        b = np.ascontiguousarray(bc, np.float32)
        self.L.orc_init_reference(self.h, b, b.size)

    def set_boundary(self, verts, vels, elems):
        v = np.ascontiguousarray(verts, np.float32).ravel()
        u = np.ascontiguousarray(vels, np.float32).ravel()
        e = np.ascontiguousarray(elems, np.uint32).ravel()
        self.nv = v.size // self.dim
        self.L.orc_set_boundary(self.h, self.nv, v, u, e.size // self.dim, e)

    def update_boundary(self, verts, vels):
        self.L.orc_update_boundary(self.h, np.ascontiguousarray(verts, np.float32).ravel(),
                                   np.ascontiguousarray(vels, np.float32).ravel())

    def vertex_forces(self):
        out = np.empty(self.nv * self.dim, np.float64)
        self.L.orc_get_vertex_forces(self.h, out)
        return out.reshape(self.nv, self.dim)

    def zero_vertex_forces(self):
        self.L.orc_zero_vertex_forces(self.h)

    def hit_table(self):
        q = 9 if self.dim == 2 else 27
        out = np.empty(self.n * q, np.uint32)
        self.L.orc_get_hit_table(self.h, out)
        return out.reshape(self.n, q)

    def step(self, n=1):
        self.L.orc_step(self.h, n)


def set_threads(n):
    return lib().orc_set_threads(n)
