"""ff_get_step_vertex_forces_noalloc against ff_get_step_vertex_forces.

The no-allocation variant stages the step's vertex forces in a static
per-GPU array instead of a device allocation per call; it runs the same
kernel on the same inputs, so its output must equal the allocating entry
point's bit for bit.  Checked on consecutive calls of different sizes (a
larger domain, then a smaller one, then the larger one again, so a stale
tail of the shared array would show), in 2D and 3D, with an output buffer
longer than needed (the variant must write exactly nv * dim values), and
the size limit: a domain with nv * dim above the array is rejected with
FF_ERR_CAPACITY before any device work.  Both entry points are called
through the C ABI.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from freeflow_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu

SCRATCH_FLOATS = 1 << 20
FF_OK, FF_ERR_CAPACITY = 0, 4


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _both(d):
    """(allocating, no-allocation) outputs of one domain, the latter written
    into a buffer 64 floats longer than nv * dim and filled with a sentinel."""
    L = capi.lib()
    n = d.nv * d.dim
    ref = np.empty(n, np.float32)
    assert L.ff_get_step_vertex_forces(d.h, _ptr(ref)) == FF_OK, L.ff_last_error().decode()
    got = np.full(n + 64, np.float32(-7.25), np.float32)
    assert L.ff_get_step_vertex_forces_noalloc(d.h, _ptr(got)) == FF_OK, L.ff_last_error().decode()
    assert np.all(got[n:] == np.float32(-7.25)), "wrote past nv * dim"
    return ref, got[:n]


def _swimmer_domain(nx, ny, n_seg):
    dx, dt = 0.005, 0.0025
    d = capi.Domain(2, nx, ny, viscosity=0.005, dx=dx, dt=dt, fluid_density=10.0)
    d.set_flags(wl.outlet_box_flags(nx, ny))
    d.set_moments(wl.equilibrium_moments(np.tile([0.01, 0.0], (nx * ny, 1))))
    sw = wl.Swimmer2D(n_seg, 0.3 * nx * dx, 0.5 * ny * dx, 0.2 * nx * dx, 0.05 * ny * dx, A=0.02, f=2.0, phi=0.4)
    v0, vel0 = sw.state(0.0)
    d.set_boundary(v0, vel0, sw.elems)
    d.set_wave(sw.rest, sw.s, sw.params[None, :])
    return d


def test_noalloc_step_vertex_forces_matches_allocating_entry_point():
    note("case no-alloc vertex forces: 2D 240x120 swimmer (400 segments), then 96x64 (60 segments)")
    big = _swimmer_domain(240, 120, 400)
    small = _swimmer_domain(96, 64, 60)
    assert small.nv < big.nv
    # before any step both return zeros
    ref, got = _both(small)
    assert np.all(ref == 0) and np.array_equal(ref, got)
    big.step(5)
    ref, got = _both(big)
    assert np.abs(ref).max() > 0
    assert np.array_equal(ref, got)
    small.step(3)  # second call in a row, smaller input
    ref, got = _both(small)
    assert np.abs(ref).max() > 0
    assert np.array_equal(ref, got)
    big.step(1)
    ref, got = _both(big)
    assert np.array_equal(ref, got)
    big.close()
    small.close()


def test_noalloc_step_vertex_forces_3d_and_size_limit():
    nx, ny, nz = 40, 32, 32
    note(f"case no-alloc vertex forces: 3D sphere {nx}x{ny}x{nz}, then the size limit")
    d = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True)
    u = np.zeros((nx * ny * nz, 3))
    u[:, 0] = 0.03
    d.set_moments(np.asarray(wl.equilibrium_moments(u), np.float32))
    c = np.array([19.3, 15.6, 16.2])
    v, tri = wl.icosphere(c, 6.1, subdiv=2)
    d.set_boundary(v, 0.02 * np.cross([0.0, 0.0, 1.0], v - c), tri)
    d.step(4)
    ref, got = _both(d)
    assert np.abs(ref).max() > 0
    assert np.array_equal(ref, got)
    # nv * dim just above the array: rejected with FF_ERR_CAPACITY, no device work
    nv = SCRATCH_FLOATS // 3 + 1
    verts = np.tile(v[:3].astype(np.float32), (nv // 3 + 1, 1))[:nv]
    big = capi.Domain(3, nx, ny, nz, viscosity=0.01, periodic=True)
    big.set_boundary(verts, np.zeros_like(verts), np.array([[0, 1, 2]], np.uint32))
    out = np.empty(nv * 3, np.float32)
    L = capi.lib()
    rc = L.ff_get_step_vertex_forces_noalloc(big.h, _ptr(out))
    msg = L.ff_last_error().decode()
    note(f"  nv * dim = {nv * 3}: rc {rc}, {msg}")
    assert rc == FF_ERR_CAPACITY and "scratch" in msg
    assert L.ff_get_step_vertex_forces(big.h, _ptr(out)) == FF_OK  # the allocating entry point has no limit
    # the domain under the limit still reads back after the rejected call
    d.step(1)
    ref, got = _both(d)
    assert np.array_equal(ref, got)
    big.close()
    d.close()
