"""Precision of the storage formats against the FP64 oracle.

FP32 arithmetic with FP32, FP16 or BF16 moment storage (rho-1, scaled u and
Pi) is compared with the oracle run entirely in FP64 on the same seeded
inputs: relative L2 error of the velocity field and relative mass drift.
The errors measured here are the ones quoted for the storage formats.
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


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def tg3d_fp64():
    nx, ny, nz, steps = 64, 32, 32, 300
    m0 = wl.taylor_green_ring(nx, ny, nz, amp=0.04, ring_r=8.0, seed=11)
    o = ol.OracleDomain(3, nx, ny, nz, fp64=True, stage_fp16=False, periodic=True, nu=0.01)
    o.set_moments(m0)
    o.step(steps)
    return (nx, ny, nz, steps, m0, o.moments(0))


# FP32 storage: rounding of every operation; 16-bit storage: one rounding of
# every stored moment per step (11 / 8 significant bits; the split velocity
# keeps about 22 for u, which dominates the FP16 error in 3D).
BOUNDS3D = {"f32": 1e-5, "f16": 1e-3, "bf16": 1e-2, "f16u": 5e-5}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "f16u"])
def test_3d_storage_error_vs_fp64(tg3d_fp64, storage):
    nx, ny, nz, steps, m0, ref = tg3d_fp64
    d = capi.Domain(3, nx, ny, nz, viscosity=0.01, storage=storage, periodic=True)
    d.set_moments(m0)
    mass0 = d.mass()
    d.step(steps)
    got = d.moments(0).astype(np.float64)
    err = rel_l2(got[:, 1:4], ref[:, 1:4])
    drift = abs(d.mass() - mass0) / mass0
    note(f"3D 64x32x32 TG+ring, {steps} steps, storage {storage}: velocity rel-L2 vs FP64 {err:.3e}, "
         f"mass drift {drift:.3e}")
    assert err < BOUNDS3D[storage]
    assert drift < (1e-5 if storage == "f32" else 1e-4)


# In 2D the swimmer's flow is sensitive to rho - 1 and Pi too: the CPU oracle
# with only the velocity kept at FP32 precision is 3.5e-4 from FP64 after 200
# steps (FP16 7.0e-4; rounding only rho - 1: 1.9e-4; only Pi: 7.5e-5).
BOUNDS2D = {"f32": 1e-5, "f16": 1e-3, "bf16": 1e-2, "f16u": 5e-4}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "f16u"])
def test_2d_swimmer_storage_error_vs_fp64(storage):
    sc = wl.batched_scenes(1, 77, (256, 256), (128, 128), (200, 200))[0]
    steps = 200
    o = ol.OracleDomain(2, sc.nx, sc.ny, fp64=True, stage_fp16=False, nu=sc.viscosity, dx=sc.dx, dt=sc.dt,
                        cf=capi.cf_of(2, sc.rho, sc.dx, sc.dt))
    o.set_flags(sc.flags)
    o.set_moments(wl.equilibrium_moments(np.zeros((sc.nx * sc.ny, 2))))
    d = capi.Domain(**sc.domain_args(), storage=storage, stage_fp16=False)
    sc.setup(d)
    d.set_moments(wl.equilibrium_moments(np.zeros((sc.nx * sc.ny, 2))))  # the oracle's initial state exactly
    b = capi.Batch(2, storage, stage_fp16=False)
    b.add(d)
    v, vel = d.boundary()
    o.set_boundary(v, vel, sc.swimmer.elems)
    for _ in range(steps):
        b.step(1)
        v, vel = d.boundary()  # kinematics the device used for this step
        o.update_boundary(v, vel)
        o.step(1)
    got = d.moments(0).astype(np.float64)
    ref = o.moments(0)
    upd = (sc.flags == 1)
    err = rel_l2(got[upd, 1:3], ref[upd, 1:3])
    note(f"2D swimmer {sc.nx}x{sc.ny}, {steps} steps, batched, storage {storage}: velocity rel-L2 vs FP64 {err:.3e}")
    assert err < BOUNDS2D[storage]
