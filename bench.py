"""Benchmark of the HOME-LBM fluid step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tg3d]

Default workload (the configuration whose bytes exceed the 256 MB MALL, so
the roofline is HBM's): 3D D3Q27 on a 256x128x128 periodic grid, FP32
storage and arithmetic, a seeded Taylor-Green field of amplitude 0.04 plus a
vortex ring of radius 20 cells, viscosity 0.01.  A step is one stream-collide
update of every cell; K steps are timed with the moments resident in HBM.

Prints one JSON line.  `value` is MLUPS (cells x steps / s / 1e6) of the
column tiles in FF_OPT_FAST_MATH form (the same update with reassociated FP32
arithmetic and FMA; `--reference-order` times the reference's operation
order instead).  `roofline` prices that kernel against 8 TB/s HBM with 81
algorithmic bytes per cell update (10 FP32 moments read + 10 written + 1
flag byte), from HIP events on the kernel's own launches.  The line also
carries:
  reference_order     the same workload in the reference's operation order
                      (bit-exact with the CPU oracle): MLUPS, kernel time,
                      fraction of HBM peak;
  accuracy            both forms after S steps on the full grid against the
                      CPU oracle in FP32 (the reference order: fraction of
                      bit-identical values) and in FP64 (velocity rel-L2),
                      with the mass drift of the timed run;
  cpu_baseline        the FP32 CPU oracle (OpenMP) on the same grid, S steps;
  swimmer2d           2D FSI 1200x400 with a 600-segment swimmer (plain FP32,
                      per-step FP64 wrench): MLUPS and the per-step latency
                      with a host sync after every step;
  batched             64 uneven 2D swimmer domains, FP16 storage, one graph
                      replay per step, against the plain per-domain FP32
                      form, with one member's velocity error vs FP64;
  served              Poisson rollout episodes at 50/s through the
                      continuous-batching server: p50/p99 step latency and
                      MLUPS against the plain per-domain FP32 form.
`--no-extras` leaves out the last three.

Other workloads (one JSON line each, same keys):
  --workload swimmer2d 2D FSI plain FP32 1200x400 swimmer alone.
  --workload batched   64 uneven 2D swimmer domains (256x128 .. 1200x400,
                       200-800 segments) in one batch, FP16 storage, one graph
                       replay per step; also times the plain per-domain FP32
                       form of the same domains and reports the ratio.
  --workload served    Poisson-arriving 2D/3D rollout episodes at --rate per s
                       for --duration s through the continuous-batching server.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + % HBM peak; batched/served vs plain at ≤1e-3 rel err; p99 step ms"
HBM_PEAK_GBS = 8000.0
# tile shapes by index, as in the library (ff_set_tile)
TILES2 = ["32x8", "32x16", "64x8", "32x32", "64x16"]
TILES3 = ["8x8x4 box", "16x4x4 box", "16x8 ring column x32", "16x8 ring column x16", "16x8 ring column x8",
          "16x16 ring column x32", "16x16 ring column x16", "16x12 staged column x8", "16x26 staged column x8",
          "64x4 staged column x16", "16x12 staged column x8, 2 planes in an LDS ring",
          "16x8 staged column x8, 3 planes in an LDS ring", "32x8 staged column x8, 2 planes in an LDS ring",
          "16x12 staged column x16, 3 planes in an LDS ring", "64x4 staged column x32", "32x12 staged column x16"]


def bytes_per_lup(dim, storage_bytes, split_velocity=False):
    """Algorithmic bytes per cell update: read and write every stored plane
    once (the split-velocity storage has dim more planes), read the flag."""
    nm = (6 if dim == 2 else 10) + (dim if split_velocity else 0)
    return 2 * nm * storage_bytes + 1


def setup_swimmer2d(capi, wl, fast=False):
    c = wl.SWIMMER2D
    nx, ny = c["nx"], c["ny"]
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.zeros((nx * ny, 2)))
    sw = wl.swimmer2d_swimmer()
    d = capi.Domain(**wl.swimmer2d_domain_args(), storage="f32", stage_fp16=not fast, fast_math=fast)
    d.set_flags(flags)
    d.set_moments(m)
    v0, vel0 = sw.state(0.0)
    d.set_boundary(v0, vel0, sw.elems)
    d.set_wave(sw.rest, sw.s, sw.params[None, :])
    return d, (flags, m, sw)


def cpu_baseline_swimmer2d(state, budget_s=12.0):
    """Time the CPU oracle on the same workload for ~budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from freeflow_amd import capi, workloads as wl

    flags, m, sw = state
    c = wl.SWIMMER2D
    nx, ny = c["nx"], c["ny"]
    threads = ol.set_threads(0)
    o = ol.OracleDomain(2, nx, ny, fp64=False, stage_fp16=True, nu=c["viscosity"], dx=c["dx"], dt=c["dt"],
                        cf=capi.cf_of(2, c["rho"], c["dx"], c["dt"]))
    o.set_flags(flags)
    o.set_moments(m)
    v, vel = sw.state(0.0)
    o.set_boundary(v, vel, sw.elems)
    steps = 0
    t0 = time.perf_counter()
    while True:
        v, vel = sw.state(steps * c["dt"])
        o.update_boundary(v.astype(np.float32), vel.astype(np.float32))
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 4000:
            break
    mlups = nx * ny * steps / el / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (FP32 restatement of the reference update, OpenMP {threads} threads) on the "
                      f"same 1200x400 swimmer workload, first {steps} steps ({el:.1f} s)"}


# (tx, ty, cells per thread) of the 2D tiles and (tx, ty, planes, kind, threads, planes in flight, LDS DMA)
# of the 3D tiles, by tile index (ff_set_tile)
SHAPES2 = [(32, 8, 1), (32, 8, 2), (64, 4, 2), (32, 8, 4), (64, 4, 4)]
SHAPES3 = [(8, 8, 4, 0, 256, 1, 0), (16, 4, 4, 0, 256, 1, 0), (16, 8, 32, 1, 128, 1, 0), (16, 8, 16, 1, 128, 1, 0),
           (16, 8, 8, 1, 128, 1, 0), (16, 16, 32, 1, 256, 1, 0), (16, 16, 16, 1, 256, 1, 0), (16, 12, 8, 2, 256, 1, 0),
           (16, 26, 8, 2, 512, 1, 0), (64, 4, 16, 2, 448, 1, 0), (16, 12, 8, 2, 256, 2, 1), (16, 8, 8, 2, 192, 3, 1),
           (32, 8, 8, 2, 384, 2, 1), (16, 12, 16, 2, 256, 3, 1), (64, 4, 32, 2, 448, 1, 0), (32, 12, 16, 2, 512, 1, 0)]


def kernel2d(prefix, s, stage, periodic, tile, fast=False):
    tx, ty, cpt = SHAPES2[tile]
    return (f"ff::{prefix}<{s}, {str(stage).lower()}, {str(periodic).lower()}, {tx}, {ty}, {cpt}, "
            f"{str(fast and not stage).lower()}>")


def kernel3d(s, periodic, tile, fast=False):
    tx, ty, tz, kind, nt, pf, dma = SHAPES3[tile]
    return (f"ff::k_bulk3d<{s}, {str(periodic).lower()}, {str(fast).lower()}, {kind}, {tx}, {ty}, {tz}, {nt}, {pf}, "
            f"{str(bool(dma)).lower()}>")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (profiles/traffic.json, counters
    scaled by a 256 MiB copy), or None when that kernel was not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
    except OSError:
        return None
    for k in t.get("kernels", []):
        if k["kernel"] == kernel:
            return k["traffic_bytes"]
    return None


def pick_tile(d, args):
    if args.tile >= 0:
        d.set_tile(args.tile)
        return args.tile
    return 0 if args.no_tune else d.autotune()


def roofline(achieved_gbs, kernel, avg_ms, bpl, cells):
    return {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel), "kernel": kernel,
            "avg_kernel_us": round(avg_ms * 1e3, 3), "bytes_per_lup": bpl, "cells_per_launch": cells}


def base_line(value, K, W, ms_per_step, dtype, config, roof, cpu=None, **extra):
    out = {"metric": METRIC, "value": round(value, 2), "unit": "MLUPS", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": config, "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    return out


def batched_summary(capi, wl, K, W, plain_steps, tile=-1, fp64_check=True):
    """The 64-domain 2D swimmer batch (FP16 storage) against the plain per-domain FP32
    form of the same domains; optionally one member's velocity error vs FP64."""
    c = wl.BATCH64
    scenes = wl.batched_scenes(c["n_domains"], c["seed"])
    cells = sum(s.nx * s.ny for s in scenes)
    # plain per-domain FP32 form: the reference's storage and FP16 tile, one
    # domain per stream, stepped round-robin
    plain = []
    for sc in scenes:
        d = capi.Domain(**sc.domain_args(), storage="f32", stage_fp16=True)
        sc.setup(d)
        plain.append(d)
    for d in plain:
        d.step(5)
    for d in plain:
        d.sync()
    print(f"[bench] plain FP32 form: {len(plain)} domains, {cells} cells, {plain_steps} steps", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for d in plain:
            d.step(1)
    for d in plain:
        d.sync()
    t_plain = time.perf_counter() - t0
    plain_mlups = cells * plain_steps / t_plain / 1e6
    for d in plain:
        d.close()
    del plain
    # batched FP16-storage form, FP32 fast-math arithmetic
    b = capi.Batch(2, c["storage"], stage_fp16=False, fast_math=True)
    doms = []
    for sc in scenes:
        d = capi.Domain(**sc.domain_args(), storage=c["storage"], stage_fp16=False, fast_math=True)
        sc.setup(d)
        b.add(d)
        doms.append(d)
    if tile >= 0:
        b.set_tile(tile)
    b.step(W)
    b.sync()
    print(f"[bench] batched FP16 form: warmup {W} done; timing {K} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b.step(K)
    b.sync()
    el = time.perf_counter() - t0
    mlups = cells * K / el / 1e6
    t_idx, tx, ty = b.tile()
    avg_ms = b.time_bulk_kernel(100)
    wr = np.concatenate([d.body_wrench() for d in doms])
    out = {"mlups": mlups, "ms_per_step": el * 1e3 / K, "steps": K, "warmup": W, "cells": cells,
           "domains": len(scenes), "tile": t_idx, "tile_shape": [tx, ty], "avg_kernel_ms": avg_ms,
           "plain_fp32_mlups": plain_mlups, "plain_steps": plain_steps, "speedup_vs_plain_fp32": mlups / plain_mlups,
           "healthy": bool(np.all(np.isfinite(wr)))}
    b.close()
    for d in doms:
        d.close()
    # the same batch with the split-velocity FP16 storage (velocity kept at
    # about FP32 precision)
    b = capi.Batch(2, "f16u", stage_fp16=False, fast_math=True)
    doms = []
    for sc in scenes:
        d = capi.Domain(**sc.domain_args(), storage="f16u", stage_fp16=False, fast_math=True)
        sc.setup(d)
        b.add(d)
        doms.append(d)
    b.step(W)
    b.sync()
    print(f"[bench] batched split-velocity FP16 form: timing {K} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b.step(K)
    b.sync()
    el_u = time.perf_counter() - t0
    out["f16u"] = {"mlups": cells * K / el_u / 1e6, "ms_per_step": el_u * 1e3 / K,
                   "speedup_vs_plain_fp32": cells * K / el_u / 1e6 / plain_mlups, "tile": b.tile()[0]}
    b.close()
    for d in doms:
        d.close()
    if fp64_check:
        out["fp16_member_error"] = cpu_baseline_batched_member(capi, wl, scenes, 1000)
        out["f16u"]["member_error"] = cpu_baseline_batched_member(capi, wl, scenes, 1000, "f16u")
    return out


def cpu_baseline_batched_member(capi, wl, scenes, steps, storage="f16"):
    """CPU leg of the batched workload: the FP64 oracle on the smallest batch
    member, fed each step the boundary the device used, against the member
    stepped in a 16-bit-storage batch (`storage`: f16 or f16u); returns the
    velocity rel-L2 over the updated cells after 100 and `steps` steps
    against the 1e-3 bar."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    k = min(range(len(scenes)), key=lambda i: scenes[i].nx * scenes[i].ny)
    sc = scenes[k]
    t0 = time.perf_counter()
    o = ol.OracleDomain(2, sc.nx, sc.ny, fp64=True, stage_fp16=False, nu=sc.viscosity, dx=sc.dx, dt=sc.dt,
                        cf=capi.cf_of(2, sc.rho, sc.dx, sc.dt))
    o.set_flags(sc.flags)
    # the FP32 oracle (no storage rounding) on the same kinematics: how far
    # plain FP32 arithmetic drifts from FP64 on this flow
    o32 = ol.OracleDomain(2, sc.nx, sc.ny, fp64=False, stage_fp16=False, nu=sc.viscosity, dx=sc.dx, dt=sc.dt,
                          cf=capi.cf_of(2, sc.rho, sc.dx, sc.dt))
    o32.set_flags(sc.flags)
    m0 = wl.equilibrium_moments(np.zeros((sc.nx * sc.ny, 2)))
    o.set_moments(m0)
    o32.set_moments(m0)
    d = capi.Domain(**sc.domain_args(), storage=storage, stage_fp16=False, fast_math=True)
    sc.setup(d)
    d.set_moments(m0)
    b = capi.Batch(2, storage, stage_fp16=False, fast_math=True)
    b.add(d)
    v, vel = d.boundary()
    o.set_boundary(v, vel, sc.swimmer.elems)
    o32.set_boundary(v, vel, sc.swimmer.elems)
    upd = sc.flags == 1

    def rel(a, r):
        return float(np.linalg.norm(a[upd, 1:3] - r[upd, 1:3]) / np.linalg.norm(r[upd, 1:3]))

    at = {}
    for s in range(steps):
        b.step(1)
        v, vel = d.boundary()  # the kinematics the device used for this step
        o.update_boundary(v, vel)
        o.step(1)
        o32.update_boundary(v, vel)
        o32.step(1)
        if s + 1 in (100, steps):
            ref = o.moments(0)
            at[s + 1] = (rel(d.moments(0).astype(np.float64), ref), rel(o32.moments(0), ref))
    b.close()
    d.close()
    return {"member": k, "nx": sc.nx, "ny": sc.ny, "steps": steps, "velocity_rel_l2_vs_fp64": at[steps][0],
            "velocity_rel_l2_vs_fp64_at_100": at[100][0], "fp32_oracle_velocity_rel_l2_vs_fp64": at[steps][1],
            "fp32_oracle_velocity_rel_l2_vs_fp64_at_100": at[100][1], "bar": 1e-3,
            "note": "storage rounding of every moment each step accumulates; the FP32 oracle's own drift from "
                    "FP64 on the same kinematics is given beside it",
            "seconds": round(time.perf_counter() - t0, 1)}


def run_batched(args, capi, wl):
    c = wl.BATCH64
    K = args.steps if args.steps is not None else c["steps"]
    W = min(args.warmup, 50)
    r = batched_summary(capi, wl, K, W, args.plain_steps, args.tile, fp64_check=not args.no_cpu)
    bpl = bytes_per_lup(2, 2)
    roof = roofline(bpl * r["cells"] / (r["avg_kernel_ms"] * 1e-3) / 1e9,
                    kernel2d("kb_step2d", "__half", False, False, r["tile"], fast=True), r["avg_kernel_ms"], bpl,
                    r["cells"])
    out = base_line(r["mlups"], K, W, r["ms_per_step"], "f32 compute, f16 storage",
                    {"workload": "64 uneven 2D swimmer domains (256x128..1200x400, 200-800 segments) in one batch, "
                                 "FP16 storage, one graph replay per step", "domains": r["domains"],
                     "cells": r["cells"], "tile": r["tile_shape"], "parallelism": "single-gpu"},
                    roof, None, plain_fp32_mlups=round(r["plain_fp32_mlups"], 2),
                    speedup_vs_plain_fp32=round(r["speedup_vs_plain_fp32"], 3), healthy=r["healthy"],
                    fp16_member_error=r.get("fp16_member_error"), f16u=r.get("f16u"))
    print(json.dumps(out), flush=True)
    return 0


def tg3d_time(capi, wl, m, storage, fast, K, W, tile, tune=True):
    """Time K steps of the 3D workload (after W warm-up steps) in one form;
    returns the rate, the chosen tile and the stream-collide launch time."""
    c = wl.TG3D
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    cells = nx * ny * nz
    d = capi.Domain(3, nx, ny, nz, viscosity=c["viscosity"], storage=storage, periodic=True, fast_math=fast)
    d.set_moments(m)
    if tile >= 0:
        d.set_tile(tile)
        t = tile
    else:
        t = d.autotune() if tune else 0
    m0 = d.mass()
    d.step(W)
    d.sync()
    print(f"[bench] tg3d {storage} {'fast' if fast else 'reference-order'} tile {t}: timing {K} steps",
          file=sys.stderr, flush=True)
    # the timed region: K steps, each step's stream launch timed by HIP events
    # attached to the launch (the kernel as stepping runs it: each launch reads
    # what the previous one wrote), the region itself by events on the stream
    t0 = time.perf_counter()
    total_ms, stream_ms = d.step_timed(K)
    host_el = time.perf_counter() - t0
    el = total_ms * 1e-3
    avg_ms = stream_ms / K
    drift = abs(d.mass() - m0) / m0
    d.close()
    bpl = bytes_per_lup(3, 4 if storage == "f32" else 2, storage == "f16u")
    sname = {"f32": "float", "f16": "__half", "bf16": "__hip_bfloat16", "f16u": "ff::h16u"}[storage]
    return {"mlups": cells * K / el / 1e6, "ms_per_step": el * 1e3 / K, "host_ms_per_step": host_el * 1e3 / K,
            "tile": t, "avg_kernel_ms": avg_ms,
            "kernel": kernel3d(sname, True, t, fast), "bytes_per_lup": bpl, "cells": cells,
            "frac": bpl * cells / (avg_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), "mass_drift": drift, "steps": K + W}


def cpu_baseline_tg3d(capi, wl, m, steps, forms):
    """CPU leg of the 3D workload: the FP32 CPU oracle (OpenMP, timed: the
    cpu_baseline) and the FP64 oracle, `steps` steps on the full grid from
    the same state, against the GPU forms timed (`forms`: name -> (storage,
    fast math, tile))."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    c = wl.TG3D
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    cells = nx * ny * nz
    threads = ol.set_threads(0)
    m32 = np.asarray(m, np.float32)
    gpu = {}
    for name, (storage, fast, tile) in forms.items():
        d = capi.Domain(3, nx, ny, nz, viscosity=c["viscosity"], storage=storage, periodic=True, fast_math=fast)
        d.set_tile(tile)
        d.set_moments(m32)
        d.step(steps)
        gpu[name] = d.moments(0).astype(np.float64)
        d.close()
    o = ol.OracleDomain(3, nx, ny, nz, fp64=False, stage_fp16=False, periodic=True, nu=c["viscosity"])
    o.set_moments(m32.astype(np.float64))
    print(f"[bench] CPU oracle FP32, {threads} threads, {steps} steps of {nx}x{ny}x{nz}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    o.step(steps)
    el = time.perf_counter() - t0
    r32 = o.moments(0)
    del o
    o = ol.OracleDomain(3, nx, ny, nz, fp64=True, stage_fp16=False, periodic=True, nu=c["viscosity"])
    o.set_moments(m32.astype(np.float64))
    o.step(steps)
    r64 = o.moments(0)
    del o
    vel = np.linalg.norm(r64[:, 1:4])
    acc = {"steps": steps, "grid": [nx, ny, nz]}
    for key, g in gpu.items():
        acc[key] = {"rel_l2_vs_fp32_oracle": float(np.linalg.norm(g - r32) / np.linalg.norm(r32)),
                    "bit_identical_fraction": float(np.mean(g == r32)),
                    "velocity_rel_l2_vs_fp64": float(np.linalg.norm(g[:, 1:4] - r64[:, 1:4]) / vel)}
    acc["fp32_oracle_velocity_rel_l2_vs_fp64"] = float(np.linalg.norm(r32[:, 1:4] - r64[:, 1:4]) / vel)
    cpu = {"value": round(cells * steps / el / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"CPU oracle (FP32 restatement of the reference update, OpenMP {threads} threads) on the same "
                     f"{nx}x{ny}x{nz} grid, {steps} steps ({el:.1f} s)"}
    return cpu, acc


def swimmer2d_summary(capi, wl, K=2000, W=100, lat_steps=200):
    """The 1200x400 2D swimmer alone: K steps with the per-step FP64 wrench kept in a device
    ring read back after the run, then per-step latency with a host sync and
    a wrench readback after every step."""
    d, _ = setup_swimmer2d(capi, wl)
    t = d.autotune()
    c = wl.SWIMMER2D
    cells = c["nx"] * c["ny"]
    d.set_wrench_history(K)
    d.step(W)
    d.sync()
    t0 = time.perf_counter()
    d.step(K)
    hist = d.wrench_history(K)
    el = time.perf_counter() - t0
    lat = []
    for _ in range(lat_steps):
        a = time.perf_counter()
        d.step(1)
        d.body_wrench()  # synchronises
        lat.append((time.perf_counter() - a) * 1e3)
    _, stream_ms = d.step_timed(400)  # the stream launch as stepping runs it
    avg_ms = stream_ms / 400
    d.close()
    # the same workload in the fast-math form (no FP16 staging, FMA, FP32 kinematics)
    df, _ = setup_swimmer2d(capi, wl, fast=True)
    tf = df.autotune()
    df.step(W)
    df.sync()
    t0 = time.perf_counter()
    df.step(K)
    df.sync()
    el_f = time.perf_counter() - t0
    df.close()
    fast = {"mlups": round(cells * K / el_f / 1e6, 2), "ms_per_step": round(el_f * 1e3 / K, 5), "tile": TILES2[tf]}
    return {"mlups": round(cells * K / el / 1e6, 2), "ms_per_step": round(el * 1e3 / K, 5), "steps": K,
            "fast_math": fast,
            "tile": TILES2[t], "stream_launch_us": round(avg_ms * 1e3, 3),
            "step_latency_ms_with_sync": {"p50": round(float(np.percentile(lat, 50)), 4),
                                          "p99": round(float(np.percentile(lat, 99)), 4)},
            "healthy": bool(np.all(np.isfinite(hist))),
            "note": "1200x400 is 23.5 MB of FP32 moments: cache-resident (MALL), latency-bound, not an HBM figure"}


def served_summary(rates=(50.0, 100.0), duration=4.0):
    return {f"rate_{int(r)}": served_at(r, duration) for r in rates}


def served_at(rate, duration):
    from freeflow_amd import serve
    out = {}
    for batched in (True, False):
        eps = serve.make_episodes(rate, duration, seed=1234)
        # batched: FP16 storage, fast math; plain: the reference's FP32 form, one domain per episode
        srv = serve.Server(storage="f16" if batched else "f32", batched=batched, fast_math=batched)
        print(f"[bench] served {'batched FP16' if batched else 'plain FP32'}: {len(eps)} episodes at {rate}/s",
              file=sys.stderr, flush=True)
        r = srv.run(eps, realtime=True)
        srv.close()
        out["batched_fp16" if batched else "plain_fp32"] = {
            "mlups": round(r["aggregate_mlups"], 2), "p50_step_ms": round(r["p50_step_ms"], 4),
            "p99_step_ms": round(r["p99_step_ms"], 4), "episodes_done": r["episodes_done"],
            "episodes_rejected": r["episodes_rejected"], "peak_active": r["peak_active"], "wall_s": round(r["wall_s"], 3),
            "phase_s": r["phase_s"]}
    out["speedup_vs_plain_fp32"] = round(out["batched_fp16"]["mlups"] / max(out["plain_fp32"]["mlups"], 1e-9), 3)
    out["rate_per_s"] = rate
    out["arrival_window_s"] = duration
    return out


def run_tg3d(args, capi, wl):
    c = wl.TG3D
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    st = args.storage
    K = args.steps if args.steps is not None else 1000
    W = min(args.warmup, 20)
    m = wl.taylor_green_ring(nx, ny, nz, c["amp"], c["ring_r"], c["seed"])
    fast = st == "f32" and not args.reference_order or args.fast_math
    main_ = tg3d_time(capi, wl, m, st, fast, K, W, args.tile, not args.no_tune)
    extra = {}
    forms = {("fast_math" if fast else "reference_order"): (st, fast, main_["tile"])}

    def summary(r):
        return {"mlups": round(r["mlups"], 2), "ms_per_step": round(r["ms_per_step"], 5), "tile": TILES3[r["tile"]],
                "kernel": r["kernel"], "avg_kernel_us": round(r["avg_kernel_ms"] * 1e3, 3),
                "bytes_per_lup": r["bytes_per_lup"], "frac": round(r["frac"], 4)}
    if st == "f32":
        ref = tg3d_time(capi, wl, m, st, not fast, K, W, -1, not args.no_tune)
        key = "reference_order" if fast else "fast_math"
        forms[key] = (st, not fast, ref["tile"])
        extra[key] = summary(ref)
        # the 16-bit storage forms of the same workload (fast math)
        extra["storage_forms"] = {}
        for s16 in ("f16", "f16u"):
            r = tg3d_time(capi, wl, m, s16, True, K, W, -1, not args.no_tune)
            forms[s16 + "_fast_math"] = (s16, True, r["tile"])
            extra["storage_forms"][s16] = summary(r)
    cpu = None
    if not args.no_cpu and st == "f32":
        cpu, acc = cpu_baseline_tg3d(capi, wl, m, args.check_steps, forms)
        acc["mass_drift_timed_run"] = main_["mass_drift"]
        extra["accuracy"] = acc
    if not args.no_extras:
        print("[bench] extras: swimmer2d", file=sys.stderr, flush=True)
        extra["swimmer2d"] = swimmer2d_summary(capi, wl)
        print("[bench] extras: batched", file=sys.stderr, flush=True)
        extra["batched"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in
                            batched_summary(capi, wl, 1000, 20, 100, fp64_check=not args.no_cpu).items()}
        print("[bench] extras: served", file=sys.stderr, flush=True)
        extra["served"] = served_summary()
    roof = roofline(main_["bytes_per_lup"] * main_["cells"] / (main_["avg_kernel_ms"] * 1e-3) / 1e9, main_["kernel"],
                    main_["avg_kernel_ms"], main_["bytes_per_lup"], main_["cells"])
    arith = ("FP32, reassociated with FMA (FF_OPT_FAST_MATH)" if fast else "FP32 in the reference's operation order")
    out = base_line(main_["mlups"], K, W, main_["ms_per_step"], f"f32 compute, {st} storage",
                    {"workload": f"3D D3Q27 {nx}x{ny}x{nz} Taylor-Green (amplitude 0.04) + vortex ring r=20, "
                                 "periodic, viscosity 0.01", "nx": nx, "ny": ny, "nz": nz, "storage": st,
                     "arithmetic": arith, "tile": TILES3[main_["tile"]], "tile_index": main_["tile"],
                     "parallelism": "single-gpu"},
                    roof, cpu, mass_drift=main_["mass_drift"], **extra)
    print(json.dumps(out), flush=True)
    return 0


def run_served(args):
    from freeflow_amd import serve
    eps = serve.make_episodes(args.rate, args.duration, seed=1234)
    srv = serve.Server(storage="f16")
    print(f"[bench] served: {len(eps)} episodes at {args.rate}/s over {args.duration} s", file=sys.stderr, flush=True)
    summ = srv.run(eps, realtime=True)
    srv.close()
    extra = {}
    if args.plain_steps > 0:
        # the same arrivals served in the plain per-domain FP32 form
        plain = serve.Server(storage="f32", batched=False, fast_math=False)
        print("[bench] served, plain per-domain FP32 form", file=sys.stderr, flush=True)
        ps = plain.run(serve.make_episodes(args.rate, args.duration, seed=1234), realtime=True)
        plain.close()
        extra = dict(plain_fp32_mlups=round(ps["aggregate_mlups"], 2),
                     plain_fp32_step_latency_ms={"p50": ps["p50_step_ms"], "p99": ps["p99_step_ms"]},
                     speedup_vs_plain_fp32=round(summ["aggregate_mlups"] / max(ps["aggregate_mlups"], 1e-9), 3))
    out = base_line(summ["aggregate_mlups"], summ["iterations"], 0, summ["wall_s"] * 1e3 / max(summ["iterations"], 1),
                    "f32 compute, f16 storage",
                    {"workload": f"Poisson rollout episodes at {args.rate}/s for {args.duration} s: 70% 2D "
                                 "(600x200..1200x400 swimmer), 30% 3D (96..192 x 96 x 96 sphere), 500 steps, "
                                 "action every 10 steps", "rate": args.rate, "duration": args.duration,
                     "parallelism": "single-gpu"}, None, None,
                    step_latency_ms={"p50": summ["p50_step_ms"], "p99": summ["p99_step_ms"]},
                    served=summ, **extra)
    print(json.dumps(out), flush=True)
    return 0


def copy_probe(mib):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings turn
    the counters of the stream-collide kernel into bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "copies": 5}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: tg3d 1000, swimmer2d 4000, batched 1000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tg3d")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shape instead of timing the candidates")
    ap.add_argument("--tile", type=int, default=-1, help="use this tile index instead of timing the candidates")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--storage", default="f32", help="tg3d: f32, f16, bf16 or f16u")
    ap.add_argument("--fast-math", action="store_true",
                    help="tg3d: reassociated FP32 arithmetic with FMA in the column tiles (FF_OPT_FAST_MATH); the "
                         "default for FP32 storage")
    ap.add_argument("--reference-order", action="store_true",
                    help="tg3d: time the reference's operation order as the headline (FP32)")
    ap.add_argument("--check-steps", type=int, default=20,
                    help="tg3d: steps of the full-grid oracle comparison (the CPU leg)")
    ap.add_argument("--no-extras", action="store_true",
                    help="tg3d: leave out the swimmer2d, batched and served figures")
    ap.add_argument("--rate", type=float, default=50.0, help="served: episode arrivals per second")
    ap.add_argument("--duration", type=float, default=5.0, help="served: arrival window in seconds")
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="batched: steps of the per-domain FP32 form; served: 0 skips that form")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the reference kernel that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(independent domains shard across GPUs without communication); use --gpus 1",
              file=sys.stderr)
        return 2
    from freeflow_amd import capi, workloads as wl

    if args.workload == "batched":
        return run_batched(args, capi, wl)
    if args.workload == "tg3d":
        return run_tg3d(args, capi, wl)
    if args.workload == "served":
        return run_served(args)
    if args.workload != "swimmer2d":
        print(f"unknown workload {args.workload}", file=sys.stderr)
        return 2
    d, state = setup_swimmer2d(capi, wl)
    tile = pick_tile(d, args)
    c = wl.SWIMMER2D
    cells = c["nx"] * c["ny"]
    K = args.steps if args.steps is not None else 4000
    W = args.warmup
    d.set_wrench_history(K)
    d.step(W)
    d.sync()
    print(f"[bench] warmup {W} done; timing {K} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    d.step(K)
    hist = d.wrench_history(K)  # per-step FP64 force/torque readback (synchronises)
    d.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = elapsed * 1e3 / K
    mlups = cells * K / elapsed / 1e6
    healthy = bool(np.all(np.isfinite(hist)))
    # per-step latency (launch + completion of one step)
    lat = []
    for _ in range(200):
        a = time.perf_counter()
        d.step(1)
        d.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    # roofline of the dominant kernel (stream+collide), HIP events attached to
    # each launch of real steps
    iters = 400
    _, stream_ms = d.step_timed(iters)
    avg_ms = stream_ms / iters
    bpl = bytes_per_lup(2, 4)
    achieved = bpl * cells / (avg_ms * 1e-3) / 1e9
    kname = kernel2d("k_step2d", "float", True, False, tile)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname),
            "kernel": kname + " (stream-collide tiles + boundary cells + next-step kinematics)",
            "avg_kernel_us": round(avg_ms * 1e3, 3),
            "bytes_per_lup": bpl, "cells_per_launch": cells}
    cpu = None if args.no_cpu else cpu_baseline_swimmer2d(state, args.cpu_budget)
    out = {
        "metric": METRIC, "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D FSI plain FP32: D2Q9 1200x400, 600-segment elliptical swimmer with seeded "
                               "travelling wave, outlets on all sides, per-step FP64 wrench kept in a device ring "
                               "read back once after the K steps (moments cache-resident: not an HBM figure)",
                   "nx": c["nx"], "ny": c["ny"], "dx": c["dx"], "dt": c["dt"], "viscosity": c["viscosity"],
                   "segments": c["n_seg"], "storage": "f32", "stage_fp16": True, "tile": TILES2[tile],
                   "parallelism": "single-gpu"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                            "p99": round(float(np.percentile(lat, 99)), 4)},
        "wrench_records": int(hist.shape[0]),
        "healthy": healthy,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
