"""Benchmark of the HOME-LBM fluid step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload swimmer2d]

Default workload (the configuration the headline metric is quoted on): 2D FSI,
plain FP32, D2Q9 1200x400, dx=0.005, dt=0.0025, an elliptical swimmer
(semi-axes 0.3 x 0.08 m, 600 boundary segments) driven by a seeded travelling
wave, all four sides outlets, with the per-step FP64 body force/torque kept in
a device ring that is read back inside the timed region.  A step is one fluid
step (wave kinematics, link detection, stream+collide, moving-wall links,
force reduction, outlet update) over the whole grid.

Prints one JSON line.  `value` is MLUPS (cells x steps / s / 1e6) with all
inputs resident on the GPU.  `roofline` prices the stream-collide kernel
(the dominant one) against 8 TB/s HBM with 49 algorithmic bytes per cell
update (6 FP32 moments read + 6 written + 1 flag byte); `cpu_baseline` is the
CPU oracle (a C++ restatement of the reference update, OpenMP) timed on this
host on a bounded number of steps of the same workload.

Other workloads (one JSON line each, same keys):
  --workload batched   64 uneven 2D swimmer domains (256x128 .. 1200x400,
                       200-800 segments) in one batch, FP16 storage, one graph
                       replay per step; also times the plain per-domain FP32
                       form of the same domains and reports the ratio.
  --workload tg3d   3D D3Q27 256x128x128 Taylor-Green + vortex ring,
                       periodic, storage from --storage (f32, f16, bf16).
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
TILES3 = ["8x8x4 box", "16x4x4 box", "16x8 column x32", "16x8 column x16", "16x8 column x8", "16x16 column x32",
          "16x16 column x16"]


def bytes_per_lup(dim, storage_bytes):
    nm = 6 if dim == 2 else 10
    return 2 * nm * storage_bytes + 1


def setup_swimmer2d(capi, wl):
    c = wl.SWIMMER2D
    nx, ny = c["nx"], c["ny"]
    flags = wl.outlet_box_flags(nx, ny)
    m = wl.equilibrium_moments(np.zeros((nx * ny, 2)))
    sw = wl.swimmer2d_swimmer()
    d = capi.Domain(**wl.swimmer2d_domain_args(), storage="f32", stage_fp16=True)
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


# (tx, ty, cells per thread) of the 2D tiles and (tx, ty, planes, ring, threads)
# of the 3D tiles, by tile index (ff_set_tile)
SHAPES2 = [(32, 8, 1), (32, 8, 2), (64, 4, 2), (32, 8, 4), (64, 4, 4)]
SHAPES3 = [(8, 8, 4, 0, 256), (16, 4, 4, 0, 256), (16, 8, 32, 1, 128), (16, 8, 16, 1, 128), (16, 8, 8, 1, 128),
           (16, 16, 32, 1, 256), (16, 16, 16, 1, 256)]


def kernel2d(prefix, s, stage, periodic, tile):
    tx, ty, cpt = SHAPES2[tile]
    return f"ff::{prefix}<{s}, {str(stage).lower()}, {str(periodic).lower()}, {tx}, {ty}, {cpt}>"


def kernel3d(s, periodic, tile):
    tx, ty, tz, ring, nt = SHAPES3[tile]
    return f"ff::k_bulk3d<{s}, {str(periodic).lower()}, {ring}, {tx}, {ty}, {tz}, {nt}>"


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


def run_batched(args, capi, wl):
    c = wl.BATCH64
    scenes = wl.batched_scenes(c["n_domains"], c["seed"])
    cells = sum(s.nx * s.ny for s in scenes)
    K = args.steps if args.steps != 4000 else c["steps"]
    W = min(args.warmup, 50)
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
    Kp = args.plain_steps
    print(f"[bench] plain FP32 form: {len(plain)} domains, {cells} cells, {Kp} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(Kp):
        for d in plain:
            d.step(1)
    for d in plain:
        d.sync()
    t_plain = time.perf_counter() - t0
    plain_mlups = cells * Kp / t_plain / 1e6
    for d in plain:
        d.close()
    del plain
    # batched FP16-storage form, FP32 arithmetic
    b = capi.Batch(2, c["storage"], stage_fp16=False)
    doms = []
    for sc in scenes:
        d = capi.Domain(**sc.domain_args(), storage=c["storage"], stage_fp16=False)
        sc.setup(d)
        b.add(d)
        doms.append(d)
    if args.tile >= 0:
        b.set_tile(args.tile)
    b.step(W)
    b.sync()
    print(f"[bench] batched FP16 form: warmup {W} done; timing {K} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b.step(K)
    b.sync()
    el = time.perf_counter() - t0
    mlups = cells * K / el / 1e6
    tile, tx, ty = b.tile()
    avg_ms = b.time_bulk_kernel(100)
    bpl = bytes_per_lup(2, 2)
    roof = roofline(bpl * cells / (avg_ms * 1e-3) / 1e9, kernel2d("kb_step2d", "__half", False, False, tile), avg_ms,
                    bpl, cells)
    wr = np.concatenate([d.body_wrench() for d in doms])
    out = base_line(mlups, K, W, el * 1e3 / K, "f32 compute, f16 storage",
                    {"workload": "64 uneven 2D swimmer domains (256x128..1200x400, 200-800 segments) in one batch, "
                                 "FP16 storage, one graph replay per step", "domains": len(scenes),
                     "cells": cells, "tile": [tx, ty], "parallelism": "single-gpu"},
                    roof, None, plain_fp32_mlups=round(plain_mlups, 2),
                    speedup_vs_plain_fp32=round(mlups / plain_mlups, 3),
                    healthy=bool(np.all(np.isfinite(wr))))
    print(json.dumps(out), flush=True)
    return 0


def run_tg3d(args, capi, wl):
    c = wl.TG3D
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    cells = nx * ny * nz
    st = args.storage
    d = capi.Domain(3, nx, ny, nz, viscosity=c["viscosity"], storage=st, periodic=True)
    d.set_moments(wl.taylor_green_ring(nx, ny, nz, c["amp"], c["ring_r"], c["seed"]))
    K = args.steps if args.steps != 4000 else 1000
    W = min(args.warmup, 20)
    tile = pick_tile(d, args)
    m0 = d.mass()
    d.step(W)
    d.sync()
    print(f"[bench] tg3d {st}: timing {K} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    d.step(K)
    d.sync()
    el = time.perf_counter() - t0
    drift = abs(d.mass() - m0) / m0
    avg_ms = d.time_bulk_kernel(50)
    bpl = bytes_per_lup(3, 4 if st == "f32" else 2)
    sname = {"f32": "float", "f16": "__half", "bf16": "__hip_bfloat16"}[st]
    roof = roofline(bpl * cells / (avg_ms * 1e-3) / 1e9, kernel3d(sname, True, tile), avg_ms, bpl, cells)
    out = base_line(cells * K / el / 1e6, K, W, el * 1e3 / K, f"f32 compute, {st} storage",
                    {"workload": "3D D3Q27 256x128x128 Taylor-Green (amplitude 0.04) + vortex ring r=20, periodic",
                     "nx": nx, "ny": ny, "nz": nz, "storage": st, "tile": TILES3[tile], "parallelism": "single-gpu"},
                    roof, None, mass_drift=drift)
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
        plain = serve.Server(storage="f32", batched=False)
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
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="swimmer2d")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shape instead of timing the candidates")
    ap.add_argument("--tile", type=int, default=-1, help="use this tile index instead of timing the candidates")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--storage", default="f16", help="tg3d: f32, f16 or bf16")
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
    K, W = args.steps, args.warmup
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
    # roofline of the dominant kernel (stream+collide), HIP events per launch
    iters = 400
    avg_ms = d.time_bulk_kernel(iters)
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
                               "travelling wave, outlets on all sides, per-step FP64 wrench readback",
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
