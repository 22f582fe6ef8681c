#!/usr/bin/env python3
"""Throughput of the meshing path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload csg64-512] [--gpus 1]

A step meshes one job (SDF lattice evaluation, contouring, normals, quads) with the
program and parameters already on the device; the mesh stays in HBM.  Mvoxels/s =
res^3 / 1e6 / seconds per step.  Prints one JSON line with:

  roofline      the SDF evaluation kernel (k_sp_tiles on the tile-pruned path: the
                pruned program at the tile's lattice points fused with contouring, one
                launch per pass for left-fold programs; k_eval_lattice on the dense path,
                one per slab): algorithmic fp32 flops executed per launch (static op
                count of the program each tile runs x the lattice points it evaluates) /
                mean launch time from HIP events recorded around the tile launches on the
                library's stream, against the fp32 peak (vector ALU: the path has no
                matrix work).  traffic: HBM bytes per launch from the rocprofv3
                FETCH_SIZE / WRITE_SIZE summary in profiles/hbm_traffic.json
  cpu_baseline  the oracle (CPU restatement of the reference algorithm: 8 corner
                evaluations per cell, OpenMP) on a bounded sample of z-slices of the
                same job
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X fp32 (vector and matrix) peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E


def workload(name: str):
    import sdf2mesh_amd as s2m
    from sdf2mesh_amd import tape

    if name == "csg64-512":
        return tape.csg64(), s2m.Params.reference(512, 2.0), {
            "workload": "csg64-512", "sdf": "smooth union (k=0.05) of 64 seeded primitives",
            "res": 512, "bounds": 2.0, "corners": "exact (reference per-cell coordinates)"}
    if name == "csg64-512-lattice":
        return tape.csg64(), s2m.Params.reference(512, 2.0, flags=s2m.COMPAT_POW2), {
            "workload": "csg64-512-lattice", "res": 512, "bounds": 2.0,
            "corners": "shared lattice"}
    if name == "torus-128":
        return tape.torus_example(), s2m.Params.reference(128, 2.0), {
            "workload": "torus-128", "res": 128, "bounds": 2.0}
    if name == "mandelbulb-1024":
        return tape.mandelbulb(8), s2m.Params.reference(1024, 2.5), {
            "workload": "mandelbulb-1024", "res": 1024, "bounds": 2.5,
            "sdf": "power-8 Mandelbulb, 8 iterations"}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(node, params, seconds: float) -> dict:
    import oracle

    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    res = [int(r) for r in params.res]
    mid = res[2] // 2
    t0 = time.perf_counter()
    cells, _ = oracle.bench_slices(node, params.bb_min, params.bb_max, res, mid, mid + 1,
                                   params.eps)
    one = time.perf_counter() - t0
    n = max(1, min(res[2] - mid, int(seconds / max(one, 1e-6))))
    t0 = time.perf_counter()
    cells, active = oracle.bench_slices(node, params.bb_min, params.bb_max, res, mid, mid + n,
                                        params.eps)
    dt = time.perf_counter() - t0
    return {"value": cells / dt / 1e6, "unit": "Mvoxels/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", threads)), "kind": "port",
            "sample": f"{n} z-slices ({cells} cells, {active} active) from z={mid} of the "
                      f"{res[0]}^3 job, reference algorithm (8 corner evaluations + contour + "
                      f"normal per cell), {dt:.1f} s"}


def run_batch(args) -> dict:
    """Batching workload: 256 seeded random CSG jobs of mixed resolution.  A step meshes
    the whole batch in shared device passes (s2m_mesh_batch_device); the comparison is
    the same 256 jobs meshed one at a time on the plain GPU path (s2m_mesh_device)."""
    import sdf2mesh_amd as s2m
    from sdf2mesh_amd.workloads import batch_jobs

    n = 256
    nodes, params = batch_jobs(n)
    voxels = sum(int(p.res[0]) * int(p.res[1]) * int(p.res[2]) for p in
                 [s2m.Params(q.bb_min, q.bb_max, tuple(s2m.round_resolution(r) for r in q.res))
                  for q in params])
    ctx = s2m.Context(0)
    progs = [ctx.program(x) for x in nodes]
    for _ in range(args.warmup):
        ctx.mesh_batch_device(progs, params)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ctx.mesh_batch_device(progs, params)
    dt_b = (time.perf_counter() - t0) / args.steps
    for q, p in zip(progs, params):  # warm the single-job path
        ctx.mesh_device(q, p)
    solo_lat = []
    t0 = time.perf_counter()
    for q, p in zip(progs, params):
        t1 = time.perf_counter()
        ctx.mesh_device(q, p)
        solo_lat.append(time.perf_counter() - t1)
    dt_s = time.perf_counter() - t0
    ctx.close()
    jobs_b, jobs_s = n / dt_b, n / dt_s
    lat = np.sort(np.asarray(solo_lat)) * 1e3
    return {
        "metric": "batched jobs/s (256 random CSG jobs, res 32-256)",
        "value": round(jobs_b, 1),
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_b * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded CSG programs)",
        "config": {"workload": "batch256", "jobs": n, "voxels_per_batch": voxels,
                   "vertices": int(sum(o["nvtx"] for o in out)),
                   "triangles": int(sum(o["ntri"] for o in out))},
        "one_at_a_time_jobs_per_s": round(jobs_s, 1),
        "speedup_vs_one_at_a_time": round(jobs_b / jobs_s, 2),
        "one_at_a_time_latency_ms": {"p50": round(float(lat[len(lat) // 2]), 3),
                                     "p99": round(float(lat[int(len(lat) * 0.99)]), 3)},
        "batched_mvoxels_per_s": round(voxels / 1e6 / dt_b, 1),
    }


def run_serve(args) -> dict:
    """Serving workload: Poisson request streams (rates 10-2000 jobs/s) of the seeded
    request mix (resolution log-uniform in [32, 512], 80 % random CSG, 20 % Mandelbulb
    with 4-8 iterations) through the continuous-batching server.  The comparison is the
    plain GPU path serving the same stream one request at a time (FIFO): its latencies
    are replayed from each request's measured service time on one context."""
    import sdf2mesh_amd as s2m
    from sdf2mesh_amd.workloads import poisson_arrivals, serve_jobs

    n = args.serve_requests
    nodes, params = serve_jobs(n)
    tapes = [x.tape() for x in nodes]
    # one at a time: measured service time of every request (mesh to host memory)
    ctx = s2m.Context(0)
    progs = [ctx.program(t) for t in tapes]
    for q, p in zip(progs[:8], params[:8]):
        ctx.mesh(q, p)
    service = []
    for q, p in zip(progs, params):
        t0 = time.perf_counter()
        ctx.mesh(q, p)
        service.append(time.perf_counter() - t0)
    for q in progs:
        q.close()
    ctx.close()
    service = np.asarray(service)
    solo_jobs_s = n / service.sum()

    rates = [float(r) for r in args.serve_rates.split(",")]
    per_rate = []
    with s2m.Server(0, workers=args.serve_workers, max_batch=args.serve_batch) as srv:
        for t, p in zip(tapes[:8], params[:8]):  # warm every worker's buffers
            srv.take(srv.submit(t, p))
        for rate in rates:
            arr = np.asarray(poisson_arrivals(n, rate, seed=int(rate)))
            tickets = []
            t0 = time.perf_counter()
            for a, t, p in zip(arr, tapes, params):
                now = time.perf_counter() - t0
                if a > now:
                    time.sleep(a - now)
                tickets.append(srv.submit(t, p))
            res = [srv.take(k) for k in tickets]
            elapsed = time.perf_counter() - t0
            lat = np.asarray([st["latency_ms"] for _, st in res])
            bs = np.asarray([st["batch_size"] for _, st in res])
            # FIFO replay of the one-at-a-time server on the same arrivals
            done, solo_lat = 0.0, []
            for a, sv in zip(arr, service):
                done = max(done, a) + sv
                solo_lat.append((done - a) * 1e3)
            solo_lat = np.asarray(solo_lat)
            per_rate.append({
                "rate": rate, "jobs_per_s": round(n / elapsed, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "mean_batch": round(float(bs.mean()), 2),
                "one_at_a_time_jobs_per_s": round(n / max(done, arr[-1]), 1),
                "one_at_a_time_p50_ms": round(float(np.percentile(solo_lat, 50)), 2),
                "one_at_a_time_p99_ms": round(float(np.percentile(solo_lat, 99)), 2),
            })
    top = per_rate[-1]
    return {
        "metric": "served jobs/s at the highest offered rate, with p50/p99 latency",
        "value": top["jobs_per_s"],
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": 8,
        "ms_per_step": round(1e3 / top["jobs_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": "serve", "requests_per_rate": n, "workers": args.serve_workers,
                   "max_batch": args.serve_batch},
        "p50_ms": top["p50_ms"],
        "p99_ms": top["p99_ms"],
        "one_at_a_time_jobs_per_s": round(solo_jobs_s, 1),
        "speedup_vs_one_at_a_time": round(top["jobs_per_s"] / solo_jobs_s, 2),
        "per_rate": per_rate,
    }


def run(args) -> dict:
    import sdf2mesh_amd as s2m

    node, params, cfg = workload(args.workload)
    if args.dense:
        params.flags |= s2m.DENSE_EVAL
    voxels = int(params.res[0]) * int(params.res[1]) * int(params.res[2])
    ctx = s2m.Context(0)
    prog = ctx.program(node)
    for _ in range(args.warmup):
        ctx.mesh_device(prog, params)
    ctx.set_profiling(1)  # events around the whole call and the evaluation launches only
    agg = {"prune_ms": 0.0, "eval_ms": 0.0, "contour_ms": 0.0, "normal_ms": 0.0, "quad_ms": 0.0,
           "total_ms": 0.0, "eval_launches": 0, "lattice_points": 0, "slabs": 0,
           "tiles_total": 0, "tiles_active": 0, "eval_flops": 0}
    raw = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dm = ctx.mesh_device(prog, params)
        raw.append(ctx.last_timing_raw())  # converted after the timed region
    dt = time.perf_counter() - t0
    for r in raw:
        t = s2m.timing_dict(r)
        for k in agg:
            agg[k] += t[k]
    ms = dt / args.steps * 1e3
    value = voxels / 1e6 / (ms / 1e3)
    # stage split (every stage bracketed by events), outside the timed region
    ctx.set_profiling(2)
    stages = {k: 0.0 for k in ("prune_ms", "eval_ms", "contour_ms", "normal_ms", "quad_ms", "total_ms")}
    n_stage = 3
    for _ in range(n_stage):
        ctx.mesh_device(prog, params)
        t = ctx.last_timing()
        for k in stages:
            stages[k] += t[k]

    flops = node.flops()
    launches = max(1, agg["eval_launches"])
    eval_avg_ms = agg["eval_ms"] / launches
    pts_per_launch = agg["lattice_points"] / launches
    flops_per_launch = agg["eval_flops"] / launches   # executed algorithmic flops
    achieved = flops_per_launch / (eval_avg_ms / 1e3) / 1e12
    tiled = agg["tiles_total"] > 0
    cfg["eval"] = ("tile-pruned (exact)" if tiled else
                   "dense" if args.dense else "dense (the program is not interval-prunable)")
    out = {
        "metric": "Mvoxels/s meshed at 512^3 fp32 (CSG-64, reference-exact corners)"
        if args.workload == "csg64-512" else f"Mvoxels/s meshed ({args.workload})",
        "value": round(value, 1),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SDF program)",
        "config": {**cfg, "nvtx": dm["nvtx"], "ntri": dm["ntri"],
                   "invalid_quads": dm["invalid_quads"], "flops_per_point": flops,
                   "lattice_points_per_step": agg["lattice_points"] // args.steps,
                   "tiles_active_per_step": agg["tiles_active"] // args.steps,
                   "tiles_per_step": agg["tiles_total"] // args.steps,
                   "slabs": agg["slabs"] // args.steps},
        "roofline": {
            "bound": "valu",
            "kernel": "k_sp_tiles" if tiled else "k_eval_lattice",
            "achieved": round(achieved, 2),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": measured_traffic(args.workload + ("-dense" if args.dense else ""),
                                        "k_sp_tiles" if tiled else "k_eval_lattice"),
            "launch_ms": round(eval_avg_ms, 4),
            "flops_per_launch": flops_per_launch,
            "points_per_launch": pts_per_launch,
        },
        "device_ms_per_step": round(agg["total_ms"] / args.steps, 3),
        "stages_ms_per_step": {k: round(v / n_stage, 3) for k, v in stages.items()},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        cb = cpu_baseline(node, params, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_baseline"] = round(value / cb["value"], 1)
    return out


def run_copy_calibration(args) -> dict:
    """A plain device copy of known size (1 GiB read + 1 GiB written per launch), whose
    FETCH_SIZE / WRITE_SIZE counters scale the counters of the meshing kernels to bytes."""
    import torch

    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"metric": "copy calibration", "launches": args.steps, "bytes_read_per_launch": 4 * n,
            "bytes_written_per_launch": 4 * n}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/hbm_traffic.json (scaled by the copy calibration), or None."""
    f = ROOT / "profiles" / "hbm_traffic.json"
    if not f.exists():
        return None
    entry = json.loads(f.read_text()).get(workload, {})
    if entry.get("kernel") != kernel:
        return None
    return entry.get("bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="csg64-512")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="10,30,100,300,1000,2000")
    ap.add_argument("--serve-workers", type=int, default=2)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--dense", action="store_true",
                    help="evaluate every lattice point with the full program (no tile pruning)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; jobs shard by job or by z-slab "
                                   "across GPUs only by design, --gpus must be 1"}))
        return 2
    if args.workload == "copy-calibration":
        out = run_copy_calibration(args)
    elif args.workload == "batch256":
        out = run_batch(args)
    elif args.workload == "serve":
        out = run_serve(args)
    else:
        out = run(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
