#!/usr/bin/env python3
"""Throughput of the meshing path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload NAME] [--only]

Headline (``value``): CSG-64 at 512^3, bounds 2, reference-exact corners.  A step is one
meshing job, timed from the call with the program and parameters already on the device
until the vertex (position + normal) and triangle index buffers are in host memory
(s2m_mesh: the library copies them into page-locked host memory it recycles), the span
of the reference's own t0 -> t1 (main.rs:292-357, which includes the per-slice readback).
Mvoxels/s = res^3 / 1e6 / median step time.  ``device_resident`` gives the same job with
the mesh left in HBM (s2m_mesh_device).

The same JSON line carries the other configurations under ``workloads``: torus-128,
mandelbulb-1024 (host landing, median step), batch256 (256 random CSG jobs per step,
s2m_mesh_batch, against the same jobs one at a time with s2m_mesh) and serve (Poisson
request streams through the continuous-batching server at 10-2000 requests/s, p50/p99
latency, against one-at-a-time meshing of the same stream).  ``--workload NAME --only``
runs one of them alone (profiling runs).

  roofline      the SDF evaluation kernel (k_sp_tiles on the tile-pruned path: the pruned
                program at each live tile's lattice points fused with contouring;
                k_eval_lattice, or k_eval_mandelbulb for a one-op Mandelbulb tape, on the
                dense path): executed pruned-program fp32 flops per
                launch (static op count of the program each tile runs x the lattice points
                it evaluates, counted on the device) / mean launch time from HIP events
                recorded around the launches on the library's stream during the timed
                steps.  Priced against the fp32 peak (157.3 TFLOP/s) and, as
                frac_of_issue_ceiling, against 39.3 T lane-ops/s (one non-FMA, non-packed
                fp32 op per lane per clock: the ceiling of code compiled without
                contraction or packed fp32).  traffic: HBM bytes per launch from the
                rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/hbm_traffic.json.
  cpu_baseline  the oracle (CPU restatement of the reference algorithm: 8 corner
                evaluations per cell, OpenMP) on a bounded sample of z-slices of the same
                job, median rate over 10 chunks
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import queue
import sys
import threading
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X fp32 peak (FMA, packed)
ISSUE_CEILING_TOPS = 39.3  # 256 CUs x 64 lanes x 2.4 GHz: one plain fp32 op per lane per clock
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E

SINGLE = ("csg64-512", "csg64-512-lattice", "torus-128", "mandelbulb-1024")


class CopyProbe:
    """The device-to-host copy rate right now: two 16 MB copies from device memory into
    page-locked host memory (HIP runtime through ctypes), timed on the host.

    Device memory that is freed is cleared by the kernel driver in the background on a
    DMA engine, at about 10 GB/s; meanwhile copies to host memory run at half rate (1 GB
    freed: ~60 ms, 8 GB: ~0.8 s).  Tuning regrows the library's pass buffers once per
    candidate shape, so a workload's timed steps start only after ``settle`` has seen the
    rate back near the link rate measured when the bench started (the library frees no
    device memory in steady state).  Tuning waits for it too: the contexts of earlier
    workloads are closed before the next one tunes."""

    def __init__(self):
        import ctypes as C

        self.C = C
        hip = C.CDLL("libamdhip64.so")
        for f, a in {"hipMalloc": [C.POINTER(C.c_void_p), C.c_size_t],
                     "hipHostMalloc": [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint],
                     "hipMemcpyAsync": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p],
                     "hipStreamSynchronize": [C.c_void_p]}.items():
            getattr(hip, f).restype = C.c_int
            getattr(hip, f).argtypes = a
        self.hip, self.n = hip, 16 << 20
        self.dev, self.host = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(self.dev), self.n) or hip.hipHostMalloc(C.byref(self.host), self.n, 0):
            raise RuntimeError("copy probe: allocation failed")
        self.rate()
        # the largest rate seen, once it has not grown by 5 % for 0.5 s (at most 3 s: memory
        # freed by a process that ended just before may still be being cleared)
        t0 = t_up = time.perf_counter()
        self.link = 0.0
        while time.perf_counter() - t_up < 0.5 and time.perf_counter() - t0 < 3.0:
            r = self.rate()
            if r > 1.05 * self.link:
                t_up = time.perf_counter()
            self.link = max(self.link, r)

    def rate(self) -> float:
        t0 = time.perf_counter()
        for _ in range(2):
            if self.hip.hipMemcpyAsync(self.host, self.dev, self.n, 2, None):  # device to host
                raise RuntimeError("copy probe: hipMemcpyAsync failed")
        if self.hip.hipStreamSynchronize(None):
            raise RuntimeError("copy probe: hipStreamSynchronize failed")
        return 2 * self.n / (time.perf_counter() - t0) / 1e9

    def settle(self, max_s: float = 5.0) -> dict:
        """Wait (at most max_s) until three rates in a row reach 85 % of the link rate."""
        t0, ok, r = time.perf_counter(), 0, 0.0
        while ok < 3 and time.perf_counter() - t0 < max_s:
            r = self.rate()
            ok = ok + 1 if r >= 0.85 * self.link else 0
        return {"link_GB_per_s": round(self.link, 1), "waited_ms": round((time.perf_counter() - t0) * 1e3, 1),
                "GB_per_s": round(r, 1)}


_PROBE: CopyProbe | None = None


def settle_copies() -> dict | None:
    return _PROBE.settle() if _PROBE is not None else None


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
            "workload": "torus-128", "sdf": "examples/torus.sdf3d", "res": 128, "bounds": 2.0}
    if name == "mandelbulb-1024":
        return tape.mandelbulb(8), s2m.Params.reference(1024, 2.5), {
            "workload": "mandelbulb-1024", "res": 1024, "bounds": 2.5,
            "sdf": "power-8 Mandelbulb, 8 iterations"}
    raise SystemExit(f"unknown workload {name}")


def cpu_info() -> dict:
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": allowed}


def cpu_baseline(node, params, seconds: float) -> dict:
    """The oracle on z-slices through the middle of the job: ~`seconds` of CPU work cut
    into 10 equal chunks, median chunk rate.  Threads: OMP_NUM_THREADS when set (the
    CPU share given to this process), else every CPU this process may run on."""
    import oracle

    info = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["affinity_cpus"]
    os.environ["OMP_NUM_THREADS"] = str(threads)
    res = [int(r) for r in params.res]
    mid = res[2] // 2
    t0 = time.perf_counter()
    oracle.bench_slices(node, params.bb_min, params.bb_max, res, mid, mid + 1, params.eps)
    one = time.perf_counter() - t0
    per_chunk = max(1, min((res[2] - mid) // 10, int(seconds / 10 / max(one, 1e-6))))
    rates, total_cells, total_active, total_s = [], 0, 0, 0.0
    for k in range(10):
        z0 = mid + k * per_chunk
        t0 = time.perf_counter()
        cells, active = oracle.bench_slices(node, params.bb_min, params.bb_max, res, z0,
                                            z0 + per_chunk, params.eps)
        dt = time.perf_counter() - t0
        rates.append(cells / dt / 1e6)
        total_cells += cells
        total_active += active
        total_s += dt
    return {"value": float(np.median(rates)), "unit": "Mvoxels/s", "cores": threads,
            "kind": "port", **info,
            "sample": f"10 chunks of {per_chunk} z-slices ({total_cells} cells, {total_active} "
                      f"active) from z={mid} of the {res[0]}^3 job, reference algorithm (8 corner "
                      f"evaluations + contour + normal per cell), {total_s:.1f} s in all; "
                      f"median chunk rate, {threads} OpenMP threads"}


def _median_ms(xs) -> float:
    return float(np.median(np.asarray(xs))) * 1e3


def single_job(name: str, args, ctx, headline: bool) -> dict:
    """Host-landing and device-resident step times of one single-job workload, with the
    evaluation kernel's roofline from events recorded during the host-landing steps."""
    import sdf2mesh_amd as s2m
    from sdf2mesh_amd import tape

    node, params, cfg = workload(name)
    if args.dense:
        params.flags |= s2m.DENSE_EVAL
    if args.slab:
        params.slab = args.slab
        cfg = {**cfg, "slab": args.slab}
    voxels = int(params.res[0]) * int(params.res[1]) * int(params.res[2])
    prog = ctx.program(node)
    settle_copies()  # memory freed by the workload before this one: cleared before tuning
    tuned = ctx.tune(prog, params) if not args.no_tune else None  # before the timed steps
    for _ in range(args.warmup):
        ctx.mesh_to_host(prog, params).free()
        ctx.mesh_device(prog, params)
    settle = settle_copies()
    ctx.set_profiling(1)  # events around the whole call and the evaluation launches only
    host_s, raw = [], []
    last = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        hm = ctx.mesh_to_host(prog, params)
        host_s.append(time.perf_counter() - t0)
        raw.append(ctx.last_timing_raw())  # converted after the timed steps
        if last is not None:
            last.free()
        last = hm
    # the last result stays readable: a checksum proves the data landed
    vsum = float(np.asarray(last.vertices()[:, :3], np.float64).sum()) if last.nvtx else 0.0
    tsum = int(np.asarray(last.triangles(), np.uint64).sum()) if last.ntri else 0
    nvtx, ntri, invalid = last.nvtx, last.ntri, last.invalid_quads
    last.free()
    ctx.set_profiling(0)
    dev_s = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.mesh_device(prog, params)
        dev_s.append(time.perf_counter() - t0)
    agg = {k: 0.0 for k in ("eval_ms", "total_ms", "d2h_ms", "eval_launches", "lattice_points",
                            "slabs", "tiles_total", "tiles_active", "eval_flops")}
    for r in raw:
        t = s2m.timing_dict(r)
        for k in agg:
            agg[k] += t[k]
    ms = _median_ms(host_s)
    dev_ms = _median_ms(dev_s)
    launches = max(1, agg["eval_launches"])
    eval_avg_ms = agg["eval_ms"] / launches
    flops_per_launch = agg["eval_flops"] / launches
    achieved = flops_per_launch / (eval_avg_ms / 1e3) / 1e12 if eval_avg_ms > 0 else 0.0
    tiled = agg["tiles_total"] > 0
    ops = node.tape()
    one_mandelbulb = len(ops) == 1 and int(ops[0]["code"]) == tape.MANDELBULB
    kernel = "k_sp_tiles" if tiled else ("k_eval_mandelbulb" if one_mandelbulb else "k_eval_lattice")
    d2h_bytes = nvtx * 24 + ntri * 12
    d2h_ms = agg["d2h_ms"] / args.steps
    cfg = {**cfg, "eval": "tile-pruned (exact)" if tiled else
           ("dense" if args.dense else "dense (the program is not interval-prunable)"),
           "nvtx": nvtx, "ntri": ntri, "invalid_quads": invalid,
           "flops_per_point": node.flops(),
           "lattice_points_per_step": int(agg["lattice_points"] // args.steps),
           "tiles_active_per_step": int(agg["tiles_active"] // args.steps),
           "tiles_per_step": int(agg["tiles_total"] // args.steps),
           "slabs": int(agg["slabs"] // args.steps),
           "autotuned": None if tuned is None else {
               k: tuned[k] for k in ("stream_passes", "dense_slab", "ms", "measured")}}
    out = {
        "value": round(voxels / 1e6 / (ms / 1e3), 1),
        "unit": "Mvoxels/s",
        "ms_per_step": round(ms, 3),
        "step_ms": {"median": round(ms, 3), "mean": round(float(np.mean(host_s)) * 1e3, 3),
                    "min": round(float(np.min(host_s)) * 1e3, 3),
                    "max": round(float(np.max(host_s)) * 1e3, 3)},
        "config": cfg,
        "mesh_in": "host memory (vertices pos+normal f32x6, triangles u32x3)",
        "d2h": {"bytes_per_step": int(d2h_bytes), "ms_per_step": round(d2h_ms, 3),
                "GB_per_s": round(d2h_bytes / (d2h_ms / 1e3) / 1e9, 1) if d2h_ms > 0 else None,
                "settle_before_timing": settle},
        "device_resident": {"value": round(voxels / 1e6 / (dev_ms / 1e3), 1), "unit": "Mvoxels/s",
                            "ms_per_step": round(dev_ms, 3),
                            "note": "same job, mesh left in HBM (s2m_mesh_device)"},
        "device_ms_per_step": round(agg["total_ms"] / args.steps, 3),
        "checksum": {"vertex_xyz_sum": vsum, "index_sum": tsum},
        "roofline": {
            "bound": "valu",
            "kernel": kernel,
            "achieved": round(achieved, 2),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "frac_of_issue_ceiling": round(achieved / ISSUE_CEILING_TOPS, 4),
            "traffic": measured_traffic(name + ("-dense" if args.dense else ""), kernel,
                                        {} if tuned is None or args.slab else
                                        {k: tuned[k] for k in ("stream_passes", "dense_slab")}),
            "launch_ms": round(eval_avg_ms, 4),
            "pruned_program_flops_per_launch" if tiled else "flops_per_launch": flops_per_launch,
            "points_per_launch": agg["lattice_points"] / launches,
        },
    }
    if headline:
        # stage split (every stage bracketed by events), outside the timed steps
        ctx.set_profiling(2)
        stages = {k: 0.0 for k in ("prune_ms", "eval_ms", "contour_ms", "normal_ms", "quad_ms",
                                   "total_ms")}
        n_stage = 3
        for _ in range(n_stage):
            ctx.mesh_device(prog, params)
            t = ctx.last_timing()
            for k in stages:
                stages[k] += t[k]
        ctx.set_profiling(0)
        out["stages_ms_per_step"] = {k: round(v / n_stage, 3) for k, v in stages.items()}
        if tiled and stages["eval_ms"] > 0:
            # the same kernel over the whole job in one launch (the device-resident job is
            # one pass): the streamed headline's half-size launches each carry a ramp and
            # a tail of their own
            one_ms = stages["eval_ms"] / n_stage
            job_flops = agg["eval_flops"] / args.steps
            one = job_flops / (one_ms / 1e3) / 1e12
            out["roofline"]["single_pass"] = {
                "launch_ms": round(one_ms, 4), "achieved": round(one, 2),
                "frac": round(one / PEAK_FP32_TFLOPS, 4),
                "frac_of_issue_ceiling": round(one / ISSUE_CEILING_TOPS, 4),
                "pruned_program_flops_per_launch": job_flops}
        out["roofline"]["reference_equivalent_tflops"] = round(
            node.flops() * (int(params.res[0]) + 1) ** 3 / (ms / 1e3) / 1e12, 1)
    prog.close()
    return out


def run_batch(args) -> dict:
    """256 seeded random CSG jobs of mixed resolution per step.  Batched: s2m_mesh_batch
    (shared device passes, every mesh copied to host memory); one at a time: s2m_mesh per
    job.  Device-resident variants of both are reported beside them."""
    import sdf2mesh_amd as s2m
    from sdf2mesh_amd.workloads import batch_jobs

    n = 256
    nodes, params = batch_jobs(n)
    voxels = sum(np.prod([s2m.round_resolution(int(r)) for r in q.res]) for q in params)
    ctx = s2m.Context(0)
    progs = [ctx.program(x) for x in nodes]
    steps = max(3, min(args.steps, 10))
    settle_copies()  # memory freed by the workload before this one: cleared before tuning
    tuned = ctx.tune_batch(progs, params) if not args.no_tune else None
    for _ in range(max(1, args.warmup)):
        for m in ctx.mesh_batch_to_host(progs, params):
            m.free()
        ctx.mesh_batch_device(progs, params)
    settle = settle_copies()
    bt, nv, nt = [], 0, 0
    for _ in range(steps):
        t0 = time.perf_counter()
        out = ctx.mesh_batch_to_host(progs, params)
        bt.append(time.perf_counter() - t0)
        nv, nt = sum(m.nvtx for m in out), sum(m.ntri for m in out)
        for m in out:
            m.free()
    bdev = []
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx.mesh_batch_device(progs, params)
        bdev.append(time.perf_counter() - t0)
    for q, p in zip(progs, params):  # warm the single-job path
        ctx.mesh_to_host(q, p).free()
    solo, solo_dev = [], []
    for q, p in zip(progs, params):
        t0 = time.perf_counter()
        hm = ctx.mesh_to_host(q, p)
        solo.append(time.perf_counter() - t0)
        hm.free()
    for q, p in zip(progs, params):
        t0 = time.perf_counter()
        ctx.mesh_device(q, p)
        solo_dev.append(time.perf_counter() - t0)
    ctx.close()
    dt_b = float(np.median(bt))
    jobs_b, jobs_s = n / dt_b, n / float(np.sum(solo))
    lat = np.sort(np.asarray(solo)) * 1e3
    return {
        "value": round(jobs_b, 1), "unit": "jobs/s", "ms_per_step": round(dt_b * 1e3, 3),
        "steps": steps,
        "config": {"workload": "batch256", "jobs": n, "voxels_per_batch": int(voxels),
                   "vertices": int(nv), "triangles": int(nt),
                   "mesh_in": "host memory",
                   "autotuned": None if tuned is None else {
                       k: tuned[k] for k in ("batch_cells", "ms", "measured")}},
        "d2h": {"bytes_per_step": int(nv) * 24 + int(nt) * 12, "settle_before_timing": settle},
        "one_at_a_time_jobs_per_s": round(jobs_s, 1),
        "speedup_vs_one_at_a_time": round(jobs_b / jobs_s, 2),
        "one_at_a_time_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                     "p99": round(float(np.percentile(lat, 99)), 3)},
        "batched_mvoxels_per_s": round(voxels / 1e6 / dt_b, 1),
        "device_resident": {"jobs_per_s": round(n / float(np.median(bdev)), 1),
                            "one_at_a_time_jobs_per_s": round(n / float(np.sum(solo_dev)), 1)},
    }


def run_serve(args) -> dict:
    """Poisson request streams (10-2000 requests/s) of the seeded request mix (resolution
    log-uniform in [32, 512], 80 % random CSG, 20 % Mandelbulb with 4-8 iterations)
    through the continuous-batching server; meshes land in host memory.  The comparison
    is the plain GPU path serving the same stream one request at a time (FIFO): each
    request's measured s2m_mesh time, replayed on the same arrivals."""
    import sdf2mesh_amd as s2m
    from sdf2mesh_amd.workloads import poisson_arrivals, serve_jobs

    n_max = args.serve_requests
    nodes, params = serve_jobs(n_max)
    tapes = [x.tape() for x in nodes]
    ctx = s2m.Context(0)
    progs = [ctx.program(t) for t in tapes]
    for q, p in zip(progs, params):  # every request once: buffers grown to the mix's largest
        ctx.mesh_to_host(q, p).free()
    settle_copies()
    service = []
    for q, p in zip(progs, params):
        t0 = time.perf_counter()
        hm = ctx.mesh_to_host(q, p)
        service.append(time.perf_counter() - t0)
        hm.free()
    for q in progs:
        q.close()
    ctx.close()
    service = np.asarray(service)

    rates = [float(r) for r in args.serve_rates.split(",")]
    per_rate = []
    with s2m.Server(0, workers=args.serve_workers, max_batch=args.serve_batch, want_keys=False,
                    dense_workers=args.serve_dense_workers) as srv:
        for t, p in zip(tapes[:8], params[:8]):  # warm every worker's buffers
            srv.take(srv.submit(t, p), host=True)[0].free()
        settle = settle_copies()
        for rate in rates + [float("inf")]:  # inf: every request at once (the capacity)
            n = n_max if rate == float("inf") else min(n_max, max(60, int(rate * 4)))  # ~4 s at low rates
            arr = (np.zeros(n) if rate == float("inf") else
                   np.asarray(poisson_arrivals(n, rate, seed=int(rate))))
            tickets = queue.Queue()
            stats = []

            def consume():  # the client: takes each result as it finishes, in order
                for _ in range(n):
                    hm, st = srv.take(tickets.get(), host=True)
                    stats.append(st)
                    hm.free()

            taker = threading.Thread(target=consume)
            t0 = time.perf_counter()
            taker.start()
            for a, t, p in zip(arr, tapes, params):
                now = time.perf_counter() - t0
                if a > now:
                    time.sleep(a - now)
                tickets.put(srv.submit(t, p))
            taker.join()
            elapsed = time.perf_counter() - t0
            lat = np.asarray([st["latency_ms"] for st in stats])
            bs = np.asarray([st["batch_size"] for st in stats])
            done, solo_lat = 0.0, []
            for a, sv in zip(arr, service[:n]):
                done = max(done, a) + sv
                solo_lat.append((done - a) * 1e3)
            solo_lat = np.asarray(solo_lat)
            per_rate.append({
                "rate": rate, "requests": n, "jobs_per_s": round(n / elapsed, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "mean_batch": round(float(bs.mean()), 2),
                "one_at_a_time_jobs_per_s": round(n / max(done, arr[-1]), 1),
                "one_at_a_time_p50_ms": round(float(np.percentile(solo_lat, 50)), 2),
                "one_at_a_time_p99_ms": round(float(np.percentile(solo_lat, 99)), 2),
            })
    burst = per_rate.pop()
    top = per_rate[-1]
    solo_jobs_s = top["requests"] / float(service[: top["requests"]].sum())
    return {
        "value": top["jobs_per_s"], "unit": "jobs/s", "rate": top["rate"],
        "ms_per_step": round(1e3 / top["jobs_per_s"], 3),
        "config": {"workload": "serve", "requests_per_rate": "min(%d, max(60, 4 x rate))" % n_max,
                   "workers": args.serve_workers, "dense_workers": args.serve_dense_workers,
                   "max_batch": args.serve_batch,
                   "mesh_in": "host memory"},
        "d2h": {"settle_before_timing": settle},
        "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "one_at_a_time_jobs_per_s": round(solo_jobs_s, 1),
        "speedup_vs_one_at_a_time": round(top["jobs_per_s"] / solo_jobs_s, 2),
        "per_rate": per_rate,
        "capacity": {"requests": burst["requests"], "jobs_per_s": burst["jobs_per_s"],
                     "one_at_a_time_jobs_per_s": burst["one_at_a_time_jobs_per_s"],
                     "speedup_vs_one_at_a_time": round(burst["jobs_per_s"] / burst["one_at_a_time_jobs_per_s"], 2),
                     "p50_ms": burst["p50_ms"], "p99_ms": burst["p99_ms"], "mean_batch": burst["mean_batch"]},
    }


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


def measured_traffic(workload: str, kernel: str, shape: dict):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/hbm_traffic.json (scaled by the copy calibration), or None when
    the run's pass / slab shape differs from the one measured there."""
    f = ROOT / "profiles" / "hbm_traffic.json"
    if not f.exists():
        return None
    entry = json.loads(f.read_text()).get(workload, {})
    if entry.get("kernel") != kernel:
        return None
    if any(shape.get(k) != v for k, v in entry.get("shape", {}).items()):
        return None
    return entry.get("bytes_per_launch")


def common(args, metric: str, unit: str, value, ms: float, data: str) -> dict:
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": data}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="csg64-512")
    ap.add_argument("--only", action="store_true", help="run --workload alone")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="use the library's rules, not the autotuner")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="10,30,100,300,1000,2000")
    ap.add_argument("--serve-workers", type=int, default=2)
    ap.add_argument("--serve-dense-workers", type=int, default=1)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--slab", type=int, default=0,
                    help="cell planes per dense slab or sparse pass (0: the autotuner's or the library's choice)")
    ap.add_argument("--dense", action="store_true",
                    help="evaluate every lattice point with the full program (no tile pruning)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; jobs would "
                                   "shard by z-slab or by job across GPUs by design only, so "
                                   "--gpus must be 1"}))
        return 2
    if args.workload == "copy-calibration":
        print(json.dumps(run_copy_calibration(args)), flush=True)
        return 0
    global _PROBE
    _PROBE = CopyProbe()  # the link rate, before the library frees any device memory
    import sdf2mesh_amd as s2m

    if args.only and args.workload == "batch256":
        r = run_batch(args)
        out = {**common(args, "batched jobs/s (256 random CSG jobs, res 32-256)", "jobs/s",
                        r["value"], r["ms_per_step"], "synthetic (seeded CSG programs)"), **r}
    elif args.only and args.workload == "serve":
        r = run_serve(args)
        out = {**common(args, "served jobs/s at the highest offered rate, with p50/p99 latency",
                        "jobs/s", r["value"], r["ms_per_step"], "synthetic (seeded request stream)"),
               **r}
    elif args.workload in SINGLE:
        ctx = s2m.Context(0)
        head = single_job(args.workload, args, ctx, headline=True)
        metric = ("Mvoxels/s meshed at 512^3 fp32 (CSG-64, reference-exact corners), mesh in host "
                  "memory" if args.workload == "csg64-512" else
                  f"Mvoxels/s meshed ({args.workload}), mesh in host memory")
        out = {**common(args, metric, "Mvoxels/s", head.pop("value"), head.pop("ms_per_step"),
                        "synthetic (seeded SDF program)"), **head}
        if not args.only:
            # The Mandelbulb job runs last: its context's buffers are the largest, and
            # the workloads measured right after they were freed ran slower and varied
            # from run to run (batch256 at 51,000-66,000 jobs/s, 65,000-98,000 with the
            # meshes left in HBM; alone or before it, 66,500-67,100 and ~99,800).
            others = {}
            if args.workload != "torus-128":
                others["torus-128"] = single_job("torus-128", args, ctx, headline=False)
            others["batch256"] = run_batch(args)
            others["serve"] = run_serve(args)
            if args.workload != "mandelbulb-1024":
                others["mandelbulb-1024"] = single_job("mandelbulb-1024", args, ctx, headline=False)
            ctx.close()
            out["workloads"] = others
        else:
            ctx.close()
        if not args.no_cpu_baseline:
            node, params, _ = workload(args.workload)
            cb = cpu_baseline(node, params, args.cpu_seconds)
            out["cpu_baseline"] = cb
            out["speedup_vs_cpu_baseline"] = round(out["value"] / cb["value"], 1)
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
