"""Serving (s2m_server_*): requests admitted under an HBM budget and batched
continuously; each result must be identical to meshing the request alone."""
import time

import numpy as np
import pytest

import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape
from sdf2mesh_amd.workloads import serve_jobs

pytestmark = pytest.mark.gpu


def same(a, b):
    assert len(a.vertices) == len(b.vertices)
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert a.invalid_quads == b.invalid_quads


def test_served_equals_alone(ctx):
    nodes, params = serve_jobs(48, seed=3)
    # keep the test short: cap the resolution at 128
    params = [s2m.Params.reference(min(int(p.res[0]), 128), p.bb_max[0] * 2.0) for p in params]
    with s2m.Server(0, workers=2, max_batch=8) as srv:
        tickets = []
        for i, (n, p) in enumerate(zip(nodes, params)):
            tickets.append(srv.submit(n, p))
            if i % 7 == 0:
                time.sleep(0.002)  # arrivals while earlier passes are running
        results = [srv.take(t) for t in tickets]
    batches = [st["batch_size"] for _, st in results]
    print("served 48 requests, batch sizes", sorted(set(batches)), flush=True)
    for (m, st), n, p in zip(results, nodes, params):
        same(m, ctx.mesh(ctx.program(n), p))
        assert st["latency_ms"] >= st["queue_ms"] >= 0.0


@pytest.mark.timeout(600)
def test_served_config5_mix(ctx):
    # the serving workload's request mix at full size: resolution log-uniform in [32, 512]
    # (rounded up to a power of two), 80 % random CSG trees, 20 % Mandelbulbs with 4-8
    # iterations; 120 requests submitted on a seeded Poisson schedule at 2000 requests/s
    # (the workload's highest rate) so that arrivals overlap running passes.  Every served
    # mesh must equal the request meshed alone, and the smallest CSG requests the oracle's
    # meshes.
    import oracle
    from sdf2mesh_amd.workloads import poisson_arrivals

    from test_gpu_parity import assert_same_mesh, reference_mesh

    n = 120
    nodes, params = serve_jobs(n, seed=5)
    res = [s2m.round_resolution(int(p.res[0])) for p in params]
    mandel = [nd.tape()[0]["code"] == tape.MANDELBULB for nd in nodes]
    print(f"config-5 mix: {sum(mandel)} Mandelbulbs, resolutions {sorted(set(res))}", flush=True)
    assert max(res) == 512 and min(res) <= 64 and sum(mandel) >= 12
    arrivals = poisson_arrivals(n, 2000.0, seed=5)
    with s2m.Server(0, workers=2, max_batch=16) as srv:
        tickets = []
        t0 = time.perf_counter()
        for a, nd, p in zip(arrivals, nodes, params):
            now = time.perf_counter() - t0
            if a > now:
                time.sleep(a - now)
            tickets.append(srv.submit(nd, p))
        results = [srv.take(t) for t in tickets]
    batches = [st["batch_size"] for _, st in results]
    lat = sorted(st["latency_ms"] for _, st in results)
    print(f"served {n}: batch sizes {sorted(set(batches))}, latency p50 {lat[n // 2]:.1f} ms "
          f"p99 {lat[int(n * 0.99)]:.1f} ms", flush=True)
    assert max(batches) > 1  # requests shared passes
    for i, ((m, st), nd, p) in enumerate(zip(results, nodes, params)):
        solo = ctx.mesh(ctx.program(nd), p)
        same(m, solo)
        assert st["latency_ms"] >= st["queue_ms"] >= 0.0
    # the oracle on the CSG requests of the smallest grids (and the largest at most 128^3)
    small = [i for i in range(n) if not mandel[i] and res[i] <= 64][:8]
    mid = [i for i in range(n) if not mandel[i] and res[i] == 128][:1]
    for i in small + mid:
        print(f"request {i}: res {res[i]} against the oracle", flush=True)
        assert_same_mesh(results[i][0], reference_mesh(nodes[i], params[i]))


def test_served_compiled_programs(ctx):
    # compiled programs (hiprtc) go through s2m_server_submit_program to the dense lane,
    # mixed with tape requests of both lanes; each result equals meshing it alone
    from pathlib import Path

    golden = Path(__file__).resolve().parent / "golden"
    eng = ctx.program_from_source(golden / "engraved_cube.sdf3d", path=True)
    glsl = ctx.program_from_source(golden / "mandelbulb8.frag", lang="glsl", path=True)
    assert eng.kind == glsl.kind == s2m.PROGRAM_COMPILED
    jobs = [(eng, s2m.Params.reference(96, 2.0)), (glsl, s2m.Params.reference(200, 2.5)),
            (tape.csg64(), s2m.Params.reference(64, 2.0)), (tape.mandelbulb(6), s2m.Params.reference(256, 2.5)),
            (eng, s2m.Params.reference(40, 2.0, flags=0)), (tape.random_csg(9), s2m.Params.reference(128, 2.0)),
            (glsl, s2m.Params.reference(48, 2.5))]
    with s2m.Server(0, workers=2, dense_workers=1, max_batch=8) as srv:
        tickets = [srv.submit(prog, p) for prog, p in jobs]
        results = [srv.take(t) for t in tickets]
    for (m, st), (prog, p) in zip(results, jobs):
        solo = ctx.mesh(prog if isinstance(prog, s2m.Program) else ctx.program(prog), p)
        assert len(m.vertices) > 0
        same(m, solo)


def test_budget_rejects_oversized_request():
    with s2m.Server(0, hbm_budget=1 << 20, workers=1) as srv:
        small = s2m.Params.reference(16, 2.0)
        assert s2m.Server.request_bytes(tape.torus_example(), small) < (1 << 20)
        t = srv.submit(tape.torus_example(), small)
        m, _ = srv.take(t)
        assert len(m.vertices) > 0
        with pytest.raises(s2m.S2MError) as e:
            srv.submit(tape.torus_example(), s2m.Params.reference(256, 2.0))
        assert e.value.code == s2m.E_BUDGET


def test_take_busy_then_done():
    with s2m.Server(0, workers=1) as srv:
        t = srv.submit(tape.csg64(), s2m.Params.reference(128, 2.0))
        r = srv.take(t, wait=False)
        if r is None:  # still queued or running
            r = srv.take(t, wait=True)
        assert len(r[0].vertices) > 0
        with pytest.raises(s2m.S2MError):
            srv.take(t)  # a ticket can be taken once


def test_failing_requests_do_not_stop_the_server(ctx):
    # requests the library refuses (a grid past the 65535-cell key limit, an invalid
    # tape) fail on their own, at submission or when taken; the requests around them are
    # served and equal to meshing them alone
    good = [(tape.csg64(), s2m.Params.reference(64, 2.0)), (tape.random_csg(3), s2m.Params.reference(96, 2.0)),
            (tape.mandelbulb(5), s2m.Params.reference(64, 2.5))]
    bad_grid = s2m.Params((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (70000, 8, 8), 1e-4, 0)
    bad_tape = np.zeros(3, s2m.TAPE_DTYPE)  # three zero ops: not a valid program
    with s2m.Server(0, workers=2, max_batch=8) as srv:
        tickets, failed = [], 0
        for r in range(3):
            for n, p in good:
                tickets.append((srv.submit(n, p), n, p))
            for prog, p in ((tape.torus_example(), bad_grid), (bad_tape, s2m.Params.reference(32, 2.0))):
                try:
                    t = srv.submit(prog, p)
                except s2m.S2MError:
                    failed += 1
                    continue
                with pytest.raises(s2m.S2MError):
                    srv.take(t)
                failed += 1
        assert failed == 6
        for t, n, p in tickets:
            m, _ = srv.take(t)
            same(m, ctx.mesh(ctx.program(n), p))
