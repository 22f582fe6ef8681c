"""Results in host memory (s2m_mesh, s2m_mesh_batch): the mesh streamed to host memory pass
by pass equals the one copied after the job, and the page-locked result memory is reused:
repeated calls page-lock no new memory, whatever ran before them."""
import numpy as np
import pytest

import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape
from sdf2mesh_amd.workloads import batch_jobs

pytestmark = pytest.mark.gpu


def same(a, b):
    assert len(a.vertices) == len(b.vertices) > 0
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert a.invalid_quads == b.invalid_quads


def device_mesh_copied(ctx, prog, params):
    """s2m_mesh_device, then its arrays copied with plain hipMemcpy (no streaming)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    dm = ctx.mesh_device(prog, params)
    nv, nt = dm["nvtx"], dm["ntri"]

    def get(ptr, dtype, count):
        a = np.empty(count, dtype)
        if count:
            assert hip.hipMemcpy(a.ctypes.data, ptr, a.nbytes, 2) == 0  # device to host
        return a
    return s2m.TriangleMesh(get(dm["vtx_ptr"], np.float32, 6 * nv).reshape(nv, 6),
                            get(dm["tri_ptr"], np.uint32, 3 * nt).reshape(nt, 3), dm["invalid_quads"],
                            get(dm["keys_ptr"], np.uint64, nv), get(dm["signs_ptr"], np.uint8, nv))


@pytest.mark.parametrize("node,res,bounds", [(tape.csg64(), 512, 2.0), (tape.random_csg(7), 256, 2.0),
                                             (tape.mandelbulb(8), 512, 2.5)])
def test_streamed_equals_device_resident(ctx, node, res, bounds):
    # s2m_mesh copies each pass (tile-sparse) or slab (dense) to host memory while the
    # next one runs; s2m_mesh_device meshes the same job (the tile-sparse one in one pass)
    # and leaves it in HBM, copied here afterwards
    s2m.tuning_clear()
    prog = ctx.program(node)
    p = s2m.Params.reference(res, bounds)
    streamed = ctx.mesh(prog, p)
    passes = ctx.last_timing()["slabs"]
    print(f"{res}^3: streamed in {passes} passes", flush=True)
    assert passes > 1
    same(streamed, device_mesh_copied(ctx, prog, p))


def test_pool_steady_state(ctx):
    prog = ctx.program(tape.csg64())
    big = s2m.Params.reference(512, 2.0)
    small = s2m.Params.reference(64, 2.0)
    nodes, params = batch_jobs(32, seed=3)
    progs = [ctx.program(n) for n in nodes]

    def round_():
        ctx.mesh_to_host(prog, big).free()
        for m in ctx.mesh_batch_to_host(progs, params):
            m.free()
        for q, p in zip(progs, params):
            ctx.mesh_to_host(q, p).free()
        ctx.mesh_to_host(prog, small).free()

    s2m.host_pool_trim()  # earlier tests may have left the pool at its idle limit
    round_()
    before = s2m.host_pool_stats()
    for _ in range(3):
        round_()
    after = s2m.host_pool_stats()
    print("pool", before, after, flush=True)
    assert after["pinned_allocs"] == before["pinned_allocs"]


@pytest.mark.parametrize("batch", [False, True])
def test_graph_replays_equal_eager(ctx, batch):
    # A one-pass job is captured into a graph the second time it comes and replayed after:
    # every replay, with other jobs (eager, and graphs of their own) and synchronisations
    # in between, must give the eager run's mesh bit for bit.
    s2m.tuning_clear()
    if batch:
        nodes, params = batch_jobs(24, seed=5)
        progs = [ctx.program(n) for n in nodes]

        def run():
            out = ctx.mesh_batch_to_host(progs, params, want_keys=True)
            meshes = [(m.vertices().copy(), m.triangles().copy()) for m in out]
            for m in out:
                m.free()
            return meshes
    else:
        prog = ctx.program(tape.csg64())
        p = s2m.Params.reference(256, 2.0)

        def run():
            m = device_mesh_copied(ctx, prog, p)
            return [(m.vertices, m.triangles)]
    other = ctx.program(tape.random_csg(4))
    q = s2m.Params.reference(128, 2.0)
    dense = s2m.Params.reference(96, 2.0, flags=s2m.REFERENCE_COMPAT | s2m.DENSE_EVAL)
    ref = run()  # eager (the job is seen for the first time)
    assert ctx.last_timing()["graph_launches"] == 0
    launches = []
    for i in range(6):
        ctx.mesh(other, q)       # another tile-pass job: eager, then a graph of its own
        ctx.mesh(other, dense)   # dense-path work between the replays
        got = run()
        launches.append(ctx.last_timing()["graph_launches"])
        assert len(got) == len(ref)
        for (gv, gt), (rv, rt) in zip(got, ref):
            assert np.array_equal(gv.view(np.uint32), rv.view(np.uint32))
            assert np.array_equal(gt, rt)
    print("graph launches per call:", launches, flush=True)
    assert launches[1:] == [1] * 5  # the second call captures, every later call replays


def test_graph_after_buffers_regrow():
    # A job captured into a graph, then a larger job that regrows (frees and allocates
    # again) the context's buffers, then the first job again: its graph holds the old
    # addresses, so the call must capture anew and still give the oracle's mesh.
    from test_gpu_parity import assert_same_mesh, reference_mesh

    s2m.tuning_clear()
    c = s2m.Context(0)  # a fresh context: buffers sized for the first job
    try:
        node = tape.csg64()
        prog = c.program(node)
        p = s2m.Params.reference(64, 2.0)
        ref = reference_mesh(node, p)
        launches = []
        for _ in range(3):
            assert_same_mesh(c.mesh(prog, p), ref)
            launches.append(c.last_timing()["graph_launches"])
        c.mesh(c.program(tape.random_csg(9)), s2m.Params.reference(512, 2.0))  # regrows the buffers
        for _ in range(3):
            assert_same_mesh(c.mesh(prog, p), ref)
            launches.append(c.last_timing()["graph_launches"])
        print("graph launches per call:", launches, flush=True)
        assert launches[:3] == [0, 1, 1] and launches[4:] == [1, 1]
    finally:
        c.close()


def test_threads_with_own_contexts(ctx):
    # One context per host thread (the C ABI's threading contract): four threads mesh
    # tile-sparse, dense, Mandelbulb and batched jobs at once, each result equal bit for
    # bit to the same job meshed alone, and each thread's failing call reports its own
    # error (the last-error message is per thread).
    import threading

    jobs = [(tape.csg64(), s2m.Params.reference(96, 2.0)),
            (tape.random_csg(21), s2m.Params.reference(80, 2.0, flags=0)),
            (tape.torus_example(), s2m.Params.reference(64, 2.0, flags=s2m.REFERENCE_COMPAT | s2m.DENSE_EVAL)),
            (tape.mandelbulb(5), s2m.Params.reference(48, 2.5))]
    solo = [ctx.mesh(ctx.program(n), p) for n, p in jobs]
    problems, results = {}, {}

    def work(t):
        try:
            c = s2m.Context(0)
            try:
                out = []
                for _ in range(3):
                    for k, (n, p) in enumerate(jobs):
                        out.append((k, c.mesh(c.program(n), p)))
                    bad = s2m.Params((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (70000 + t, 4, 4), 1e-4, 0)
                    try:
                        c.mesh(c.program(jobs[0][0]), bad)
                        problems[t] = "no error for an oversized grid"
                    except s2m.S2MError as e:
                        if str(70000 + t) not in str(e):
                            problems[t] = f"another thread's message: {e}"
                    b = c.mesh_batch([c.program(n) for n, _ in jobs[:2]], [p for _, p in jobs[:2]])
                    out += [(0, b[0]), (1, b[1])]
                results[t] = out
            finally:
                c.close()
        except Exception as e:  # reported below, with the thread
            problems[t] = repr(e)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=240)
    assert not problems, problems
    for t in range(4):
        assert len(results[t]) == 18
        for k, m in results[t]:
            r = solo[k]
            assert np.array_equal(m.vertices.view(np.uint32), r.vertices.view(np.uint32))
            assert np.array_equal(m.triangles, r.triangles)
            assert np.array_equal(m.cell_keys, r.cell_keys) and m.invalid_quads == r.invalid_quads
