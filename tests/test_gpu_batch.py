"""Batched meshing (s2m_mesh_batch): every job of a batch must equal the same job meshed
alone (which the parity tests tie to the oracle), bit for bit."""
import numpy as np
import pytest

import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape
from sdf2mesh_amd.workloads import batch_jobs

pytestmark = pytest.mark.gpu


def same(a, b):
    assert len(a.vertices) == len(b.vertices)
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert a.invalid_quads == b.invalid_quads


def test_batch_mixed_equals_solo(ctx):
    nodes = [tape.random_csg(s) for s in range(1, 13)]
    nodes += [tape.torus_example(), tape.csg64(), tape.mandelbulb(5)]
    res = [32, 40, 64, 24, 96, 17, 33, 128, 48, 64, 8, 1, 64, 80, 40]
    flags = [s2m.REFERENCE_COMPAT, 0, s2m.COMPAT_DROP_LAST_SLICE, s2m.CORNERS_EXACT] * 4
    progs = [ctx.program(n) for n in nodes]
    params = [s2m.Params.reference(r, 2.5 if n.tape()[0]["code"] == tape.MANDELBULB else 2.0,
                                   flags=f) for n, r, f in zip(nodes, res, flags)]
    params.append(s2m.Params.reference(0, 2.0))  # empty grid
    progs.append(progs[0])
    print(f"batch of {len(progs)}", flush=True)
    batch = ctx.mesh_batch(progs, params)
    for i, (p, q) in enumerate(zip(progs, params)):
        solo = ctx.mesh(p, q)
        print(f"job {i}: {len(solo.vertices)} vertices", flush=True)
        same(batch[i], solo)


def test_batch_config4_equals_solo(ctx):
    # the batching workload: 256 seeded random CSG trees, resolutions from
    # {32,64,96,128,192,256} with weights {4,4,3,3,2,1}, rounded as the reference does
    nodes, params = batch_jobs(256)
    progs = [ctx.program(n) for n in nodes]
    batch = ctx.mesh_batch(progs, params)
    dev = ctx.mesh_batch_device(progs, params)
    for i in range(0, 256, 5):
        solo = ctx.mesh(progs[i], params[i])
        same(batch[i], solo)
        assert dev[i]["nvtx"] == len(solo.vertices) and dev[i]["ntri"] == len(solo.triangles)
    print("config-4 batch: %d vertices, %d triangles" % (
        sum(len(m.vertices) for m in batch), sum(len(m.triangles) for m in batch)), flush=True)


def test_batch_anisotropic_jobs(ctx):
    # jobs whose grids and boxes differ per axis, and from each other, share the passes:
    # per-job tile, super-tile and row numbering with different tile counts per axis;
    # each batched mesh equals the job alone, and the first jobs the oracle's meshes
    from test_gpu_parity import ANISO, assert_same_mesh, reference_mesh

    nodes, params = [], []
    for k in range(24):
        res, bmin, bmax = ANISO[k % len(ANISO)]
        node = tape.csg64() if k % 3 == 0 else tape.random_csg(50 + k)
        flags = (0, s2m.CORNERS_EXACT, s2m.REFERENCE_COMPAT)[k % 3]
        nodes.append(node)
        params.append(s2m.Params(bmin, bmax, res, 1e-4, flags))
    progs = [ctx.program(n) for n in nodes]
    batch = ctx.mesh_batch(progs, params)
    for i in range(len(nodes)):
        same(batch[i], ctx.mesh(progs[i], params[i]))
    for i in range(4):
        assert_same_mesh(batch[i], reference_mesh(nodes[i], params[i]))
    assert sum(len(m.vertices) for m in batch) > 10000


def test_batch_repeatable(ctx):
    nodes, params = batch_jobs(32, seed=7)
    progs = [ctx.program(n) for n in nodes]
    a = ctx.mesh_batch(progs, params)
    b = ctx.mesh_batch(progs[::-1], params[::-1])[::-1]  # order of jobs does not matter
    for x, y in zip(a, b):
        same(x, y)


def device_copy(dm):
    """A device-resident mesh (vtx/tri/keys/signs addresses) copied with plain hipMemcpy."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    nv, nt = dm["nvtx"], dm["ntri"]

    def get(ptr, dtype, count):
        a = np.empty(count, dtype)
        if count:
            assert hip.hipMemcpy(a.ctypes.data, ptr, a.nbytes, 2) == 0  # device to host
        return a
    return s2m.TriangleMesh(get(dm["vtx_ptr"], np.float32, 6 * nv).reshape(nv, 6),
                            get(dm["tri_ptr"], np.uint32, 3 * nt).reshape(nt, 3), dm["invalid_quads"],
                            get(dm["keys_ptr"], np.uint64, nv), get(dm["signs_ptr"], np.uint8, nv))


def test_device_batch_with_dense_jobs(ctx):
    # s2m_mesh_batch_device with jobs for the dense path among the tile-sparse ones: a
    # compiled program, a Mandelbulb above 128^3, a CSG tape forced dense; each result,
    # read back from the device, equals the job meshed alone
    from pathlib import Path

    golden = Path(__file__).resolve().parent / "golden"
    eng = ctx.program_from_source(golden / "engraved_cube.sdf3d", path=True)
    assert eng.kind == s2m.PROGRAM_COMPILED
    mb = ctx.program(tape.mandelbulb(6))
    csg = [ctx.program(tape.random_csg(s)) for s in (2, 9)]
    progs = [csg[0], eng, mb, csg[1], csg[0], eng]
    params = [s2m.Params.reference(96, 2.0), s2m.Params.reference(64, 2.0), s2m.Params.reference(160, 2.5),
              s2m.Params.reference(48, 2.0, flags=s2m.REFERENCE_COMPAT | s2m.DENSE_EVAL),
              s2m.Params.reference(0, 2.0), s2m.Params.reference(37, 2.0, flags=0)]
    dev = ctx.mesh_batch_device(progs, params)
    got = [device_copy(d) for d in dev]  # before any other meshing call on the context
    for i, (p, q) in enumerate(zip(progs, params)):
        solo = ctx.mesh(p, q)
        print(f"device batch job {i}: {len(solo.vertices)} vertices", flush=True)
        same(got[i], solo)


def stacked_boxes(n: int, z0: float) -> tape.Node:
    """n thin boxes stacked in z over most of the xy extent: 2n surface crossings per
    lattice column, far above the first vertex estimate of a pass (6 per column)."""
    return tape.fold([tape.box((0.0, 0.0, z0 + 0.15 * i), (1.5, 1.5, 0.05)) for i in range(n)],
                     op=tape.union)


def test_batch_passes_overflow_and_repeat():
    # A fresh context, three batch passes (four 256^3 jobs fill one at the default cut of
    # 64M cells): the first passes overflow their first vertex, triangle and slot
    # estimates and are queued again after the next pass has been queued.  Every job
    # equals the job alone; a second batch (capacities grown) and a profiled batch (every
    # pass run to completion in turn) equal the first.
    print("batch overflow", flush=True)
    with s2m.Context(0) as fresh:
        nodes = [stacked_boxes(12, -0.85 + 0.01 * s) for s in range(9)]
        nodes.insert(3, tape.torus_example())
        progs = [fresh.program(x) for x in nodes]
        params = [s2m.Params.reference(256, 2.0) for _ in progs]
        params[3] = s2m.Params.reference(64, 2.0)
        first = fresh.mesh_batch(progs, params)
        second = fresh.mesh_batch(progs, params)
        fresh.set_profiling(2)
        profiled = fresh.mesh_batch(progs, params)
        fresh.set_profiling(0)
        for i, (p, q) in enumerate(zip(progs, params)):
            solo = fresh.mesh(p, q)
            print(f"overflow batch job {i}: {len(solo.vertices)} vertices", flush=True)
            if q.res[0] == 256:  # above its share of the pass's first estimate
                assert len(solo.vertices) > 6 * 256 * 256
            same(first[i], solo)
            same(second[i], solo)
            same(profiled[i], solo)


def test_batch_nan_results_equal_solo(ctx):
    # A 3-iteration Mandelbulb whose estimator gives NaN at some lattice points, so some
    # vertex positions and normals are NaN: meshed alone it takes the dense path, in a
    # batch the shared tile pass; both store NaN results as the canonical quiet NaN, so
    # the meshes are equal bit for bit, NaNs included.
    half = (0.8948788543249019, 0.9494887433921423, 0.6620873067782631)
    p = s2m.Params(tuple(-h for h in half), half, (9, 33, 33), 1e-4, 0)
    prog = ctx.program(tape.mandelbulb(3))
    others = [ctx.program(tape.random_csg(5)), ctx.program(tape.csg64())]
    q = s2m.Params.reference(48, 2.0)
    solo = ctx.mesh(prog, p)
    assert np.isnan(solo.vertices).any()
    canonical = np.float32(np.nan).view(np.uint32)
    assert np.all(solo.vertices.view(np.uint32)[np.isnan(solo.vertices)] == canonical)
    batch = ctx.mesh_batch([others[0], prog, others[1]], [q, p, q])
    same(batch[1], solo)


def test_randomized_batches_equal_solo(ctx):
    # Seeded random batches of 2-40 jobs: random CSG trees, CSG-64, the operator zoo,
    # clustered folds and Mandelbulbs on random grids (often anisotropic), boxes and flag
    # sets; each batched mesh equals the job meshed alone, bit for bit.
    import random

    from test_gpu_parity import clustered_fold, op_zoo

    rng = random.Random(3)
    for bi in range(24):
        progs, params = [], []
        for _ in range(rng.randint(2, 40)):
            seed = rng.randrange(1, 10 ** 6)
            u = rng.random()
            if u < 0.75:
                node = tape.random_csg(seed)
            elif u < 0.8:
                node = tape.csg64()
            elif u < 0.88:
                node = op_zoo(rng.choice([0.0, 0.02, 0.3]))
            elif u < 0.96:
                node = clustered_fold(rng.randint(20, 120), seed, rng.choice([0.05, 0.5]))
            else:
                node = tape.mandelbulb(rng.randint(3, 8))
            res = tuple(rng.choice([9, 16, 24, 33, 48, 64, 96]) for _ in range(3))
            half = tuple(rng.uniform(0.4, 1.3) for _ in range(3))
            flags = rng.choice([0, s2m.CORNERS_EXACT, s2m.REFERENCE_COMPAT])
            progs.append(ctx.program(node))
            params.append(s2m.Params(tuple(-h for h in half), half, res, 1e-4, flags))
        print(f"random batch {bi}: {len(progs)} jobs", flush=True)
        for pr, p, m in zip(progs, params, ctx.mesh_batch(progs, params)):
            same(m, ctx.mesh(pr, p))
