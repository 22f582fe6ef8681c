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


@pytest.mark.parametrize("node,res,bounds", [(tape.csg64(), 512, 2.0), (tape.random_csg(7), 256, 2.0),
                                             (tape.mandelbulb(8), 512, 2.5)])
def test_streamed_equals_copied_after(ctx, node, res, bounds):
    # slab beyond the grid: one pass / slab, copied after the job; slab 0: the passes or
    # slabs the library picks, each copied while the next one runs
    s2m.tuning_clear()
    prog = ctx.program(node)
    streamed = ctx.mesh(prog, s2m.Params.reference(res, bounds))
    passes = ctx.last_timing()["slabs"]
    whole = ctx.mesh(prog, s2m.Params.reference(res, bounds, slab=4096))
    assert ctx.last_timing()["slabs"] == 1
    print(f"{res}^3: streamed in {passes} passes", flush=True)
    if node.tape()[0]["code"] != tape.MANDELBULB:
        assert passes > 1
    same(streamed, whole)


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

    round_()
    before = s2m.host_pool_stats()
    for _ in range(3):
        round_()
    after = s2m.host_pool_stats()
    print("pool", before, after, flush=True)
    assert after["pinned_allocs"] == before["pinned_allocs"]
