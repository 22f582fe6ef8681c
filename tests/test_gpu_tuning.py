"""The autotuner on the device: s2m_tune times the candidate shapes of a job and every
later meshing call of that key uses the chosen one; a shape never changes the mesh."""
import numpy as np
import pytest

import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape
from sdf2mesh_amd.workloads import batch_jobs

pytestmark = pytest.mark.gpu


def same(a, b):
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert a.invalid_quads == b.invalid_quads


@pytest.mark.parametrize("node,res,bounds,knob", [
    (tape.csg64(), 256, 2.0, "stream_passes"),      # tile-sparse job streamed to host memory
    (tape.mandelbulb(6), 160, 2.5, "dense_slab"),    # dense slabs (rounded up to 256^3)
])
def test_tune_applies_choice(ctx, tmp_path, node, res, bounds, knob):
    s2m.tuning_clear()
    prog = ctx.program(node)
    p = s2m.Params.reference(res, bounds)
    ref = ctx.mesh(prog, p)  # the rule's shape
    t = ctx.tune(prog, p, reps=3)
    print(f"tuned {knob}: {t}", flush=True)
    assert t["measured"] >= 2 and t[knob] > 0 and t["ms"] > 0
    again = ctx.tune(prog, p)  # the table holds the key now: no timing
    assert again["measured"] == 0 and again[knob] == t[knob]
    m = ctx.mesh(prog, p)
    slabs = ctx.last_timing()["slabs"]
    if knob == "stream_passes":
        assert slabs == t[knob]
    else:
        assert slabs == -(-(s2m.round_resolution(res) - 1) // t[knob])
    same(m, ref)
    # the choice survives a save / clear / load round trip
    f = tmp_path / "tune.json"
    s2m.tuning_save(f)
    s2m.tuning_clear()
    s2m.tuning_load(f)
    assert ctx.tune(prog, p)["measured"] == 0
    s2m.tuning_clear()


def test_tune_batch(ctx):
    s2m.tuning_clear()
    nodes, params = batch_jobs(48, seed=11)
    progs = [ctx.program(n) for n in nodes]
    ref = ctx.mesh_batch(progs, params)
    t = ctx.tune_batch(progs, params, reps=3)
    print(f"tuned batch: {t}", flush=True)
    assert t["measured"] == 7 and 16 << 20 <= t["batch_cells"] <= 512 << 20
    for a, b in zip(ctx.mesh_batch(progs, params), ref):
        same(a, b)
    s2m.tuning_clear()
