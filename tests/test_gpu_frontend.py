"""GPU parity of programs from SDF source (s2m_program_from_source).

  * a program compiled with hiprtc (tests/golden/engraved_cube.sdf3d: swizzles, loops,
    branches, user functions) against the oracle's C restatement of the same program:
    lattice values and whole meshes bit-exact;
  * a CSG program forced down the compiled path against the same program as a tape:
    identical meshes (the compiled code calls the same library functions);
  * the GLSL Mandelbulb compiled against the MANDELBULB tape op: identical meshes (both
    use the device math library, so no fp64 tolerance is needed);
  * batches mixing compiled and tape programs equal meshing each job alone;
  * a user-defined sdf3d_normal replaces the library normal.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape

from test_gpu_parity import assert_same_mesh, reference_mesh

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
ENGRAVED = GOLDEN / "engraved_cube.sdf3d"


def same_mesh(a: s2m.TriangleMesh, b: s2m.TriangleMesh):
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert a.invalid_quads == b.invalid_quads


@pytest.fixture(scope="module")
def engraved(ctx):
    prog = ctx.program_from_source(ENGRAVED, path=True)
    assert prog.kind == s2m.PROGRAM_COMPILED
    return prog


@pytest.mark.parametrize("res,flags", [(32, s2m.REFERENCE_COMPAT), (48, 0), (64, s2m.CORNERS_EXACT)])
def test_compiled_lattice_matches_oracle(ctx, engraved, res, flags):
    print(f"engraved lattice res={res} flags={flags}", flush=True)
    p = s2m.Params.reference(res, 2.0, flags=flags)   # box [-1, 1]^3
    got = ctx.eval_lattice(engraved, p)
    exact = bool(flags & s2m.CORNERS_EXACT)
    r = s2m.round_resolution(res) if flags & s2m.COMPAT_POW2 else res
    ax = s2m.axis_table(p.bb_min[0], p.bb_max[0], r, exact)[0]
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1)  # (x, y, z, 3)
    pts = g.transpose(2, 1, 0, 3).reshape(-1, 3)               # z slowest
    want = oracle.eval_f32(oracle.example_tape(oracle.ORA_OP_ENGRAVED_CUBE), pts)
    assert np.array_equal(got.reshape(-1).view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("res,flags", [(32, s2m.REFERENCE_COMPAT), (64, s2m.REFERENCE_COMPAT),
                                       (100, 0)])
def test_compiled_mesh_matches_oracle(ctx, engraved, res, flags):
    print(f"engraved mesh res={res} flags={flags}", flush=True)
    p = s2m.Params.reference(res, 2.0, flags=flags)
    assert_same_mesh(ctx.mesh(engraved, p),
                     reference_mesh(oracle.example_tape(oracle.ORA_OP_ENGRAVED_CUBE), p))


@pytest.mark.parametrize("flags", [0, s2m.REFERENCE_COMPAT])
def test_compiled_anisotropic_matches_oracle(ctx, engraved, flags):
    # the compiled lattice kernel on grids and boxes that differ per axis, over several
    # dense slabs: mesh and lattice against the oracle's restatement
    from test_gpu_parity import ANISO

    ora = oracle.example_tape(oracle.ORA_OP_ENGRAVED_CUBE)
    for res, bmin, bmax in ANISO:
        print(f"engraved anisotropic res={res} flags={flags}", flush=True)
        p = s2m.Params(bmin, bmax, res, 1e-4, flags, 8)
        assert_same_mesh(ctx.mesh(engraved, p), reference_mesh(ora, p))
        got = ctx.eval_lattice(engraved, p)
        cells = [s2m.round_resolution(r) for r in res] if flags & s2m.COMPAT_POW2 else list(res)
        exact = bool(flags & s2m.CORNERS_EXACT)
        ax = [s2m.axis_table(bmin[a], bmax[a], cells[a], exact)[0] for a in range(3)]
        g = np.stack(np.meshgrid(ax[0], ax[1], ax[2], indexing="ij"), -1)  # (x, y, z, 3)
        pts = g.transpose(2, 1, 0, 3).reshape(-1, 3)                        # z slowest
        want = oracle.eval_f32(ora, pts)
        assert got.shape == (len(ax[2]), len(ax[1]), len(ax[0]))
        assert np.array_equal(got.reshape(-1).view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("name", ["csg64.sdf3d", "random_csg2.sdf3d"])
def test_forced_compiled_equals_tape(ctx, name):
    print(f"forced compiled {name}", flush=True)
    t = ctx.program_from_source(GOLDEN / name, path=True)
    c = ctx.program_from_source(GOLDEN / name, path=True, force_compiled=True)
    assert t.kind == s2m.PROGRAM_TAPE and c.kind == s2m.PROGRAM_COMPILED
    for res, flags in [(128, s2m.REFERENCE_COMPAT), (72, 0)]:
        p = s2m.Params.reference(res, 2.0, flags=flags)
        same_mesh(ctx.mesh(c, p), ctx.mesh(t, p))


def test_glsl_mandelbulb_equals_tape_op(ctx):
    # the compiled GLSL takes the same device math calls as the tape op evaluated as
    # written (S2M_MANDEL_TRIG): identical meshes
    print("glsl mandelbulb", flush=True)
    c = ctx.program_from_source(GOLDEN / "mandelbulb8.frag", lang="glsl", path=True)
    assert c.kind == s2m.PROGRAM_COMPILED
    t = ctx.program(tape.mandelbulb(8, trig=True))
    p = s2m.Params.reference(64, 2.5)
    same_mesh(ctx.mesh(c, p), ctx.mesh(t, p))


GYROID = """use sdf3d::*;

fn sdf3d(p: vec3f) -> f32 {
    let q = p * 20.0;
    return sin(q.x) * cos(q.y) + sin(q.y) * cos(q.z) + sin(q.z) * cos(q.x);
}
"""


def test_capacity_regrow():
    # a fresh context's first vertex capacity at 64^3 is 65536; this gyroid has more
    # vertices, so the first job overflows and repeats with exact sizes (on the dense
    # path, vertices past the capacity form no quads until then).  It must equal the same
    # job run again in the grown context.
    print("capacity regrow", flush=True)
    with s2m.Context(0) as fresh:
        prog = fresh.program_from_source(GYROID)
        p = s2m.Params.reference(64, 2.0)
        first = fresh.mesh(prog, p)
        again = fresh.mesh(prog, p)
        prog.close()
    assert len(first.vertices) > 65536
    same_mesh(first, again)
    assert len(first.triangles) > len(first.vertices)


def test_batch_mixes_compiled_and_tape(ctx, engraved):
    print("mixed batch", flush=True)
    progs = [ctx.program(tape.torus_example()), engraved, ctx.program(tape.random_csg(5))]
    params = [s2m.Params.reference(48, 2.0), s2m.Params.reference(40, 2.0),
              s2m.Params.reference(32, 2.0)]
    batch = ctx.mesh_batch(progs, params)
    for prog, p, m in zip(progs, params, batch):
        same_mesh(m, ctx.mesh(prog, p))


def test_user_normal_function(ctx):
    print("user normal", flush=True)
    # without `use sdf3d::normal` the program supplies sdf3d_normal itself
    src = """use sdf3d::primitives;
fn sdf3d(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.6); }
fn sdf3d_normal(p: vec3f, eps: f32) -> vec3f { return vec3(p.x, 2.0 * p.y, 0.0); }
"""
    prog = ctx.program_from_source(src)
    assert prog.kind == s2m.PROGRAM_COMPILED
    m = ctx.mesh(prog, s2m.Params.reference(32, 2.0))
    v = m.vertices.astype(np.float32)
    assert len(v) > 100
    gx, gy, gz = v[:, 0], np.float32(2.0) * v[:, 1], np.zeros_like(v[:, 0])
    with np.errstate(invalid="ignore", divide="ignore"):
        ln = np.sqrt((gx * gx + gy * gy) + gz * gz).astype(np.float32)
        want = np.stack([gx / ln, gy / ln, gz / ln], 1).astype(np.float32)
    nan = np.isnan(want)  # g = 0 on the z axis: NaN on both sides (payload may differ)
    assert np.array_equal(np.isnan(v[:, 3:]), nan)
    assert np.array_equal(v[:, 3:][~nan].view(np.uint32), want[~nan].view(np.uint32))
