"""GPU parity: the HIP path through the C ABI against the oracle, on the same inputs.

Integer/index outputs (cell keys, sign bits, triangle indices, invalid-quad counts) and
vertex positions must be bit-exact: both sides compute fp32 in the reference's
operation order with correctly rounded division and square root and no contraction.
Normals (4-tap gradient, PLY only) are compared within 1e-5 absolute.
"""
import numpy as np
import pytest

import oracle
import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape

pytestmark = pytest.mark.gpu

NORMAL_ATOL = 1e-5


def assert_same_mesh(m, r, normal_atol=NORMAL_ATOL):
    assert len(m.vertices) == len(r.vertices), (len(m.vertices), len(r.vertices))
    assert np.array_equal(m.cell_keys, r.keys)
    assert np.array_equal(m.signs, r.signs)
    assert np.array_equal(m.vertices[:, :3].view(np.uint32), r.vertices[:, :3].view(np.uint32))
    assert len(m.triangles) == len(r.triangles), (len(m.triangles), len(r.triangles))
    assert np.array_equal(m.triangles, r.triangles)
    assert m.invalid_quads == r.invalid_quads
    if normal_atol is not None and len(m.vertices):
        np.testing.assert_allclose(m.vertices[:, 3:], r.vertices[:, 3:], atol=normal_atol, rtol=0)


def reference_mesh(node, params: s2m.Params):
    res = [s2m.round_resolution(r) for r in params.res] if params.flags & s2m.COMPAT_POW2 \
        else list(params.res)
    return oracle.mesh(node, params.bb_min, params.bb_max, res, params.eps,
                       lattice=not (params.flags & s2m.CORNERS_EXACT),
                       lag=bool(params.flags & s2m.COMPAT_DROP_LAST_SLICE))


@pytest.mark.parametrize("res", [2, 3, 16, 32, 64, 128])
def test_torus_reference_compat(ctx, res):
    print(f"torus res={res}", flush=True)
    node = tape.torus_example()
    p = s2m.Params.reference(res, 2.0)
    prog = ctx.program(node)
    assert_same_mesh(ctx.mesh(prog, p), reference_mesh(node, p))


@pytest.mark.parametrize("flags", [0, s2m.COMPAT_DROP_LAST_SLICE, s2m.CORNERS_EXACT,
                                   s2m.REFERENCE_COMPAT])
def test_csg64_modes(ctx, flags):
    print(f"csg64 res=48 flags={flags}", flush=True)
    node = tape.csg64()
    p = s2m.Params.reference(48, 2.0, flags=flags)
    prog = ctx.program(node)
    assert_same_mesh(ctx.mesh(prog, p), reference_mesh(node, p))


@pytest.mark.parametrize("seed", range(1, 9))
def test_random_csg(ctx, seed):
    print(f"random_csg seed={seed}", flush=True)
    node = tape.random_csg(seed)
    p = s2m.Params.reference(40, 2.0)
    prog = ctx.program(node)
    assert_same_mesh(ctx.mesh(prog, p), reference_mesh(node, p))


@pytest.mark.parametrize("slab", [1, 3, 7])
def test_slab_invariance(ctx, slab):
    print(f"slab={slab}", flush=True)
    node = tape.csg64()
    prog = ctx.program(node)
    a = ctx.mesh(prog, s2m.Params.reference(64, 2.0))
    b = ctx.mesh(prog, s2m.Params.reference(64, 2.0, slab=slab))
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert a.invalid_quads == b.invalid_quads


def test_lattice_values_bit_exact(ctx):
    node = tape.csg64()
    p = s2m.Params.reference(64, 2.0)
    lat = ctx.eval_lattice(ctx.program(node), p)
    cs = [s2m.axis_table(p.bb_min[a], p.bb_max[a], 64, True)[0] for a in range(3)]
    Z, Y, X = np.meshgrid(cs[2], cs[1], cs[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    ref = oracle.eval_f32(node, pts).reshape(lat.shape)
    assert lat.shape == (len(cs[2]), len(cs[1]), len(cs[0]))
    assert np.array_equal(lat, ref)  # equal as IEEE values (+0 == -0)


def test_contour_cells_bit_exact(ctx):
    node = tape.random_csg(3)
    res, b = 24, 2.0
    lo = np.float32(0) - np.float32(b) * np.float32(0.5)
    hi = np.float32(0) + np.float32(b) * np.float32(0.5)
    corners = oracle.cell_corners(node, [lo] * 3, [hi] * 3, [res] * 3)
    blo, bhi = oracle.axis_bounds(lo, hi, res)
    for lag in (False, True):
        m = ctx.contour_cells(corners, [blo] * 3, [bhi] * 3,
                              s2m.COMPAT_DROP_LAST_SLICE if lag else 0)
        r = oracle.contour_corners(corners, [blo] * 3, [bhi] * 3, lag=lag)
        assert_same_mesh(m, r, normal_atol=None)


def test_mandelbulb_lattice_stats(ctx):
    node = tape.mandelbulb(8)
    p = s2m.Params.reference(32, 2.5, flags=s2m.COMPAT_DROP_LAST_SLICE)
    lat = ctx.eval_lattice(ctx.program(node), p)
    cs = [s2m.axis_table(p.bb_min[a], p.bb_max[a], 32, False)[0] for a in range(3)]
    Z, Y, X = np.meshgrid(cs[2], cs[1], cs[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    r32 = oracle.eval_f32(node, pts)
    r64 = oracle.eval_f64(node, pts)
    g = lat.ravel().astype(np.float64)
    fin = np.isfinite(r64) & np.isfinite(g)
    err = np.abs(g[fin] - r64[fin])
    print("mandelbulb gpu-vs-f64 max", err.max(), "p99", np.quantile(err, 0.99),
          "cpu32-vs-f64 max", np.abs(r32[fin] - r64[fin]).max(),
          "gpu==cpu32", np.mean(lat.ravel() == r32), "nonfinite", (~fin).sum(), flush=True)
