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
@pytest.mark.parametrize("dense", [False, True])
def test_random_csg(ctx, seed, dense):
    print(f"random_csg seed={seed} dense={dense}", flush=True)
    node = tape.random_csg(seed)
    p = s2m.Params.reference(40, 2.0, flags=s2m.REFERENCE_COMPAT | (s2m.DENSE_EVAL if dense else 0))
    prog = ctx.program(node)
    assert_same_mesh(ctx.mesh(prog, p), reference_mesh(node, p))


@pytest.mark.parametrize("flags", [0, s2m.REFERENCE_COMPAT])
def test_csg64_dense_path(ctx, flags):
    print(f"csg64 dense flags={flags}", flush=True)
    node = tape.csg64()
    p = s2m.Params.reference(56, 2.0, flags=flags | s2m.DENSE_EVAL)
    assert_same_mesh(ctx.mesh(ctx.program(node), p), reference_mesh(node, p))


@pytest.mark.parametrize("res,flags", [(512, s2m.REFERENCE_COMPAT), (384, 0)])
def test_pruned_equals_dense_full_size(ctx, res, flags):
    # size-independent check at the workload size: the pruned path (tile culling +
    # pruned tapes) returns exactly the dense path's mesh
    print(f"csg64 pruned vs dense res={res}", flush=True)
    prog = ctx.program(tape.csg64())
    a = ctx.mesh(prog, s2m.Params.reference(res, 2.0, flags=flags))
    ta = ctx.last_timing()
    b = ctx.mesh(prog, s2m.Params.reference(res, 2.0, flags=flags | s2m.DENSE_EVAL))
    assert ta["tiles_active"] < ta["tiles_total"]
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert np.array_equal(a.vertices[:, :3].view(np.uint32), b.vertices[:, :3].view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert a.invalid_quads == b.invalid_quads
    # pruning is exact where the normal taps fall too: the normals agree bit for bit
    assert np.array_equal(a.vertices[:, 3:].view(np.uint32), b.vertices[:, 3:].view(np.uint32))


@pytest.mark.parametrize("dense", [False, True])
def test_no_normals(ctx, dense):
    # S2M_NO_NORMALS: the same mesh with zero normals
    print(f"csg64 no normals dense={dense}", flush=True)
    prog = ctx.program(tape.csg64())
    flags = s2m.REFERENCE_COMPAT | (s2m.DENSE_EVAL if dense else 0)
    a = ctx.mesh(prog, s2m.Params.reference(64, 2.0, flags=flags))
    b = ctx.mesh(prog, s2m.Params.reference(64, 2.0, flags=flags | s2m.NO_NORMALS))
    assert len(a.vertices) > 0
    assert np.array_equal(a.vertices[:, :3].view(np.uint32), b.vertices[:, :3].view(np.uint32))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert not np.any(b.vertices[:, 3:])


@pytest.mark.parametrize("slab", [0, 1, 3, 7])
def test_slab_invariance(ctx, slab):
    # the dense path in z-slabs of `slab` cell planes (0: automatic): the column count
    # pass, the activity bits cleared between slabs and the map plane carried from one slab
    # to the next must give the tile-pruned path's mesh, also on a size that is not a
    # multiple of 64 (rows that end inside a wave, bit words shared by two rows)
    node = tape.csg64()
    prog = ctx.program(node)
    for res in (64, 61):
        print(f"slab={slab} res={res}", flush=True)
        a = ctx.mesh(prog, s2m.Params.reference(res, 2.0))
        b = ctx.mesh(prog, s2m.Params.reference(res, 2.0, flags=s2m.REFERENCE_COMPAT | s2m.DENSE_EVAL, slab=slab))
        assert len(a.vertices) > 0
        assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
        assert np.array_equal(a.cell_keys, b.cell_keys)
        assert np.array_equal(a.signs, b.signs)
        assert np.array_equal(a.triangles, b.triangles)
        assert a.invalid_quads == b.invalid_quads


def test_sincos_matches_sin_cos(ctx):
    # the Mandelbulb's escape step takes sine and cosine with one sincosf; over every
    # float input its bits must be sinf's and cosf's (the calls the reference's shader
    # makes), NaN payloads included
    print("sincos over every float", flush=True)
    assert ctx.check_sincos() == (0, 0, 0, 0)


def test_bailout_threshold_exact(ctx):
    # the power-8 step tests the bailout on the squared length; over bailouts spread
    # across every float bit pattern, each squared-length test must agree with the
    # correctly rounded root's (the written form's `r > bailout`)
    checked, bad = ctx.check_bailout()
    print(f"bailout threshold: {checked} comparisons, {bad} disagreements", flush=True)
    assert checked > 40 * (1 << 20) and bad == 0


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


def test_contour_cells_special_values(ctx):
    # corners drawn half from special values (+-0, +-smallest denormal, +-1e-38, +-1e-30,
    # +-1, +-2e38, +-3e38) and half from N(0, 1): edge differences that overflow and edge
    # parameters that underflow to 0 (such an edge from corner 000 does not count toward
    # the cell's vertex).  The count pass's verdict and the emitted mesh must equal the
    # oracle's.
    rng = np.random.default_rng(1)
    pool = np.array([0.0, -0.0, 1e-45, -1e-45, 1e-38, -1e-38, 3e38, -3e38, 1.0, -1.0, 1e-30, -1e-30,
                     2e38, -2.5e38], np.float32)
    res = 24
    corners = rng.standard_normal((res, res, res, 8)).astype(np.float32)
    pick = rng.random(corners.shape) < 0.5
    corners[pick] = pool[rng.integers(0, len(pool), int(pick.sum()))]
    with np.errstate(all="ignore"):
        t = (np.float32(0) - corners[..., 0]) / (corners[..., 4] - corners[..., 0])
    straddle = (corners[..., 0] > 0) != (corners[..., 4] > 0)
    assert (straddle & ~(t > 0)).sum() > 100  # edges whose parameter is 0
    blo, bhi = oracle.axis_bounds(np.float32(-1), np.float32(1), res)
    for lag in (False, True):
        m = ctx.contour_cells(corners, [blo] * 3, [bhi] * 3, s2m.COMPAT_DROP_LAST_SLICE if lag else 0)
        r = oracle.contour_corners(corners, [blo] * 3, [bhi] * 3, lag=lag)
        assert len(r.vertices) > 1000
        assert_same_mesh(m, r, normal_atol=None)


def lattice_corners(lat, params, res):
    """Per-cell corners (rz, ry, rx, 8) and per-axis cell bounds read from a corner
    lattice evaluated on the library's coordinate tables (res: cells per axis, one
    number for a cube or (rx, ry, rz))."""
    res = (res,) * 3 if np.isscalar(res) else tuple(res)
    exact = bool(params.flags & s2m.CORNERS_EXACT)
    tabs = [s2m.axis_table(params.bb_min[a], params.bb_max[a], res[a], exact) for a in range(3)]
    (cx, lx, hx), (cy, ly, hy), (cz, lz, hz) = tabs
    idx = [(lx, hx), (ly, hy), (lz, hz)]
    corners = np.empty((res[2], res[1], res[0], 8), np.float32)
    for k in range(8):
        xi = idx[0][k & 1]
        yi = idx[1][(k >> 1) & 1]
        zi = idx[2][(k >> 2) & 1]
        corners[..., k] = lat[np.ix_(zi, yi, xi)]
    lo = [c[l] for c, l, _ in tabs]
    hi = [c[h] for c, _, h in tabs]
    return corners, lo, hi


@pytest.mark.parametrize("iters,res,flags", [(8, 48, s2m.REFERENCE_COMPAT), (5, 40, 0),
                                             (8, 64, s2m.COMPAT_DROP_LAST_SLICE)])
def test_mandelbulb_contour_parity(ctx, iters, res, flags):
    # transcendental SDF: feed the GPU's own corner values to the oracle's contouring;
    # masks, active cells, quads and positions must then match bit for bit
    print(f"mandelbulb iters={iters} res={res} flags={flags}", flush=True)
    node = tape.mandelbulb(iters)
    p = s2m.Params.reference(res, 2.5, flags=flags)
    if flags & s2m.COMPAT_POW2:
        res = s2m.round_resolution(res)  # 48 -> 64
    prog = ctx.program(node)
    lat = ctx.eval_lattice(prog, p)
    corners, lo, hi = lattice_corners(lat, p, res)
    m = ctx.mesh(prog, p)
    r = oracle.contour_corners(corners, lo, hi, lag=bool(flags & s2m.COMPAT_DROP_LAST_SLICE))
    assert len(m.vertices) > 100
    assert_same_mesh(m, r, normal_atol=None)


@pytest.mark.parametrize("iters,res,flags", [(8, 160, s2m.REFERENCE_COMPAT), (5, 97, s2m.CORNERS_EXACT)])
def test_mandelbulb_kernel_equals_interpreter(ctx, iters, res, flags):
    # A tape that is one Mandelbulb op is evaluated by k_eval_mandelbulb (lattice) and
    # k_normals<1> (normals), which call the estimator without the tape interpreter.
    # union(M, M) = min(d, d) = d takes the interpreter: on the dense path (S2M_DENSE_EVAL)
    # both must give the same mesh bit for bit, normals included, over several slabs.
    print(f"mandelbulb kernel vs interpreter iters={iters} res={res}", flush=True)
    one = tape.mandelbulb(iters)
    two = tape.union(tape.mandelbulb(iters), tape.mandelbulb(iters))
    assert len(one.tape()) == 1 and len(two.tape()) == 3
    p = s2m.Params.reference(res, 2.5, flags=flags | s2m.DENSE_EVAL, slab=24)
    a = ctx.mesh(ctx.program(one), p)
    b = ctx.mesh(ctx.program(two), p)
    assert len(a.vertices) > 1000
    assert a.invalid_quads == b.invalid_quads
    assert np.array_equal(a.cell_keys, b.cell_keys) and np.array_equal(a.signs, b.signs)
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))  # positions and normals
    assert np.array_equal(a.triangles, b.triangles)
    la = ctx.eval_lattice(ctx.program(one), p)
    lb = ctx.eval_lattice(ctx.program(two), p)
    assert np.array_equal(la.view(np.uint32), lb.view(np.uint32))


MANDEL_ULP_FACTOR = 32    # C in |f32 - f64| <= C * (ulp(f) + s)
MANDEL_BAIL_MARGIN = 1e-5  # points whose bailout tests all clear the threshold by this


@pytest.mark.parametrize("form", ["squaring", "trig", "glsl"])
@pytest.mark.parametrize("iters,res,flags", [(8, 32, s2m.COMPAT_DROP_LAST_SLICE),
                                             (8, 48, s2m.REFERENCE_COMPAT),
                                             (5, 40, s2m.REFERENCE_COMPAT)])
def test_mandelbulb_lattice_vs_fp64(ctx, iters, res, flags, form):
    # GPU fp32 against the fp64 oracle, at every lattice point.
    # Bound: |f32 - f64| <= C * (ulp_f32(f) + s) with C = 32, where s is the first-order
    # error scale of the evaluation (oracle.mandel_cond: the summed effect on the fp64
    # result of one relative rounding 2^-24 of each intermediate - z per step, r, theta,
    # phi, r^8, dr, the final log, product and quotient).  It holds at every point whose
    # bailout tests all clear |r - 2| / 2 >= 1e-5; the rest (a rounding there can take the
    # other branch of `if (r > 2.0) continue;`, a jump no error bound covers) are counted
    # and must stay under 0.1 % of the points.  Three forms of the estimator: the tape op's
    # power-8 step by complex squaring with hardware reciprocal square roots (the
    # default), the tape op evaluated as written with ocml's transcendentals
    # (S2M_MANDEL_TRIG), and the GLSL source compiled at run time.
    node = tape.mandelbulb(iters)
    if form == "glsl":
        prog = ctx.program_from_source(tape.mandelbulb_glsl(iters), lang="glsl")
        assert prog.kind == s2m.PROGRAM_COMPILED
    else:
        prog = ctx.program(tape.mandelbulb(iters, trig=form == "trig"))
    p = s2m.Params.reference(res, 2.5, flags=flags)
    exact = bool(flags & s2m.CORNERS_EXACT)
    res = s2m.round_resolution(res) if flags & s2m.COMPAT_POW2 else res
    lat = ctx.eval_lattice(prog, p)
    cs = [s2m.axis_table(p.bb_min[a], p.bb_max[a], res, exact)[0] for a in range(3)]
    Z, Y, X = np.meshgrid(cs[2], cs[1], cs[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    g = lat.ravel().astype(np.float64)
    r64 = oracle.eval_f64(node, pts)
    cond = oracle.mandel_cond(node, pts)
    fin = np.isfinite(r64)
    assert np.array_equal(np.isfinite(g), fin)
    keep = fin & (cond[:, 0] >= MANDEL_BAIL_MARGIN)
    ulp = np.spacing(np.abs(r64).astype(np.float32)).astype(np.float64)
    bound = MANDEL_ULP_FACTOR * (ulp + cond[:, 2])
    err = np.abs(g - r64)
    ratio = err[keep] / (ulp + cond[:, 2])[keep]
    print(f"mandelbulb {form} iters={iters} res={res}: {keep.sum()} of {fin.sum()} points checked, "
          f"{fin.sum() - keep.sum()} within {MANDEL_BAIL_MARGIN} of a bailout test; "
          f"max err / (ulp + s) = {ratio.max():.3g}, max err {err[fin].max():.3g}", flush=True)
    assert fin.sum() - keep.sum() <= 0.001 * fin.sum()
    bad = np.nonzero(keep & (err > bound))[0]
    assert len(bad) == 0, (pts[bad[:5]], g[bad[:5]], r64[bad[:5]], cond[bad[:5]])


@pytest.mark.parametrize("res,flags", [(37, 0), (61, s2m.CORNERS_EXACT), (99, s2m.CORNERS_EXACT),
                                       (45, s2m.COMPAT_DROP_LAST_SLICE)])
@pytest.mark.parametrize("seed", [4, 11])
def test_partial_tiles(ctx, res, flags, seed):
    # resolutions that are not multiples of the 8-cell tile: partial tiles, partial
    # bricks and partial super-tiles at the grid's far faces
    print(f"partial tiles res={res} flags={flags} seed={seed}", flush=True)
    node = tape.random_csg(seed)
    p = s2m.Params.reference(res, 2.0, flags=flags)
    assert_same_mesh(ctx.mesh(ctx.program(node), p), reference_mesh(node, p))


@pytest.mark.parametrize("res", [24, 40])
def test_csg64_linear_vs_dense_small(ctx, res):
    # the left-fold path (brick culling, two points per lane) against the dense path on
    # grids whose bricks straddle the surface everywhere
    print(f"csg64 linear vs dense res={res}", flush=True)
    prog = ctx.program(tape.csg64())
    for flags in (0, s2m.REFERENCE_COMPAT):
        a = ctx.mesh(prog, s2m.Params.reference(res, 1.3, flags=flags))
        b = ctx.mesh(prog, s2m.Params.reference(res, 1.3, flags=flags | s2m.DENSE_EVAL))
        assert np.array_equal(a.vertices[:, :3].view(np.uint32), b.vertices[:, :3].view(np.uint32))
        assert np.array_equal(a.triangles, b.triangles)
        assert a.invalid_quads == b.invalid_quads
        np.testing.assert_allclose(a.vertices[:, 3:], b.vertices[:, 3:], atol=1e-6, rtol=0)


# grids and boxes that differ per axis (the uniform's dims and bounds are per-axis vectors,
# dualcontour.wgsl:22-27; the CLI only makes cubes)
ANISO = [((40, 24, 56), (-1.0, -0.6, -1.2), (1.0, 0.6, 1.2)),
         ((17, 64, 9), (-0.9, -1.0, -0.3), (0.9, 1.0, 0.3)),
         ((64, 8, 33), (-1.0, -0.2, -1.0), (1.0, 0.2, 1.0)),
         ((48, 36, 20), (-0.3, -1.1, -0.2), (1.1, 0.4, 0.9))]  # not centred on the origin


@pytest.mark.parametrize("grid", range(len(ANISO)))
@pytest.mark.parametrize("flags", [0, s2m.CORNERS_EXACT, s2m.REFERENCE_COMPAT])
def test_anisotropic_grids(ctx, grid, flags):
    # per-axis coordinate tables, tile counts and partial tiles on each axis separately:
    # the tile-sparse path (a left fold and a general program), its multi-pass form and
    # the dense path, each against the oracle
    res, bmin, bmax = ANISO[grid]
    for name, node in (("csg64", tape.csg64()), ("random_csg(5)", tape.random_csg(5)),
                       ("torus", tape.torus_example())):
        prog = ctx.program(node)
        r = None
        for extra, slab in ((0, 0), (0, 16), (s2m.DENSE_EVAL, 0), (s2m.DENSE_EVAL, 8)):
            p = s2m.Params(bmin, bmax, res, 1e-4, flags | extra, slab)
            print(f"anisotropic {name} res={res} flags={flags | extra} slab={slab}", flush=True)
            if r is None:
                r = reference_mesh(node, p)
                assert len(r.vertices) > 100
            assert_same_mesh(ctx.mesh(prog, p), r)


@pytest.mark.parametrize("res,flags", [((33, 20, 27), s2m.REFERENCE_COMPAT), ((24, 40, 31), 0)])
def test_anisotropic_mandelbulb(ctx, res, flags):
    # the Mandelbulb kernel on a grid and box that differ per axis, over several slabs:
    # the oracle's contouring of the GPU's own corners must give the same mesh
    print(f"anisotropic mandelbulb res={res} flags={flags}", flush=True)
    p = s2m.Params((-1.25, -0.9, -1.1), (1.25, 0.9, 1.1), res, 1e-4, flags, 8)
    cells = [s2m.round_resolution(r) for r in res] if flags & s2m.COMPAT_POW2 else list(res)
    prog = ctx.program(tape.mandelbulb(6))
    lat = ctx.eval_lattice(prog, p)
    corners, lo, hi = lattice_corners(lat, p, cells)
    m = ctx.mesh(prog, p)
    r = oracle.contour_corners(corners, lo, hi, lag=bool(flags & s2m.COMPAT_DROP_LAST_SLICE))
    assert len(m.vertices) > 100
    assert_same_mesh(m, r, normal_atol=None)


# boxes far smaller or larger than the programs' features: cells a few dozen float
# spacings wide around a point of the surface (the interval pruning's margins and the
# crossing parameters at the limit of fp32), and whole primitives within a few cells
SCALES = [("torus", (0.69, -0.01, -0.01), (0.71, 0.01, 0.01), (64, 64, 64)),
          ("torus", (0.6999, -1e-4, -1e-4), (0.7001, 1e-4, 1e-4), (48, 48, 48)),
          ("csg64", (-0.05, -0.05, -0.05), (0.05, 0.05, 0.05), (64, 64, 64)),
          ("csg64", (-6.0, -6.0, -6.0), (6.0, 6.0, 6.0), (128, 128, 128)),
          ("torus", (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), (256, 256, 256))]


@pytest.mark.parametrize("case", range(len(SCALES)))
@pytest.mark.parametrize("flags", [0, s2m.REFERENCE_COMPAT])
def test_tiny_and_large_boxes(ctx, case, flags):
    name, bmin, bmax, res = SCALES[case]
    node = tape.csg64() if name == "csg64" else tape.torus_example()
    prog = ctx.program(node)
    r = None
    for extra in (0, s2m.DENSE_EVAL):
        p = s2m.Params(bmin, bmax, res, 1e-4, flags | extra)
        print(f"scale {name} box {bmin}..{bmax} res={res} flags={flags | extra}", flush=True)
        if r is None:
            r = reference_mesh(node, p)
            assert len(r.vertices) > 100
        assert_same_mesh(ctx.mesh(prog, p), r)


def op_zoo(k: float):
    """Every operator (the smooth ones with radius k) over every primitive, nested so
    that each operator's operands are themselves composite somewhere."""
    sph = tape.sphere((0.1, -0.05, 0.0), 0.45)
    box = tape.box((-0.2, 0.1, 0.05), (0.6, 0.35, 0.5))
    tor = tape.torus((0.05, 0.0, -0.1), (0.4, 0.12))
    cyl = tape.cylinder((0.0, 0.2, 0.15), 0.7, 0.18)
    cap = tape.capsule((-0.5, -0.3, -0.2), (0.45, 0.35, 0.3), 0.12)
    a = tape.smooth_intersection(tape.smooth_union(sph, cyl, k), box, k)
    b = tape.smooth_subtraction(tape.intersection(tor, tape.sphere((0.3, 0.0, 0.0), 0.5)), cap, k)
    c = tape.subtraction(tape.sphere((0.0, 0.0, 0.0), 0.2), tape.union(a, b))
    return tape.smooth_union(c, tape.smooth_subtraction(cyl, tor, k), k)


@pytest.mark.parametrize("k", [0.2, 0.05, 1e-3, 0.0])
@pytest.mark.parametrize("res,flags", [(48, s2m.REFERENCE_COMPAT), (61, 0), (80, s2m.CORNERS_EXACT)])
def test_operator_zoo(ctx, k, res, flags):
    # all six operators over all five library primitives in one general (not left-fold)
    # program: the interval rules of every operator in the pruning (k = 0 disables the
    # smooth ones' pruning; their division by k then gives the same infinities and NaNs
    # on both sides), the tape interpreter's dispatch, sparse and dense paths
    node = op_zoo(k)
    prog = ctx.program(node)
    r = None
    for extra in (0, s2m.DENSE_EVAL):
        p = s2m.Params.reference(res, 2.0, flags=flags | extra)
        print(f"operator zoo k={k} res={res} flags={flags | extra}", flush=True)
        if r is None:
            r = reference_mesh(node, p)
            assert len(r.vertices) > 200
        assert_same_mesh(ctx.mesh(prog, p), r)


def clustered_fold(n: int, seed: int, k: float):
    """A smooth-union fold of n overlapping spheres (2n - 1 ops); with a large radius k
    no primitive can be pruned, so every tile keeps the whole program."""
    rng = tape.Rng(seed)
    prims = [tape.sphere((rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3)),
                         rng.uniform(0.12, 0.3)) for _ in range(n)]
    return tape.fold(prims, tape.smooth_union, k)


def balanced_tree(n: int, seed: int):
    """Pairwise reduction of n mixed primitives under cycling operators (2n - 1 ops,
    operand stack depth log2(n) + 1): a general program, not a left fold."""
    rng = tape.Rng(seed)
    level = []
    for i in range(n):
        c = (rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6))
        s = rng.uniform(0.05, 0.15)
        level.append([tape.sphere(c, s), tape.box(c, (2 * s, s, 1.5 * s)), tape.torus(c, (s, 0.4 * s))][i % 3])
    ops = [tape.union, lambda a, b: tape.smooth_union(a, b, 0.03), tape.union,
           lambda a, b: tape.smooth_subtraction(a, b, 0.02), tape.subtraction]
    i = 0
    while len(level) > 1:
        nxt = []
        for j in range(0, len(level) - 1, 2):
            nxt.append(ops[i % len(ops)](level[j], level[j + 1]))
            i += 1
        if len(level) % 2:
            nxt.append(level[-1])
        level = nxt
    return level[0]


def right_chain(n: int, seed: int):
    """P0 op (P1 op (... op P(n-1))): operand stack depth n (16 is the accepted maximum)."""
    rng = tape.Rng(seed)
    prims = [tape.sphere((rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)),
                         rng.uniform(0.1, 0.35)) for _ in range(n)]
    d = prims[-1]
    for p in reversed(prims[:-1]):
        d = tape.smooth_union(p, d, 0.04)
    return d


LONG = {"fold300_k0.05": lambda: clustered_fold(300, 3, 0.05),
        "fold300_k0.5": lambda: clustered_fold(300, 3, 0.5),
        "tree256": lambda: balanced_tree(256, 5),
        "chain16": lambda: right_chain(16, 7)}


@pytest.mark.parametrize("name", list(LONG))
@pytest.mark.parametrize("res,flags", [(64, 0), (57, s2m.CORNERS_EXACT), (48, s2m.REFERENCE_COMPAT)])
def test_long_and_deep_programs(ctx, name, res, flags):
    # programs past the 128 ops staged in LDS (the pruning's unstaged walk, the tile and
    # emission kernels' interpreter over global memory) and the deepest operand stack
    node = LONG[name]()
    prog = ctx.program(node)
    r = None
    for extra in (0, s2m.DENSE_EVAL):
        p = s2m.Params.reference(res, 2.0, flags=flags | extra)
        print(f"program {name} ({len(node.tape())} ops) res={res} flags={flags | extra}", flush=True)
        if r is None:
            r = reference_mesh(node, p)
            assert len(r.vertices) > 500
        assert_same_mesh(ctx.mesh(prog, p), r)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_largest_resolution_per_axis(ctx, axis):
    # 65535 cells on one axis, the most a u16 cell-key component holds (mesh.rs:214): a
    # capsule through the whole length of a 2000-long box, so that the last cell index
    # carries vertices; 8192 tiles in a row along x, 65535 rows along y, 65535 cell planes
    # (passes and slabs) along z.  Rounding 65535 up to a power of two is refused.
    res = [4, 4, 4]
    res[axis] = 65535
    a, b, lo, hi = [0.0] * 3, [0.0] * 3, [-0.5] * 3, [0.5] * 3
    a[axis], b[axis], lo[axis], hi[axis] = -1100.0, 1100.0, -1000.0, 1000.0
    node = tape.capsule(tuple(a), tuple(b), 0.3)
    prog = ctx.program(node)
    for flags in (0, s2m.CORNERS_EXACT | s2m.COMPAT_DROP_LAST_SLICE):
        want = None
        for extra in (0, s2m.DENSE_EVAL):
            p = s2m.Params(tuple(lo), tuple(hi), tuple(res), 1e-4, flags | extra)
            print(f"largest resolution axis={axis} res={res} flags={flags | extra}", flush=True)
            if want is None:
                want = reference_mesh(node, p)
                top = (want.keys.astype(np.uint64) >> np.uint64(16 * axis)) & np.uint64(0xFFFF)
                assert int(top.max()) >= 65533  # 65534, or 65533 under the readback lag on z
            assert_same_mesh(ctx.mesh(prog, p), want)
    with pytest.raises(s2m.S2MError):
        ctx.mesh(prog, s2m.Params(tuple(lo), tuple(hi), tuple(res), 1e-4, s2m.COMPAT_POW2))


def degenerate_capsule_programs():
    """A capsule whose segment has length 0 (its clamp takes 0 / 0 = NaN, which the IEEE
    max of the clamp turns into 0: a sphere) under every operator, on both sides."""
    deg = tape.capsule((0.1, 0.0, 0.0), (0.1, 0.0, 0.0), 0.3)
    sph = tape.sphere((0.0, 0.0, 0.0), 0.5)
    tor = tape.torus_example()
    return {"union": tape.union(deg, sph), "union_rev": tape.union(sph, deg),
            "intersection": tape.intersection(tor, deg), "subtraction": tape.subtraction(deg, sph),
            "smooth_union": tape.smooth_union(deg, sph, 0.1),
            "smooth_intersection": tape.smooth_intersection(sph, deg, 0.1),
            "mixed": tape.union(tape.smooth_union(tor, deg, 0.05), tape.intersection(sph, tape.union(deg, tor)))}


@pytest.mark.parametrize("name", list(degenerate_capsule_programs()))
@pytest.mark.parametrize("res,flags", [(48, 0), (61, s2m.REFERENCE_COMPAT)])
def test_degenerate_capsule(ctx, name, res, flags):
    # the NaN inside the capsule's clamp: the pruning's interval of the primitive, the
    # interpreters and the oracle must agree on its value everywhere
    node = degenerate_capsule_programs()[name]
    prog = ctx.program(node)
    r = None
    for extra in (0, s2m.DENSE_EVAL):
        p = s2m.Params.reference(res, 2.0, flags=flags | extra)
        print(f"degenerate capsule {name} res={res} flags={flags | extra}", flush=True)
        if r is None:
            r = reference_mesh(node, p)
            assert len(r.vertices) > 100
        assert_same_mesh(ctx.mesh(prog, p), r)


def test_randomized_sweep(ctx):
    # Seeded random cases: random CSG trees, CSG-64, the operator zoo with random radii,
    # clustered folds of random length and radius and balanced trees, on random grids
    # (often anisotropic), boxes, flag sets and pass sizes, sparse and dense, each vs the
    # oracle.
    import random

    rng = random.Random(20261203)
    for it in range(200):
        seed = rng.randrange(1, 10 ** 6)
        u = rng.random()
        if u < 0.7:
            node, name = tape.random_csg(seed), f"random_csg({seed})"
        elif u < 0.75:
            node, name = tape.csg64(), "csg64"
        elif u < 0.85:
            k = rng.choice([0.0, 1e-3, 0.02, 0.1, 0.3])
            node, name = op_zoo(k), f"op_zoo({k})"
        elif u < 0.95:
            m, k = rng.randint(20, 150), rng.choice([0.01, 0.05, 0.2, 0.5])
            node, name = clustered_fold(m, seed, k), f"clustered_fold({m}, {seed}, {k})"
        else:
            m = rng.randint(8, 200)
            node, name = balanced_tree(m, seed), f"balanced_tree({m}, {seed})"
        res = tuple(rng.choice([9, 16, 17, 24, 31, 33, 40, 48, 57, 64, 80, 96]) for _ in range(3))
        if rng.random() < 0.4:
            res = (res[0],) * 3
        half = [rng.uniform(0.3, 1.3) for _ in range(3)]
        ctr = [rng.uniform(-0.3, 0.3) for _ in range(3)]
        bmin = tuple(c - h for c, h in zip(ctr, half))
        bmax = tuple(c + h for c, h in zip(ctr, half))
        flags = rng.choice([0, s2m.CORNERS_EXACT, s2m.REFERENCE_COMPAT, s2m.COMPAT_DROP_LAST_SLICE,
                            s2m.COMPAT_POW2])
        slab = rng.choice([0, 0, 0, 8, 16])
        print(f"sweep case {it}: {name} res={res} box={bmin}..{bmax} flags={flags} slab={slab}", flush=True)
        prog = ctx.program(node)
        want = None
        for extra in (0, s2m.DENSE_EVAL):
            p = s2m.Params(bmin, bmax, res, 1e-4, flags | extra, slab)
            if want is None:
                want = reference_mesh(node, p)
            assert_same_mesh(ctx.mesh(prog, p), want)
