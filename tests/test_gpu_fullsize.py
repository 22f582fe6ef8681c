"""GPU parity at the workloads' full sizes, and of the multi-pass / multi-slab paths.

  * CSG-64 at 512^3, reference-exact corners and the readback lag (the headline job):
    the HIP mesh against the oracle's mesh of the same job, bit for bit (the oracle takes
    about a minute on 16 cores);
  * the tile-pruned path cut into several z-passes (forced with `slab`, and the automatic
    split of a 1024^3 job whose pass data exceeds 1 GiB): the cell -> vertex map of each
    pass's last plane carried to the next pass must give the single-pass mesh, the dense
    mesh and the oracle's mesh;
  * the Mandelbulb at 1024^3 (32 dense slabs): slab-size invariance, keys strictly
    ascending, every vertex inside its cell, and one z-window contoured by the oracle from
    the GPU's own corner values equal to the same window of the GPU mesh;
  * Mandelbulb and compiled programs across several dense slabs.
"""
import numpy as np
import pytest

import oracle
import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape

from test_gpu_parity import assert_same_mesh, reference_mesh

pytestmark = pytest.mark.gpu


def same_mesh(a, b, normals=True):
    assert len(a.vertices) == len(b.vertices), (len(a.vertices), len(b.vertices))
    cols = slice(0, 6) if normals else slice(0, 3)
    assert np.array_equal(a.vertices[:, cols].view(np.uint32), b.vertices[:, cols].view(np.uint32))
    assert np.array_equal(a.cell_keys, b.cell_keys)
    assert np.array_equal(a.signs, b.signs)
    assert np.array_equal(a.triangles, b.triangles)
    assert a.invalid_quads == b.invalid_quads


def key_xyz(keys):
    keys = keys.astype(np.uint64)
    return ((keys & np.uint64(0xFFFF)).astype(np.int64),
            ((keys >> np.uint64(16)) & np.uint64(0xFFFF)).astype(np.int64),
            (keys >> np.uint64(32)).astype(np.int64))


@pytest.mark.timeout(900)
def test_csg64_512_reference_compat_vs_oracle(ctx):
    print("csg64 512^3 reference-compat: HIP mesh, then the oracle (~1 min)", flush=True)
    node = tape.csg64()
    p = s2m.Params.reference(512, 2.0)
    m = ctx.mesh(ctx.program(node), p)
    r = reference_mesh(node, p)
    print(f"csg64 512^3: {len(m.vertices)} vertices, {len(m.triangles)} triangles, "
          f"{m.invalid_quads} invalid quads", flush=True)
    assert len(m.vertices) > 900_000
    assert_same_mesh(m, r)


@pytest.mark.parametrize("res,flags", [(64, s2m.REFERENCE_COMPAT),
                                       (99, s2m.CORNERS_EXACT | s2m.COMPAT_DROP_LAST_SLICE),
                                       (61, 0), (128, s2m.REFERENCE_COMPAT)])
@pytest.mark.parametrize("which", ["csg64", "random5"])
def test_sparse_multipass(ctx, which, res, flags):
    # slab = 8 / 24 / 40 cell planes: 1, 3 or 5 tile layers per pass
    node = tape.csg64() if which == "csg64" else tape.random_csg(5)
    prog = ctx.program(node)
    p1 = s2m.Params.reference(res, 2.0, flags=flags)
    one = ctx.mesh(prog, p1)
    assert ctx.last_timing()["slabs"] == 1
    assert len(one.vertices) > 0
    if res < 128:
        assert_same_mesh(one, reference_mesh(node, p1))
    nz = res - (1 if flags & s2m.COMPAT_DROP_LAST_SLICE else 0)  # cell planes meshed
    tile_layers = -(-nz // 8)
    for slab in (8, 24, 40):
        print(f"sparse multipass {which} res={res} flags={flags} slab={slab}", flush=True)
        m = ctx.mesh(prog, s2m.Params.reference(res, 2.0, flags=flags, slab=slab))
        want_passes = -(-tile_layers // ((slab + 7) // 8))
        assert ctx.last_timing()["slabs"] == want_passes > 1
        same_mesh(m, one)


def test_sparse_auto_multipass_1024(ctx):
    # a 1024^3 CSG-64 job's pass data exceeds 1 GiB, so the job is split into z-passes
    # automatically; the mesh must equal the dense path's
    print("csg64 1024^3 sparse (automatic passes) vs dense", flush=True)
    prog = ctx.program(tape.csg64())
    a = ctx.mesh(prog, s2m.Params.reference(1024, 2.0))
    passes = ctx.last_timing()["slabs"]
    b = ctx.mesh(prog, s2m.Params.reference(1024, 2.0, flags=s2m.REFERENCE_COMPAT | s2m.DENSE_EVAL))
    print(f"csg64 1024^3: {passes} passes, {len(a.vertices)} vertices", flush=True)
    assert passes > 1
    same_mesh(a, b)


@pytest.mark.timeout(600)
def test_mandelbulb_1024_properties(ctx):
    node = tape.mandelbulb(8)
    prog = ctx.program(node)
    p = s2m.Params.reference(1024, 2.5)
    print("mandelbulb 1024^3: automatic slabs", flush=True)
    m = ctx.mesh(prog, p)
    slabs_auto = ctx.last_timing()["slabs"]
    print(f"mandelbulb 1024^3: {len(m.vertices)} vertices in {slabs_auto} slabs; slab=16", flush=True)
    m16 = ctx.mesh(prog, s2m.Params.reference(1024, 2.5, slab=16))
    assert ctx.last_timing()["slabs"] > slabs_auto > 1
    same_mesh(m, m16)
    del m16
    assert len(m.vertices) > 1_000_000

    keys = m.cell_keys
    assert np.all(np.diff(keys.astype(np.int64)) > 0)  # strictly ascending: VertexList order
    x, y, z = key_xyz(keys)
    assert z.max() <= 1022  # the lag drops the top cell plane
    # every vertex inside its cell [cmin, cmax] (one ulp of slack for the rounding of
    # cmin + ((cmax - cmin) * avg) / count)
    for a, idx in enumerate((x, y, z)):
        coord, lo, hi = s2m.axis_table(p.bb_min[a], p.bb_max[a], 1024, True)
        cmin, cmax = coord[lo][idx], coord[hi][idx]
        pa = m.vertices[:, a]
        assert np.all(pa >= np.nextafter(cmin, np.float32(-np.inf)))
        assert np.all(pa <= np.nextafter(cmax, np.float32(np.inf)))

    # one z-window: the GPU's corner values of cell planes [z0, z1), contoured by the oracle
    z0, z1 = 512, 520
    print(f"mandelbulb 1024^3: window z in [{z0}, {z1}) through the oracle", flush=True)
    lat, lz0 = ctx.eval_lattice_range(prog, p, z0, z1)
    tabs = [s2m.axis_table(p.bb_min[a], p.bb_max[a], 1024, True) for a in range(3)]
    (cx, lx, hx), (cy, ly, hy), (cz, lzt, hzt) = tabs
    zi = (lzt[z0:z1].astype(np.int64) - lz0, hzt[z0:z1].astype(np.int64) - lz0)
    corners = np.empty((z1 - z0, 1024, 1024, 8), np.float32)
    for k in range(8):
        corners[..., k] = lat[np.ix_(zi[(k >> 2) & 1], (ly, hy)[(k >> 1) & 1], (lx, hx)[k & 1])]
    lo = [cx[lx], cy[ly], cz[lzt[z0:z1]]]
    hi = [cx[hx], cy[hy], cz[hzt[z0:z1]]]
    w = oracle.contour_corners(corners, lo, hi, lag=False)
    del corners, lat
    sel = (z >= z0) & (z < z1)
    assert sel.sum() == len(w.vertices) > 10_000
    assert np.array_equal(keys[sel], w.keys + (np.uint64(z0) << np.uint64(32)))
    assert np.array_equal(m.signs[sel], w.signs)
    assert np.array_equal(m.vertices[sel, :3].view(np.uint32), w.vertices[:, :3].view(np.uint32))
    # quads whose four vertices lie in the window: the GPU's, in order, against the oracle's
    # (each valid quad is two consecutive triangles (q2,q1,q0), (q0,q3,q2))
    gq = m.triangles.reshape(-1, 2, 3)
    inside = sel[gq].all(axis=(1, 2))
    g_keys = keys[gq[inside]]
    w_keys = w.keys[w.triangles.reshape(-1, 2, 3)] + (np.uint64(z0) << np.uint64(32))
    assert len(g_keys) == len(w_keys) > 10_000
    assert np.array_equal(g_keys, w_keys)


@pytest.mark.parametrize("slab", [1, 3, 7])
def test_multislab_mandelbulb_and_compiled(ctx, slab):
    from pathlib import Path

    eng = ctx.program_from_source(Path(__file__).resolve().parent / "golden" / "engraved_cube.sdf3d",
                                  path=True)
    assert eng.kind == s2m.PROGRAM_COMPILED
    glsl = ctx.program_from_source(Path(__file__).resolve().parent / "golden" / "mandelbulb8.frag",
                                   lang="glsl", path=True)
    cases = [(ctx.program(tape.mandelbulb(8)), 2.5, 44), (ctx.program(tape.mandelbulb(5)), 2.5, 37),
             (eng, 2.0, 40), (glsl, 2.5, 40)]
    for prog, bounds, res in cases:
        print(f"multislab slab={slab} kind={prog.kind} res={res}", flush=True)
        for flags in (s2m.REFERENCE_COMPAT, 0):
            a = ctx.mesh(prog, s2m.Params.reference(res, bounds, flags=flags))
            b = ctx.mesh(prog, s2m.Params.reference(res, bounds, flags=flags, slab=slab))
            assert ctx.last_timing()["slabs"] > 1
            assert len(a.vertices) > 0
            same_mesh(a, b)
