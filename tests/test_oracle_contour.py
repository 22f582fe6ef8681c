"""Pins for the oracle's cell rule (dualcontour.wgsl:57-131).

The check is an independent geometric formulation: the 12 cube edges are the corner
pairs that differ in one coordinate; a sign change puts a crossing at a + t (b - a),
t = -v_a / (v_b - v_a); a crossing counts unless it is the cube origin; a vertex needs
two counted crossings and sits at their centroid.  It is brute-forced over all 3^8
sign classes of the corner values, and linear fields check the vertex lies on the plane.
"""
import itertools

import numpy as np
import pytest

import oracle

CORNERS = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]  # 000,100,010,...


def geometric_vertex(c):
    pts = []
    for a in range(8):
        for axis in range(3):
            if (a >> axis) & 1:
                continue
            b = a | (1 << axis)
            va, vb = float(c[a]), float(c[b])
            if (va > 0) == (vb > 0):
                continue
            t = (0.0 - va) / (vb - va)
            p = list(CORNERS[a])
            p[axis] = t
            if any(v > 0 for v in p):
                pts.append(p)
    if len(pts) <= 1:
        return None
    return np.mean(np.asarray(pts), axis=0)


def corner_values(classes, rng, zero=0.0):
    v = np.empty(8, np.float32)
    for i, k in enumerate(classes):
        if k == 0:
            v[i] = -rng.uniform(0.01, 1.0)
        elif k == 1:
            v[i] = zero
        else:
            v[i] = rng.uniform(0.01, 1.0)
    return v


@pytest.mark.parametrize("zero", [0.0, -0.0])
def test_all_sign_classes(zero):
    rng = np.random.default_rng(11)
    lo = np.zeros(3, np.float32)
    hi = np.ones(3, np.float32)
    n_active = 0
    for classes in itertools.product(range(3), repeat=8):
        c = corner_values(classes, rng, zero)
        act, pos = oracle.cell_vertex(c, lo, hi)
        want = geometric_vertex(c)
        assert act == (want is not None), classes
        if act:
            n_active += 1
            np.testing.assert_allclose(pos, want, atol=2e-6)
        sb = oracle.cell_signs(c)
        assert sb == (int(c[1] > 0) | int(c[2] > 0) << 1 | int(c[4] > 0) << 2 | int(c[0] > 0) << 3)
    assert 0 < n_active < 3 ** 8


def test_t_zero_exclusion():
    # c000 == 0 (class <= 0) with every other corner > 0: the three crossings are all at
    # the origin and are not counted, so there is no vertex (dualcontour.wgsl:120)
    c = np.array([0, 1, 1, 1, 1, 1, 1, 1], np.float32)
    act, _ = oracle.cell_vertex(c, np.zeros(3, np.float32), np.ones(3, np.float32))
    assert not act
    # the same sign pattern with c000 slightly negative has 3 counted crossings
    c[0] = -1e-3
    act, pos = oracle.cell_vertex(c, np.zeros(3, np.float32), np.ones(3, np.float32))
    assert act and np.all(pos > 0) and np.all(pos < 0.01)


def test_cell_scaling():
    # pos = cmin + (cmax - cmin) * avg / count
    rng = np.random.default_rng(5)
    c = np.array([-1, 1, -1, 1, -1, 1, -1, 1], np.float32) * np.float32(0.5)
    lo = np.array([2.0, -3.0, 0.5], np.float32)
    hi = np.array([2.5, -2.0, 0.75], np.float32)
    act, pos = oracle.cell_vertex(c, lo, hi)
    assert act
    # crossings at x = 0.5 on the 4 x-edges: centroid (0.5, 0.5, 0.5) of the unit cell
    np.testing.assert_allclose(pos, lo + (hi - lo) * 0.5, atol=1e-6)
    del rng


def test_linear_fields_vertex_on_plane():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        lo = rng.uniform(-1, 1, 3).astype(np.float32)
        h = np.float32(rng.uniform(0.01, 0.2))
        hi = (lo + h).astype(np.float32)
        p0 = lo + rng.uniform(0, 1, 3).astype(np.float32) * h
        pts = np.array([[lo[0] if not x else hi[0], lo[1] if not y else hi[1],
                         lo[2] if not z else hi[2]] for (x, y, z) in CORNERS], np.float64)
        c = ((pts - p0) @ n).astype(np.float32)
        act, pos = oracle.cell_vertex(c, lo, hi)
        if not act:
            continue
        assert abs(float((pos.astype(np.float64) - p0) @ n)) < 1e-5 * max(1.0, float(h) * 10)
        assert np.all(pos >= lo) and np.all(pos <= hi)


def test_cell_bounds_match_reference_formula():
    # dualcontour.wgsl:22-27 in fp32: size = (max-min)/f32(res-1); cmin = min + size*i;
    # cmax = cmin + size; the shared-lattice max is min + size*(i+1)
    for b, res in [(2.0, 128), (2.0, 512), (2.5, 1024), (5.0, 7)]:
        mn, mx = np.float32(-b / 2), np.float32(b / 2)
        size = (mx - mn) / np.float32(res - 1)
        for i in [0, 1, res // 3, res - 2, res - 1]:
            lo, hi = oracle.cell_bounds(mn, mx, res, i)
            assert lo == mn + size * np.float32(i)
            assert hi == lo + size
            llo, lhi = oracle.cell_bounds(mn, mx, res, i, lattice=True)
            assert llo == lo and lhi == mn + size * np.float32(i + 1)
        # the last cell's max corner is one h past bb_max (the grid has res cells of
        # size (max-min)/(res-1))
        _, hi = oracle.cell_bounds(mn, mx, res, res - 1)
        assert abs(float(hi) - (b / 2 + b / (res - 1))) < 1e-5
