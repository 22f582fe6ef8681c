"""The C ABI library loads and exports every entry point include/sdf2mesh.h declares;
host-side logic of the boundary (params, tape validation, coordinate tables) needs no
GPU."""
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape as T

HEADER = Path(__file__).resolve().parent.parent / "include" / "sdf2mesh.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(s2m_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    lib = s2m.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers every one of them
    assert set(names) <= set(s2m._SIGNATURES)


def test_params_reference():
    p = s2m.Params.reference(100, 5.0)
    assert p.bb_min == (-2.5, -2.5, -2.5) and p.bb_max == (2.5, 2.5, 2.5)
    assert p.res == (100, 100, 100)
    assert np.float32(p.eps) == np.float32(1e-4)
    assert p.flags == s2m.REFERENCE_COMPAT


def test_tape_validation():
    good = T.csg64().tape()
    assert s2m.tape_validate(good)
    assert not s2m.tape_validate(good[:-1])
    assert not s2m.tape_validate(good[1:])
    bad = good.copy()
    bad[5]["code"] = 42
    assert not s2m.tape_validate(bad)
    assert not s2m.tape_validate(good[:0])
    deep = T.sphere((0, 0, 0), 1.0)
    for _ in range(20):  # right-leaning tree: stack depth 21 > 16
        deep = T.union(T.sphere((0, 0, 0), 0.5), deep)
    assert not s2m.tape_validate(deep.tape())
    for seed in range(1, 30):
        t = T.random_csg(seed).tape()
        assert s2m.tape_validate(t) == oracle.tape_check(t)


@pytest.mark.parametrize("b,res", [(2.0, 128), (2.0, 512), (2.5, 1024), (2.0, 7), (5.0, 96),
                                   (2.0, 2), (3.0, 1)])
@pytest.mark.parametrize("exact", [True, False])
def test_axis_tables_match_cell_bounds(b, res, exact):
    mn = np.float32(0) - np.float32(b) * np.float32(0.5)
    mx = np.float32(0) + np.float32(b) * np.float32(0.5)
    coord, lo, hi = s2m.axis_table(mn, mx, res, exact)
    olo, ohi = oracle.axis_bounds(mn, mx, res, lattice=not exact)
    assert np.array_equal(coord[lo], olo, equal_nan=True)
    assert np.array_equal(coord[hi], ohi, equal_nan=True)
    if exact and res > 1:
        # one evaluation per distinct coordinate: no two table entries are equal
        assert len(np.unique(coord)) == len(coord)
        assert np.all(np.diff(lo.astype(np.int64)) > 0) and np.all(hi > lo)


def test_exact_table_sizes():
    # distinct corner coordinates per axis at the workload configurations
    for b, res, n in [(2.0, 128, 156), (2.0, 512, 713), (2.5, 1024, 1100)]:
        mn, mx = np.float32(-b / 2), np.float32(b / 2)
        assert len(s2m.axis_table(mn, mx, res, True)[0]) == n


def test_context_without_gpu_fails_cleanly():
    try:
        ctx = s2m.Context(0)
    except s2m.S2MError as e:
        assert e.code == s2m.E_HIP
        return
    ctx.close()
    pytest.skip("a GPU is present")


def test_tape_flops_match_python_count():
    for node in [T.torus_example(), T.csg64(), T.mandelbulb(8), T.mandelbulb(5)] + \
            [T.random_csg(s) for s in range(1, 20)]:
        assert s2m.tape_flops(node) == node.flops()
    assert T.torus_example().flops() == 10
    assert T.csg64().flops() == 1774
    assert T.mandelbulb(8).flops() == 245
