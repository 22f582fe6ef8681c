"""Oracle: CPU restatement of the reference meshing path (refcpu.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker / CPU baseline; the product (sdf2mesh_amd) never
imports it.  Build: ``make -C oracle`` (or oracle.build()).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "librefcpu.so"

# s2m_op layout (include/sdf2mesh.h)
TAPE_DTYPE = np.dtype([("code", "<u4"), ("iarg", "<u4"), ("f", "<f4", (10,))])

# Oracle-only primitive codes for whole example programs (examples.inc): a one-op tape
# evaluates examples/martin_cube.sdf3d, examples/p_key.sdf3d (the reference's examples)
# or tests/golden/engraved_cube.sdf3d (this repository's compiled-path test program).
ORA_OP_MARTIN_CUBE = 100
ORA_OP_P_KEY = 101
ORA_OP_ENGRAVED_CUBE = 102


def example_tape(code: int) -> np.ndarray:
    t = np.zeros(1, TAPE_DTYPE)
    t[0]["code"] = code
    return t


_h: Optional[C.CDLL] = None


class _OraMesh(C.Structure):
    _fields_ = [("nvtx", C.c_uint64), ("vtx", C.POINTER(C.c_float)),
                ("keys", C.POINTER(C.c_uint64)), ("signs", C.POINTER(C.c_uint8)),
                ("ntri", C.c_uint64), ("tri", C.POINTER(C.c_uint32)),
                ("invalid_quads", C.c_uint64)]


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _h
    if _h is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        vp, u32, u64, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
        sig = {
            "ora_tape_check": (C.c_int, [vp, u32]),
            "ora_eval_f32": (None, [vp, u32, vp, u64, vp]),
            "ora_eval_f64": (None, [vp, u32, vp, u64, vp]),
            "ora_mandel_cond": (None, [vp, u32, vp, u64, vp]),
            "ora_cell_bounds": (None, [f32, f32, u32, u32, C.c_int, C.POINTER(f32), C.POINTER(f32)]),
            "ora_cell_signs": (u32, [vp]),
            "ora_cell_vertex": (C.c_int, [vp, vp, vp, vp]),
            "ora_normal": (None, [vp, u32, vp, f32, vp]),
            "ora_mesh": (C.POINTER(_OraMesh), [vp, u32, vp, vp, vp, f32, C.c_int, C.c_int, C.c_int]),
            "ora_contour_corners": (C.POINTER(_OraMesh), [vp, vp, vp, vp, C.c_int]),
            "ora_cell_corners": (None, [vp, u32, vp, vp, vp, u32, u32, C.c_int, vp]),
            "ora_bench_slices": (u64, [vp, u32, vp, vp, vp, f32, u32, u32, C.POINTER(u64)]),
            "ora_mesh_free": (None, [C.POINTER(_OraMesh)]),
            "ora_format_f32": (C.c_int, [f32, C.c_char_p]),
            "ora_write_stl": (C.c_int, [vp, u64, vp, u64, C.c_char_p]),
            "ora_write_ply": (C.c_int, [vp, u64, vp, u64, C.c_char_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype, fn.argtypes = res, args
        _h = h
    return _h


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _tape(t) -> np.ndarray:
    t = t.tape() if hasattr(t, "tape") else t
    return np.ascontiguousarray(t, TAPE_DTYPE)


@dataclass
class Mesh:
    vertices: np.ndarray   # (V, 6) float32
    keys: np.ndarray       # (V,) uint64, real cell coordinates
    signs: np.ndarray      # (V,) uint8
    triangles: np.ndarray  # (T, 3) uint32
    invalid_quads: int


def _take(ptr) -> Mesh:
    if not ptr:
        raise ValueError("oracle rejected the input (invalid tape?)")
    m = ptr.contents
    nv, nt = int(m.nvtx), int(m.ntri)
    try:
        v = np.ctypeslib.as_array(m.vtx, (max(nv, 1) * 6,))[: 6 * nv].reshape(nv, 6).copy()
        k = np.ctypeslib.as_array(m.keys, (max(nv, 1),))[:nv].copy()
        s = np.ctypeslib.as_array(m.signs, (max(nv, 1),))[:nv].copy()
        t = np.ctypeslib.as_array(m.tri, (max(nt, 1) * 3,))[: 3 * nt].reshape(nt, 3).copy()
    finally:
        lib().ora_mesh_free(ptr)
    return Mesh(v, k, s, t, int(m.invalid_quads))


def eval_f32(t, pts: np.ndarray) -> np.ndarray:
    t = _tape(t)
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.empty(len(pts), np.float32)
    lib().ora_eval_f32(_p(t), len(t), _p(pts), len(pts), _p(out))
    return out


def eval_f64(t, pts: np.ndarray) -> np.ndarray:
    t = _tape(t)
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.empty(len(pts), np.float64)
    lib().ora_eval_f64(_p(t), len(t), _p(pts), len(pts), _p(out))
    return out


def mandel_cond(t, pts: np.ndarray) -> np.ndarray:
    """(N, 3) conditioning of the tape's Mandelbulb at each point (fp64 trajectory, see
    ora_mandel_cond): [:, 0] the smallest relative distance |r - bailout| / bailout of any
    bailout test, [:, 1] the smallest rho / r of any escape step, [:, 2] the first-order
    error scale (summed effect of one 2^-24 relative rounding of each intermediate)."""
    t = _tape(t)
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.empty((len(pts), 3), np.float64)
    lib().ora_mandel_cond(_p(t), len(t), _p(pts), len(pts), _p(out))
    return out


def tape_check(t) -> bool:
    t = _tape(t)
    return bool(lib().ora_tape_check(_p(t), len(t)))


def cell_bounds(bmin: float, bmax: float, res: int, i: int, lattice: bool = False):
    lo, hi = C.c_float(), C.c_float()
    lib().ora_cell_bounds(float(np.float32(bmin)), float(np.float32(bmax)), res, i, int(lattice),
                          C.byref(lo), C.byref(hi))
    return np.float32(lo.value), np.float32(hi.value)


def cell_vertex(c8, cmin, cmax):
    """(active, pos) of one cell from its 8 corner values."""
    c = np.ascontiguousarray(c8, np.float32)
    lo = np.ascontiguousarray(cmin, np.float32)
    hi = np.ascontiguousarray(cmax, np.float32)
    pos = np.zeros(3, np.float32)
    act = lib().ora_cell_vertex(_p(c), _p(lo), _p(hi), _p(pos))
    return bool(act), pos


def cell_signs(c8) -> int:
    c = np.ascontiguousarray(c8, np.float32)
    return int(lib().ora_cell_signs(_p(c)))


def normal(t, p, eps: float = 1e-4) -> np.ndarray:
    t = _tape(t)
    p = np.ascontiguousarray(p, np.float32)
    n = np.zeros(3, np.float32)
    lib().ora_normal(_p(t), len(t), _p(p), float(np.float32(eps)), _p(n))
    return n


def _box(bmin, bmax, res):
    return (np.ascontiguousarray(bmin, np.float32), np.ascontiguousarray(bmax, np.float32),
            np.ascontiguousarray(res, np.uint32))


def mesh(t, bmin, bmax, res, eps: float = 1e-4, lattice: bool = False, lag: bool = True,
         normals: bool = True) -> Mesh:
    """The reference driver: 8 corner evaluations per cell, slice by slice, host scan,
    binary-search quad assembly.  lag models the one-slice readback lag."""
    t = _tape(t)
    bmin, bmax, res = _box(bmin, bmax, res)
    return _take(lib().ora_mesh(_p(t), len(t), _p(bmin), _p(bmax), _p(res),
                                float(np.float32(eps)), int(lattice), int(lag), int(normals)))


def contour_corners(corners: np.ndarray, cell_lo, cell_hi, lag: bool = False) -> Mesh:
    corners = np.ascontiguousarray(corners, np.float32)
    rz, ry, rx = corners.shape[:3]
    res = np.array([rx, ry, rz], np.uint32)
    lo = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32) for a in cell_lo]))
    hi = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32) for a in cell_hi]))
    return _take(lib().ora_contour_corners(_p(corners), _p(res), _p(lo), _p(hi), int(lag)))


def cell_corners(t, bmin, bmax, res, z0: int = 0, z1: Optional[int] = None,
                 lattice: bool = False) -> np.ndarray:
    """Per-cell corner dump (rz, ry, rx, 8) at the reference's corner coordinates."""
    t = _tape(t)
    bmin, bmax, res = _box(bmin, bmax, res)
    z1 = int(res[2]) if z1 is None else z1
    out = np.empty((z1 - z0, int(res[1]), int(res[0]), 8), np.float32)
    lib().ora_cell_corners(_p(t), len(t), _p(bmin), _p(bmax), _p(res), z0, z1, int(lattice),
                           _p(out))
    return out


def axis_bounds(bmin: float, bmax: float, res: int, lattice: bool = False):
    """Per-axis arrays of cell min / max coordinates (cell_bounds per index)."""
    lo = np.empty(res, np.float32)
    hi = np.empty(res, np.float32)
    for i in range(res):
        lo[i], hi[i] = cell_bounds(bmin, bmax, res, i, lattice)
    return lo, hi


def bench_slices(t, bmin, bmax, res, z0: int, z1: int, eps: float = 1e-4):
    """Run the reference algorithm on slices [z0, z1); returns (cells, active)."""
    t = _tape(t)
    bmin, bmax, res = _box(bmin, bmax, res)
    act = C.c_uint64()
    n = lib().ora_bench_slices(_p(t), len(t), _p(bmin), _p(bmax), _p(res),
                               float(np.float32(eps)), z0, z1, C.byref(act))
    return int(n), int(act.value)


def format_f32(x: float) -> str:
    buf = C.create_string_buffer(64)
    n = lib().ora_format_f32(float(np.float32(x)), buf)
    return buf.raw[:n].decode()


def write_stl(vertices: np.ndarray, triangles: np.ndarray, path) -> None:
    v = np.ascontiguousarray(vertices, np.float32)
    t = np.ascontiguousarray(triangles, np.uint32)
    if lib().ora_write_stl(_p(v), len(v), _p(t), len(t), str(path).encode()) != 0:
        raise OSError(f"cannot write {path}")


def write_ply(vertices: np.ndarray, triangles: np.ndarray, path) -> None:
    v = np.ascontiguousarray(vertices, np.float32)
    t = np.ascontiguousarray(triangles, np.uint32)
    if lib().ora_write_ply(_p(v), len(v), _p(t), len(t), str(path).encode()) != 0:
        raise OSError(f"cannot write {path}")
