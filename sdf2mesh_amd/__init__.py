"""sdf2mesh_amd - MI355X-native meshing path of sdf2mesh.

Python mirror of the reference's host interface for this path, over the C ABI of
libsdf2mesh.so (include/sdf2mesh.h):

    ctx  = Context(0)                                 # wgpu device set-up (main.rs:180-196)
    prog = ctx.program(tape.torus_example())          # Sdf3DShader + shader module
    mesh = ctx.mesh(prog, Params.reference(128, 2.0)) # slice loop + TriangleMesh::from
    mesh.write_to_file("torus.stl")                   # TriangleMesh::write_to_file

The library must be built (``python -m sdf2mesh_amd.build``); nothing here falls back
to a CPU path: every meshing call runs the HIP kernels or raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

from . import tape
from .tape import TAPE_DTYPE

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsdf2mesh.so"

# ---- status codes (s2m_status) ---------------------------------------------------------
OK, E_REQUEST, E_SHADER, E_PARSE, E_WGSL, E_VALIDATION, E_MISSING_SDF = range(7)
E_HIP, E_OOM, E_BUDGET, E_ARG, E_IO, E_BUSY = range(7, 13)
STATUS_NAMES = {
    0: "Ok", 1: "RequestError", 2: "ShaderError", 3: "ParseErrors", 4: "WgslError",
    5: "ValidationError", 6: "MissingSdf", 7: "HipError", 8: "OutOfMemory",
    9: "BudgetExceeded", 10: "InvalidArgument", 11: "IoError", 12: "Busy",
}

# ---- flags -------------------------------------------------------------------------------
COMPAT_POW2 = 0x1
COMPAT_DROP_LAST_SLICE = 0x2
CORNERS_EXACT = 0x4
NO_NORMALS = 0x8
DENSE_EVAL = 0x10
REFERENCE_COMPAT = COMPAT_POW2 | COMPAT_DROP_LAST_SLICE | CORNERS_EXACT

# source programs (s2m_program_from_source / s2m_translate)
LANG_SDF3D = 0
LANG_GLSL = 1
SOURCE_IS_PATH = 0x100
FORCE_COMPILED = 0x200
PROGRAM_TAPE = 0
PROGRAM_COMPILED = 1
STAGE_SOURCE = 0
STAGE_TAPE = 1
STAGE_HIP = 2
STAGE_CODE = 3


class S2MError(RuntimeError):
    """A failed library call; .code is the s2m_status value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


# ---- C structs -----------------------------------------------------------------------------
class _Params(C.Structure):
    _fields_ = [("bb_min", C.c_float * 3), ("bb_max", C.c_float * 3), ("res", C.c_uint32 * 3),
                ("eps", C.c_float), ("flags", C.c_uint32), ("slab", C.c_uint32)]


class _MeshOut(C.Structure):
    _fields_ = [("vtx", C.POINTER(C.c_float)), ("nvtx", C.c_uint64),
                ("tri", C.POINTER(C.c_uint32)), ("ntri", C.c_uint64),
                ("invalid_quads", C.c_uint64), ("cell_keys", C.POINTER(C.c_uint64)),
                ("signs", C.POINTER(C.c_uint8))]


class _DeviceMesh(C.Structure):
    _fields_ = [("vtx", C.c_void_p), ("tri", C.c_void_p), ("cell_keys", C.c_void_p),
                ("signs", C.c_void_p), ("nvtx", C.c_uint64), ("ntri", C.c_uint64),
                ("invalid_quads", C.c_uint64)]


class _MeshBuffers(C.Structure):
    _fields_ = [("vtx", C.c_void_p), ("cell_keys", C.c_void_p), ("signs", C.c_void_p),
                ("vcap", C.c_uint64), ("tri", C.c_void_p), ("tcap", C.c_uint64)]


class _Tuning(C.Structure):
    _fields_ = [("stream_passes", C.c_uint32), ("dense_slab", C.c_uint32), ("batch_cells", C.c_uint64),
                ("ms", C.c_double), ("measured", C.c_uint32)]


class _ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_uint64), ("max_batch", C.c_uint32),
                ("workers", C.c_uint32), ("want_keys", C.c_int), ("dense_workers", C.c_uint32)]


class _RequestStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("latency_ms", C.c_double), ("batch_size", C.c_uint32),
                ("worker", C.c_uint32)]


class _Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("eval_ms", C.c_double), ("contour_ms", C.c_double),
                ("normal_ms", C.c_double), ("quad_ms", C.c_double), ("d2h_ms", C.c_double),
                ("eval_launches", C.c_uint64), ("lattice_points", C.c_uint64),
                ("slabs", C.c_uint64), ("tiles_total", C.c_uint64), ("tiles_active", C.c_uint64),
                ("eval_flops", C.c_uint64), ("prune_ms", C.c_double), ("graph_launches", C.c_uint64)]


_lib_handle: Optional[C.CDLL] = None

# name -> (restype, argtypes); every entry point of include/sdf2mesh.h
_SIGNATURES = {
    "s2m_create": (C.c_void_p, [C.c_int]),
    "s2m_destroy": (None, [C.c_void_p]),
    "s2m_last_error": (C.c_char_p, []),
    "s2m_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "s2m_last_timing": (C.c_int, [C.c_void_p, C.POINTER(_Timing)]),
    "s2m_program_from_tape": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32,
                                        C.POINTER(C.c_void_p)]),
    "s2m_program_free": (None, [C.c_void_p]),
    "s2m_program_from_source": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_char_p,
                                          C.POINTER(C.c_void_p)]),
    "s2m_program_kind": (C.c_int, [C.c_void_p]),
    "s2m_program_tape": (C.c_uint32, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "s2m_translate": (C.c_uint64, [C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_void_p,
                                   C.c_uint64, C.POINTER(C.c_int)]),
    "s2m_tape_validate": (C.c_int, [C.c_void_p, C.c_uint32]),
    "s2m_tape_flops": (C.c_uint64, [C.c_void_p, C.c_uint32]),
    "s2m_params_reference": (None, [C.POINTER(_Params), C.c_uint32, C.c_float]),
    "s2m_round_resolution": (C.c_uint32, [C.c_uint32]),
    "s2m_mesh": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params), C.c_int,
                           C.POINTER(_MeshOut)]),
    "s2m_mesh_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params),
                                  C.POINTER(_DeviceMesh)]),
    "s2m_mesh_free": (None, [C.POINTER(_MeshOut)]),
    "s2m_mesh_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params),
                                          C.POINTER(_MeshBuffers), C.POINTER(_DeviceMesh)]),
    "s2m_mesh_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(_Params),
                                 C.c_int, C.POINTER(_MeshOut)]),
    "s2m_mesh_batch_device": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p),
                                        C.POINTER(_Params), C.POINTER(_DeviceMesh)]),
    "s2m_tune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params), C.c_uint32, C.POINTER(_Tuning)]),
    "s2m_tune_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(_Params),
                                 C.c_uint32, C.POINTER(_Tuning)]),
    "s2m_tuning_save": (C.c_int, [C.c_char_p]),
    "s2m_tuning_load": (C.c_int, [C.c_char_p]),
    "s2m_tuning_clear": (None, []),
    "s2m_tune_pick": (C.c_uint32, [C.POINTER(C.c_double), C.c_uint32]),
    "s2m_server_create": (C.c_void_p, [C.POINTER(_ServerConfig)]),
    "s2m_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(_Params),
                                    C.POINTER(C.c_uint64)]),
    "s2m_server_submit_program": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params),
                                            C.POINTER(C.c_uint64)]),
    "s2m_server_take": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(_MeshOut),
                                  C.POINTER(_RequestStats)]),
    "s2m_server_request_bytes": (C.c_uint64, [C.c_void_p, C.c_uint32, C.POINTER(_Params)]),
    "s2m_server_destroy": (None, [C.c_void_p]),
    "s2m_write_stl": (C.c_int, [C.POINTER(_MeshOut), C.c_char_p]),
    "s2m_write_ply": (C.c_int, [C.POINTER(_MeshOut), C.c_char_p]),
    "s2m_write_to_file": (C.c_int, [C.POINTER(_MeshOut), C.c_char_p]),
    "s2m_format_f32": (C.c_int, [C.c_float, C.c_char_p]),
    "s2m_check_sincos": (C.c_int, [C.c_void_p, C.c_void_p]),
    "s2m_check_bailout": (C.c_int, [C.c_void_p, C.c_void_p]),
    "s2m_host_pool_stats": (None, [C.c_void_p]),
    "s2m_host_pool_trim": (None, []),
    "s2m_axis_table": (C.c_uint32, [C.c_float, C.c_float, C.c_uint32, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "s2m_eval_lattice": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params), C.c_void_p,
                                   C.c_uint64, C.POINTER(C.c_uint32)]),
    "s2m_eval_lattice_range": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_Params), C.c_uint32,
                                         C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32)]),
    "s2m_contour_cells": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32),
                                    C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(_MeshOut)]),
}


def lib() -> C.CDLL:
    """Load libsdf2mesh.so (built in-tree); raises if it is missing."""
    global _lib_handle
    if _lib_handle is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m sdf2mesh_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
    return _lib_handle


def _check(rc: int) -> None:
    if rc != OK:
        raise S2MError(rc, lib().s2m_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


# ---- params ------------------------------------------------------------------------------
@dataclass
class Params:
    """AppState of main.rs:28-33 as plain fields (see s2m_params)."""
    bb_min: tuple = (-1.0, -1.0, -1.0)
    bb_max: tuple = (1.0, 1.0, 1.0)
    res: tuple = (128, 128, 128)
    eps: float = 1e-4
    flags: int = REFERENCE_COMPAT
    slab: int = 0

    @classmethod
    def reference(cls, resolution: int = 256, bounds: float = 2.0, flags: int = REFERENCE_COMPAT,
                  slab: int = 0) -> "Params":
        """From<&Arguments> (main.rs:139-175): cube of edge `bounds` at the origin."""
        p = _Params()
        lib().s2m_params_reference(C.byref(p), resolution, bounds)
        return cls(tuple(p.bb_min), tuple(p.bb_max), tuple(p.res), p.eps, flags, slab)

    def c(self) -> _Params:
        """The s2m_params struct (built again only when a field has changed)."""
        key = (tuple(self.bb_min), tuple(self.bb_max), tuple(self.res), self.eps, self.flags, self.slab)
        cached = self.__dict__.get("_c")
        if cached is not None and cached[0] == key:
            return cached[1]
        p = _Params()
        p.bb_min[:] = [float(np.float32(v)) for v in self.bb_min]
        p.bb_max[:] = [float(np.float32(v)) for v in self.bb_max]
        p.res[:] = [int(v) for v in self.res]
        p.eps = self.eps
        p.flags = self.flags
        p.slab = self.slab
        self.__dict__["_c"] = (key, p)
        return p


def timing_dict(t: "_Timing") -> dict:
    return {k: getattr(t, k) for k, _ in _Timing._fields_}


def round_resolution(res: int) -> int:
    """main.rs:141-148: `2 << ilog2(res)` unless res is a power of two."""
    return int(lib().s2m_round_resolution(res))


# ---- results -----------------------------------------------------------------------------
@dataclass
class TriangleMesh:
    """TriangleMesh (mesh.rs:143-211) plus the VertexList metadata."""
    vertices: np.ndarray        # (V, 6) float32: pos.xyz, normal.xyz
    triangles: np.ndarray       # (T, 3) uint32
    invalid_quads: int
    cell_keys: Optional[np.ndarray] = None  # (V,) uint64 x | y<<16 | z<<32
    signs: Optional[np.ndarray] = None      # (V,) uint8

    def _c(self) -> _MeshOut:
        m = _MeshOut()
        self._v = np.ascontiguousarray(self.vertices, np.float32)
        self._t = np.ascontiguousarray(self.triangles, np.uint32)
        m.vtx = self._v.ctypes.data_as(C.POINTER(C.c_float))
        m.nvtx = len(self._v)
        m.tri = self._t.ctypes.data_as(C.POINTER(C.c_uint32))
        m.ntri = len(self._t)
        m.invalid_quads = self.invalid_quads
        return m

    def write_stl_to_file(self, path) -> None:
        _check(lib().s2m_write_stl(C.byref(self._c()), str(path).encode()))

    def write_ply_to_file(self, path) -> None:
        _check(lib().s2m_write_ply(C.byref(self._c()), str(path).encode()))

    def write_to_file(self, path) -> None:
        _check(lib().s2m_write_to_file(C.byref(self._c()), str(path).encode()))


def _take_mesh(out: _MeshOut, want_keys: bool) -> TriangleMesh:
    nv, nt = int(out.nvtx), int(out.ntri)
    try:
        v = np.ctypeslib.as_array(out.vtx, (max(nv, 1) * 6,))[: nv * 6].reshape(nv, 6).copy()
        t = np.ctypeslib.as_array(out.tri, (max(nt, 1) * 3,))[: nt * 3].reshape(nt, 3).copy()
        keys = signs = None
        if want_keys:
            keys = np.ctypeslib.as_array(out.cell_keys, (max(nv, 1),))[:nv].copy()
            signs = np.ctypeslib.as_array(out.signs, (max(nv, 1),))[:nv].copy()
        return TriangleMesh(v, t, int(out.invalid_quads), keys, signs)
    finally:
        lib().s2m_mesh_free(C.byref(out))


# s2m_mesh_out as a numpy record (pointers as integers): the counts of a whole batch's
# results read at once
_MESH_OUT_DTYPE = np.dtype([("vtx", "<u8"), ("nvtx", "<u8"), ("tri", "<u8"), ("ntri", "<u8"),
                            ("invalid_quads", "<u8"), ("cell_keys", "<u8"), ("signs", "<u8")])
assert _MESH_OUT_DTYPE.itemsize == C.sizeof(_MeshOut)


class HostMesh:
    """A mesh in library-owned host memory (s2m_mesh_out); views, no copies."""

    __slots__ = ("_out", "nvtx", "ntri", "invalid_quads")

    def __init__(self, out: _MeshOut, nvtx: int | None = None, ntri: int | None = None,
                 invalid_quads: int | None = None):
        self._out = out
        self.nvtx = int(out.nvtx) if nvtx is None else nvtx
        self.ntri = int(out.ntri) if ntri is None else ntri
        self.invalid_quads = int(out.invalid_quads) if invalid_quads is None else invalid_quads

    def vertices(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._out.vtx, (max(self.nvtx, 1) * 6,))[: self.nvtx * 6] \
            .reshape(self.nvtx, 6)

    def triangles(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._out.tri, (max(self.ntri, 1) * 3,))[: self.ntri * 3] \
            .reshape(self.ntri, 3)

    def free(self) -> None:
        if self._out is not None:
            lib().s2m_mesh_free(C.byref(self._out))
            self._out = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def tuning_save(path) -> None:
    """Write the autotuner's table as JSON (s2m_tuning_save)."""
    _check(lib().s2m_tuning_save(str(path).encode()))


def tuning_load(path) -> None:
    """Merge a JSON file's entries into the autotuner's table (s2m_tuning_load)."""
    _check(lib().s2m_tuning_load(str(path).encode()))


def tuning_clear() -> None:
    lib().s2m_tuning_clear()


def host_pool_stats() -> dict:
    """s2m_host_pool_stats: page-locked blocks allocated so far, bytes held, idle blocks
    and idle bytes of the result memory pool."""
    st = (C.c_uint64 * 4)()
    lib().s2m_host_pool_stats(st)
    return {"pinned_allocs": int(st[0]), "pinned_bytes": int(st[1]), "idle_blocks": int(st[2]),
            "idle_bytes": int(st[3])}


def host_pool_trim() -> None:
    """s2m_host_pool_trim: return the pool's idle page-locked blocks to the system."""
    lib().s2m_host_pool_trim()


def tune_pick(ms) -> int:
    """The autotuner's choice rule: index of the chosen timing (s2m_tune_pick)."""
    a = (C.c_double * len(ms))(*[float(x) for x in ms])
    return int(lib().s2m_tune_pick(a, len(ms)))


def _tuning_dict(t: "_Tuning") -> dict:
    return {k: getattr(t, k) for k, _ in _Tuning._fields_}


def format_f32(x: float) -> str:
    """Rust `{}` Display of an f32, as the writers print it."""
    buf = C.create_string_buffer(64)
    n = lib().s2m_format_f32(float(np.float32(x)), buf)
    return buf.raw[:n].decode()


def axis_table(bmin: float, bmax: float, res: int, exact: bool):
    """Per-axis corner coordinates and each cell's (lo, hi) coordinate index."""
    coord = np.zeros(2 * res + 2, np.float32)
    lo = np.zeros(max(res, 1), np.uint32)
    hi = np.zeros(max(res, 1), np.uint32)
    n = lib().s2m_axis_table(float(np.float32(bmin)), float(np.float32(bmax)), res, int(exact),
                             _ptr(coord), _ptr(lo), _ptr(hi))
    return coord[:n].copy(), lo[:res].copy(), hi[:res].copy()


def tape_flops(t) -> int:
    """Algorithmic fp32 operations of one evaluation of the program (s2m_tape_flops)."""
    t = t.tape() if isinstance(t, tape.Node) else np.ascontiguousarray(t, TAPE_DTYPE)
    return int(lib().s2m_tape_flops(_ptr(t), len(t)))


def _lang_flags(lang: str, path: bool, force_compiled: bool) -> int:
    if lang not in ("sdf3d", "glsl"):
        raise ValueError(f"unknown source language {lang!r}")
    return ((LANG_GLSL if lang == "glsl" else LANG_SDF3D) | (SOURCE_IS_PATH if path else 0)
            | (FORCE_COMPILED if force_compiled else 0))


def translate(source, stage: int, lang: str = "sdf3d", sdf_name: Optional[str] = None,
              path: bool = False) -> bytes:
    """Run the front end without a device: STAGE_SOURCE (text after use/include
    expansion), STAGE_TAPE (tape bytes), STAGE_HIP (generated HIP C++) or STAGE_CODE
    (gfx950 code object).  `source` is the text, or a file path with path=True."""
    src = str(source).encode()
    flags = _lang_flags(lang, path, False)
    name = sdf_name.encode() if sdf_name else None
    st = C.c_int(0)
    n = lib().s2m_translate(src, flags, name, stage, None, 0, C.byref(st))
    if n == 0:
        if st.value != OK:
            _check(st.value)
        return b""
    buf = C.create_string_buffer(n)
    lib().s2m_translate(src, flags, name, stage, buf, n, C.byref(st))
    _check(st.value)
    return buf.raw


def translate_tape(source, lang: str = "sdf3d", sdf_name: Optional[str] = None,
                   path: bool = False) -> np.ndarray:
    """The tape a source program lowers to (S2MError E_WGSL if it does not)."""
    return np.frombuffer(translate(source, STAGE_TAPE, lang, sdf_name, path), TAPE_DTYPE).copy()


def tape_validate(t: np.ndarray) -> bool:
    t = np.ascontiguousarray(t, TAPE_DTYPE)
    return lib().s2m_tape_validate(_ptr(t), len(t)) == OK


class Program:
    """A program uploaded to a context's device."""

    def __init__(self, ctx: "Context", handle: int, tape_array: Optional[np.ndarray], flops: int):
        self._ctx = ctx
        self.handle = handle
        self.tape = tape_array  # None for a compiled program
        self.flops_per_point = flops

    @property
    def kind(self) -> int:
        """PROGRAM_TAPE or PROGRAM_COMPILED."""
        return lib().s2m_program_kind(C.c_void_p(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().s2m_program_free(C.c_void_p(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device, one HIP stream and its reusable device buffers."""

    def __init__(self, device: int = 0):
        h = lib().s2m_create(device)
        if not h:
            raise S2MError(E_HIP, lib().s2m_last_error().decode(errors="replace"))
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            lib().s2m_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, level: int = 1) -> None:
        """0 off; 1 (or True): whole call and evaluation launches; 2: every stage."""
        _check(lib().s2m_set_profiling(C.c_void_p(self.handle), int(level)))

    def last_timing(self) -> dict:
        return timing_dict(self.last_timing_raw())

    def last_timing_raw(self) -> "_Timing":
        """s2m_last_timing as the C struct (timing_dict converts it)."""
        t = _Timing()
        _check(lib().s2m_last_timing(C.c_void_p(self.handle), C.byref(t)))
        return t

    def program(self, src) -> Program:
        """Upload a tape (numpy TAPE_DTYPE array) or a tape.Node expression."""
        t = src.tape() if isinstance(src, tape.Node) else np.ascontiguousarray(src, TAPE_DTYPE)
        flops = tape_flops(t)
        out = C.c_void_p()
        _check(lib().s2m_program_from_tape(C.c_void_p(self.handle), _ptr(t), len(t), C.byref(out)))
        return Program(self, out.value, t, flops)

    def program_from_source(self, source, lang: str = "sdf3d", sdf_name: Optional[str] = None,
                            path: bool = False, force_compiled: bool = False) -> Program:
        """A program from .sdf3d WGSL or GLSL text (or a file with path=True): a tape when
        it lowers to library primitives and operators, else code compiled by hiprtc."""
        out = C.c_void_p()
        _check(lib().s2m_program_from_source(
            C.c_void_p(self.handle), str(source).encode(), _lang_flags(lang, path, force_compiled),
            sdf_name.encode() if sdf_name else None, C.byref(out)))
        h = out.value
        if lib().s2m_program_kind(C.c_void_p(h)) == PROGRAM_TAPE:
            n = lib().s2m_program_tape(C.c_void_p(h), None, 0)
            t = np.zeros(n, TAPE_DTYPE)
            lib().s2m_program_tape(C.c_void_p(h), _ptr(t), n)
            return Program(self, h, t, tape_flops(t))
        return Program(self, h, None, 0)

    def mesh(self, prog: Program, params: Params, want_keys: bool = True) -> TriangleMesh:
        """Mesh the program; results copied to host memory."""
        p = params.c()
        out = _MeshOut()
        _check(lib().s2m_mesh(C.c_void_p(self.handle), C.c_void_p(prog.handle), C.byref(p),
                              int(want_keys), C.byref(out)))
        return _take_mesh(out, want_keys)

    def mesh_to_host(self, prog: Program, params: Params, want_keys: bool = False) -> "HostMesh":
        """s2m_mesh without the copy into numpy arrays: the result stays in the library's
        host memory (page-locked, recycled) until HostMesh.free()."""
        p = params.c()
        out = _MeshOut()
        rc = lib().s2m_mesh(C.c_void_p(self.handle), C.c_void_p(prog.handle), C.byref(p),
                            int(want_keys), C.byref(out))
        if rc != OK:
            lib().s2m_mesh_free(C.byref(out))
            _check(rc)
        return HostMesh(out)

    def mesh_batch(self, progs, params, want_keys: bool = True) -> list:
        """Mesh many jobs in shared device passes; one TriangleMesh per job."""
        n = len(progs)
        hp = (C.c_void_p * n)(*[p.handle for p in progs])
        pa = (_Params * n)(*[q.c() for q in params])
        outs = (_MeshOut * n)()
        rc = lib().s2m_mesh_batch(C.c_void_p(self.handle), n, hp, pa, int(want_keys), outs)
        if rc != OK:
            for o in outs:
                lib().s2m_mesh_free(C.byref(o))
            _check(rc)
        return [_take_mesh(outs[i], want_keys) for i in range(n)]

    def mesh_batch_to_host(self, progs, params, want_keys: bool = False) -> list:
        """s2m_mesh_batch without copies into numpy arrays: one HostMesh per job, in the
        library's host memory until each is freed."""
        n = len(progs)
        hp = (C.c_void_p * n)(*[p.handle for p in progs])
        pa = (_Params * n)(*[q.c() for q in params])
        outs = (_MeshOut * n)()
        rc = lib().s2m_mesh_batch(C.c_void_p(self.handle), n, hp, pa, int(want_keys), outs)
        if rc != OK:
            for o in outs:
                lib().s2m_mesh_free(C.byref(o))
            _check(rc)
        rec = np.frombuffer(outs, _MESH_OUT_DTYPE) if n else None
        nv, nt, iq = ((rec["nvtx"].tolist(), rec["ntri"].tolist(), rec["invalid_quads"].tolist())
                      if n else ([], [], []))
        return [HostMesh(outs[i], nv[i], nt[i], iq[i]) for i in range(n)]

    def mesh_batch_device(self, progs, params) -> list:
        """One shared pass; results stay in device memory owned by the context."""
        n = len(progs)
        hp = (C.c_void_p * n)(*[p.handle for p in progs])
        pa = (_Params * n)(*[q.c() for q in params])
        outs = (_DeviceMesh * n)()
        _check(lib().s2m_mesh_batch_device(C.c_void_p(self.handle), n, hp, pa, outs))
        return [{"nvtx": int(o.nvtx), "ntri": int(o.ntri), "invalid_quads": int(o.invalid_quads),
                 "vtx_ptr": o.vtx, "tri_ptr": o.tri, "keys_ptr": o.cell_keys, "signs_ptr": o.signs}
                for o in outs]

    def mesh_device(self, prog: Program, params: Params) -> dict:
        """Mesh the program; results stay in device memory owned by the context."""
        p = params.c()
        dm = _DeviceMesh()
        _check(lib().s2m_mesh_device(C.c_void_p(self.handle), C.c_void_p(prog.handle),
                                     C.byref(p), C.byref(dm)))
        return {"nvtx": int(dm.nvtx), "ntri": int(dm.ntri), "invalid_quads": int(dm.invalid_quads),
                "vtx_ptr": dm.vtx, "tri_ptr": dm.tri, "keys_ptr": dm.cell_keys, "signs_ptr": dm.signs}

    def mesh_device_noalloc(self, prog: Program, params: Params, vtx: int, cell_keys: int,
                            signs: int, vcap: int, tri: int, tcap: int) -> dict:
        """s2m_mesh_device_noalloc into caller-owned device buffers (addresses): the
        library allocates nothing.  Returns the counts; raises S2MError (E_OOM with
        .needed = (vertices, triangles) when the buffers are too small)."""
        p = params.c()
        bufs = _MeshBuffers(vtx, cell_keys, signs, vcap, tri, tcap)
        dm = _DeviceMesh()
        rc = lib().s2m_mesh_device_noalloc(C.c_void_p(self.handle), C.c_void_p(prog.handle),
                                           C.byref(p), C.byref(bufs), C.byref(dm))
        if rc != OK:
            err = S2MError(rc, lib().s2m_last_error().decode(errors="replace"))
            err.needed = (int(dm.nvtx), int(dm.ntri))
            raise err
        return {"nvtx": int(dm.nvtx), "ntri": int(dm.ntri), "invalid_quads": int(dm.invalid_quads)}

    def tune(self, prog: Program, params: Params, reps: int = 5) -> dict:
        """s2m_tune: time the candidate shapes of this job, keep the fastest for its key."""
        t = _Tuning()
        _check(lib().s2m_tune(C.c_void_p(self.handle), C.c_void_p(prog.handle), C.byref(params.c()),
                              reps, C.byref(t)))
        return _tuning_dict(t)

    def tune_batch(self, progs, params, reps: int = 5) -> dict:
        """s2m_tune_batch: the cells per pass of this batch."""
        n = len(progs)
        hp = (C.c_void_p * n)(*[p.handle for p in progs])
        pa = (_Params * n)(*[q.c() for q in params])
        t = _Tuning()
        _check(lib().s2m_tune_batch(C.c_void_p(self.handle), n, hp, pa, reps, C.byref(t)))
        return _tuning_dict(t)

    def check_sincos(self) -> tuple:
        """s2m_check_sincos: (sine mismatches, of them NaN pairs, cosine mismatches, of
        them NaN pairs) of sincosf against sinf / cosf over every float."""
        counts = (C.c_uint64 * 4)()
        _check(lib().s2m_check_sincos(C.c_void_p(self.handle), counts))
        return tuple(int(c) for c in counts)

    def check_bailout(self) -> tuple:
        """s2m_check_bailout: (comparisons, disagreements) of the Mandelbulb's squared
        bailout test r^2 > t against sqrtf(r^2) > bailout."""
        counts = (C.c_uint64 * 2)()
        _check(lib().s2m_check_bailout(C.c_void_p(self.handle), counts))
        return tuple(int(c) for c in counts)

    def eval_lattice(self, prog: Program, params: Params) -> np.ndarray:
        """SDF values on the corner lattice, shape (nz, ny, nx)."""
        p = params.c()
        dims = (C.c_uint32 * 3)()
        # first call with no buffer reports the extent
        rc = lib().s2m_eval_lattice(C.c_void_p(self.handle), C.c_void_p(prog.handle), C.byref(p),
                                    None, 0, dims)
        n = int(dims[0]) * int(dims[1]) * int(dims[2])
        if n == 0:
            return np.zeros((int(dims[2]), int(dims[1]), int(dims[0])), np.float32)
        out = np.empty(n, np.float32)
        _check(lib().s2m_eval_lattice(C.c_void_p(self.handle), C.c_void_p(prog.handle), C.byref(p),
                                      _ptr(out), n, dims))
        return out.reshape(int(dims[2]), int(dims[1]), int(dims[0]))

    def eval_lattice_range(self, prog: Program, params: Params, z0: int, z1: int):
        """(lattice, lz0): SDF values on the lattice planes holding the corners of cell
        planes [z0, z1), shape (planes, ny, nx); lz0 is the first plane's index in the
        whole corner lattice."""
        p = params.c()
        dims = (C.c_uint32 * 3)()
        lz0 = C.c_uint32()
        lib().s2m_eval_lattice_range(C.c_void_p(self.handle), C.c_void_p(prog.handle), C.byref(p),
                                     z0, z1, None, 0, dims, C.byref(lz0))
        n = int(dims[0]) * int(dims[1]) * int(dims[2])
        out = np.empty(max(n, 1), np.float32)
        _check(lib().s2m_eval_lattice_range(C.c_void_p(self.handle), C.c_void_p(prog.handle),
                                            C.byref(p), z0, z1, _ptr(out), n, dims, C.byref(lz0)))
        return out[:n].reshape(int(dims[2]), int(dims[1]), int(dims[0])), int(lz0.value)

    def contour_cells(self, corners: np.ndarray, cell_lo, cell_hi, flags: int = 0) -> TriangleMesh:
        """Contour per-cell corner values, corners shape (rz, ry, rx, 8); cell_lo/cell_hi
        are 3 per-axis arrays of cell min/max coordinates."""
        corners = np.ascontiguousarray(corners, np.float32)
        rz, ry, rx = corners.shape[:3]
        res = (C.c_uint32 * 3)(rx, ry, rz)
        lo = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32) for a in cell_lo]))
        hi = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32) for a in cell_hi]))
        out = _MeshOut()
        _check(lib().s2m_contour_cells(C.c_void_p(self.handle), _ptr(corners), res, _ptr(lo),
                                       _ptr(hi), flags, C.byref(out)))
        return _take_mesh(out, True)


class Server:
    """Continuous-batching request loop (s2m_server_*): submit tapes, take meshes."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_batch: int = 0, workers: int = 0,
                 want_keys: bool = True, dense_workers: int = 0):
        cfg = _ServerConfig(device, hbm_budget, max_batch, workers, int(want_keys), dense_workers)
        h = lib().s2m_server_create(C.byref(cfg))
        if not h:
            raise S2MError(E_HIP, lib().s2m_last_error().decode(errors="replace"))
        self.handle = h
        self.want_keys = want_keys

    def submit(self, program, params: Params) -> int:
        """Queue a tape (tape.Node or TAPE_DTYPE array) or a Program (a compiled one must
        stay open until its request has been taken)."""
        p = params.c()
        ticket = C.c_uint64()
        if isinstance(program, Program):
            _check(lib().s2m_server_submit_program(C.c_void_p(self.handle), C.c_void_p(program.handle),
                                                   C.byref(p), C.byref(ticket)))
            return int(ticket.value)
        t = program.tape() if isinstance(program, tape.Node) else np.ascontiguousarray(program, TAPE_DTYPE)
        _check(lib().s2m_server_submit(C.c_void_p(self.handle), _ptr(t), len(t), C.byref(p),
                                       C.byref(ticket)))
        return int(ticket.value)

    def take(self, ticket: int, wait: bool = True, host: bool = False):
        """(TriangleMesh, stats dict), or None if not finished and wait is False.  With
        host=True the mesh is a HostMesh: views of the library's host memory, no copies
        (free() hands the memory back for later results)."""
        out = _MeshOut()
        st = _RequestStats()
        rc = lib().s2m_server_take(C.c_void_p(self.handle), ticket, int(wait), C.byref(out),
                                   C.byref(st))
        if rc == E_BUSY:
            return None
        if rc != OK:
            lib().s2m_mesh_free(C.byref(out))
            _check(rc)
        stats = {k: getattr(st, k) for k, _ in _RequestStats._fields_}
        return (HostMesh(out) if host else _take_mesh(out, self.want_keys)), stats

    @staticmethod
    def request_bytes(program, params: Params) -> int:
        t = program.tape() if isinstance(program, tape.Node) else np.ascontiguousarray(program, TAPE_DTYPE)
        p = params.c()
        return int(lib().s2m_server_request_bytes(_ptr(t), len(t), C.byref(p)))

    def close(self) -> None:
        if self.handle:
            lib().s2m_server_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
