"""The no-allocation entry point (s2m_mesh_device_noalloc) against s2m_mesh_device, the
entry point it varies: same mesh, bit for bit, written into caller-owned device buffers
with every working buffer taken from the library's static scratch array.  Two calls in a
row, the second on a smaller grid (the scratch keeps the first call's contents, so what
must start at zero has to be cleared), then the error paths: buffers too small (the sizes
needed are reported, a second call with them succeeds) and a grid whose cell planes do
not fit the scratch."""
import ctypes as C

import numpy as np
import pytest

import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """libamdhip64 for the caller's buffers (the test allocates; the library must not)."""
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceBuffers:
    def __init__(self, vcap: int, tcap: int):
        self.vcap, self.tcap = vcap, tcap
        self.ptr = {}
        for name, nbytes in (("vtx", 24 * vcap), ("keys", 8 * vcap), ("signs", vcap),
                             ("tri", 12 * tcap)):
            p = C.c_void_p()
            assert hip().hipMalloc(C.byref(p), max(nbytes, 256)) == 0
            self.ptr[name] = p.value

    def read(self, n: int, m: int):
        def get(name, dtype, count):
            a = np.empty(count, dtype)
            if count:
                assert hip().hipMemcpy(a.ctypes.data, self.ptr[name], a.nbytes, 2) == 0  # D2H
            return a
        return (get("vtx", np.float32, 6 * n).reshape(n, 6), get("keys", np.uint64, n),
                get("signs", np.uint8, n), get("tri", np.uint32, 3 * m).reshape(m, 3))

    def free(self):
        for p in self.ptr.values():
            hip().hipFree(p)
        self.ptr = {}


def noalloc(ctx, prog, params, bufs):
    return ctx.mesh_device_noalloc(prog, params, bufs.ptr["vtx"], bufs.ptr["keys"], bufs.ptr["signs"],
                                   bufs.vcap, bufs.ptr["tri"], bufs.tcap)


def check_same(ctx, prog, params, bufs):
    got = noalloc(ctx, prog, params, bufs)
    got["slabs"] = ctx.last_timing()["slabs"]
    want = ctx.mesh(prog, params)  # s2m_mesh: s2m_mesh_device, then a copy to host memory
    v, k, sg, t = bufs.read(got["nvtx"], got["ntri"])
    assert got["nvtx"] == len(want.vertices) > 0
    assert got["ntri"] == len(want.triangles)
    assert got["invalid_quads"] == want.invalid_quads
    assert np.array_equal(v.view(np.uint32), want.vertices.view(np.uint32))
    assert np.array_equal(k, want.cell_keys)
    assert np.array_equal(sg, want.signs)
    assert np.array_equal(t, want.triangles)
    return got


def test_noalloc_equals_mesh_device(ctx):
    big = ctx.program(tape.csg64())
    small = ctx.program(tape.mandelbulb(5))
    bufs = DeviceBuffers(2_000_000, 4_000_000)
    try:
        print("noalloc: csg64 at 256^3 (REFERENCE_COMPAT, 200 rounded up)", flush=True)
        a = check_same(ctx, big, s2m.Params.reference(200, 2.0), bufs)
        print(f"noalloc: {a['nvtx']} vertices; then the Mandelbulb at 64^3", flush=True)
        b = check_same(ctx, small, s2m.Params.reference(64, 2.5, flags=0), bufs)
        assert b["nvtx"] < a["nvtx"]
    finally:
        bufs.free()


def test_noalloc_1024_in_small_slabs(ctx):
    # a 1024^3 Mandelbulb: the scratch holds a few cell planes per slab, so the job runs in
    # hundreds of slabs, each carrying its last map plane to the next inside the scratch
    print("noalloc: a 1024^3 Mandelbulb job in small slabs", flush=True)
    bufs = DeviceBuffers(12_000_000, 26_000_000)
    try:
        got = check_same(ctx, ctx.program(tape.mandelbulb(8)), s2m.Params.reference(1024, 2.5), bufs)
        print(f"noalloc: {got['nvtx']} vertices in {got['slabs']} slabs", flush=True)
        assert got["slabs"] > 100
    finally:
        bufs.free()


def test_noalloc_reports_sizes_and_limits(ctx):
    prog = ctx.program(tape.csg64())
    p = s2m.Params.reference(128, 2.0)
    tiny = DeviceBuffers(1000, 1000)
    try:
        with pytest.raises(s2m.S2MError) as e:
            noalloc(ctx, prog, p, tiny)
        assert e.value.code == s2m.E_OOM
        nv, nt = e.value.needed
        assert nv > 1000
    finally:
        tiny.free()
    fit = DeviceBuffers(nv, max(nt, 3 * nv))
    try:
        got = check_same(ctx, prog, p, fit)
        assert got["nvtx"] == nv
    finally:
        fit.free()
    # one cell plane of an 8192 x 8192 grid needs more working memory than the scratch
    flat = s2m.Params((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (8192, 8192, 4), 1e-4, 0)
    one = DeviceBuffers(1, 1)
    try:
        with pytest.raises(s2m.S2MError) as e:
            noalloc(ctx, prog, flat, one)
        assert e.value.code == s2m.E_ARG
    finally:
        one.free()
