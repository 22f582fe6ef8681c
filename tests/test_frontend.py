"""Front end (.sdf3d WGSL and GLSL -> tape or HIP C++), on the CPU.

Pins:
  * tape lowering: the generated .sdf3d sources lower to exactly the tapes they were
    written from (CSG-64, random CSG trees), the reference's torus example to the
    TORUS op it states;
  * generated HIP C++: compiled for the host with g++ (the same text hiprtc receives,
    minus the kernels) and evaluated at many points, it reproduces the oracle bit for
    bit: the oracle's hand restatements of examples/martin_cube.sdf3d and
    examples/p_key.sdf3d, the oracle tape evaluator for CSG programs forced down the
    compiled path, and the oracle Mandelbulb for the GLSL Mandelbulb (host libm on
    both sides);
  * WGSL constant evaluation: abstract expressions fold in double and round once,
    f32 expressions round per operation;
  * error behaviour of shader.rs / naga: codes for parse errors, unknown names,
    library functions without `use`, GLSL without main or without the sdf function;
  * hiprtc builds a gfx950 code object for every example (no device needed).
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle
import sdf2mesh_amd as s2m
from sdf2mesh_amd import tape as T

REPO = Path(__file__).resolve().parents[1]
GOLDEN = REPO / "tests" / "golden"
CSRC = REPO / "sdf2mesh_amd" / "csrc"
# The reference's example programs are read where the reference is checked out (they
# are its sources, not fixtures); tests that need them are skipped elsewhere.
EXAMPLES = Path("/root/reference/examples")
needs_examples = pytest.mark.skipif(not (EXAMPLES / "martin_cube.sdf3d").exists(),
                                    reason="reference examples not available")

# ---------------------------------------------------------------------------------
# host execution of the generated code
# ---------------------------------------------------------------------------------
_HOST_PREFIX = f"""
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#define __HIPCC_RTC__ 1
#define __device__
#define __forceinline__ inline
#define __builtin_amdgcn_rsqf(x) (1.0f / sqrtf(x))
#define __builtin_amdgcn_sqrtf(x) sqrtf(x)
#define __builtin_amdgcn_logf(x) log2f(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
static inline uint32_t __float_as_uint(float f) {{ uint32_t u; memcpy(&u, &f, 4); return u; }}
static inline float __uint_as_float(uint32_t u) {{ float f; memcpy(&f, &u, 4); return f; }}
#include "{CSRC / 'device_sdf.h'}"
#include "{CSRC / 'rtc_prelude.h'}"
"""
_HOST_SUFFIX = """
extern "C" void host_eval(const float* p, long n, float* out)
{
    for (long i = 0; i < n; ++i)
        out[i] = s2m_entry(s2m::rtc::vf3{{p[3 * i], p[3 * i + 1], p[3 * i + 2]}});
}
"""
_host_cache: dict = {}


def host_program(hip: str):
    """Compile the device functions of generated HIP C++ for the host; returns
    f(points[n,3]) -> float32[n]."""
    body = hip[: hip.index("\nstruct s2m_user_fn")]
    if body in _host_cache:
        return _host_cache[body]
    tmp = Path(tempfile.mkdtemp(prefix="s2m_fe_"))
    src = tmp / "prog.cpp"
    so = tmp / "prog.so"
    src.write_text(_HOST_PREFIX + body + _HOST_SUFFIX)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-msse2",
                    "-mfpmath=sse", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    h = C.CDLL(str(so))
    h.host_eval.argtypes = [C.c_void_p, C.c_long, C.c_void_p]

    def f(pts: np.ndarray) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float32)
        out = np.empty(len(pts), np.float32)
        h.host_eval(pts.ctypes.data, len(pts), out.ctypes.data)
        return out

    _host_cache[body] = f
    return f


def sample_points(n: int, bound: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-bound, bound, (n, 3)).astype(np.float32)
    # plus lattice points of a reference grid (the values meshing actually uses)
    res = 24
    ax = np.linspace(-bound, bound, res, dtype=np.float32)
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    return np.concatenate([pts, g])


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def hip_of(source, lang="sdf3d", path=False, sdf_name=None) -> str:
    return s2m.translate(source, s2m.STAGE_HIP, lang, sdf_name, path).decode()


def status_of(source, lang="sdf3d", stage=s2m.STAGE_TAPE, path=False, sdf_name=None) -> int:
    try:
        s2m.translate(source, stage, lang, sdf_name, path)
    except s2m.S2MError as e:
        return e.code
    return s2m.OK


# ---------------------------------------------------------------------------------
# tape lowering
# ---------------------------------------------------------------------------------
def test_csg64_source_lowers_to_its_tape():
    t = s2m.translate_tape(GOLDEN / "csg64.sdf3d", path=True)
    assert t.tobytes() == T.csg64().tape().tobytes()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_csg_sources_lower_to_their_tapes(seed):
    t = s2m.translate_tape(GOLDEN / f"random_csg{seed}.sdf3d", path=True)
    assert t.tobytes() == T.random_csg(seed).tape().tobytes()


@pytest.mark.parametrize("seed", range(10, 30))
def test_random_csg_text_round_trip(seed):
    node = T.random_csg(seed, depth=4)
    t = s2m.translate_tape(node.sdf3d_source())
    assert t.tobytes() == node.tape().tobytes()


@needs_examples
def test_reference_torus_example():
    t = s2m.translate_tape(EXAMPLES / "torus.sdf3d", path=True)
    assert len(t) == 1 and t[0]["code"] == T.TORUS
    assert list(t[0]["f"][:5]) == [0.0, 0.0, 0.0, np.float32(0.5), np.float32(0.2)]


def test_inlining_and_constants_lower_to_a_tape():
    src = """use sdf3d::*;
use sdf::op;
const R = 0.25;
const OFF = vec3(0.5, 0.0, 0.0);
fn ball(p: vec3f, c: vec3f) -> f32 { return sdf3d_sphere(p - c, R * 2.0); }
fn sdf3d(p: vec3f) -> f32 {
    let a = ball(p, OFF);
    let b = sdf3d_box(p + vec3(0.25, 0.0, 0.0), vec3(0.5));
    return max(-a, sdf_op_smooth_union(b, sdf3d_capsule(p, vec3(0.0), vec3(0.0, 1.0, 0.0), 0.1), 0.05));
}
"""
    t = s2m.translate_tape(src)
    # p + c is p - (-c) bit for bit, so the box offset keeps the signs of -c: (-0.25, -0, -0)
    want = T.subtraction(T.sphere([0.5, 0, 0], 0.5),
                         T.smooth_union(T.box([-0.25, -0.0, -0.0], [0.5] * 3),
                                        T.capsule([0, 0, 0], [0, 1, 0], 0.1), 0.05)).tape()
    assert t.tobytes() == want.tobytes()


@pytest.mark.parametrize("body,why", [
    ("return sdf3d_sphere(p.xzy, 0.5);", "swizzle"),
    ("return sdf3d_sphere(p * 2.0, 0.5);", "arithmetic"),
    ("return sdf3d_sphere(p, 0.5) - 0.1;", "arithmetic"),
    ("var d = 1.0; for (var i = 0; i < 2; i++) { d = min(d, sdf3d_sphere(p, 0.5)); } return d;", "loop"),
])
def test_non_tape_programs(body, why):
    src = f"use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 {{ {body} }}\n"
    with pytest.raises(s2m.S2MError) as e:
        s2m.translate_tape(src)
    assert e.value.code == s2m.E_WGSL and why in str(e.value)
    assert "s2m_c_eval_lattice" in hip_of(src)


def test_abstract_and_f32_constant_evaluation():
    # abstract: (16777217 + 1) folds in double to 16777218, exact in f32
    # f32: f32(16777217) = 16777216, + 1 rounds back to 16777216 (ties to even)
    a = s2m.translate_tape("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { return sdf3d_sphere(p, (16777217.0 + 1.0) * 1e-7); }")
    b = s2m.translate_tape("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { let x: f32 = 16777217.0; return sdf3d_sphere(p, (x + 1.0) * 1e-7); }")
    assert a[0]["f"][3] == np.float32(16777218.0 * 1e-7)
    assert b[0]["f"][3] != a[0]["f"][3]
    x = np.float32(16777217.0)
    assert b[0]["f"][3] == np.float32(np.float32(x + np.float32(1.0)) * np.float32(1e-7))


# ---------------------------------------------------------------------------------
# generated HIP C++ executed on the host vs the oracle
# ---------------------------------------------------------------------------------
@pytest.mark.parametrize("path,code,bound", [
    pytest.param(EXAMPLES / "martin_cube.sdf3d", oracle.ORA_OP_MARTIN_CUBE, 1.6, marks=needs_examples),
    pytest.param(EXAMPLES / "p_key.sdf3d", oracle.ORA_OP_P_KEY, 10.0, marks=needs_examples),
    (GOLDEN / "engraved_cube.sdf3d", oracle.ORA_OP_ENGRAVED_CUBE, 1.0),
])
def test_programs_compiled_match_oracle(path, code, bound):
    f = host_program(hip_of(path, path=True))
    pts = sample_points(40000, bound, 7)
    got = f(pts)
    want = oracle.eval_f32(oracle.example_tape(code), pts)
    np.testing.assert_array_equal(bits(got), bits(want))
    # the restatement agrees with the fp64 evaluation of the same expressions
    ref64 = oracle.eval_f64(oracle.example_tape(code), pts)
    assert np.max(np.abs(want - ref64)) < 1e-5 * bound


@pytest.mark.parametrize("src", ["csg64.sdf3d", "random_csg1.sdf3d", "random_csg2.sdf3d",
                                 "random_csg3.sdf3d"])
def test_tape_programs_compiled_match_tape(src):
    f = host_program(hip_of(GOLDEN / src, path=True))
    t = s2m.translate_tape(GOLDEN / src, path=True)
    pts = sample_points(20000, 1.0, 3)
    np.testing.assert_array_equal(bits(f(pts)), bits(oracle.eval_f32(t, pts)))


@pytest.mark.parametrize("iters", [4, 8])
def test_glsl_mandelbulb_compiled_matches_oracle(iters):
    f = host_program(hip_of(T.mandelbulb_glsl(iters), lang="glsl"))
    pts = sample_points(20000, 1.25, 5)
    want = oracle.eval_f32(T.mandelbulb(iters).tape(), pts)
    np.testing.assert_array_equal(bits(f(pts)), bits(want))


@needs_examples
def test_reference_mandelmesh_example_matches_oracle():
    # examples/mandelmesh.frag: 5 iterations, power 8, bailout 2, offset 0.003
    f = host_program(hip_of(EXAMPLES / "mandelmesh.frag", lang="glsl", path=True))
    pts = sample_points(20000, 1.25, 9)
    want = oracle.eval_f32(T.mandelbulb(5).tape(), pts)
    np.testing.assert_array_equal(bits(f(pts)), bits(want))


def test_builtins_and_control_flow():
    """Loops, breaks, swizzle assignment, integer arithmetic and the GLSL builtins,
    against a numpy fp32 restatement of the same program."""
    glsl = """#version 450
float sdf(vec3 p) {
    vec3 q = abs(p) - vec3(0.3);
    q.xy = q.yx * 2.0;
    int n = 0;
    float acc = 0.0;
    for (int i = 0; i < 10; i++) {
        if (i == 7) break;
        if (i % 2 == 1) continue;
        acc += float(i) * 0.125;
        n += 1;
    }
    float s = clamp(q.x, -0.5, 0.5) + mix(q.y, q.z, 0.25) + step(0.0, q.z) * 0.01;
    return length(max(q, 0.0)) + min(max(q.x, max(q.y, q.z)), 0.0) + s * 0.001 + acc * 0.0001 + float(n);
}
void main() {}
"""
    f = host_program(hip_of(glsl, lang="glsl"))
    pts = sample_points(5000, 1.0, 11)
    f32 = np.float32
    x, y, z = (pts[:, k].astype(f32) for k in range(3))
    qx0, qy0, qz = np.abs(x) - f32(0.3), np.abs(y) - f32(0.3), np.abs(z) - f32(0.3)
    qx, qy = qy0 * f32(2.0), qx0 * f32(2.0)
    acc, n = f32(0.0), 0
    for i in range(10):
        if i == 7:
            break
        if i % 2 == 1:
            continue
        acc = f32(acc + f32(f32(i) * f32(0.125)))
        n += 1
    s = (np.minimum(np.maximum(qx, f32(-0.5)), f32(0.5)) + (qy * (f32(1) - f32(0.25)) + qz * f32(0.25))
         + np.where(qz >= 0, f32(1), f32(0)) * f32(0.01)).astype(f32)
    mx, my, mz = np.maximum(qx, f32(0)), np.maximum(qy, f32(0)), np.maximum(qz, f32(0))
    length = np.sqrt(((mx * mx + my * my) + mz * mz).astype(f32)).astype(f32)
    inner = np.minimum(np.maximum(qx, np.maximum(qy, qz)), f32(0))
    want = ((((length + inner) + s * f32(0.001)) + acc * f32(0.0001)) + f32(n)).astype(f32)
    np.testing.assert_array_equal(bits(f(pts)), bits(want))


def test_glsl_loop_expression_is_any_expression():
    # GLSL's for-loop expression may be any expression, not only an assignment: a bare
    # `i` is a no-op step (this loop never ends by its condition; the emitted trip-count
    # guard ends it), and the generated code evaluates it as an expression statement
    glsl = ("#version 450\nfloat sdf(vec3 p) { float d = length(p) - 0.5; "
            "for (int i = 0; i < 8; i) { if (d > 10.0) break; } return d; }\nvoid main() {}\n")
    hip = s2m.translate(glsl, s2m.STAGE_HIP, lang="glsl").decode()
    assert "(void)(u_i)" in hip
    assert len(s2m.translate(glsl, s2m.STAGE_CODE, lang="glsl")) > 0  # hiprtc accepts it


def test_unbounded_loop_terminates():
    src = "use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { var i = 0; while (true) { i = i + 1; } return sdf3d_sphere(p, 0.5); }"
    f = host_program(hip_of(src))
    pts = np.zeros((1, 3), np.float32)
    assert f(pts)[0] == np.float32(-0.5)


# ---------------------------------------------------------------------------------
# preprocessing and errors
# ---------------------------------------------------------------------------------
def test_use_and_include_expansion(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    (tmp_path / "part.wgsl").write_text("fn helper(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.5); }\ninclude \"part.wgsl\";\n")
    (tmp_path / "main.sdf3d").write_text("use sdf3d::*;\nuse unknown::module;\ninclude \"part.wgsl\";\nfn sdf3d(p: vec3f) -> f32 { return helper(p); }\n")
    text = s2m.translate(tmp_path / "main.sdf3d", s2m.STAGE_SOURCE, path=True).decode()
    # use lines removed, the include spliced in once (a file that includes itself is skipped)
    assert text == ("fn helper(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.5); }\n"
                    "fn sdf3d(p: vec3f) -> f32 { return helper(p); }\n")
    t = s2m.translate_tape(tmp_path / "main.sdf3d", path=True)
    assert len(t) == 1 and t[0]["code"] == T.SPHERE


@pytest.mark.parametrize("src,lang,code", [
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.5) }", "sdf3d", s2m.E_PARSE),
    ("use sdf3d::*;\nfn other(p: vec3f) -> f32 { return 0.0; }", "sdf3d", s2m.E_MISSING_SDF),
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { return nope(p); }", "sdf3d", s2m.E_VALIDATION),
    ("fn sdf3d(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.5); }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::primitives;\nfn sdf3d(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.5); }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { return sdf_op_smooth_union(1.0, 2.0, 0.1); }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { return sdf3d(p); }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::*;\nfn sdf3d_sphere(p: vec3f, s: f32) -> f32 { return 0.0; }\nfn sdf3d(p: vec3f) -> f32 { return 0.0; }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { let a = 1.0; a = 2.0; return a; }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> vec3f { return p; }", "sdf3d", s2m.E_VALIDATION),
    ("use sdf3d::*;\nfn sdf3d(p: vec3f) -> f32 { var d; d = 1.0; return d; }", "sdf3d", s2m.E_PARSE),
    ("#version 450\nfloat sdf(vec3 p) { return length(p) - 1.0; }\n", "glsl", s2m.E_SHADER),
    ("#version 450\nfloat other(vec3 p) { return length(p) - 1.0; }\nvoid main() {}\n", "glsl", s2m.E_MISSING_SDF),
    ("#version 450\nfloat sdf(vec3 p) { return length(p) - 1.0 }\nvoid main() {}\n", "glsl", s2m.E_PARSE),
])
def test_error_codes(src, lang, code):
    assert status_of(src, lang=lang, stage=s2m.STAGE_HIP) == code


def test_missing_file():
    assert status_of("/nonexistent/x.sdf3d", path=True) == s2m.E_IO


def test_glsl_named_sdf_and_sdf3d_wrapper():
    glsl = "#version 450\nfloat ball(vec3 p) { return length(p) - 0.5; }\nvoid main() {}\n"
    f = host_program(hip_of(glsl, lang="glsl", sdf_name="ball"))
    assert f(np.array([[1.0, 0.0, 0.0]], np.float32))[0] == np.float32(0.5)
    # a function named sdf3d is the SDF even when --glsl-sdf names another one
    glsl2 = ("#version 450\nfloat ball(vec3 p) { return length(p) - 0.5; }\n"
             "float sdf3d(vec3 p) { return length(p) - 0.25; }\nvoid main() {}\n")
    f2 = host_program(hip_of(glsl2, lang="glsl", sdf_name="ball"))
    assert f2(np.array([[1.0, 0.0, 0.0]], np.float32))[0] == np.float32(0.75)


@pytest.mark.parametrize("path,lang", [
    pytest.param(EXAMPLES / "martin_cube.sdf3d", "sdf3d", marks=needs_examples),
    pytest.param(EXAMPLES / "p_key.sdf3d", "sdf3d", marks=needs_examples),
    pytest.param(EXAMPLES / "mandelmesh.frag", "glsl", marks=needs_examples),
    (GOLDEN / "engraved_cube.sdf3d", "sdf3d"),
    (GOLDEN / "mandelbulb8.frag", "glsl"),
])
def test_hiprtc_builds_code_objects(path, lang):
    code = s2m.translate(path, s2m.STAGE_CODE, lang=lang, path=True)
    assert code[:4] == b"\x7fELF" and len(code) > 1000


@pytest.mark.parametrize("src,lang", [
    # vector-valued user functions, a user normal, integer vectors, every builtin family
    ("use sdf3d::primitives;\nfn sdf3d(p: vec3f) -> f32 { return sdf3d_sphere(p, 0.6); }\n"
     "fn sdf3d_normal(p: vec3f, eps: f32) -> vec3f { return vec3(p.x, 2.0 * p.y, 0.0); }\n", "sdf3d"),
    ("use sdf3d::*;\nfn f(q: vec2f) -> vec2f { return abs(q) - vec2(0.1); }\n"
     "fn sdf3d(p: vec3f) -> f32 { let i = vec3<i32>(1, 2, 3); var u = 7u; u = u / 2u + u % 3u;\n"
     "  let a = f(p.xy); let s = sin(p.x) + cos(p.y) + tan(0.1 * p.z) + exp(-p.x) + log(2.0 + p.y)\n"
     "  + pow(abs(p.z), 1.5) + sqrt(abs(p.x)) + inverseSqrt(1.0 + p.y * p.y) + floor(p.x) + ceil(p.y)\n"
     "  + fract(p.z) + sign(p.x) + step(0.0, p.y) + smoothstep(0.0, 1.0, p.z) + atan2(p.y, p.x)\n"
     "  + trunc(p.x) + round(p.y) + exp2(p.z) + log2(2.0 + p.x) + asin(0.5) + acos(0.5);\n"
     "  return length(a) + dot(cross(p, vec3(0.0, 1.0, 0.0)), normalize(vec3(1.0))) + distance(p, vec3(0.5))\n"
     "  + f32(i.y) * 0.01 + f32(u) * 0.001 + s * 1e-6 + select(0.0, 1.0, p.x > 0.0) - 0.5; }\n", "sdf3d"),
    ("#version 450\nvec2 g(vec2 q) { return q * 2.0; }\nfloat sdf(vec3 p) { ivec3 k = ivec3(1, 2, 3);\n"
     "  float m = mod(p.x, 0.3) + (p.y > 0.0 ? 1.0 : 0.0) + atan(p.y, p.x) + inversesqrt(2.0);\n"
     "  int n = 0; while (n < 3) { n++; } return length(g(p.xy)) - 0.5 + m * 0.01 + float(k.x + n) * 0.0; }\n"
     "void main() {}\n", "glsl"),
])
def test_hiprtc_builds_varied_programs(src, lang):
    code = s2m.translate(src, s2m.STAGE_CODE, lang=lang)
    assert code[:4] == b"\x7fELF"


_MUTATE = r'''
import random, sys
from pathlib import Path
import sdf2mesh_amd as s2m
G = Path(sys.argv[1])
seeds = [(p.read_text(), "sdf3d") for p in sorted(G.glob("*.sdf3d"))] + [((G / "mandelbulb8.frag").read_text(), "glsl")]
alph = list("(){}[];:,.+-*/<>=!&|?^%~ \n0123456789exyz_") + ["fn ", "let ", "var ", "return ", "if ", "else ",
       "for ", "while ", "vec3f(", "sdf3d", "1e38", "1e-45", "i32", "u32", "f32", "break;", "float ", "vec3 "]
rng = random.Random(int(sys.argv[2]))
for it in range(int(sys.argv[3])):
    src, lang = rng.choice(seeds)
    for _ in range(rng.randint(1, 6)):
        if not src:
            break
        i = rng.randrange(len(src)); j = min(len(src), i + rng.randint(1, 40)); k = rng.randrange(6)
        if k == 0: src = src[:i] + src[j:]
        elif k == 1: src = src[:i] + src[i:j] + src[i:]
        elif k == 2: src = src[:i] + rng.choice(alph) + src[i:]
        elif k == 3: src = src[:i] + chr(rng.randrange(1, 128)) + src[i + 1:]
        elif k == 4: src = src[:i]
        else: src = src[:i] + src[j:] + src[i:j]
    for stage in (s2m.STAGE_SOURCE, s2m.STAGE_TAPE, s2m.STAGE_HIP):
        try:
            s2m.translate(src, stage, lang)
        except s2m.S2MError:
            pass
print("ok")
'''


@pytest.mark.parametrize("seed", [1, 6, 7])
def test_mutated_programs_fail_cleanly(seed):
    # random edits of the test programs through every host stage: each is translated or
    # rejected with an error code, never a crash (a child process, so a crash fails here)
    r = subprocess.run([sys.executable, "-c", _MUTATE, str(GOLDEN), str(seed), "1500"], capture_output=True,
                       text=True, timeout=600, cwd=str(REPO))
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.returncode, r.stderr[-2000:])
