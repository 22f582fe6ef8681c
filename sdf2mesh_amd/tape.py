"""SDF programs as tapes over the reference's primitive/operator library.

A program is an expression tree of library primitives (sdf3d_primitives.wgsl:7-36)
and operators (sdf_op.wgsl:7-23, plus min/max unions).  It lowers to

* a postfix tape (numpy array of ``TAPE_DTYPE``, the ``s2m_op`` layout of
  include/sdf2mesh.h) that the device interpreter evaluates, and
* the equivalent ``.sdf3d`` WGSL text the reference CLI would read
  (``use sdf3d::*; use sdf::op;`` + ``fn sdf3d(p: vec3f) -> f32``).

Operands are stored as float32; WGSL literals are printed so that they convert to
exactly the same float32 values.

Workload generators (seeded, platform independent):
  torus_example()    examples/torus.sdf3d: torus (0.5, 0.2) at the origin
  csg64()            64 smooth-unioned random primitives (seed 0x5DF2)
  mandelbulb()       examples/mandelmesh.frag distance estimator (power 8)
  random_csg(seed)   random CSG tree of depth 1-4 over the five library primitives
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

TAPE_DTYPE = np.dtype([("code", "<u4"), ("iarg", "<u4"), ("f", "<f4", (10,))])
assert TAPE_DTYPE.itemsize == 48

SPHERE, BOX, TORUS, CYLINDER, CAPSULE, MANDELBULB = 1, 2, 3, 4, 5, 6
UNION, INTERSECTION, SUBTRACTION = 16, 17, 18
SMOOTH_UNION, SMOOTH_INTERSECTION, SMOOTH_SUBTRACTION = 19, 20, 21
MANDEL_SWAP_YZ = 1 << 16
MANDEL_TRIG = 1 << 17  # the escape step as written (atan, asin, pow, sin, cos)

_OP_WGSL = {
    SMOOTH_UNION: "sdf_op_smooth_union",
    SMOOTH_INTERSECTION: "sdf_op_smooth_intersection",
    SMOOTH_SUBTRACTION: "sdf_op_smooth_subtraction",
}


def f32(x) -> np.float32:
    return np.float32(x)


def wgsl_float(x) -> str:
    """A WGSL float literal whose conversion to f32 gives exactly float32(x)."""
    v = np.float32(x)
    s = np.format_float_positional(v, unique=True, trim="-")
    if np.float32(float(s)) != v or not np.isfinite(v):  # guard against double rounding
        s = repr(float(v))
    if "." not in s and "e" not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


def _vec(xs) -> str:
    return f"vec{len(xs)}(" + ", ".join(wgsl_float(x) for x in xs) + ")"


class Node:
    """Expression node of an SDF program."""

    def postfix(self, out: list) -> None:
        raise NotImplementedError

    def wgsl(self) -> str:
        raise NotImplementedError

    def flops(self) -> int:
        """Static fp32 operation count of the WGSL expression (algorithmic work per
        point): + - * / min max abs sqrt and each transcendental = 1, clamp = 2,
        mix = 3, length(vec3) = 6, length(vec2) = 4, dot(vec3) = 5; constant
        subexpressions are folded."""
        raise NotImplementedError

    def tape(self) -> np.ndarray:
        rows: list = []
        self.postfix(rows)
        t = np.zeros(len(rows), TAPE_DTYPE)
        for i, (code, iarg, fs) in enumerate(rows):
            t[i]["code"] = code
            t[i]["iarg"] = iarg
            t[i]["f"][: len(fs)] = np.asarray(fs, np.float32)
        return t

    def sdf3d_source(self) -> str:
        """The program as an .sdf3d file for the reference CLI.  The left spine of
        operators (a left fold such as CSG-64) becomes a `var d` update chain, which
        evaluates the same operations in the same order as the nested expression."""
        spine = []
        n = self
        while isinstance(n, Op):
            spine.append(n)
            n = n.a
        lines = ["use sdf3d::*;", "use sdf::op;", "", "fn sdf3d(p: vec3f) -> f32 {"]
        if len(spine) < 2:
            lines.append(f"    return {self.wgsl()};")
        else:
            lines.append(f"    var d = {n.wgsl()};")
            for op in reversed(spine):
                lines.append(f"    d = {op.wgsl_with_left('d')};")
            lines.append("    return d;")
        lines.append("}")
        return "\n".join(lines) + "\n"


@dataclass
class Prim(Node):
    code: int
    f: List[float]
    iarg: int = 0

    def __post_init__(self):
        self.f = [float(np.float32(v)) for v in self.f]

    def postfix(self, out):
        out.append((self.code, self.iarg, self.f))

    def _p(self) -> str:
        c = self.f[:3]
        return "p" if all(v == 0.0 and not np.signbit(v) for v in c) else f"(p - {_vec(c)})"

    def _translated(self) -> bool:
        return not all(v == 0.0 and not np.signbit(v) for v in self.f[:3])

    def wgsl(self):
        f = self.f
        if self.code == SPHERE:
            return f"sdf3d_sphere({self._p()}, {wgsl_float(f[3])})"
        if self.code == BOX:
            return f"sdf3d_box({self._p()}, {_vec(f[3:6])})"
        if self.code == TORUS:
            return f"sdf3d_torus({self._p()}, {_vec(f[3:5])})"
        if self.code == CYLINDER:
            return f"sdf3d_cylinder({self._p()}, {wgsl_float(f[3])}, {wgsl_float(f[4])})"
        if self.code == CAPSULE:
            return (f"sdf3d_capsule(p, {_vec(f[0:3])}, {_vec(f[3:6])}, "
                    f"{wgsl_float(f[6])})")
        raise ValueError("the Mandelbulb is not a library call; use mandelbulb_glsl()")

    def flops(self):
        t = 3 if self.code in (SPHERE, BOX, TORUS, CYLINDER) and self._translated() else 0
        if self.code == SPHERE:      # length3, -s
            return t + 7
        if self.code == BOX:         # abs*3, -*3, max*3, length3, max*2, min, +
            return t + 3 + 3 + 3 + 6 + 2 + 1 + 1
        if self.code == TORUS:       # length2, -, length2, -
            return t + 4 + 1 + 4 + 1
        if self.code == CYLINDER:    # length2, abs*2, -*2, max, min, +, max*2, length2
            return t + 4 + 2 + 2 + 1 + 1 + 1 + 2 + 4
        if self.code == CAPSULE:     # pa(3), dot(5), /, clamp(2), ba*h(3), pa-(3), length3, -r
            return 3 + 5 + 1 + 2 + 3 + 3 + 6 + 1
        if self.code == MANDELBULB:
            iters = self.iarg & 0xFFFF
            # per iteration: length3 6, atan+div 2, asin+div 2, pow*dr*power+1 4, pow 1,
            # theta/phi scale 2, 5 trig + 2 mul + r*vec 3 + p 3 = 13
            return iters * (6 + 2 + 2 + 4 + 1 + 2 + 13) + 5
        raise ValueError(self.code)


@dataclass
class Op(Node):
    code: int
    a: Node
    b: Node
    k: float = 0.0

    def __post_init__(self):
        self.k = float(np.float32(self.k))

    def postfix(self, out):
        self.a.postfix(out)
        self.b.postfix(out)
        out.append((self.code, 0, [self.k]))

    def wgsl(self):
        return self.wgsl_with_left(self.a.wgsl())

    def wgsl_with_left(self, A: str) -> str:
        B = self.b.wgsl()
        if self.code == UNION:
            return f"min({A}, {B})"
        if self.code == INTERSECTION:
            return f"max({A}, {B})"
        if self.code == SUBTRACTION:
            return f"max(-({A}), {B})"
        return f"{_OP_WGSL[self.code]}({A}, {B}, {wgsl_float(self.k)})"

    def flops(self):
        own = {UNION: 1, INTERSECTION: 1, SUBTRACTION: 2}.get(self.code)
        if own is None:  # smooth ops: d2-d1, 0.5*, /k, 0.5+-, clamp 2, mix 3, k*h, 1-h, *, -+
            own = 1 + 1 + 1 + 1 + 2 + 3 + 1 + 1 + 1 + 1
        return self.a.flops() + self.b.flops() + own


# ---- constructors ------------------------------------------------------------------
def sphere(c, s) -> Prim:
    return Prim(SPHERE, [*c, s])


def box(c, b) -> Prim:
    return Prim(BOX, [*c, *b])


def torus(c, t) -> Prim:
    return Prim(TORUS, [*c, *t])


def cylinder(c, h, r) -> Prim:
    return Prim(CYLINDER, [*c, h, r])


def capsule(a, b, r) -> Prim:
    return Prim(CAPSULE, [*a, *b, r])


def mandelbulb(iterations: int = 8, power: float = 8.0, bailout: float = 2.0,
               offset: float = 0.003, swap_yz: bool = True, trig: bool = False) -> Prim:
    """The estimator of examples/mandelmesh.frag; trig=True evaluates its escape step as
    written (else a power-8 step is computed without transcendentals, same function)."""
    return Prim(MANDELBULB, [power, bailout, offset],
                iarg=(iterations & 0xFFFF) | (MANDEL_SWAP_YZ if swap_yz else 0)
                | (MANDEL_TRIG if trig else 0))


def union(a, b) -> Op:
    return Op(UNION, a, b)


def intersection(a, b) -> Op:
    return Op(INTERSECTION, a, b)


def subtraction(a, b) -> Op:
    return Op(SUBTRACTION, a, b)


def smooth_union(a, b, k) -> Op:
    return Op(SMOOTH_UNION, a, b, k)


def smooth_intersection(a, b, k) -> Op:
    return Op(SMOOTH_INTERSECTION, a, b, k)


def smooth_subtraction(a, b, k) -> Op:
    return Op(SMOOTH_SUBTRACTION, a, b, k)


def fold(nodes: Sequence[Node], op=smooth_union, k: float = 0.05) -> Node:
    """Left fold d = op(d, d_i) over the nodes."""
    d = nodes[0]
    for n in nodes[1:]:
        d = op(d, n, k) if op in (smooth_union, smooth_intersection, smooth_subtraction) else op(d, n)
    return d


# ---- deterministic RNG (splitmix64) -------------------------------------------------
class Rng:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * ((self.u64() >> 11) * (1.0 / (1 << 53)))

    def below(self, n: int) -> int:
        return self.u64() % n

    def shuffle(self, xs: list) -> list:
        for i in range(len(xs) - 1, 0, -1):
            j = self.below(i + 1)
            xs[i], xs[j] = xs[j], xs[i]
        return xs

    def choice_weighted(self, items, weights):
        t = self.uniform(0.0, float(sum(weights)))
        for it, w in zip(items, weights):
            if t < w:
                return it
            t -= w
        return items[-1]


# ---- workloads ----------------------------------------------------------------------
def torus_example() -> Node:
    """examples/torus.sdf3d: sdf3d_torus(p, vec2(0.5, 0.2))."""
    return torus((0.0, 0.0, 0.0), (0.5, 0.2))


TORUS_SDF3D = """use sdf3d::*;

fn sdf3d(p: vec3f) -> f32 {
    return sdf3d_torus(p, vec2(0.5, 0.2));
}
"""


def csg64(seed: int = 0x5DF2, n: int = 64, k: float = 0.05) -> Node:
    """Smooth union (k) of n primitives: spheres, boxes and tori in equal proportion
    (22/21/21 for n = 64) in shuffled order, centres U[-0.6, 0.6]^3, size
    s ~ U[0.05, 0.2]: sphere radius s, box edges 2s, torus (s, 0.4 s)."""
    rng = Rng(seed)
    kinds = [SPHERE] * (n - 2 * (n // 3)) + [BOX] * (n // 3) + [TORUS] * (n // 3)
    rng.shuffle(kinds)
    prims = []
    for kind in kinds:
        c = [rng.uniform(-0.6, 0.6) for _ in range(3)]
        s = rng.uniform(0.05, 0.2)
        if kind == SPHERE:
            prims.append(sphere(c, s))
        elif kind == BOX:
            prims.append(box(c, (2 * s, 2 * s, 2 * s)))
        else:
            prims.append(torus(c, (s, 0.4 * s)))
    return fold(prims, smooth_union, k)


def random_csg(seed: int, depth: int | None = None, k: float = 0.05) -> Node:
    """Random CSG tree of depth 1-4 over sphere/box/torus/cylinder/capsule with
    union / intersection / subtraction / smooth union, inside [-0.8, 0.8]^3."""
    rng = Rng(0xC56 ^ (seed * 0x9E3779B1))
    if depth is None:
        depth = 1 + rng.below(4)

    def prim() -> Node:
        kind = rng.below(5)
        c = [rng.uniform(-0.45, 0.45) for _ in range(3)]
        s = rng.uniform(0.12, 0.35)
        if kind == 0:
            return sphere(c, s)
        if kind == 1:
            return box(c, [s * rng.uniform(1.0, 2.0) for _ in range(3)])
        if kind == 2:
            return torus(c, (s, s * rng.uniform(0.25, 0.5)))
        if kind == 3:
            return cylinder(c, s, s * rng.uniform(0.4, 0.9))
        a = [ci - s * 0.5 for ci in c]
        b = [ci + s * 0.5 for ci in c]
        return capsule(a, b, s * 0.4)

    def tree(d: int) -> Node:
        if d == 0:
            return prim()
        op = rng.below(4)
        a, b = tree(d - 1), tree(d - 1 if rng.below(2) else max(0, d - 2))
        if op == 0:
            return union(a, b)
        if op == 1:
            # keep intersections non-empty: intersect with a larger primitive
            return intersection(a, union(b, sphere((0.0, 0.0, 0.0), 0.6)))
        if op == 2:
            return subtraction(b, a)
        return smooth_union(a, b, k)

    return tree(depth)


_MANDELBULB_GLSL = """#version 450 core
// Power-8 Mandelbulb distance estimator, ITER iterations, bailout radius 2.
float sdf(vec3 p) {
    p.xyz = p.xzy;
    vec3 z = p;
    float power = 8.0;
    float r, theta, phi;
    float dr = 1.0;
    for (int i = 0; i < ITER; ++i) {
        r = length(z);
        if (r > 2.0) continue;
        theta = atan(z.y / z.x);
        phi = asin(z.z / r);
        dr = pow(r, power - 1.0) * dr * power + 1.0;
        r = pow(r, power);
        theta = theta * power;
        phi = phi * power;
        z = r * vec3(cos(theta) * cos(phi), sin(theta) * cos(phi), sin(phi)) + p;
    }
    return 0.5 * log(r) * r / dr - 0.003;
}

void main() {}
"""


def mandelbulb_glsl(iterations: int = 8) -> str:
    """GLSL fragment shader with `float sdf(vec3)`: the distance estimator of
    examples/mandelmesh.frag:3-28 with the given loop bound."""
    return _MANDELBULB_GLSL.replace("ITER", str(int(iterations)))
