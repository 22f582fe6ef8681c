// device_sdf.h - the SDF primitive/operator library and the tape interpreter, as
// device functions for gfx950.
//
// fp32 throughout, each vector expression of the WGSL library expanded per component in
// the order the WGSL text evaluates it.  The library is compiled with
// -ffp-contract=off and correctly rounded division and square root, so every value
// below is the IEEE result of the written operation sequence.
//   sdf3d_primitives.wgsl:7-36   box, cylinder, capsule, sphere, torus
//   sdf_op.wgsl:7-23             smooth union / intersection / subtraction
//   examples/mandelmesh.frag:3-28 the Mandelbulb distance estimator
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "../../include/sdf2mesh.h"

namespace s2m {

__device__ __forceinline__ float len2(float x, float y) { return sqrtf(x * x + y * y); }
__device__ __forceinline__ float len3(float x, float y, float z)
{
    return sqrtf((x * x + y * y) + z * z);
}
// WGSL clamp(e, low, high) = min(max(e, low), high)
__device__ __forceinline__ float wclamp(float e, float lo, float hi)
{
    return fminf(fmaxf(e, lo), hi);
}
// WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3
__device__ __forceinline__ float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// sdf3d_sphere: length(p) - s
__device__ __forceinline__ float sd_sphere(float x, float y, float z, float s)
{
    return len3(x, y, z) - s;
}

// sdf3d_box: q = abs(p) - 0.5*b; length(max(q,0)) + min(max(q.x,max(q.y,q.z)),0)
__device__ __forceinline__ float sd_box(float x, float y, float z, float bx, float by, float bz)
{
    const float qx = fabsf(x) - 0.5f * bx;
    const float qy = fabsf(y) - 0.5f * by;
    const float qz = fabsf(z) - 0.5f * bz;
    const float outside = len3(fmaxf(qx, 0.0f), fmaxf(qy, 0.0f), fmaxf(qz, 0.0f));
    return outside + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
}

// sdf3d_cylinder: d = abs(vec2(length(p.xz), p.y)) - vec2(r, h);
// min(max(d.x,d.y),0) + length(max(d,0))
__device__ __forceinline__ float sd_cylinder(float x, float y, float z, float h, float r)
{
    const float dx = fabsf(len2(x, z)) - r;
    const float dy = fabsf(y) - h;
    return fminf(fmaxf(dx, dy), 0.0f) + len2(fmaxf(dx, 0.0f), fmaxf(dy, 0.0f));
}

// sdf3d_capsule: pa = p - a; ba = b - a; h = clamp(dot(pa,ba)/dot(ba,ba), 0, 1);
// length(pa - ba*h) - r
__device__ __forceinline__ float sd_capsule(float x, float y, float z, float ax, float ay, float az,
                                            float bx, float by, float bz, float r)
{
    const float pax = x - ax, pay = y - ay, paz = z - az;
    const float bax = bx - ax, bay = by - ay, baz = bz - az;
    const float h = wclamp(((pax * bax + pay * bay) + paz * baz) / ((bax * bax + bay * bay) + baz * baz),
                           0.0f, 1.0f);
    return len3(pax - bax * h, pay - bay * h, paz - baz * h) - r;
}

// sdf3d_torus: q = vec2(length(p.xz) - t.x, p.y); length(q) - t.y
__device__ __forceinline__ float sd_torus(float x, float y, float z, float tx, float ty)
{
    return len2(len2(x, z) - tx, y) - ty;
}

// sdf_op_smooth_union: h = clamp(0.5 + 0.5*(d2-d1)/k, 0, 1); mix(d2,d1,h) - k*h*(1-h)
__device__ __forceinline__ float op_smooth_union(float d1, float d2, float k)
{
    const float h = wclamp(0.5f + (0.5f * (d2 - d1)) / k, 0.0f, 1.0f);
    return wmix(d2, d1, h) - (k * h) * (1.0f - h);
}
// sdf_op_smooth_intersection: h = clamp(0.5 - 0.5*(d2-d1)/k, 0, 1); mix(d2,d1,h) + k*h*(1-h)
__device__ __forceinline__ float op_smooth_intersection(float d1, float d2, float k)
{
    const float h = wclamp(0.5f - (0.5f * (d2 - d1)) / k, 0.0f, 1.0f);
    return wmix(d2, d1, h) + (k * h) * (1.0f - h);
}
// sdf_op_smooth_subtraction: h = clamp(0.5 - 0.5*(d2+d1)/k, 0, 1); mix(d2,-d1,h) + k*h*(1-h)
__device__ __forceinline__ float op_smooth_subtraction(float d1, float d2, float k)
{
    const float h = wclamp(0.5f - (0.5f * (d2 + d1)) / k, 0.0f, 1.0f);
    return wmix(d2, -d1, h) + (k * h) * (1.0f - h);
}

// 1 / sqrt(x) from the hardware reciprocal square root (about 1 ulp).  Inputs below
// 2^-96 are scaled by 2^64 first (and the result by 2^32), so that denormal inputs get
// their true reciprocal root instead of a flushed one; 0 gives +inf.
__device__ __forceinline__ float rsqrt_full(float x)
{
    const bool tiny = x < 0x1p-96f;
    const float y = __builtin_amdgcn_rsqf(tiny ? x * 0x1p64f : x);
    return tiny ? y * 0x1p32f : y;
}

// The largest float t with sqrtf(t) <= b (correctly rounded square root), so that
// sqrtf(x) > b exactly when x > t: b * b rounded, moved to that boundary by at most a few
// ulp (the boundary (b + ulp(b) / 2)^2 lies within 1.5 ulp of b^2).  b < 0 gives a
// negative threshold (every length exceeds b), b = +-0 gives b (every nonzero length
// exceeds it) and NaN gives NaN (no length exceeds it), as the comparisons of lengths do.
__device__ __forceinline__ float bailout_threshold(float b)
{
    if (!(b > 0.0f)) return b < 0.0f ? -1.0f : b;
    float t = b * b;
    for (int k = 0; k < 4 && sqrtf(t) > b; ++k) t = __uint_as_float(__float_as_uint(t) - 1u);
    for (int k = 0; k < 4; ++k) {
        const float up = __uint_as_float(__float_as_uint(t) + 1u);
        if (!(sqrtf(up) <= b)) break;
        t = up;
    }
    return t;
}

// examples/mandelmesh.frag:3-28 with power, bailout, offset and the iteration count as
// operands.  A bailed-out point keeps r = length(z): the loop `continue`s, so r is
// recomputed from the unchanged z in every remaining iteration.  t2 =
// bailout_threshold(bailout), computed once per tape upload (k_tape_prepare) rather than
// at every point.
__device__ __forceinline__ float sd_mandelbulb(float px, float py, float pz, float power,
                                               float bailout, float offset, uint32_t iarg, float t2)
{
    if (iarg & S2M_MANDEL_SWAP_YZ) {  // p.xyz = p.xzy
        const float t = py;
        py = pz;
        pz = t;
    }
    const uint32_t iters = iarg & 0xffffu;
    float zx = px, zy = py, zz = pz;
    float r = 0.0f, dr = 1.0f;
    if (power == 8.0f && !(iarg & S2M_MANDEL_TRIG)) {
        // The power-8 step without transcendentals or divisions.  theta = atan(zy / zx)
        // differs from the direction angle of (zx, zy) by 0 or pi, which 8 * theta turns
        // into a multiple of 8 pi: cos / sin of 8 theta are the components of
        // ((zx, zy) / rho)^8 as complex numbers (rho = |(zx, zy)|), three complex squarings
        // of the unit vector (1 / rho from a reciprocal square root).  phi = asin(zz / r)
        // lies in [-pi/2, pi/2], so r (cos phi, sin phi) = (rho, zz), and
        // r^8 (cos 8 phi, sin 8 phi) = (rho + i zz)^8: its first square is
        // (rho^2 - zz^2, 2 rho zz) with rho^2 taken as zx^2 + zy^2, then two more squarings.
        // So z' = (r^8 cos 8 phi (cos 8 theta, sin 8 theta), r^8 sin 8 phi) + p without 1 / r;
        // r^7 for dr is r^6 sqrt(r^2).  Fused multiply-adds where a product meets a sum.
        // The same function as the written form, with its own roundings (its error against
        // the fp64 estimator is bounded as for the written form).  The bailout test
        // compares r^2 with the threshold whose correctly rounded square root first exceeds
        // the bailout, so every point leaves the loop where the written form does, with
        // r = |z| (hardware square root); a point that never bails out leaves with
        // r = |z|^8 of the last step, as the written form's r = pow(r, 8).  rho = 0 gives
        // NaN, as atan(0 / 0) does.
        float r2_out = -1.0f;  // r^2 where the point bailed out (the root is taken once, below)
        float r2_last = 0.0f;  // r^2 of the last step taken
        for (uint32_t i = 0; i < iters; ++i) {
            const float xy2 = zx * zx + zy * zy;
            const float r2 = xy2 + zz * zz;
            if (r2 > t2) {  // as below: z no longer changes past the bailout
                r2_out = r2;
                break;
            }
            r2_last = r2;
            const float irho = rsqrt_full(xy2);
            const float rho = xy2 * irho;
            float ct = zx * irho, st = zy * irho;
            float cp = __builtin_fmaf(-zz, zz, xy2), sp = (rho + rho) * zz;
#pragma unroll
            for (int k = 0; k < 3; ++k) {  // (c + i s)^2 = (c^2 - s^2) + i 2cs, one rounding for the real part
                const float c2 = __builtin_fmaf(ct, ct, -(st * st)), s2 = (ct + ct) * st;
                ct = c2;
                st = s2;
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const float p2 = __builtin_fmaf(cp, cp, -(sp * sp)), q2 = (cp + cp) * sp;
                cp = p2;
                sp = q2;
            }
            const float r7 = ((r2 * r2) * r2) * __builtin_amdgcn_sqrtf(r2);
            dr = __builtin_fmaf(r7 * dr, 8.0f, 1.0f);
            zx = __builtin_fmaf(cp, ct, px);
            zy = __builtin_fmaf(cp, st, py);
            zz = sp + pz;
        }
        if (r2_out >= 0.0f) {
            r = __builtin_amdgcn_sqrtf(r2_out);
        } else {
            const float r4 = r2_last * r2_last;
            r = r4 * r4;
        }
        // 0.5 ln(r) r / dr from the hardware log2 (times ln 2; a denormal r scaled by 2^24
        // first) and reciprocal, each within about an ulp: dr >= 1, and r = 0 gives NaN as
        // log(0) * 0 does
        const bool tiny = r < 0x1p-126f;
        const float lr = (__builtin_amdgcn_logf(tiny ? r * 0x1p24f : r) - (tiny ? 24.0f : 0.0f)) * 0.693147182f;
        return ((0.5f * lr) * r) * __builtin_amdgcn_rcpf(dr) - offset;
    }
    for (uint32_t i = 0; i < iters; ++i) {
        r = len3(zx, zy, zz);
        // past the bailout z no longer changes, so every later step would recompute this
        // r and skip the rest: leaving the loop gives the same r and dr
        if (r > bailout) break;
        float theta = atanf(zy / zx);
        float phi = asinf(zz / r);
        dr = (powf(r, power - 1.0f) * dr) * power + 1.0f;
        r = powf(r, power);
        theta = theta * power;
        phi = phi * power;
        // sincosf gives sinf's and cosf's bits (checked over every float input) and reduces
        // the argument once
        float st, ct, sp, cp;
        sincosf(theta, &st, &ct);
        sincosf(phi, &sp, &cp);
        zx = r * (ct * cp) + px;
        zy = r * (st * cp) + py;
        zz = r * sp + pz;
    }
    return ((0.5f * logf(r)) * r) / dr - offset;
}

// One primitive of the tape at p.
__device__ __forceinline__ float eval_primitive(uint32_t code, uint32_t iarg, const float* f,
                                                float px, float py, float pz)
{
    switch (code) {
    case S2M_OP_SPHERE: return sd_sphere(px - f[0], py - f[1], pz - f[2], f[3]);
    case S2M_OP_BOX: return sd_box(px - f[0], py - f[1], pz - f[2], f[3], f[4], f[5]);
    case S2M_OP_TORUS: return sd_torus(px - f[0], py - f[1], pz - f[2], f[3], f[4]);
    case S2M_OP_CYLINDER: return sd_cylinder(px - f[0], py - f[1], pz - f[2], f[3], f[4]);
    case S2M_OP_CAPSULE: return sd_capsule(px, py, pz, f[0], f[1], f[2], f[3], f[4], f[5], f[6]);
    default: return sd_mandelbulb(px, py, pz, f[0], f[1], f[2], iarg, f[3]);  // f[3]: k_tape_prepare
    }
}

__device__ __forceinline__ float apply_op(uint32_t code, float d1, float d2, float k)
{
    switch (code) {
    case S2M_OP_UNION: return fminf(d1, d2);
    case S2M_OP_INTERSECTION: return fmaxf(d1, d2);
    case S2M_OP_SUBTRACTION: return fmaxf(-d1, d2);
    case S2M_OP_SMOOTH_UNION: return op_smooth_union(d1, d2, k);
    case S2M_OP_SMOOTH_INTERSECTION: return op_smooth_intersection(d1, d2, k);
    default: return op_smooth_subtraction(d1, d2, k);
    }
}

// Tape interpreter.  The tape index is wave-uniform, so opcode and operands are read
// through the scalar cache and the dispatch is a scalar branch.  The top of the operand
// stack lives in `acc`; deeper entries in a small register array indexed by the
// (uniform) stack depth.
__device__ __forceinline__ float tape_eval(const s2m_op* __restrict__ ops, uint32_t n, float px,
                                           float py, float pz)
{
    float st[S2M_MAX_STACK];
    float acc = 0.0f;
    uint32_t sp = 0;  // entries below acc
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t code = ops[i].code;
        const float* f = ops[i].f;
        if (code < S2M_OP_UNION) {
            const float d = eval_primitive(code, ops[i].iarg, f, px, py, pz);
            if (i != 0) st[sp++] = acc;
            acc = d;
        } else {
            const float d1 = st[--sp];
            acc = apply_op(code, d1, acc, f[0]);
        }
    }
    return acc;
}

// The same interpreter over a pruned tape: `list` holds the indices of the ops to run,
// in order (a valid postfix program whose value equals the full tape's on its box).
__device__ __forceinline__ float tape_eval_list(const s2m_op* __restrict__ ops,
                                                const uint16_t* __restrict__ list, uint32_t n,
                                                float px, float py, float pz)
{
    float st[S2M_MAX_STACK];
    float acc = 0.0f;
    uint32_t sp = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const s2m_op* o = ops + list[k];
        const uint32_t code = o->code;
        if (code < S2M_OP_UNION) {
            const float d = eval_primitive(code, o->iarg, o->f, px, py, pz);
            if (k != 0) st[sp++] = acc;
            acc = d;
        } else {
            const float d1 = st[--sp];
            acc = apply_op(code, d1, acc, o->f[0]);
        }
    }
    return acc;
}

}  // namespace s2m
