// device_interval.h - interval bounds of the SDF tape over a box of points, and exact
// per-box tape pruning.
//
// For a box of lattice points B (and anything within `pad` of it), iv_eval bounds the
// fp32 value the tape interpreter computes at any point of B.  Bounds are evaluated in
// fp64 from the real-valued function (every library primitive is a distance, every
// operator is monotone in each operand) and widened by kMargin x the magnitudes
// involved, which covers the rounding of the few fp32 operations of each op (a handful
// of ulps, ~6e-8 relative each).
//
// Pruning keeps the postfix tape's value identical, not approximately equal: an operator
// is replaced by one operand only when the interval bounds force the fp32 result to be
// exactly that operand on the whole box:
//   min(d1,d2), max(d1,d2): bounds of one operand entirely on one side of the other;
//   max(-d1,d2):            d2 >= -d1 everywhere  -> d2;
//   smooth union:  d2 >= d1 + k  -> h = 1, mix(d2,d1,1) - k*1*0 = d1;
//                  d2 <= d1 - k  -> h = 0, result d2;
//   smooth intersection: the mirror image; smooth subtraction: d1 + d2 >= k -> d2.
// (The replaced result can differ from the operand only in the sign of a zero, which no
// comparison `> 0`, subtraction or division downstream can observe.)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sdf2mesh.h"

namespace s2m {

constexpr double kMargin = 64.0 / 16777216.0;  // 64 fp32 ulps, relative
constexpr double kHuge = 1e30;

struct Iv {
    double lo, hi;
    bool known;
};

__device__ __forceinline__ Iv iv_unknown() { return {-kHuge, kHuge, false}; }

__device__ __forceinline__ Iv iv_make(double lo, double hi, double mag)
{
    const double m = kMargin * (fabs(mag) + 1e-30);
    Iv r{lo - m, hi + m, true};
    if (!(r.lo > -kHuge && r.hi < kHuge)) return iv_unknown();  // also catches NaN
    return r;
}

// range of |q| for q in the box [lo, hi] (per component)
__device__ __forceinline__ void abs_range(double lo, double hi, double& near, double& far)
{
    near = lo > 0 ? lo : (hi < 0 ? -hi : 0.0);
    far = fmax(fabs(lo), fabs(hi));
}

// distance range from the origin to the points of an n-D box
__device__ __forceinline__ void norm_range(const double* lo, const double* hi, int n, double& dmin,
                                           double& dmax)
{
    double s0 = 0, s1 = 0;
    for (int a = 0; a < n; ++a) {
        double nr, fr;
        abs_range(lo[a], hi[a], nr, fr);
        s0 += nr * nr;
        s1 += fr * fr;
    }
    dmin = sqrt(s0);
    dmax = sqrt(s1);
}

// Real-valued smooth operators (sdf_op.wgsl:7-23) evaluated in fp64.
__device__ __forceinline__ double r_clamp01(double x) { return fmin(fmax(x, 0.0), 1.0); }
__device__ __forceinline__ double r_su(double d1, double d2, double k)
{
    const double h = r_clamp01(0.5 + 0.5 * (d2 - d1) / k);
    return d2 * (1 - h) + d1 * h - k * h * (1 - h);
}
__device__ __forceinline__ double r_si(double d1, double d2, double k)
{
    const double h = r_clamp01(0.5 - 0.5 * (d2 - d1) / k);
    return d2 * (1 - h) + d1 * h + k * h * (1 - h);
}
__device__ __forceinline__ double r_ss(double d1, double d2, double k)
{
    const double h = r_clamp01(0.5 - 0.5 * (d2 + d1) / k);
    return d2 * (1 - h) - d1 * h + k * h * (1 - h);
}

// Bounds of one primitive over the box [bl, bh].
__device__ __forceinline__ Iv iv_primitive(uint32_t code, const float* f, const double* bl,
                                           const double* bh)
{
    double lo[3], hi[3];
    if (code == S2M_OP_CAPSULE || code == S2M_OP_MANDELBULB) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = bl[a];
            hi[a] = bh[a];
        }
    } else {  // translated primitives: q = p - c
        for (int a = 0; a < 3; ++a) {
            lo[a] = bl[a] - (double)f[a];
            hi[a] = bh[a] - (double)f[a];
        }
    }
    switch (code) {
    case S2M_OP_SPHERE: {  // length(q) - s
        double dmin, dmax;
        norm_range(lo, hi, 3, dmin, dmax);
        const double s = f[3];
        return iv_make(dmin - s, dmax - s, dmax + fabs(s));
    }
    case S2M_OP_BOX: {  // monotone in each q_a = |p_a - c_a| - b_a/2
        double ql[3], qh[3], mag = 0;
        for (int a = 0; a < 3; ++a) {
            double nr, fr;
            abs_range(lo[a], hi[a], nr, fr);
            const double hb = 0.5 * (double)f[3 + a];
            ql[a] = nr - hb;
            qh[a] = fr - hb;
            mag += fr + fabs(hb);
        }
        auto g = [](const double* q) {
            const double x = fmax(q[0], 0.0), y = fmax(q[1], 0.0), z = fmax(q[2], 0.0);
            return sqrt(x * x + y * y + z * z) + fmin(fmax(q[0], fmax(q[1], q[2])), 0.0);
        };
        return iv_make(g(ql), g(qh), mag);
    }
    case S2M_OP_TORUS: {  // length(vec2(length(q.xz) - R, q.y)) - r
        const double xl[2] = {lo[0], lo[2]}, xh[2] = {hi[0], hi[2]};
        double rmin, rmax;
        norm_range(xl, xh, 2, rmin, rmax);
        const double ql[2] = {rmin - (double)f[3], lo[1]}, qh[2] = {rmax - (double)f[3], hi[1]};
        double dmin, dmax;
        norm_range(ql, qh, 2, dmin, dmax);
        const double r = f[4];
        return iv_make(dmin - r, dmax - r, rmax + fabs((double)f[3]) + fabs(lo[1]) + fabs(hi[1]) +
                                               fabs(r));
    }
    case S2M_OP_CYLINDER: {  // monotone in dx = length(q.xz) - r, dy = |q.y| - h
        const double xl[2] = {lo[0], lo[2]}, xh[2] = {hi[0], hi[2]};
        double rmin, rmax, ny, fy;
        norm_range(xl, xh, 2, rmin, rmax);
        abs_range(lo[1], hi[1], ny, fy);
        const double h = f[3], r = f[4];
        auto g = [](double dx, double dy) {
            const double px = fmax(dx, 0.0), py = fmax(dy, 0.0);
            return fmin(fmax(dx, dy), 0.0) + sqrt(px * px + py * py);
        };
        return iv_make(g(rmin - r, ny - h), g(rmax - r, fy - h), rmax + fy + fabs(h) + fabs(r));
    }
    case S2M_OP_CAPSULE: {  // distance to segment ab, minus r
        const double a[3] = {f[0], f[1], f[2]}, b[3] = {f[3], f[4], f[5]};
        double half = 0, ml[3], mh[3], al[3], ah[3], bl2[3], bh2[3];
        for (int i = 0; i < 3; ++i) {
            const double m = 0.5 * (a[i] + b[i]);
            half += (b[i] - a[i]) * (b[i] - a[i]);
            ml[i] = lo[i] - m;
            mh[i] = hi[i] - m;
            al[i] = lo[i] - a[i];
            ah[i] = hi[i] - a[i];
            bl2[i] = lo[i] - b[i];
            bh2[i] = hi[i] - b[i];
        }
        half = 0.5 * sqrt(half);
        double mmin, mmax, amin, amax, bmin, bmax;
        norm_range(ml, mh, 3, mmin, mmax);
        norm_range(al, ah, 3, amin, amax);
        norm_range(bl2, bh2, 3, bmin, bmax);
        const double r = f[6];
        return iv_make(fmax(mmin - half, 0.0) - r, fmin(amax, bmax) - r, mmax + 2 * half + fabs(r));
    }
    default: return iv_unknown();  // Mandelbulb: iterated, may produce NaN
    }
}

enum Keep : int { kKeepBoth = 0, kKeepA = 1, kKeepB = 2 };

// Interval of op(A, B) and whether one operand can replace it exactly on the box.
__device__ __forceinline__ Iv iv_op(uint32_t code, Iv A, Iv B, float kf, int& keep)
{
    keep = kKeepBoth;
    if (!A.known || !B.known) return iv_unknown();
    const double k = kf;
    const double tiny = 1e-12 * (fabs(A.lo) + fabs(A.hi) + fabs(B.lo) + fabs(B.hi) + fabs(k));
    switch (code) {
    case S2M_OP_UNION:
        if (B.lo >= A.hi + tiny) { keep = kKeepA; return A; }
        if (B.hi + tiny <= A.lo) { keep = kKeepB; return B; }
        return {fmin(A.lo, B.lo), fmin(A.hi, B.hi), true};
    case S2M_OP_INTERSECTION:
        if (B.hi + tiny <= A.lo) { keep = kKeepA; return A; }
        if (A.hi + tiny <= B.lo) { keep = kKeepB; return B; }
        return {fmax(A.lo, B.lo), fmax(A.hi, B.hi), true};
    case S2M_OP_SUBTRACTION:  // max(-A, B)
        if (B.lo >= -A.lo + tiny) { keep = kKeepB; return B; }
        return {fmax(-A.hi, B.lo), fmax(-A.lo, B.hi), true};
    default: break;
    }
    if (!(k > 0)) return iv_unknown();  // degenerate smoothing radius: no pruning
    const double mag = fmax(fabs(A.lo), fabs(A.hi)) + fmax(fabs(B.lo), fabs(B.hi)) + k;
    switch (code) {
    case S2M_OP_SMOOTH_UNION:
        if (B.lo >= A.hi + k + tiny) { keep = kKeepA; return A; }
        if (B.hi + k + tiny <= A.lo) { keep = kKeepB; return B; }
        return iv_make(r_su(A.lo, B.lo, k), r_su(A.hi, B.hi, k), mag);
    case S2M_OP_SMOOTH_INTERSECTION:
        if (B.lo >= A.hi + k + tiny) { keep = kKeepB; return B; }
        if (B.hi + k + tiny <= A.lo) { keep = kKeepA; return A; }
        return iv_make(r_si(A.lo, B.lo, k), r_si(A.hi, B.hi, k), mag);
    default:  // smooth subtraction: decreasing in A, increasing in B
        if (A.lo + B.lo >= k + tiny) { keep = kKeepB; return B; }
        return iv_make(r_ss(A.hi, B.lo, k), r_ss(A.lo, B.hi, k), mag);
    }
}

}  // namespace s2m
