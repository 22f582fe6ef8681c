// rtc_prelude.h - vector types and WGSL / GLSL builtins for SDF programs compiled at run
// time by hiprtc (the front end's HIP code generator emits calls into this header).
//
// Every operation is written per component, in the order WGSL defines it, with the same
// fp32 expressions as the library (device_sdf.h): a compiled program therefore computes
// the same values as the tape interpreter for a program both can run.  Scalar operands of
// vector operations are broadcast by the code generator (splat), so the overloads below
// take operands of one shape.
#pragma once

namespace s2m::rtc {

template <class T, int N>
struct V {
    T v[N];
};
using vf2 = V<float, 2>;
using vf3 = V<float, 3>;
using vf4 = V<float, 4>;

template <class T, int N>
__device__ __forceinline__ V<T, N> splat(T x)
{
    V<T, N> r;
    for (int k = 0; k < N; ++k) r.v[k] = x;
    return r;
}

// swizzle / component selection: sw<1, 0>(a) = a.yx
template <int... K, class T, int N>
__device__ __forceinline__ V<T, (int)sizeof...(K)> sw(const V<T, N>& a)
{
    return V<T, (int)sizeof...(K)>{{a.v[K]...}};
}

// component-wise operators (WGSL vector arithmetic)
#define S2M_VBIN(OP)                                                                     \
    template <class T, int N>                                                            \
    __device__ __forceinline__ V<T, N> operator OP(const V<T, N>& a, const V<T, N>& b)   \
    {                                                                                    \
        V<T, N> r;                                                                       \
        for (int k = 0; k < N; ++k) r.v[k] = a.v[k] OP b.v[k];                           \
        return r;                                                                        \
    }
S2M_VBIN(+)
S2M_VBIN(-)
S2M_VBIN(*)
S2M_VBIN(/)
#undef S2M_VBIN

template <class T, int N>
__device__ __forceinline__ V<T, N> operator-(const V<T, N>& a)
{
    V<T, N> r;
    for (int k = 0; k < N; ++k) r.v[k] = -a.v[k];
    return r;
}

// Builtins on vectors apply the scalar builtin per component.
#define S2M_MAP1(F)                                                        \
    template <class T, int N>                                              \
    __device__ __forceinline__ V<T, N> F(const V<T, N>& a)                 \
    {                                                                      \
        V<T, N> r;                                                         \
        for (int k = 0; k < N; ++k) r.v[k] = F(a.v[k]);                    \
        return r;                                                          \
    }
#define S2M_MAP2(F)                                                               \
    template <class T, int N>                                                     \
    __device__ __forceinline__ V<T, N> F(const V<T, N>& a, const V<T, N>& b)      \
    {                                                                             \
        V<T, N> r;                                                                \
        for (int k = 0; k < N; ++k) r.v[k] = F(a.v[k], b.v[k]);                   \
        return r;                                                                 \
    }
#define S2M_MAP3(F)                                                                          \
    template <class T, int N>                                                                \
    __device__ __forceinline__ V<T, N> F(const V<T, N>& a, const V<T, N>& b, const V<T, N>& c) \
    {                                                                                        \
        V<T, N> r;                                                                           \
        for (int k = 0; k < N; ++k) r.v[k] = F(a.v[k], b.v[k], c.v[k]);                      \
        return r;                                                                            \
    }

// integer division and remainder with WGSL's defined results for a zero divisor and
// for INT_MIN / -1 (the dividend, resp. 0)
__device__ __forceinline__ int b_div(int a, int b)
{
    return (b == 0 || (a == (-2147483647 - 1) && b == -1)) ? a : a / b;
}
__device__ __forceinline__ unsigned b_div(unsigned a, unsigned b) { return b == 0 ? a : a / b; }
__device__ __forceinline__ int b_rem(int a, int b)
{
    return (b == 0 || (a == (-2147483647 - 1) && b == -1)) ? 0 : a % b;
}
__device__ __forceinline__ unsigned b_rem(unsigned a, unsigned b) { return b == 0 ? 0u : a % b; }
// WGSL float remainder: e1 - e2 * trunc(e1 / e2)
__device__ __forceinline__ float b_rem(float a, float b) { return a - b * truncf(a / b); }
S2M_MAP2(b_div)
S2M_MAP2(b_rem)

__device__ __forceinline__ float b_abs(float x) { return fabsf(x); }
__device__ __forceinline__ int b_abs(int x) { return x < 0 ? -x : x; }
__device__ __forceinline__ unsigned b_abs(unsigned x) { return x; }
__device__ __forceinline__ float b_min(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ int b_min(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ unsigned b_min(unsigned a, unsigned b) { return a < b ? a : b; }
__device__ __forceinline__ float b_max(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ int b_max(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned b_max(unsigned a, unsigned b) { return a > b ? a : b; }
// clamp(e, lo, hi) = min(max(e, lo), hi)
__device__ __forceinline__ float b_clamp(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }
__device__ __forceinline__ int b_clamp(int e, int lo, int hi) { return b_min(b_max(e, lo), hi); }
__device__ __forceinline__ unsigned b_clamp(unsigned e, unsigned lo, unsigned hi)
{
    return b_min(b_max(e, lo), hi);
}
// mix(a, b, t) = a * (1 - t) + b * t
__device__ __forceinline__ float b_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
// step(edge, x) = x >= edge ? 1 : 0
__device__ __forceinline__ float b_step(float edge, float x) { return x >= edge ? 1.0f : 0.0f; }
// smoothstep: t = clamp((x - e0) / (e1 - e0), 0, 1); t * t * (3 - 2 * t)
__device__ __forceinline__ float b_smoothstep(float e0, float e1, float x)
{
    const float t = b_clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float b_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ int b_sign(int x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
__device__ __forceinline__ float b_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ float b_inverse_sqrt(float x) { return 1.0f / sqrtf(x); }
__device__ __forceinline__ float b_sin(float x) { return sinf(x); }
__device__ __forceinline__ float b_cos(float x) { return cosf(x); }
__device__ __forceinline__ float b_tan(float x) { return tanf(x); }
__device__ __forceinline__ float b_asin(float x) { return asinf(x); }
__device__ __forceinline__ float b_acos(float x) { return acosf(x); }
__device__ __forceinline__ float b_atan(float x) { return atanf(x); }
__device__ __forceinline__ float b_atan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ float b_sinh(float x) { return sinhf(x); }
__device__ __forceinline__ float b_cosh(float x) { return coshf(x); }
__device__ __forceinline__ float b_tanh(float x) { return tanhf(x); }
__device__ __forceinline__ float b_exp(float x) { return expf(x); }
__device__ __forceinline__ float b_exp2(float x) { return exp2f(x); }
__device__ __forceinline__ float b_log(float x) { return logf(x); }
__device__ __forceinline__ float b_log2(float x) { return log2f(x); }
__device__ __forceinline__ float b_pow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ float b_floor(float x) { return floorf(x); }
__device__ __forceinline__ float b_ceil(float x) { return ceilf(x); }
__device__ __forceinline__ float b_trunc(float x) { return truncf(x); }
__device__ __forceinline__ float b_round(float x) { return rintf(x); }  // ties to even
__device__ __forceinline__ float b_fract(float x) { return x - floorf(x); }
// GLSL mod(x, y) = x - y * floor(x / y)
__device__ __forceinline__ float b_mod(float x, float y) { return x - y * floorf(x / y); }
__device__ __forceinline__ float b_radians(float x) { return x * 0.017453292519943295f; }
__device__ __forceinline__ float b_degrees(float x) { return x * 57.29577951308232f; }
S2M_MAP1(b_abs)
S2M_MAP2(b_min)
S2M_MAP2(b_max)
S2M_MAP3(b_clamp)
S2M_MAP3(b_mix)
S2M_MAP2(b_step)
S2M_MAP3(b_smoothstep)
S2M_MAP1(b_sign)
S2M_MAP1(b_sqrt)
S2M_MAP1(b_inverse_sqrt)
S2M_MAP1(b_sin)
S2M_MAP1(b_cos)
S2M_MAP1(b_tan)
S2M_MAP1(b_asin)
S2M_MAP1(b_acos)
S2M_MAP1(b_atan)
S2M_MAP2(b_atan2)
S2M_MAP1(b_sinh)
S2M_MAP1(b_cosh)
S2M_MAP1(b_tanh)
S2M_MAP1(b_exp)
S2M_MAP1(b_exp2)
S2M_MAP1(b_log)
S2M_MAP1(b_log2)
S2M_MAP2(b_pow)
S2M_MAP1(b_floor)
S2M_MAP1(b_ceil)
S2M_MAP1(b_trunc)
S2M_MAP1(b_round)
S2M_MAP1(b_fract)
S2M_MAP2(b_mod)
S2M_MAP1(b_radians)
S2M_MAP1(b_degrees)
#undef S2M_MAP1
#undef S2M_MAP2
#undef S2M_MAP3

// dot / length summed left to right: ((x*x' + y*y') + z*z') + w*w'
template <class T, int N>
__device__ __forceinline__ T b_dot(const V<T, N>& a, const V<T, N>& b)
{
    T s = a.v[0] * b.v[0];
    for (int k = 1; k < N; ++k) s = s + a.v[k] * b.v[k];
    return s;
}
__device__ __forceinline__ float b_length(float x) { return fabsf(x); }
template <int N>
__device__ __forceinline__ float b_length(const V<float, N>& a)
{
    return sqrtf(b_dot(a, a));
}
__device__ __forceinline__ float b_distance(float a, float b) { return fabsf(a - b); }
template <int N>
__device__ __forceinline__ float b_distance(const V<float, N>& a, const V<float, N>& b)
{
    return b_length(a - b);
}
// normalize(v) = v / length(v)
template <int N>
__device__ __forceinline__ V<float, N> b_normalize(const V<float, N>& a)
{
    return a / splat<float, N>(b_length(a));
}
__device__ __forceinline__ float b_normalize(float x) { return x / fabsf(x); }
__device__ __forceinline__ vf3 b_cross(const vf3& a, const vf3& b)
{
    return vf3{{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
                a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}

// the library functions (sdf3d_primitives.wgsl, sdf_op.wgsl) on vector arguments
__device__ __forceinline__ float lib_sphere(const vf3& p, float s) { return s2m::sd_sphere(p.v[0], p.v[1], p.v[2], s); }
__device__ __forceinline__ float lib_box(const vf3& p, const vf3& b)
{
    return s2m::sd_box(p.v[0], p.v[1], p.v[2], b.v[0], b.v[1], b.v[2]);
}
__device__ __forceinline__ float lib_torus(const vf3& p, const vf2& t)
{
    return s2m::sd_torus(p.v[0], p.v[1], p.v[2], t.v[0], t.v[1]);
}
__device__ __forceinline__ float lib_cylinder(const vf3& p, float h, float r)
{
    return s2m::sd_cylinder(p.v[0], p.v[1], p.v[2], h, r);
}
__device__ __forceinline__ float lib_capsule(const vf3& p, const vf3& a, const vf3& b, float r)
{
    return s2m::sd_capsule(p.v[0], p.v[1], p.v[2], a.v[0], a.v[1], a.v[2], b.v[0], b.v[1], b.v[2], r);
}
__device__ __forceinline__ float lib_smooth_union(float a, float b, float k) { return s2m::op_smooth_union(a, b, k); }
__device__ __forceinline__ float lib_smooth_intersection(float a, float b, float k)
{
    return s2m::op_smooth_intersection(a, b, k);
}
__device__ __forceinline__ float lib_smooth_subtraction(float a, float b, float k)
{
    return s2m::op_smooth_subtraction(a, b, k);
}

}  // namespace s2m::rtc
