// device_staged.h - a tile's pruned program, staged in LDS, run with wave-uniform dispatch.
//
// Every lane of a block evaluates the same pruned program (one tile per block), so the
// opcode is read once per op for the whole wave (readfirstlane) and the dispatch is a
// scalar branch; operands are LDS broadcasts.  Programs whose postfix form is a left fold
// of library primitives (P0 P1 op1 P2 op2 ...: CSG built by accumulation, and every
// pruned sub-program of one) run without an operand stack, several points per lane
// sharing each op's decode; other programs run on a uniform-depth stack.  Values are
// bit-identical to tape_eval (device_sdf.h): the same functions in the same order.
#pragma once
#include "device_sdf.h"

namespace s2m {

constexpr uint32_t kStageOps = 128;            // longest program staged in LDS
constexpr uint32_t kLinearFlag = 0x80000000u;  // SpPass::lens: the program is a left fold

template <uint32_t CAP>
struct StagedTapeN {
    uint32_t code[CAP];
    uint32_t iarg[CAP];
    float f[CAP][8];  // primitives use f[0..6], operators f[0]

    __device__ __forceinline__ uint32_t code_at(uint32_t k) const { return __builtin_amdgcn_readfirstlane(code[k]); }
    __device__ __forceinline__ uint32_t iarg_at(uint32_t k) const { return __builtin_amdgcn_readfirstlane(iarg[k]); }
    __device__ __forceinline__ const float* f_at(uint32_t k) const { return f[k]; }
};
using StagedTape = StagedTapeN<kStageOps>;

// A left fold (P0 P1 op1 P2 op2 ...) staged in 44 % of StagedTapeN's space: entry j holds
// primitive 2j - 1 (primitive 0 for j = 0) in f[j][0..6] and operator 2j in f[j][7]
// (primitives read at most 7 operands, operators one).  Left folds hold no Mandelbulb, so
// there is no integer operand.
template <uint32_t CAP>
struct StagedFoldN {
    uint8_t code[CAP];
    float f[CAP / 2 + 1][8];

    __device__ __forceinline__ uint32_t code_at(uint32_t k) const { return __builtin_amdgcn_readfirstlane(code[k]); }
    __device__ __forceinline__ uint32_t iarg_at(uint32_t) const { return 0u; }
    __device__ __forceinline__ const float* f_at(uint32_t k) const
    {
        return (k & 1u) || k == 0 ? f[(k + 1u) >> 1] : &f[k >> 1][7];
    }
};
using StagedFold = StagedFoldN<kStageOps>;

// A left fold in StagedFoldN's layout read in place from global memory: its address and
// the op index are wave-uniform, so the reads are scalar loads (through the scalar data
// cache, into SGPRs) and no LDS copy is made.  The memory is written by an earlier kernel.
struct ScalarFold {
    using cu32 = const __attribute__((address_space(4))) uint32_t;
    using cf32 = const __attribute__((address_space(4))) float;
    cu32* code4;  // the code bytes, four per word
    cf32* f;      // the rows of 8 floats

    __device__ static ScalarFold at(const void* staged_fold)
    {
        const char* base = static_cast<const char*>(staged_fold);
        return ScalarFold{(cu32*)base, (cf32*)(base + offsetof(StagedFold, f))};
    }
    __device__ __forceinline__ uint32_t code_at(uint32_t k) const { return (code4[k >> 2] >> ((k & 3u) * 8u)) & 0xffu; }
    __device__ __forceinline__ uint32_t iarg_at(uint32_t) const { return 0u; }
    __device__ __forceinline__ cf32* f_at(uint32_t k) const
    {
        return (k & 1u) || k == 0 ? f + ((k + 1u) >> 1) * 8u : f + (k >> 1) * 8u + 7u;
    }
};

// Copy ops[L[k]], k < n <= CAP, into T: worker t of nt copies entries t, t + nt, ... (by
// default the threads of the block; the caller synchronises before and after).
template <uint32_t CAP>
__device__ __forceinline__ void stage_tape(StagedTapeN<CAP>& T, const s2m_op* __restrict__ ops,
                                           const uint16_t* __restrict__ L, uint32_t n,
                                           uint32_t t = threadIdx.x, uint32_t nt = blockDim.x)
{
    for (uint32_t k = t; k < n; k += nt) {
        const s2m_op& o = ops[L[k]];
        T.code[k] = o.code;
        T.iarg[k] = o.iarg;
#pragma unroll
        for (int j = 0; j < 8; ++j) T.f[k][j] = o.f[j];
    }
}

// The same for a left fold (n <= CAP, n odd, primitives at 0 and the odd positions).
template <uint32_t CAP>
__device__ __forceinline__ void stage_tape(StagedFoldN<CAP>& T, const s2m_op* __restrict__ ops,
                                           const uint16_t* __restrict__ L, uint32_t n,
                                           uint32_t t = threadIdx.x, uint32_t nt = blockDim.x)
{
    for (uint32_t k = t; k < n; k += nt) {
        const s2m_op& o = ops[L[k]];
        T.code[k] = (uint8_t)o.code;
        if ((k & 1u) || k == 0) {
#pragma unroll
            for (int j = 0; j < 7; ++j) T.f[(k + 1u) >> 1][j] = o.f[j];
        } else {
            T.f[k >> 1][7] = o.f[0];
        }
    }
}

template <class TT>
__device__ __forceinline__ uint32_t uniform_code(const TT& T, uint32_t k)
{
    return T.code_at(k);
}

// Primitive `code` with operands f (iarg: Mandelbulb only) at NP points (code is
// wave-uniform).  MANDEL = false leaves the Mandelbulb out of the code (left folds never
// contain one).
template <int NP, bool MANDEL, class TT>
__device__ __forceinline__ void staged_prim(const TT& T, uint32_t k, uint32_t code, const float (&x)[NP],
                                            const float (&y)[NP], const float (&z)[NP], float (&d)[NP])
{
    const auto* f = T.f_at(k);
    switch (code) {
    case S2M_OP_SPHERE: {
        const float c0 = f[0], c1 = f[1], c2 = f[2], s = f[3];
#pragma unroll
        for (int p = 0; p < NP; ++p) d[p] = sd_sphere(x[p] - c0, y[p] - c1, z[p] - c2, s);
        break;
    }
    case S2M_OP_BOX: {
        const float c0 = f[0], c1 = f[1], c2 = f[2], b0 = f[3], b1 = f[4], b2 = f[5];
#pragma unroll
        for (int p = 0; p < NP; ++p) d[p] = sd_box(x[p] - c0, y[p] - c1, z[p] - c2, b0, b1, b2);
        break;
    }
    case S2M_OP_TORUS: {
        const float c0 = f[0], c1 = f[1], c2 = f[2], t0 = f[3], t1 = f[4];
#pragma unroll
        for (int p = 0; p < NP; ++p) d[p] = sd_torus(x[p] - c0, y[p] - c1, z[p] - c2, t0, t1);
        break;
    }
    case S2M_OP_CYLINDER: {
        const float c0 = f[0], c1 = f[1], c2 = f[2], h = f[3], r = f[4];
#pragma unroll
        for (int p = 0; p < NP; ++p) d[p] = sd_cylinder(x[p] - c0, y[p] - c1, z[p] - c2, h, r);
        break;
    }
    case S2M_OP_CAPSULE: {
        const float a0 = f[0], a1 = f[1], a2 = f[2], b0 = f[3], b1 = f[4], b2 = f[5], r = f[6];
#pragma unroll
        for (int p = 0; p < NP; ++p) d[p] = sd_capsule(x[p], y[p], z[p], a0, a1, a2, b0, b1, b2, r);
        break;
    }
    default:
        if (MANDEL) {
            const float pw = f[0], bo = f[1], off = f[2], t2 = f[3];  // f[3]: k_tape_prepare
            const uint32_t ia = T.iarg_at(k);
#pragma unroll
            for (int p = 0; p < NP; ++p) d[p] = sd_mandelbulb(x[p], y[p], z[p], pw, bo, off, ia, t2);
        }
        break;
    }
}

// acc = op(acc, d) at NP points (code is wave-uniform)
template <int NP>
__device__ __forceinline__ void staged_op(uint32_t code, float k, float (&acc)[NP], const float (&d)[NP])
{
    switch (code) {
    case S2M_OP_UNION:
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] = fminf(acc[p], d[p]);
        break;
    case S2M_OP_INTERSECTION:
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] = fmaxf(acc[p], d[p]);
        break;
    case S2M_OP_SUBTRACTION:
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] = fmaxf(-acc[p], d[p]);
        break;
    case S2M_OP_SMOOTH_UNION:
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] = op_smooth_union(acc[p], d[p], k);
        break;
    case S2M_OP_SMOOTH_INTERSECTION:
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] = op_smooth_intersection(acc[p], d[p], k);
        break;
    default:
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] = op_smooth_subtraction(acc[p], d[p], k);
        break;
    }
}

// A left-fold program (kLinearFlag): acc = P0; acc = op_i(acc, P_i) for i = 1, 2, ...
template <int NP, class TT>
__device__ __forceinline__ void run_linear(const TT& T, uint32_t n, const float (&x)[NP], const float (&y)[NP],
                                           const float (&z)[NP], float (&acc)[NP])
{
    staged_prim<NP, false>(T, 0, uniform_code(T, 0), x, y, z, acc);
    for (uint32_t k = 1; k + 1 < n; k += 2) {
        float d[NP];
        staged_prim<NP, false>(T, k, uniform_code(T, k), x, y, z, d);
        staged_op<NP>(uniform_code(T, k + 1), T.f_at(k + 1)[0], acc, d);
    }
}

// Any valid program: the postfix stack machine of tape_eval with a wave-uniform depth.
template <class TT>
__device__ __forceinline__ float run_general(const TT& T, uint32_t n, float px, float py, float pz)
{
    float st[S2M_MAX_STACK];
    float acc = 0.0f;
    uint32_t sp = 0;
    const float x[1] = {px}, y[1] = {py}, z[1] = {pz};
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t code = uniform_code(T, k);
        if (code < S2M_OP_UNION) {
            float d[1];
            staged_prim<1, true>(T, k, code, x, y, z, d);
            if (k != 0) st[sp++] = acc;
            acc = d[0];
        } else {
            float a[1] = {st[--sp]};
            const float d[1] = {acc};
            staged_op<1>(code, T.f_at(k)[0], a, d);
            acc = a[0];
        }
    }
    return acc;
}

}  // namespace s2m
