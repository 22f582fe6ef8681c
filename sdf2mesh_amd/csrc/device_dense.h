// device_dense.h - per-thread bodies of the dense lattice-evaluation and vertex-normal
// kernels, templated on the SDF.  The tape interpreter's kernels (kernels.hip) and the
// kernels of programs compiled at run time by hiprtc instantiate the same bodies, so both
// paths run identical per-point code around their SDF.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace s2m {

// Indices of the vertex counters in the runtime's per-job counter block.
constexpr uint32_t kDenseVtxTotal = 0;  // vertices emitted so far
constexpr uint32_t kDenseVtxPrev = 1;   // vertex total before the current slab

// One lattice point per thread: 256-thread blocks of 16 x 4 x 4 points, each wave a
// 4 x 4 x 4 block (lanes close together, so data-dependent loops diverge less than along
// a row), grid (ceil(nx / 16), ceil(ny / 4), ceil(planes / 4)).
template <class F>
__device__ __forceinline__ void eval_lattice_point(F sdf, const float* __restrict__ cx,
                                                   const float* __restrict__ cy,
                                                   const float* __restrict__ cz, uint32_t nx,
                                                   uint32_t ny, uint32_t lz0, uint32_t nzl, uint32_t pitch,
                                                   float* __restrict__ lat)
{
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t ix = blockIdx.x * 16u + w * 4u + (l & 3u);
    const uint32_t iy = blockIdx.y * 4u + ((l >> 2) & 3u);
    const uint32_t iz = blockIdx.z * 4u + (l >> 4);
    if (ix >= nx || iy >= ny || iz >= nzl) return;
    lat[((size_t)iz * ny + iy) * pitch + ix] = sdf(cx[ix], cy[iy], cz[lz0 + iz]);
}

// A NaN result is stored as the canonical quiet NaN (0x7fc00000): which NaN operand an
// instruction propagates, and so the sign and payload of a NaN result, differ between
// kernels that compute the same value (the dense and the tile-sparse path), so vertex
// positions and normals that are NaN are written with one bit pattern on every path.
__device__ __forceinline__ float canonical_nan(float v) { return v != v ? __builtin_nanf("") : v; }

// normalize(g) written to o[3..5] (dualcontour.wgsl:171): g / length(g).
__device__ __forceinline__ void write_unit(float* o, float gx, float gy, float gz)
{
    const float len = sqrtf((gx * gx + gy * gy) + gz * gz);
    o[3] = canonical_nan(gx / len);
    o[4] = canonical_nan(gy / len);
    o[5] = canonical_nan(gz / len);
}

// sdf3d_normal (sdf3d_normal.wgsl:4-9): taps v1=(1,-1,-1), v2=(-1,-1,1), v3=(-1,1,-1),
// v4=(1,1,1) at p + v*eps, summed v1*d1 + v2*d2 + v3*d3 + v4*d4 left to right, then
// normalized.
template <class F>
__device__ __forceinline__ void library_normal(F sdf, float px, float py, float pz, float eps,
                                               float* o)
{
    const float d1 = sdf(px + eps, py + (-eps), pz + (-eps));
    const float d2 = sdf(px + (-eps), py + (-eps), pz + eps);
    const float d3 = sdf(px + (-eps), py + eps, pz + (-eps));
    const float d4 = sdf(px + eps, py + eps, pz + eps);
    const float gx = ((d1 + (-d2)) + (-d3)) + d4;
    const float gy = (((-d1) + (-d2)) + d3) + d4;
    const float gz = (((-d1) + d2) + (-d3)) + d4;
    write_unit(o, gx, gy, gz);
}

// Normals of the vertices [counters[prev], min(counters[total], vcap)) in a grid-stride
// loop, so the range never has to travel to the host.  vtx holds 6 floats per vertex
// (position, normal).  N(px, py, pz, eps, o) writes o[3..5].
template <class N>
__device__ __forceinline__ void normals_range(N normal, float eps, float* __restrict__ vtx,
                                              const uint64_t* __restrict__ counters, uint64_t vcap)
{
    const uint64_t beg = counters[kDenseVtxPrev];
    const uint64_t end = counters[kDenseVtxTotal] < vcap ? counters[kDenseVtxTotal] : vcap;
    for (uint64_t v = beg + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < end;
         v += (uint64_t)gridDim.x * blockDim.x) {
        float* o = vtx + 6 * v;
        normal(o[0], o[1], o[2], eps, o);
    }
}

}  // namespace s2m
