// internal.h - declarations shared by the host runtime (api.cpp) and the kernels
// (kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sdf2mesh.h"

namespace s2m {

constexpr uint32_t kBlock = 256;             // threads per block of the cell kernels
constexpr uint32_t kRounds = 4;              // cells per thread (strided by kBlock)
constexpr uint32_t kCellsPerBlock = kBlock * kRounds;
constexpr uint32_t kScanThreads = 1024;      // single-block scan of block counts
constexpr uint32_t kMaxScanBlocks = kScanThreads * 64;
constexpr uint32_t kNone = 0xffffffffu;      // "no vertex" (u32::MAX, mesh.rs:330)

// Unsigned division by a runtime constant d >= 1 (round-up multiplier method):
// n / d == (umulhi(n, m) + n) >> l for every 32-bit n, evaluated in 64 bits.
struct FastDiv {
    uint32_t d, m, l;
};
FastDiv make_fastdiv(uint32_t d);

// Device counters of one job (u64 each).
enum Counter : uint32_t {
    kVtxTotal = 0,   // vertices emitted so far (all slabs)
    kVtxPrev = 1,    // vertex total before the current slab
    kTriTotal = 2,   // triangles emitted so far
    kTriPrev = 3,
    kInvalid = 4,    // quads skipped for a missing vertex
    kNumCounters = 8
};

// Everything the cell kernels need to address one slab of cells [z0, z0+nz).
struct SlabArgs {
    const float* lat;      // corner lattice of the slab, x fastest
    uint32_t pitch;        // floats per lattice row
    uint32_t lat_ny;       // lattice rows per lattice plane
    uint32_t lz0;          // lattice z index of the slab's first lattice plane
    const float* cx;       // per-axis coordinate tables
    const float* cy;
    const float* cz;
    const uint32_t* lox;   // per-axis cell -> coordinate index of the min / max corner
    const uint32_t* hix;
    const uint32_t* loy;
    const uint32_t* hiy;
    const uint32_t* loz;
    const uint32_t* hiz;
    uint32_t rx, ry;       // cells per row, rows per plane
    uint32_t z0;           // first cell plane of the slab
    uint32_t ncells;       // rx * ry * nz
    FastDiv div_rx;
    FastDiv div_rxy;
    uint32_t* vidx;        // (nz + 1) planes of rx*ry; plane 0 holds cell plane z0-1
    uint32_t flags;        // S2M_* flags
};

struct MeshBufs {
    float* vtx;            // 6 floats per vertex
    uint64_t* keys;
    uint8_t* signs;
    uint32_t* tri;         // 3 per triangle
    uint64_t vcap, tcap;   // capacities (vertices, triangles)
    uint64_t* counters;    // Counter
    uint32_t* bcount;      // per-block counts
    uint32_t* boffset;     // per-block exclusive offsets
};

// ---- launchers (kernels.hip) ----
hipError_t launch_eval_lattice(hipStream_t s, const s2m_op* ops, uint32_t nops, const float* cx,
                               const float* cy, const float* cz, uint32_t nx, uint32_t ny,
                               uint32_t lz0, uint32_t nzl, uint32_t pitch, float* lat);
hipError_t launch_cells(hipStream_t s, const SlabArgs& a, const MeshBufs& b, int emit,
                        int write_zero_normals);
hipError_t launch_scan(hipStream_t s, const MeshBufs& b, uint32_t nblocks, uint32_t which);
hipError_t launch_normals(hipStream_t s, const s2m_op* ops, uint32_t nops, float eps,
                          const MeshBufs& b);
hipError_t launch_quads(hipStream_t s, const SlabArgs& a, const MeshBufs& b, int emit);
// Contour corner values given per cell (test hook): cells of the slab read their 8
// corners from `corners` (8 floats per cell, cells in (z,y,x) order from plane z0).
hipError_t launch_cells_from_corners(hipStream_t s, const SlabArgs& a, const float* corners,
                                     const MeshBufs& b, int emit);

}  // namespace s2m
