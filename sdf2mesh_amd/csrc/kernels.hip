// kernels.hip - HIP kernels of the meshing path for gfx950 (wave64).
//
// Per z-slab of cells the host launches, in stream order:
//   k_eval_lattice     SDF at every corner-lattice point of the slab       (VALU bound)
//   k_count_cols + k_popc   activity bits and active-cell count per block of 1024 cells
//                      (k_cells<0> for per-cell corner dumps)
//   k_scan_sums + k_scan_offsets   exclusive scan of block counts + running vertex total
//   k_cells_map + k_cell_vertices  cell->vertex map, vertex position/signs in order
//                      (k_cells<1> for corner dumps)
//   k_normals          4-tap normal of the slab's new vertices
//   k_quads<0> / scan / k_quads<1>  the same count/scan/emit scheme for triangles
// (with several slabs the evaluation runs one slab ahead on a second stream)
// Cells are numbered in (z, y, x) order inside the slab and every block owns 1024
// consecutive cells, so block offsets from the scan give each vertex and triangle its
// place in the reference's output order (VertexList order, mesh.rs:280-321).
#include <mutex>

#include "device_dense.h"
#include "device_sdf.h"
#include "internal.h"

namespace s2m {

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f)
{
    const uint64_t t = (uint64_t)__umulhi(n, f.m) + n;
    return (uint32_t)(t >> f.l);
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Exclusive prefix of v over the 256 threads of the block (4 waves); *total = sum.
// Contains two __syncthreads: every thread of the block must call it.
__device__ __forceinline__ uint32_t block_exclusive_256(uint32_t v, uint32_t* lds, uint32_t* total)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += t;
    }
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; ++k) {
        const uint32_t s = lds[k];
        woff += (k < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return woff + inc - v;
}

// ---------------------------------------------------------------------------------
// Lattice evaluation: one thread per lattice point.  Each wave covers a 4 x 4 x 4 block
// of points (a block: 16 x 4 x 4), so the lanes of a wave lie close together and a
// program with data-dependent loops (the Mandelbulb's escape iteration) runs fewer
// iterations with idle lanes than along a row of 64 points.  Held to 64 VGPRs (8 waves
// per SIMD; a 12-byte spill): 3 % faster on a 64-primitive CSG program, neutral on the
// Mandelbulb.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_eval_lattice(
    const s2m_op* __restrict__ ops, uint32_t nops,
                                                      const float* __restrict__ cx,
                                                      const float* __restrict__ cy,
                                                      const float* __restrict__ cz, uint32_t nx,
                                                      uint32_t ny, uint32_t lz0, uint32_t nzl,
                                                      uint32_t pitch, float* __restrict__ lat)
{
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t ix = blockIdx.x * 16u + w * 4u + (l & 3u);
    const uint32_t iy = blockIdx.y * 4u + ((l >> 2) & 3u);
    const uint32_t iz = blockIdx.z * 4u + (l >> 4);
    if (ix >= nx || iy >= ny || iz >= nzl) return;
    lat[((size_t)iz * ny + iy) * pitch + ix] = tape_eval(ops, nops, cx[ix], cy[iy], cz[lz0 + iz]);
}

// The same lattice for a tape that is one Mandelbulb op: the estimator called directly
// with its operands (wave-uniform loads), as tape_eval calls it, so the values are the
// interpreter's bit for bit without its per-point dispatch.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_eval_mandelbulb(
    const s2m_op* __restrict__ op, const float* __restrict__ cx, const float* __restrict__ cy,
    const float* __restrict__ cz, uint32_t nx, uint32_t ny, uint32_t lz0, uint32_t nzl, uint32_t pitch,
    float* __restrict__ lat)
{
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t ix = blockIdx.x * 16u + w * 4u + (l & 3u);
    const uint32_t iy = blockIdx.y * 4u + ((l >> 2) & 3u);
    const uint32_t iz = blockIdx.z * 4u + (l >> 4);
    if (ix >= nx || iy >= ny || iz >= nzl) return;
    const float* f = op->f;
    lat[((size_t)iz * ny + iy) * pitch + ix] =
        sd_mandelbulb(cx[ix], cy[iy], cz[lz0 + iz], f[0], f[1], f[2], op->iarg, f[3]);
}

// ---------------------------------------------------------------------------------
// One cell: dualcontour.wgsl:72-131 (crossings, centroid) and :57-69 (sign bits).
// c[] in corner order 000,100,010,110,001,101,011,111.
// ---------------------------------------------------------------------------------
struct Edge {
    uint8_t a, b;       // corner indices
    uint8_t axis;       // component that carries t
    uint8_t fx, fy, fz; // fixed components (0 or 1), the t component entry is unused
};
// z-edges 0-3, y-edges 4-7, x-edges 8-11 (dualcontour.wgsl:98-114)
__constant__ Edge kEdges[12] = {
    {0, 4, 2, 0, 0, 0}, {2, 6, 2, 0, 1, 0}, {1, 5, 2, 1, 0, 0}, {3, 7, 2, 1, 1, 0},
    {0, 2, 1, 0, 0, 0}, {4, 6, 1, 0, 0, 1}, {1, 3, 1, 1, 0, 0}, {5, 7, 1, 1, 0, 1},
    {0, 1, 0, 0, 0, 0}, {4, 5, 0, 0, 0, 1}, {2, 3, 0, 0, 1, 0}, {6, 7, 0, 0, 1, 1},
};

__device__ __forceinline__ bool contour_cell(const float c[8], const float lo[3], const float hi[3],
                                             float pos[3])
{
    // no edge changes sign when all eight corners fall on one side of `> 0`
    uint32_t pos_mask = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) pos_mask |= (uint32_t)(c[k] > 0.0f) << k;
    if (pos_mask == 0u || pos_mask == 0xffu) return false;
    float ax = 0.0f, ay = 0.0f, az = 0.0f, cnt = 0.0f;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const float v0 = c[kEdges[e].a], v1 = c[kEdges[e].b];
        const float t = (0.0f - v0) / (v1 - v0);  // _cell_adapt
        float x = kEdges[e].fx, y = kEdges[e].fy, z = kEdges[e].fz;
        if (kEdges[e].axis == 0) x = t;
        else if (kEdges[e].axis == 1) y = t;
        else z = t;
        if ((v0 > 0.0f) == (v1 > 0.0f)) x = y = z = 0.0f;  // _cell_change
        if (x > 0.0f || y > 0.0f || z > 0.0f) {
            ax = ax + x;
            ay = ay + y;
            az = az + z;
            cnt = cnt + 1.0f;
        }
    }
    if (cnt <= 1.0f) return false;
    pos[0] = canonical_nan(lo[0] + ((hi[0] - lo[0]) * ax) / cnt);
    pos[1] = canonical_nan(lo[1] + ((hi[1] - lo[1]) * ay) / cnt);
    pos[2] = canonical_nan(lo[2] + ((hi[2] - lo[2]) * az) / cnt);
    return true;
}

// contour_cell's verdict alone (whether the cell gets a vertex), without the position.
// An edge counts when its corners straddle `> 0` and its point has a component > 0.  Nine
// of the twelve edges have a fixed component 1, so they count whenever their corners
// straddle; the three edges from corner 000 (z: 0-4, y: 0-2, x: 0-1) have only the edge
// parameter t = (0 - v0) / (v1 - v0), which must be > 0 as the IEEE quotient (it is 0
// when the difference overflows or the quotient underflows).  The cell gets a vertex
// when at least two edges count.
__device__ __forceinline__ bool cell_active(const float c[8])
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) m |= (uint32_t)(c[k] > 0.0f) << k;
    if (m == 0u || m == 0xffu) return false;
    const uint32_t ch = (m ^ (m >> 4)) | ((m ^ (m >> 2)) << 8) | ((m ^ (m >> 1)) << 16);
    // bits of `ch`: z-edges a-(a+4) at a = 0..3, y-edges a-(a+2) at 8 + {0, 1, 4, 5},
    // x-edges a-(a+1) at 16 + {0, 2, 4, 6}
    const uint32_t fixed1 = ch & ((0xeu) | (0x32u << 8) | (0x54u << 16));
    uint32_t cnt = (uint32_t)__builtin_popcount(fixed1);
    auto from0 = [&](uint32_t bit, int b) -> uint32_t {
        if (!((ch >> bit) & 1u)) return 0u;
        const float t = (0.0f - c[0]) / (c[b] - c[0]);
        return t > 0.0f ? 1u : 0u;
    };
    cnt += from0(0, 4) + from0(8, 2) + from0(16, 1);
    return cnt >= 2u;
}

__device__ __forceinline__ uint32_t sign_bits(const float c[8])
{
    return (uint32_t)(c[1] > 0.0f) | ((uint32_t)(c[2] > 0.0f) << 1) |
           ((uint32_t)(c[4] > 0.0f) << 2) | ((uint32_t)(c[0] > 0.0f) << 3);
}

// Corners of cell (x, y, zl) from the slab lattice, and the cell's bounds.
__device__ __forceinline__ void gather_lattice(const SlabArgs& a, uint32_t x, uint32_t y, uint32_t zl,
                                               float c[8], float lo[3], float hi[3])
{
    const uint32_t zg = a.z0 + zl;
    const uint32_t xi[2] = {a.lox[x], a.hix[x]};
    const uint32_t yi[2] = {a.loy[y], a.hiy[y]};
    const uint32_t zi[2] = {a.loz[zg], a.hiz[zg]};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const size_t off = ((size_t)(zi[(k >> 2) & 1] - a.lz0) * a.lat_ny + yi[(k >> 1) & 1]) * a.pitch +
                           xi[k & 1];
        c[k] = a.lat[off];
    }
    lo[0] = a.cx[xi[0]];
    hi[0] = a.cx[xi[1]];
    lo[1] = a.cy[yi[0]];
    hi[1] = a.cy[yi[1]];
    lo[2] = a.cz[zi[0]];
    hi[2] = a.cz[zi[1]];
}

__device__ __forceinline__ void cell_coords(const SlabArgs& a, uint32_t i, uint32_t& x, uint32_t& y,
                                            uint32_t& zl)
{
    const uint32_t row = fdiv(i, a.div_rx);  // y + ry * zl
    x = i - row * a.rx;
    zl = fdiv(i, a.div_rxy);
    y = row - zl * a.ry;
}

// ---------------------------------------------------------------------------------
// Cell passes over per-cell corner dumps (FROM_CORNERS; the slab lattice has k_count_cols,
// k_cells_map and k_cell_vertices).  EMIT = 0: every cell's vertex rule; block active
// counts and each cell's activity bit (b.actbits, one ballot word per 64 cells).  EMIT = 1: the cell -> vertex
// index map for every cell (kNone where no vertex), its ranks from the activity bits, then the block's active
// cells, listed in LDS in vertex order, are gathered and contoured again on full waves
// (the same rule on the same corners gives the same vertex) and their vertices written.
// FROM_CORNERS selects the per-cell corner dump source.
// ---------------------------------------------------------------------------------
template <int FROM_CORNERS>
__device__ __forceinline__ bool cell_vertex(const SlabArgs& a, const float* __restrict__ corners, uint32_t i,
                                            uint32_t x, uint32_t y, uint32_t zl, float (&c)[8], float (&pos)[3])
{
    float lo[3], hi[3];
    if (FROM_CORNERS) {
        const float* src = corners + 8 * (size_t)i;
#pragma unroll
        for (int k = 0; k < 8; ++k) c[k] = src[k];
        const uint32_t zg = a.z0 + zl;
        lo[0] = a.cx[a.lox[x]];
        hi[0] = a.cx[a.hix[x]];
        lo[1] = a.cy[a.loy[y]];
        hi[1] = a.cy[a.hiy[y]];
        lo[2] = a.cz[a.loz[zg]];
        hi[2] = a.cz[a.hiz[zg]];
    } else {
        gather_lattice(a, x, y, zl, c, lo, hi);
    }
    return contour_cell(c, lo, hi, pos);
}

template <int EMIT, int FROM_CORNERS>
__global__ __launch_bounds__(256) void k_cells(SlabArgs a, MeshBufs b, const float* __restrict__ corners,
                                               int zero_normals)
{
    const uint32_t base = blockIdx.x * kCellsPerBlock;
    if constexpr (!EMIT) {
        __shared__ uint32_t lds[kBlock / 64];
        uint32_t count = 0;
        for (uint32_t j = 0; j < kRounds; ++j) {
            const uint32_t i = base + j * kBlock + threadIdx.x;
            bool act = false;
            if (i < a.ncells) {
                uint32_t x, y, zl;
                float c[8];
                cell_coords(a, i, x, y, zl);
                if (FROM_CORNERS) {
                    const float* src = corners + 8 * (size_t)i;
#pragma unroll
                    for (int k = 0; k < 8; ++k) c[k] = src[k];
                } else {
                    float lo[3], hi[3];
                    gather_lattice(a, x, y, zl, c, lo, hi);
                }
                act = cell_active(c);
            }
            const uint64_t w = __ballot(act);
            if (lane_id() == 0 && i < a.ncells) b.actbits[i >> 6] = w;
            uint32_t total;
            block_exclusive_256(act ? 1u : 0u, lds, &total);
            count += total;
        }
        if (threadIdx.x == 0) b.bcount[blockIdx.x] = count;
    } else {
        __shared__ uint32_t list[kCellsPerBlock];  // the block's active cells, in vertex order
        const uint32_t first = b.boffset[blockIdx.x];
        uint32_t before = 0;  // active cells of earlier rounds
        for (uint32_t j = 0; j < kRounds; ++j) {
            const uint32_t i0 = base + j * kBlock, i = i0 + threadIdx.x;
            const uint32_t wv = threadIdx.x >> 6, ln = lane_id();
            uint64_t wd[kBlock / 64];
#pragma unroll
            for (uint32_t k = 0; k < kBlock / 64; ++k) wd[k] = i0 + 64u * k < a.ncells ? b.actbits[(i0 >> 6) + k] : 0ull;
            uint32_t rank = before, total = 0;
#pragma unroll
            for (uint32_t k = 0; k < kBlock / 64; ++k) {
                const uint32_t pc = (uint32_t)__builtin_popcountll(wd[k]);
                total += pc;
                if (k < wv) rank += pc;
            }
            const uint64_t mine = wd[wv];
            const bool act = (mine >> ln) & 1ull;
            rank += (uint32_t)__builtin_popcountll(mine & ((1ull << ln) - 1ull));
            if (i < a.ncells) {
                uint32_t x, y, zl;
                cell_coords(a, i, x, y, zl);
                const size_t m = ((size_t)(zl + 1) * a.ry + y) * a.rx + x;
                a.vidx[m] = act ? first + rank : kNone;
                if (act) list[rank] = i;
            }
            before += total;
        }
        __syncthreads();  // the block's list is complete
        for (uint32_t q = threadIdx.x; q < before; q += kBlock) {
            const uint32_t i = list[q];
            const uint64_t v = (uint64_t)first + q;
            uint32_t x, y, zl;
            float c[8], pos[3];
            cell_coords(a, i, x, y, zl);
            cell_vertex<FROM_CORNERS>(a, corners, i, x, y, zl, c, pos);
            if (v < b.vcap) {
                float* o = b.vtx + 6 * v;
                o[0] = pos[0];
                o[1] = pos[1];
                o[2] = pos[2];
                if (zero_normals) {
                    o[3] = 0.0f;
                    o[4] = 0.0f;
                    o[5] = 0.0f;
                }
                b.keys[v] = (uint64_t)x | ((uint64_t)y << 16) | ((uint64_t)(a.z0 + zl) << 32);
                b.signs[v] = (uint8_t)sign_bits(c);
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Emit pass over the slab lattice, in two kernels.  k_cells_map: the cell -> vertex map
// for the cells with a vertex (and every cell of the slab's last plane), the ranks from
// the activity bits (as k_cells<1>), and each vertex's cell key; the bits stay for the
// quad pass (cleared after it).  k_cell_vertices: a grid-stride walk over
// the slab's vertices [kVtxPrev, min(kVtxTotal, vcap)) that gathers each vertex's cell
// from its key and writes the vertex (contour_cell, as the count pass decided) and its
// sign bits.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cells_map(SlabArgs a, MeshBufs b)
{
    // thread t: cells base + 4t .. base + 4t + 3, bits 4t % 64 .. of word t / 16; the map
    // entry of slab cell i is vidx[plane + i] (plane 0 holds the previous slab's last plane)
    constexpr uint32_t kWords = kCellsPerBlock / 64;
    const uint32_t base = blockIdx.x * kCellsPerBlock, t = threadIdx.x;
    const uint32_t first = b.boffset[blockIdx.x];
    uint64_t wd[kWords];
#pragma unroll
    for (uint32_t k = 0; k < kWords; ++k) wd[k] = base + 64u * k < a.ncells ? b.actbits[(base >> 6) + k] : 0ull;
    const uint32_t w = t >> 4, sh = (4u * t) & 63u;
    uint32_t rank = first;
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kWords; ++k) {
        if (k < w) rank += (uint32_t)__builtin_popcountll(wd[k]);
        if (k == w) mine = wd[k];
    }
    rank += (uint32_t)__builtin_popcountll(mine & ((1ull << sh) - 1ull));
    const uint32_t i0 = base + 4u * t;
    const size_t plane = (size_t)a.rx * a.ry;
    uint32_t v[4];
#pragma unroll
    for (uint32_t u = 0; u < 4u; ++u) {
        const bool act = (mine >> (sh + u)) & 1ull;
        v[u] = act ? rank : kNone;
        if (act && i0 + u < a.ncells && rank < b.vcap) {
            uint32_t x, y, zl;
            cell_coords(a, i0 + u, x, y, zl);
            b.keys[rank] = (uint64_t)x | ((uint64_t)y << 16) | ((uint64_t)(a.z0 + zl) << 32);
        }
        rank += act ? 1u : 0u;
    }
    // entries for the cells with a vertex only (k_quads tests a cell's activity bit before
    // reading its entry), except in the slab's last cell plane, which the next slab reads
    // as its previous plane without the bits: written in full, kNone where no vertex
    uint32_t* m = a.vidx + plane + i0;
    const uint32_t last = a.ncells - (uint32_t)plane;  // first cell of the last plane
    if (i0 >= last && i0 + 3u < a.ncells && (plane & 3u) == 0u) {
        *reinterpret_cast<uint4*>(m) = make_uint4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u)
            if (i0 + u < a.ncells && (v[u] != kNone || i0 + u >= last)) m[u] = v[u];
    }
}

__global__ __launch_bounds__(256) void k_cell_vertices(SlabArgs a, MeshBufs b, int zero_normals)
{
    const uint64_t beg = b.counters[kVtxPrev];
    const uint64_t end = b.counters[kVtxTotal] < b.vcap ? b.counters[kVtxTotal] : b.vcap;
    for (uint64_t v = beg + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < end;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = b.keys[v];
        const uint32_t x = (uint32_t)(key & 0xffffu), y = (uint32_t)((key >> 16) & 0xffffu);
        const uint32_t zl = (uint32_t)(key >> 32) - a.z0;
        float c[8], pos[3];
        cell_vertex<0>(a, nullptr, 0, x, y, zl, c, pos);
        float* o = b.vtx + 6 * v;
        o[0] = pos[0];
        o[1] = pos[1];
        o[2] = pos[2];
        if (zero_normals) {
            o[3] = 0.0f;
            o[4] = 0.0f;
            o[5] = 0.0f;
        }
        b.signs[v] = (uint8_t)sign_bits(c);
    }
}

// ---------------------------------------------------------------------------------
// Count pass over the slab lattice, one z-column of cells per thread: block (bx, y),
// thread x = 256 bx + t walks zl = 0 .. nzl - 1.  The walk is latency-bound (a handful
// of gathers per cell, then the rule), so the corners of kCountDepth cells are loaded
// before any of them is tested: one memory round trip per kCountDepth cells.  A cell's
// faces shared with its z-neighbour are loaded twice; the second load hits the cache.
// The vertex rule's verdict (cell_active) sets bit i % 64 of word i / 64 of `bits` (i = the cell's slab index); a wave's 64
// cells of a row span at most two words, ORed in with atomics only where a bit is set.
// `bits` is zero on entry (cleared after the slab's quad pass, launch_clear_bits).
// ---------------------------------------------------------------------------------
constexpr uint32_t kCountDepth = 8;

__global__ __launch_bounds__(256) void k_count_cols(SlabArgs a, uint32_t nzl, uint64_t* __restrict__ bits)
{
    // Workgroups go to the device's 8 XCDs in turn (block b to XCD b % 8), and each XCD
    // has its own L2.  Cell rows y and y + 1 share a lattice row (the upper face of one,
    // the lower face of the other), so each XCD walks one contiguous band of rows: block b
    // takes row-chunk (b % 8) * per + b / 8.  Neighbouring rows then run on one XCD and
    // the shared lattice row is fetched from HBM once instead of once per XCD.
    const uint32_t nxc = (a.rx + kBlock - 1) / kBlock, nblk = nxc * a.ry;
    const uint32_t per = (nblk + 7u) / 8u;
    const uint32_t lin = (blockIdx.x % 8u) * per + blockIdx.x / 8u;
    if (lin >= nblk) return;  // the launch rounds up to a multiple of 8 blocks
    const uint32_t xc = lin % nxc, y = lin / nxc;
    const uint32_t x = xc * kBlock + threadIdx.x;
    const uint32_t x0 = xc * kBlock + (threadIdx.x & ~63u);  // the wave's first cell
    if (x0 >= a.rx) return;  // whole wave past the row (wave-uniform)
    const bool in = x < a.rx;
    const uint32_t xi0 = in ? a.lox[x] : 0u, xi1 = in ? a.hix[x] : 0u;
    const size_t pl = (size_t)a.lat_ny * a.pitch;
    const float* r0 = a.lat + (size_t)a.loy[y] * a.pitch;  // lattice rows of the cell's y lo / hi
    const float* r1 = a.lat + (size_t)a.hiy[y] * a.pitch;
    auto face = [&](uint32_t zi, float* f) {  // the 4 corners of lattice plane zi
        const size_t o = (size_t)zi * pl;
        f[0] = r0[o + xi0];
        f[1] = r0[o + xi1];
        f[2] = r1[o + xi0];
        f[3] = r1[o + xi1];
    };
    for (uint32_t zb = 0; zb < nzl; zb += kCountDepth) {
        float c[kCountDepth][8];
#pragma unroll
        for (uint32_t j = 0; j < kCountDepth; ++j) {
            if (zb + j < nzl) {  // wave-uniform
                face(a.loz[a.z0 + zb + j] - a.lz0, c[j]);
                face(a.hiz[a.z0 + zb + j] - a.lz0, c[j] + 4);
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kCountDepth; ++j) {
            if (zb + j >= nzl) break;
            const bool act = in && cell_active(c[j]);
            const uint64_t w = __ballot(act);
            if (w != 0ull && lane_id() == 0) {
                const uint64_t i0 = x0 + (uint64_t)a.rx * (y + (uint64_t)a.ry * (zb + j));
                const uint32_t sh = (uint32_t)(i0 & 63u);
                unsigned long long* wp = reinterpret_cast<unsigned long long*>(bits + (i0 >> 6));
                atomicOr(wp, (unsigned long long)(w << sh));
                if (sh != 0u && (w >> (64u - sh)) != 0ull) atomicOr(wp + 1, (unsigned long long)(w >> (64u - sh)));
            }
        }
    }
}

// Active cells per 1024-cell block: the popcount of its 16 bit words (words past the
// slab's last cell are zero).
__global__ __launch_bounds__(256) void k_popc(const uint64_t* __restrict__ bits, uint32_t nb,
                                              uint32_t* __restrict__ bcount)
{
    const uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nb) return;
    const uint4* w = reinterpret_cast<const uint4*>(bits + (size_t)blk * (kCellsPerBlock / 64));
    uint32_t n = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCellsPerBlock / 128; ++k) {
        const uint4 q = w[k];
        n += __builtin_popcount(q.x) + __builtin_popcount(q.y) + __builtin_popcount(q.z) + __builtin_popcount(q.w);
    }
    bcount[blk] = n;
}

// ---------------------------------------------------------------------------------
// Exclusive scan of the per-block counts (which = 0: vertices, 1: triangles), in two
// launches of one workgroup per kDenseScanChunk counts, so that no single workgroup's latency
// is on the slab's critical path (the dense path evaluates the next slab on another stream
// at the same time).  k_scan_sums: each chunk's sum into part[g] (the words after the
// counters); workgroup 0 saves the running total of earlier slabs as the *_PREV counter.
// k_scan_offsets: each chunk's offsets from that total, the chunk sums before it and a
// scan within the workgroup; the last workgroup advances the running total.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ uint4 load_counts4(const uint32_t* __restrict__ bcount, uint32_t nb, uint32_t i)
{
    uint4 q = make_uint4(0u, 0u, 0u, 0u);
    if (i + 3 < nb) {
        q = *reinterpret_cast<const uint4*>(bcount + i);
    } else if (i < nb) {
        q.x = bcount[i];
        q.y = i + 1 < nb ? bcount[i + 1] : 0u;
        q.z = i + 2 < nb ? bcount[i + 2] : 0u;
    }
    return q;
}

// Inclusive prefix sum over a 256-thread workgroup (wsum: 4 words of LDS); *total gets
// the workgroup's sum.  Contains barriers.
__device__ __forceinline__ uint64_t block_inclusive_256(uint64_t v, uint64_t* wsum, uint64_t* total)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; ++k) {
        before += k < w ? wsum[k] : 0ull;
        all += wsum[k];
    }
    __syncthreads();  // wsum may be reused
    *total = all;
    return before + v;
}

static_assert(kDenseScanChunk == 4 * kBlock, "one 16-byte vector of counts per thread");

__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t* __restrict__ bcount, uint32_t nb,
                                                   uint64_t* __restrict__ counters, uint32_t which)
{
    __shared__ uint64_t wsum[kBlock / 64];
    const uint4 q = load_counts4(bcount, nb, blockIdx.x * kDenseScanChunk + 4u * threadIdx.x);
    uint64_t total;
    (void)block_inclusive_256((uint64_t)q.x + q.y + q.z + q.w, wsum, &total);
    if (threadIdx.x == 0) {
        counters[kNumCounters + blockIdx.x] = total;
        if (blockIdx.x == 0) counters[which ? kTriPrev : kVtxPrev] = counters[which ? kTriTotal : kVtxTotal];
    }
}

__global__ __launch_bounds__(256) void k_scan_offsets(const uint32_t* __restrict__ bcount, uint32_t nb,
                                                      uint32_t* __restrict__ boffset,
                                                      uint64_t* __restrict__ counters, uint32_t which)
{
    __shared__ uint64_t wsum[kBlock / 64];
    const uint32_t g = blockIdx.x;
    uint64_t parts;  // the chunk sums before this chunk (at most kDenseScanParts of them)
    (void)block_inclusive_256(threadIdx.x < g ? counters[kNumCounters + threadIdx.x] : 0ull, wsum, &parts);
    const uint32_t i = g * kDenseScanChunk + 4u * threadIdx.x;
    const uint4 q = load_counts4(bcount, nb, i);
    const uint64_t s = (uint64_t)q.x + q.y + q.z + q.w;
    uint64_t total;
    const uint64_t inc = block_inclusive_256(s, wsum, &total);
    const uint64_t base = counters[which ? kTriPrev : kVtxPrev] + parts;
    uint64_t run = base + inc - s;
    if (i < nb) {
        uint4 o;
        o.x = (uint32_t)run;
        run += q.x;
        o.y = (uint32_t)run;
        run += q.y;
        o.z = (uint32_t)run;
        run += q.z;
        o.w = (uint32_t)run;
        if (i + 3 < nb) {
            *reinterpret_cast<uint4*>(boffset + i) = o;
        } else {
            boffset[i] = o.x;
            if (i + 1 < nb) boffset[i + 1] = o.y;
            if (i + 2 < nb) boffset[i + 2] = o.z;
        }
    }
    if (g + 1 == gridDim.x && threadIdx.x == 0) counters[which ? kTriTotal : kVtxTotal] = base + total;
}

// ---------------------------------------------------------------------------------
// Normals of the vertices emitted by the current slab: sdf3d_normal + normalize
// (sdf3d_normal.wgsl:4-9, dualcontour.wgsl:170-171).  Grid-stride over the range
// [prev, total) read from the device counters, so no host round trip is needed.
// ---------------------------------------------------------------------------------
static_assert(kDenseVtxTotal == kVtxTotal && kDenseVtxPrev == kVtxPrev, "counter layout");

// library_normal over the vertices [counters[prev], min(counters[total], vcap)) with one
// lane per tap: lanes 4j .. 4j + 3 of a wave evaluate taps d1 .. d4 of one vertex (the
// same points as library_normal), lane 4j gathers them and writes the unit gradient.
// MANDEL = 1: the tape is one Mandelbulb op, called directly as tape_eval calls it (the
// same values without the interpreter's per-point dispatch).
template <int MANDEL>
__global__ __launch_bounds__(256) void k_normals(const s2m_op* __restrict__ ops, uint32_t nops,
                                                 float eps, float* __restrict__ vtx,
                                                 const uint64_t* __restrict__ counters, uint64_t vcap)
{
    const uint64_t beg = counters[kVtxPrev];
    const uint64_t end = counters[kVtxTotal] < vcap ? counters[kVtxTotal] : vcap;
    const uint32_t tap = threadIdx.x & 3u, l0 = lane_id() & ~3u;
    // taps v1 = (1,-1,-1), v2 = (-1,-1,1), v3 = (-1,1,-1), v4 = (1,1,1) times eps
    const float ex = (tap == 0u || tap == 3u) ? eps : -eps;
    const float ey = tap >= 2u ? eps : -eps;
    const float ez = (tap == 1u || tap == 3u) ? eps : -eps;
    // the 4 lanes of a vertex share the loop condition
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; beg + (i >> 2) < end;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float* o = vtx + 6 * (beg + (i >> 2));
        const float px = o[0] + ex, py = o[1] + ey, pz = o[2] + ez;
        const float d = MANDEL ? sd_mandelbulb(px, py, pz, ops->f[0], ops->f[1], ops->f[2], ops->iarg, ops->f[3])
                               : tape_eval(ops, nops, px, py, pz);
        const float d1 = __shfl(d, (int)l0, 64), d2 = __shfl(d, (int)l0 + 1, 64);
        const float d3 = __shfl(d, (int)l0 + 2, 64), d4 = __shfl(d, (int)l0 + 3, 64);
        if (tap == 0u) {
            const float gx = ((d1 + (-d2)) + (-d3)) + d4;
            const float gy = (((-d1) + (-d2)) + d3) + d4;
            const float gz = (((-d1) + d2) + (-d3)) + d4;
            write_unit(o, gx, gy, gz);
        }
    }
}

// ---------------------------------------------------------------------------------
// Quad pass (mesh.rs:267-324).  Per cell with a vertex: the x-, y- and z-edge quads,
// neighbour vertices from the map planes of this and the previous cell plane.  A quad
// with a missing vertex is skipped and counted (mesh.rs:270-278).  The cells with a
// vertex are the slab's vertices [kVtxPrev, kVtxTotal), in cell order, each with its
// cell key and sign bits, so block b walks vertices kVtxPrev + 1024 b ... (a grid sized
// for a vertex in every cell; blocks past the slab's vertices count zero).  EMIT = 0
// counts triangles per block, EMIT = 1 writes them in order.  A vertex beyond the
// output capacity has no key: it forms no quads (the job repeats with a larger capacity).
// ---------------------------------------------------------------------------------
template <int EMIT>
__global__ __launch_bounds__(256) void k_quads(SlabArgs a, MeshBufs b)
{
    __shared__ uint32_t lds[kBlock / 64];
    const uint64_t vbeg = b.counters[kVtxPrev] + (uint64_t)blockIdx.x * kCellsPerBlock;
    const uint64_t vend = min(b.counters[kVtxTotal], b.vcap);
    if (vbeg >= vend) {  // block-uniform
        if (!EMIT && threadIdx.x == 0) b.bcount[blockIdx.x] = 0;
        return;
    }
    uint64_t running = EMIT ? (uint64_t)b.boffset[blockIdx.x] : 0;
    uint32_t count = 0, invalid = 0;
    const size_t plane = (size_t)a.rx * a.ry;
    const bool lag = (a.flags & S2M_COMPAT_DROP_LAST_SLICE) != 0;
    for (uint32_t j = 0; j < kRounds; ++j) {
        const uint64_t vv = vbeg + j * kBlock + threadIdx.x;
        uint32_t q[3][4];
        uint32_t nq = 0;  // bit e set: quad e is valid
        if (vv < vend) {
            const uint32_t v = (uint32_t)vv;
            const uint64_t key = b.keys[v];
            const uint32_t x = (uint32_t)(key & 0xffffu), y = (uint32_t)((key >> 16) & 0xffffu);
            const uint32_t zl = (uint32_t)(key >> 32) - a.z0;
            // the vertex of the cell at offset `off` of map plane P (P = zl: the plane below,
            // P = zl + 1: this plane): plane 0 (the previous slab's last plane) is written in
            // full; in the slab's own planes only cells with a vertex have entries, so the
            // cell's activity bit is tested first
            auto at = [&](uint32_t P, size_t off) -> uint32_t {
                if (P == 0) return a.vidx[off];
                const size_t i = (size_t)(P - 1) * plane + off;  // the cell's slab index
                return ((b.actbits[i >> 6] >> (i & 63u)) & 1ull) ? a.vidx[(size_t)P * plane + off] : kNone;
            };
            const uint32_t pc = zl + 1, pp = zl;  // map planes of this and the previous cell plane
            const uint32_t s = b.signs[v];
            const bool c0 = s & 1, c1 = (s >> 1) & 1, c2 = (s >> 2) & 1, c3 = (s >> 3) & 1;
            const uint32_t zg = a.z0 + zl;
            const size_t r0 = (size_t)y * a.rx, rm = (size_t)(y - 1) * a.rx;
            // x-edge: Quad(v(x,y-1,z-1), v(x,y,z-1), v(x,y,z), v(x,y-1,z)).swap(c.0)
            if (c0 != c3 && y > 0) {
                if (zg > 0) {
                    uint32_t t[4] = {at(pp, rm + x), at(pp, r0 + x), v, at(pc, rm + x)};
                    if (c0) { q[0][0] = t[3]; q[0][1] = t[2]; q[0][2] = t[1]; q[0][3] = t[0]; }
                    else { q[0][0] = t[0]; q[0][1] = t[1]; q[0][2] = t[2]; q[0][3] = t[3]; }
                    if (t[0] != kNone && t[1] != kNone && t[3] != kNone) nq |= 1;
                    else ++invalid;
                } else if (lag) {
                    ++invalid;  // the lagged reference looks up label z-1 and misses
                }
            }
            // y-edge: Quad(v(x-1,y,z-1), v(x,y,z-1), v(x,y,z), v(x-1,y,z)).swap(!c.1)
            if (c1 != c3 && x > 0) {
                if (zg > 0) {
                    uint32_t t[4] = {at(pp, r0 + x - 1), at(pp, r0 + x), v, at(pc, r0 + x - 1)};
                    if (!c1) { q[1][0] = t[3]; q[1][1] = t[2]; q[1][2] = t[1]; q[1][3] = t[0]; }
                    else { q[1][0] = t[0]; q[1][1] = t[1]; q[1][2] = t[2]; q[1][3] = t[3]; }
                    if (t[0] != kNone && t[1] != kNone && t[3] != kNone) nq |= 2;
                    else ++invalid;
                } else if (lag) {
                    ++invalid;
                }
            }
            // z-edge: Quad(v(x-1,y-1,z), v(x,y-1,z), v(x,y,z), v(x-1,y,z)).swap(c.2)
            if (c2 != c3 && x > 0 && y > 0) {
                uint32_t t[4] = {at(pc, rm + x - 1), at(pc, rm + x), v, at(pc, r0 + x - 1)};
                if (c2) { q[2][0] = t[3]; q[2][1] = t[2]; q[2][2] = t[1]; q[2][3] = t[0]; }
                else { q[2][0] = t[0]; q[2][1] = t[1]; q[2][2] = t[2]; q[2][3] = t[3]; }
                if (t[0] != kNone && t[1] != kNone && t[3] != kNone) nq |= 4;
                else ++invalid;
            }
        }
        const uint32_t ntri = 2u * (uint32_t)__builtin_popcount(nq);
        uint32_t total;
        const uint32_t rank = block_exclusive_256(ntri, lds, &total);
        if (EMIT && nq) {
            uint64_t t = running + rank;
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                if (!(nq & (1u << e))) continue;
                // make_triangles: (q2, q1, q0), (q0, q3, q2)  (lib.rs:199-205)
                if (t + 1 < b.tcap) {
                    uint32_t* o = b.tri + 3 * t;
                    o[0] = q[e][2];
                    o[1] = q[e][1];
                    o[2] = q[e][0];
                    o[3] = q[e][0];
                    o[4] = q[e][3];
                    o[5] = q[e][2];
                }
                t += 2;
            }
        }
        running += total;
        count += total;
    }
    if (!EMIT) {
        if (invalid) atomicAdd((unsigned long long*)&b.counters[kInvalid], (unsigned long long)invalid);
        if (threadIdx.x == 0) b.bcount[blockIdx.x] = count;
    }
}

// ---------------------------------------------------------------------------------
// Every float bit pattern x: sincosf(x) against sinf(x) and cosf(x), bitwise.  cnt[0] /
// cnt[2]: sine / cosine results that differ, cnt[1] / cnt[3]: those of them where both
// are NaN.  (The Mandelbulb's escape step takes sine and cosine with one sincosf.)
// ---------------------------------------------------------------------------------
// Each side in a function of its own that is never inlined: a sinf and a cosf of the
// same argument in one function may be merged into one sincos call by the compiler,
// and the check would then compare sincosf with itself.
__device__ __attribute__((noinline)) float check_sin(float x) { return sinf(x); }
__device__ __attribute__((noinline)) float check_cos(float x) { return cosf(x); }
__device__ __attribute__((noinline)) void check_sincos(float x, float* s, float* c) { sincosf(x, s, c); }

__global__ __launch_bounds__(256) void k_sincos_check(unsigned long long* __restrict__ cnt)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32);
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float x = __uint_as_float((uint32_t)i);
        const float s1 = check_sin(x), c1 = check_cos(x);
        float s2, c2;
        check_sincos(x, &s2, &c2);
        if (__float_as_uint(s1) != __float_as_uint(s2)) atomicAdd(&cnt[(s1 != s1 && s2 != s2) ? 1 : 0], 1ull);
        if (__float_as_uint(c1) != __float_as_uint(c2)) atomicAdd(&cnt[(c1 != c1 && c2 != c2) ? 3 : 2], 1ull);
    }
}

// The Mandelbulb's bailout test r > b on r = sqrtf(r^2) taken as r^2 > t with
// t = bailout_threshold(b).  Bailouts b: 2^20 bit patterns spread over all 2^32
// (stride 4099: denormals, normals, infinities, NaNs, both signs) and 2.0; squared
// lengths x: the 65 floats around b * b and 0, the smallest denormal, 1, +inf and NaN,
// each x >= 0 or NaN as a sum of squares is.  cnt[0]: comparisons, cnt[1]: those where
// the two tests disagree.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bailout_check(unsigned long long* __restrict__ cnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > (1u << 20)) return;
    const float b = i == (1u << 20) ? 2.0f : __uint_as_float(i * 4099u);
    const float t = bailout_threshold(b);
    const uint32_t u = __float_as_uint(b * b);
    uint32_t checked = 0, bad = 0;
    for (int d = -37; d <= 32; ++d) {
        float x;
        switch (d) {
        case -37: x = 0.0f; break;
        case -36: x = __uint_as_float(1u); break;
        case -35: x = 1.0f; break;
        case -34: x = __uint_as_float(0x7f800000u); break;
        case -33: x = __uint_as_float(0x7fc00000u); break;
        default: x = __uint_as_float(u + (uint32_t)d); break;
        }
        if (x != x ? false : (__float_as_uint(x) >> 31) != 0u) continue;  // negative: no sum of squares
        ++checked;
        bad += (sqrtf(x) > b) != (x > t);
    }
    atomicAdd(&cnt[0], (unsigned long long)checked);
    if (bad) atomicAdd(&cnt[1], (unsigned long long)bad);
}

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
hipError_t launch_eval_lattice(hipStream_t s, const s2m_op* ops, uint32_t nops, const float* cx,
                               const float* cy, const float* cz, uint32_t nx, uint32_t ny,
                               uint32_t lz0, uint32_t nzl, uint32_t pitch, float* lat)
{
    if (nx == 0 || ny == 0 || nzl == 0) return hipSuccess;
    dim3 grid((nx + 15) / 16, (ny + 3) / 4, (nzl + 3) / 4);
    hipLaunchKernelGGL(k_eval_lattice, grid, dim3(256), 0, s, ops, nops, cx, cy, cz, nx, ny, lz0, nzl,
                       pitch, lat);
    return hipGetLastError();
}

hipError_t launch_eval_mandelbulb(hipStream_t s, const s2m_op* op, const float* cx, const float* cy,
                                  const float* cz, uint32_t nx, uint32_t ny, uint32_t lz0, uint32_t nzl,
                                  uint32_t pitch, float* lat)
{
    if (nx == 0 || ny == 0 || nzl == 0) return hipSuccess;
    dim3 grid((nx + 15) / 16, (ny + 3) / 4, (nzl + 3) / 4);
    hipLaunchKernelGGL(k_eval_mandelbulb, grid, dim3(256), 0, s, op, cx, cy, cz, nx, ny, lz0, nzl, pitch, lat);
    return hipGetLastError();
}

hipError_t launch_cells(hipStream_t s, const SlabArgs& a, const MeshBufs& b, int emit,
                        int write_zero_normals)
{
    const uint32_t nb = (a.ncells + kCellsPerBlock - 1) / kCellsPerBlock;
    if (nb == 0) return hipSuccess;
    if (emit) {
        hipLaunchKernelGGL(k_cells_map, dim3(nb), dim3(kBlock), 0, s, a, b);
        // about one vertex per thread (a slab's vertex count is only known on the device):
        // each vertex is a chain of dependent gathers, so many waves keep the loads in flight
        hipLaunchKernelGGL(k_cell_vertices, dim3(4096), dim3(kBlock), 0, s, a, b, write_zero_normals);
    } else {
        const uint32_t nzl = a.ncells / (a.rx * a.ry);
        const uint32_t nblk = (a.rx + kBlock - 1) / kBlock * a.ry;  // one block per 256 cells of a row
        hipLaunchKernelGGL(k_count_cols, dim3((nblk + 7u) / 8u * 8u), dim3(kBlock), 0, s, a, nzl, b.actbits);
        hipLaunchKernelGGL(k_popc, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b.actbits, nb, b.bcount);
    }
    return hipGetLastError();
}

hipError_t launch_cells_from_corners(hipStream_t s, const SlabArgs& a, const float* corners,
                                     const MeshBufs& b, int emit)
{
    const uint32_t nb = (a.ncells + kCellsPerBlock - 1) / kCellsPerBlock;
    if (nb == 0) return hipSuccess;
    if (emit)
        hipLaunchKernelGGL((k_cells<1, 1>), dim3(nb), dim3(kBlock), 0, s, a, b, corners, 1);
    else
        hipLaunchKernelGGL((k_cells<0, 1>), dim3(nb), dim3(kBlock), 0, s, a, b, corners, 1);
    return hipGetLastError();
}

hipError_t launch_sincos_check(hipStream_t s, uint64_t* cnt)
{
    hipLaunchKernelGGL(k_sincos_check, dim3(8192), dim3(256), 0, s, reinterpret_cast<unsigned long long*>(cnt));
    return hipGetLastError();
}

// One thread per op of a tape in device memory: a Mandelbulb op's f[3] (unused by the
// estimator's operands) gets bailout_threshold(f[1]), the r^2 bound of its power-8 step.
__global__ __launch_bounds__(256) void k_tape_prepare(s2m_op* __restrict__ ops, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ops[i].code == S2M_OP_MANDELBULB) ops[i].f[3] = bailout_threshold(ops[i].f[1]);
}

hipError_t launch_tape_prepare(hipStream_t s, s2m_op* ops, uint32_t n)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tape_prepare, dim3((n + 255) / 256), dim3(256), 0, s, ops, n);
    return hipGetLastError();
}

hipError_t launch_bailout_check(hipStream_t s, uint64_t* cnt)
{
    hipLaunchKernelGGL(k_bailout_check, dim3(((1u << 20) + 256u) / 256u), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(cnt));
    return hipGetLastError();
}

hipError_t launch_scan(hipStream_t s, const MeshBufs& b, uint32_t nblocks, uint32_t which)
{
    if (nblocks > kMaxScanBlocks) return hipErrorInvalidValue;  // kDenseScanParts chunk sums at most
    const uint32_t chunks = nblocks ? (nblocks + kDenseScanChunk - 1) / kDenseScanChunk : 1u;
    hipLaunchKernelGGL(k_scan_sums, dim3(chunks), dim3(kBlock), 0, s, b.bcount, nblocks, b.counters, which);
    hipLaunchKernelGGL(k_scan_offsets, dim3(chunks), dim3(kBlock), 0, s, b.bcount, nblocks, b.boffset, b.counters,
                       which);
    return hipGetLastError();
}

hipError_t launch_normals(hipStream_t s, const s2m_op* ops, uint32_t nops, float eps,
                          const MeshBufs& b, bool mandel)
{
    if (mandel)
        hipLaunchKernelGGL(k_normals<1>, dim3(4096), dim3(256), 0, s, ops, nops, eps, b.vtx, b.counters, b.vcap);
    else
        hipLaunchKernelGGL(k_normals<0>, dim3(4096), dim3(256), 0, s, ops, nops, eps, b.vtx, b.counters, b.vcap);
    return hipGetLastError();
}

// Zero `n` words (a slab's activity bits once its quad pass has read them, for the count
// pass that ORs into them two slabs later).
__global__ __launch_bounds__(256) void k_clear_words(uint64_t* __restrict__ w, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0ull;
}

hipError_t launch_clear_bits(hipStream_t s, uint64_t* bits, uint32_t ncells)
{
    const uint32_t n = (ncells + 63u) / 64u;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_words, dim3(std::min<uint32_t>((n + 255u) / 256u, 2048u)), dim3(256), 0, s, bits, n);
    return hipGetLastError();
}

hipError_t launch_quads(hipStream_t s, const SlabArgs& a, const MeshBufs& b, int emit)
{
    const uint32_t nb = (a.ncells + kCellsPerBlock - 1) / kCellsPerBlock;
    if (nb == 0) return hipSuccess;
    if (emit) hipLaunchKernelGGL((k_quads<1>), dim3(nb), dim3(kBlock), 0, s, a, b);
    else hipLaunchKernelGGL((k_quads<0>), dim3(nb), dim3(kBlock), 0, s, a, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Scratch of the no-allocation entry point (s2m_mesh_device_noalloc): one static device
// array per GPU that holds every working buffer of a dense job (coordinate tables, the
// slab lattice, cell -> vertex planes, block counts and offsets, activity bits,
// counters).  Host code never passes the array's name to a kernel or a copy (on the
// host it names a stand-in, not the device memory): it uses the address
// hipGetSymbolAddress returns for the device, taken once per device by noalloc_scratch.
// ---------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= (int)kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_done[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipErrorInvalidDevice;  // caller makes it current
        void* a = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        g_scratch_addr[device] = e == hipSuccess ? a : nullptr;
        g_scratch_err[device] = e;
        g_scratch_done[device] = true;
    }
    *ptr = g_scratch_addr[device];
    return g_scratch_err[device];
}

}  // namespace s2m
