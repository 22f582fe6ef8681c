// sparse.hip - the tile-sparse meshing pipeline, for one job or a batch of jobs.
//
// Programs that interval pruning applies to (every library primitive and operator except
// the Mandelbulb, which is kept whole) are meshed tile by tile: cells are grouped in
// tiles of 8x8x8 and super-tiles of 4x4x4 tiles, and only tiles whose SDF can change sign
// are evaluated and contoured.  A pass covers the tiles of one slab of one job, or of
// every job of a batch (each job whole); tiles, super-tiles and row segments of the pass
// are numbered job after job.  In stream order:
//   k_sp_prune_fold (left folds of unions / smooth unions) or k_sp_prune_prims +
//   k_sp_prune_super (other programs), then k_sp_prune   exact interval pruning: the
//                   job's program on each super-tile's padded box, then the super-tile's
//                   list on each of its tiles (one lane per tile); single-signed tiles are
//                   dropped, the others get a slot in a pool shared by all jobs (and a
//                   flag when their list is a left fold)
//   k_sp_bricks     per slot, one lane per brick of 2x2x2 cells: interval bounds of a
//                   left fold on the brick (bricks that keep one sign are not evaluated),
//                   and the slot's record: header, live bricks, axis tables, box
//                   coordinates and the program in k_sp_tiles' LDS layout
//   k_sp_tiles<1/0> per slot (one block): the pruned program in LDS (left folds: from the
//                   record, read while the block contours the slot before), at the
//                   lattice points the tile's live bricks use (left folds) or at all
//                   points of its box (other programs), into an LDS box; then the tile's
//                   cells: vertex and sign bits of each active cell, staged compactly, the
//                   slot's cell bits and each row segment's vertex count
//   k_sp_quadcount  per slot: valid quads of the staged cells (neighbour cells' bits),
//                   two triangles each, per row segment; invalid quads counted
//   k_sp_rowsum / k_sp_rowdown   offsets of the row segments in the reference's
//                   VertexList order (job after job, then z, y, tile x): per row, the
//                   segments of its active tiles; then a scan over rows, which also gives
//                   each job's first vertex and triangle index
//   k_sp_emit<1/0>  per slot: the staged vertices at their ordered index with their
//                   4-tap normals (the tile's pruned program), then the triangles of
//                   their quads, vertex indices relative to the job's first vertex
//   k_sp_boundary   (multi-slab single job) cell -> vertex map of the slab's last plane
// The per-slot kernels are persistent (as many blocks as the device holds) and draw slots
// from a work queue.  A cell's vertex index is never stored: it is the first index of its row segment plus
// the number of cells with a vertex before it in the segment (the slot's cell bits).
// Loops over slots read their bound on the device: no host round trip inside a pass.
#include <algorithm>
#include <mutex>
#include <vector>

#include "device_dense.h"
#include "device_interval.h"
#include "device_sdf.h"
#include "device_staged.h"
#include "internal.h"

namespace s2m {

// threadIdx.x through an opaque copy: values derived from it are recomputed where they
// are used instead of being hoisted out of the kernels' slot loops and held in registers
// for the whole kernel (which costs occupancy).
__device__ __forceinline__ uint32_t thread_id()
{
    uint32_t t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((uint32_t)threadIdx.x));
    return t;
}

namespace {

__device__ __forceinline__ uint32_t lane() { return thread_id() & 63u; }

__device__ __forceinline__ uint32_t active_slots(const SpPass& P)
{
    const uint64_t n = P.counters[kSlotsUsed];
    return (uint32_t)(n < P.slot_cap ? n : P.slot_cap);
}

// Work queues of a persistent launch: worker b of nid draws slots g, g + G, g + 2G, ... (g =
// b mod G, G = min(kQueueGroups, nid)) one at a time from its group's counter, so blocks that
// draw cheap tiles take more of them.  One counter per group, each on its own 128-byte
// line: updates of one address are serialised by the memory system, so a single counter
// for all blocks would throttle them.
__device__ __forceinline__ uint32_t queue_take(const SpPass& P, uint32_t q, uint32_t id, uint32_t nid)
{
    const uint32_t G = min(kQueueGroups, nid), g = id % G;
    const uint32_t j = atomicAdd(&P.queues[(q * kQueueGroups + g) * 32u], 1u);
    const uint64_t slot = (uint64_t)j * G + g;
    return (uint32_t)(slot < 0xffffffffull ? slot : 0xffffffffull);
}

// Row segment of row (zl, y) and tile column tx of job J.
// Per-slot row data: array k (kRowVtx, kRowTri: counts; kRowVtxOff, kRowTriOff: first
// vertex / triangle of the row segment within its row) of slot `slot`, row r = (z % 8) * 8
// + y % 8 of the tile.
enum RowArray : uint32_t { kRowVtx = 0, kRowTri = 1, kRowVtxOff = 2, kRowTriOff = 3 };
__device__ __forceinline__ uint32_t* slot_row(const SpPass& P, uint32_t k, uint32_t slot)
{
    return P.slot_rows + ((size_t)k * P.slot_cap + slot) * 64u;
}

// Row (zl, y) of job J in the pass's row numbering (rows job after job, zl then y).
__device__ __forceinline__ size_t row_index(const SpJob& J, uint32_t zl, uint32_t y)
{
    return J.row_base + (size_t)zl * J.ry + y;
}

// Tile coordinates (tx, ty, tile layer within the pass) of pass tile `tile` of job J.
__device__ __forceinline__ void tile_xyz(const SpJob& J, uint64_t tile, uint32_t& tx, uint32_t& ty,
                                         uint32_t& tzl)
{
    const uint32_t t = (uint32_t)(tile - J.tile_base);
    tx = t % J.ntx;
    ty = (t / J.ntx) % J.nty;
    tzl = t / (J.ntx * J.nty);
}

// Existence bits of row r (= z % 8 * 8 + y % 8) of a tile with a slot: bit x % 8.
__device__ __forceinline__ uint32_t slot_row_bits(const SpPass& P, int32_t slot, uint32_t r)
{
    return (P.cellbits[(size_t)slot * 16u + (r >> 2)] >> ((r & 3u) * 8u)) & 0xffu;
}

// Cells (8 * tx - 1 + i, y, z), i = 0..8, of job J (z is the job's plane index, y and z
// may be -1; cells outside the grid have no vertex): bit i of the result is set iff the
// cell has a vertex.  INDEX: idx[i] = its pass-level vertex index, kNone if none (the
// first vertex of the row segment plus the cell's rank in it).  Plane z0 - 1 comes from
// the previous slab's boundary map.
// ns: the slots of the tile (tx, ty, tz) and of its look-back neighbours, ns[(dz * 2 + dy) * 2
// + dx] for tile (tx - dx, ty - dy, tz - dz) (k_sp_bricks), -1 without one.
template <bool INDEX>
__device__ uint32_t nb_row(const SpPass& P, const SpJob& J, uint32_t tx, uint32_t ty, uint32_t tz,
                           const int32_t (&ns)[8], int y, int z, uint32_t* idx)
{
    uint32_t bits = 0;
    if (INDEX)
        for (int i = 0; i < 9; ++i) idx[i] = kNone;
    if (y < 0 || z < 0 || (uint32_t)y >= J.ry || (uint32_t)z >= J.z0 + J.nz) return 0;
    if ((uint32_t)z < J.z0) {
        for (int i = 0; i < 9; ++i) {
            const int x = (int)(tx * 8u) + i - 1;
            if (x < 0 || (uint32_t)x >= J.rx) continue;
            const uint32_t v = J.boundary[(size_t)y * J.rx + (uint32_t)x];
            if (v != kNone) bits |= 1u << i;
            if (INDEX) idx[i] = v;
        }
        return bits;
    }
    const uint32_t zl = (uint32_t)z - J.z0;
    const uint32_t nsel = ((tz - ((uint32_t)z >> 3)) * 2u + (ty - ((uint32_t)y >> 3))) * 2u;
    const uint32_t r = ((uint32_t)z & 7u) * 8u + ((uint32_t)y & 7u);
    for (uint32_t c = 0; c < 2; ++c) {  // c = 0: the tile column's own cells, 1: the cell before it
        const int32_t slot = ns[nsel + c];
        if (slot < 0) continue;
        const uint32_t b8 = slot_row_bits(P, slot, r);
        if (c ? !(b8 & 0x80u) : !b8) continue;
        bits |= c ? 1u : b8 << 1;
        if (INDEX) {
            const uint32_t first = slot_row(P, kRowVtxOff, (uint32_t)slot)[r] + P.row_off[row_index(J, zl, (uint32_t)y)];
            if (c) {
                idx[0] = first + (uint32_t)__builtin_popcount(b8 & 0x7fu);
            } else {
                for (uint32_t i = 0; i < 8; ++i)
                    if ((b8 >> i) & 1u) idx[i + 1] = first + (uint32_t)__builtin_popcount(b8 & ((1u << i) - 1u));
            }
        }
    }
    return bits;
}

// Pass-level vertex index of cell (x, y, z) of job J (z is the job's plane index), kNone
// if it has no vertex.  Plane z0 - 1 comes from the previous slab's boundary map.
__device__ __forceinline__ uint32_t vertex_at(const SpPass& P, const SpJob& J, uint32_t x, uint32_t y,
                                              uint32_t z)
{
    if (z < J.z0) return J.boundary[(size_t)y * J.rx + x];
    const uint64_t tile =
        J.tile_base + ((uint64_t)((z >> 3) - J.tz0) * J.nty + (y >> 3)) * J.ntx + (x >> 3);
    const int32_t slot = P.slot_of_tile[tile];
    if (slot < 0) return kNone;
    const uint32_t b8 = slot_row_bits(P, slot, (z & 7u) * 8u + (y & 7u)), i = x & 7u;
    if (!((b8 >> i) & 1u)) return kNone;
    return slot_row(P, kRowVtxOff, (uint32_t)slot)[(z & 7u) * 8u + (y & 7u)] + P.row_off[row_index(J, z - J.z0, y)] +
           (uint32_t)__builtin_popcount(b8 & ((1u << i) - 1u));
}

// dualcontour.wgsl:72-131: crossings on the 12 edges in the reference's order, counted
// iff the sign changes and the crossing is not at the cell origin, centroid if >= 2.
__device__ __forceinline__ bool contour_cell_sp(const float c[8], const float lo[3],
                                                const float hi[3], float pos[3])
{
    uint32_t pm = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) pm |= (uint32_t)(c[k] > 0.0f) << k;
    if (pm == 0u || pm == 0xffu) return false;
    // (corner a, corner b, axis of t, fixed x, fixed y, fixed z): z-, y-, then x-edges
    constexpr uint8_t E[12][6] = {{0, 4, 2, 0, 0, 0}, {2, 6, 2, 0, 1, 0}, {1, 5, 2, 1, 0, 0},
                                  {3, 7, 2, 1, 1, 0}, {0, 2, 1, 0, 0, 0}, {4, 6, 1, 0, 0, 1},
                                  {1, 3, 1, 1, 0, 0}, {5, 7, 1, 1, 0, 1}, {0, 1, 0, 0, 0, 0},
                                  {4, 5, 0, 0, 0, 1}, {2, 3, 0, 0, 1, 0}, {6, 7, 0, 0, 1, 1}};
    float ax = 0.0f, ay = 0.0f, az = 0.0f, cnt = 0.0f;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const float v0 = c[E[e][0]], v1 = c[E[e][1]];
        const float t = (0.0f - v0) / (v1 - v0);
        float x = E[e][3], y = E[e][4], z = E[e][5];
        if (E[e][2] == 0) x = t;
        else if (E[e][2] == 1) y = t;
        else z = t;
        if ((v0 > 0.0f) == (v1 > 0.0f)) x = y = z = 0.0f;
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

}  // namespace

namespace {

// Exact pruning of postfix programs on boxes (device_interval.h).  The kept op indices
// are written in order; the interval of the result is `top` when the walk ends.  The top
// of the interval stack is kept in registers, deeper entries in LDS columns (entry d of
// lane l at [d * 64 + l]).
struct PruneStack {
    double* lo;        // [depth][64]
    double* hi;        // [depth][64]
    uint32_t* sum;     // [depth][64] PruneWalk::top_sum of the entry
    uint16_t* start;   // [depth][64]
    uint16_t* single;  // [depth][64] op index of a one-primitive subtree, else kComposite

    // LDS bytes of a stack of `depth` entries per lane, and its layout from `base`.
    static constexpr size_t bytes(uint32_t depth)
    {
        return (size_t)depth * 64u * (2 * sizeof(double) + sizeof(uint32_t) + 2 * sizeof(uint16_t));
    }
    __device__ static PruneStack at(double* base, uint32_t depth)
    {
        uint32_t* s32 = reinterpret_cast<uint32_t*>(base + 128u * depth);
        uint16_t* s16 = reinterpret_cast<uint16_t*>(s32 + 64u * depth);
        return PruneStack{base, base + 64u * depth, s32, s16, s16 + 64u * depth};
    }
};

constexpr uint32_t kComposite = 0xffffu;
constexpr uint32_t kFoldBit = 0x80000000u;  // PruneWalk sums: the subtree is a left fold

struct PruneWalk {
    Iv top = iv_unknown();
    uint32_t top_start = 0, depth = 0, len = 0;  // depth: entries below `top`
    uint32_t top_single = kComposite;  // op index when `top` is one primitive
    // flops of the kept ops of the top subtree (op_flops), | kFoldBit when that subtree is a
    // left fold of library primitives (a primitive other than the Mandelbulb, or a left
    // fold, such a primitive and an operator): the tile's flops and left-fold flag without
    // reading its list back
    uint32_t top_sum = 0;

    // The primitive with op index i, interval v (q: its position in the program; sum: its
    // flops | kFoldBit unless it is a Mandelbulb).
    __device__ __forceinline__ void prim(uint32_t q, uint32_t i, const Iv& v, uint16_t* out, const PruneStack& S,
                                         uint32_t ln, uint32_t sum = 0)
    {
        if (q != 0) {  // push the previous top
            S.lo[depth * 64u + ln] = top.lo;
            S.hi[depth * 64u + ln] = top.hi;
            S.sum[depth * 64u + ln] = top_sum;
            S.start[depth * 64u + ln] = (uint16_t)top_start;
            S.single[depth * 64u + ln] = (uint16_t)top_single;
            ++depth;
        }
        top = v;
        top_start = len;
        top_single = i;
        top_sum = sum;
        out[len++] = (uint16_t)i;
    }

    // Operator i (code, smoothing radius k, flops fl) on the two topmost values.
    __device__ __forceinline__ void op(uint32_t i, uint32_t code, float k, uint16_t* out, const PruneStack& S,
                                       uint32_t ln, uint32_t fl = 0)
    {
        const Iv B = top;
        const uint32_t sB = top_start;
        --depth;
        const double alo = S.lo[depth * 64u + ln], ahi = S.hi[depth * 64u + ln];
        const Iv A{alo, ahi, alo > -kHuge};  // iv_make / iv_unknown: known iff lo > -kHuge
        const uint32_t sA = S.start[depth * 64u + ln], singleA = S.single[depth * 64u + ln];
        const uint32_t sumA = S.sum[depth * 64u + ln];
        int keep;
        const Iv r = iv_op(code, A, B, k, keep);
        if (keep == kKeepBoth) {
            out[len++] = (uint16_t)i;
            const bool fold = (sumA & kFoldBit) && top_single != kComposite && (top_sum & kFoldBit);
            top_sum = ((sumA & ~kFoldBit) + (top_sum & ~kFoldBit) + fl) | (fold ? kFoldBit : 0u);
            top = r;
            top_single = kComposite;
        } else if (keep == kKeepA) {
            len = sB;  // drop B's subtree and the operator
            top = A;
            top_single = singleA;
            top_sum = sumA;
        } else {       // drop A's subtree and the operator
            if (top_single != kComposite) {  // B is one primitive: no read-back of the list
                out[sA] = (uint16_t)top_single;
                len = sA + 1;
            } else {
                const uint32_t nb = len - sB;
                for (uint32_t j = 0; j < nb; ++j) out[sA + j] = out[sB + j];
                len = sA + nb;
            }
            top = B;
        }
        top_start = sA;
    }

    __device__ __forceinline__ bool one_sign() const { return top.known && (top.lo > 0.0 || top.hi < 0.0); }
};

// PruneWalk for a left fold (P0 P1 op1 P2 op2 ...): the same decisions, list and sums with
// the accumulated value and the latest primitive in registers instead of the LDS stack
// (every operator's left operand is the accumulated value, which starts the list).
struct FoldWalk {
    Iv top = iv_unknown();
    uint32_t len = 0;
    uint32_t top_single = kComposite;
    uint32_t top_sum = 0;
    Iv b = iv_unknown();  // the latest primitive (the right operand of the next operator)
    uint32_t b_i = 0, b_sum = 0, b_start = 0;

    __device__ __forceinline__ void prim(uint32_t q, uint32_t i, const Iv& v, uint16_t* out, const PruneStack&,
                                         uint32_t, uint32_t sum = 0)
    {
        if (q == 0) {
            top = v;
            top_single = i;
            top_sum = sum;
        } else {
            b = v;
            b_i = i;
            b_sum = sum;
            b_start = len;
        }
        out[len++] = (uint16_t)i;
    }

    __device__ __forceinline__ void op(uint32_t i, uint32_t code, float k, uint16_t* out, const PruneStack&,
                                       uint32_t, uint32_t fl = 0)
    {
        int keep;
        const Iv r = iv_op(code, top, b, k, keep);
        if (keep == kKeepBoth) {
            out[len++] = (uint16_t)i;
            const bool fold = (top_sum & kFoldBit) && (b_sum & kFoldBit);
            top_sum = ((top_sum & ~kFoldBit) + (b_sum & ~kFoldBit) + fl) | (fold ? kFoldBit : 0u);
            top = r;
            top_single = kComposite;
        } else if (keep == kKeepA) {
            len = b_start;  // drop the primitive and the operator
        } else {            // drop the accumulated value and the operator: the list is the primitive
            out[0] = (uint16_t)b_i;
            len = 1;
            top = b;
            top_single = b_i;
            top_sum = b_sum;
        }
    }

    __device__ __forceinline__ bool one_sign() const { return top.known && (top.lo > 0.0 || top.hi < 0.0); }
};

__device__ __forceinline__ uint32_t job_of(const SpPass& P, uint64_t index, bool super)
{
    uint32_t lo = 0, hi = P.njobs - 1;  // last job whose first tile (super-tile) <= index
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if ((super ? P.jobs[mid].super_base : P.jobs[mid].tile_base) <= index) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

// Padded box of the tiles [t0, t1] (per axis, t1 inclusive; z: tile layers of the job).
__device__ __forceinline__ void tiles_box(const SpJob& J, const uint32_t t0[3], const uint32_t t1[3], double bl[3],
                                          double bh[3])
{
    const float* e[3] = {J.ext, J.ext + 2 * J.ntx, J.ext + 2 * J.ntx + 2 * J.nty};
    const uint32_t n[3] = {J.ntx, J.nty, J.ntz};
    for (int a = 0; a < 3; ++a) {
        bl[a] = (double)e[a][t0[a]] - J.pad;
        bh[a] = (double)e[a][n[a] + t1[a]] + J.pad;
    }
}

// Tiles of super-tile s of job J: first and last tile per axis (z: tile layer of the job).
__device__ __forceinline__ void super_tiles(const SpJob& J, uint64_t s, uint32_t t0[3], uint32_t t1[3])
{
    const uint32_t t = (uint32_t)(s - J.super_base);
    const uint32_t sx = t % J.nsx, sy = (t / J.nsx) % J.nsy, szl = t / (J.nsx * J.nsy);
    const uint32_t layers = (J.nz + kTile - 1) / kTile;  // tile layers of this pass
    t0[0] = sx * kSuper;
    t1[0] = min(t0[0] + kSuper, J.ntx) - 1;
    t0[1] = sy * kSuper;
    t1[1] = min(t0[1] + kSuper, J.nty) - 1;
    t0[2] = J.tz0 + szl * kSuper;
    t1[2] = J.tz0 + min(szl * kSuper + kSuper, layers) - 1;
}

// ---------------------------------------------------------------------------------
// Pruning, in three launches:
//  k_sp_prune_prims  one lane per (super-tile, primitive): the primitive's interval on the
//                    super-tile's padded box (super_iv)
//  k_sp_prune_super  one lane per super-tile: the job's program combined in order from
//                    those intervals, exactly pruned (super_lists), or kSingleSign when
//                    the SDF keeps one sign on the super-tile (super_lens)
//  k_sp_prune        one wave per super-tile, lane l = its tile (l % 4, l / 4 % 4, l / 16):
//                    the super-tile's list pruned on the tile's padded box (every lane walks
//                    the same list: wave-uniform op dispatch).  Tiles that keep one sign are
//                    dropped, the others get a slot in the pool shared by all jobs (and a
//                    flag when their list is a left fold).
// Dynamic LDS of the last two: the prune stack (P.prune_depth entries per lane).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sp_prune_prims(SpPass P, uint32_t chunks)
{
    const uint64_t s = blockIdx.x / chunks;
    const uint32_t k = (blockIdx.x % chunks) * 64u + (thread_id() & 63u);
    const SpJob& J = P.jobs[__builtin_amdgcn_readfirstlane(job_of(P, s, true))];
    if (J.union_fold || k >= J.nops || J.ops[k].code >= S2M_OP_UNION) return;  // folds: k_sp_prune_fold
    uint32_t t0[3], t1[3];
    super_tiles(J, s, t0, t1);
    double bl[3], bh[3];
    tiles_box(J, t0, t1, bl, bh);
    const Iv v = iv_primitive(J.ops[k].code, J.ops[k].f, bl, bh);
    P.super_iv[s * P.list_stride + k] = make_double2(v.lo, v.hi);
}

// ---------------------------------------------------------------------------------
// k_sp_prune_fold: the super-tile lists of programs that are left folds of unions and
// smooth unions (P0 P1 op1 P2 op2 ..., op in {UNION, SMOOTH_UNION with k > 0}), one wave
// per super-tile, without a sequential walk: lane q evaluates op q's interval, and
// primitive P_m (m >= 1) is dropped with its operator when
//     lo(P_m) >= U_{m-1} + k_m + slack,   U_{m-1} = min over j < m of hi(P_j),
// a prefix minimum over the lanes.  The accumulated value after m - 1 steps never exceeds
// min_j p_j by more than the rounding of its fp32 operations (smooth union <= min of its
// operands), which the slack bounds: kMargin (64 ulps) x the magnitude of every step so
// far, each bounded by |p_j| + k_j + max_i |p_i| + sum_i k_i / 4 (|acc| never exceeds the
// last two terms); so p_m >= acc + k_m everywhere on the box, h = 1 and
// the operator returns acc bit for bit (union: p_m >= acc returns acc).  Dropping never
// changes acc, so later bounds stay valid.  The super-tile keeps one sign when
// min_j hi(P_j) + slack < 0 (acc < 0), or when min over the kept P_j of lo(P_j) - the kept
// smooth radii / 4 - slack > 0 (smooth union >= min of its operands - k / 4).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sp_prune_fold(SpPass P)
{
    const uint64_t s = blockIdx.x;
    const uint32_t ln = thread_id() & 63u;
    const SpJob& J = P.jobs[__builtin_amdgcn_readfirstlane(job_of(P, s, true))];
    if (!J.union_fold) return;
    uint32_t t0[3], t1[3];
    super_tiles(J, s, t0, t1);
    double bl[3], bh[3];
    tiles_box(J, t0, t1, bl, bh);
    const s2m_op* ops = J.ops;
    const uint32_t nops = J.nops;
    uint16_t* out = P.super_lists + s * P.list_stride;
    double carry_u = kHuge;    // min hi(P_j) over the primitives so far
    double carry_mag = 0.0;    // sum of |P_j| + k_j over the primitives so far
    double carry_max = 0.0;    // max of |P_j| + k_j so far
    double carry_ka = 0.0;     // sum of k_j / 4 so far
    double carry_cnt = 0.0;    // primitives so far
    double carry_lo = kHuge;   // min lo(P_j) over the kept primitives
    double carry_k4 = 0.0;     // sum of the kept smooth radii / 4
    uint32_t nkept = 0;
    bool prev_kept = true;     // was the last primitive of the previous chunk kept?
    for (uint32_t c0 = 0; c0 < nops; c0 += 64u) {
        const uint32_t q = c0 + ln;
        const bool valid = q < nops;
        const bool prim = valid && (q == 0 || (q & 1u));  // a left fold: P0, then P_m at odd q
        // the operator after primitive q (q >= 1) and its smoothing radius
        double kk = 0.0;
        if (prim && q >= 1 && ops[q + 1].code == S2M_OP_SMOOTH_UNION) kk = (double)ops[q + 1].f[0];
        Iv B = iv_unknown();
        if (prim) B = iv_primitive(ops[q].code, ops[q].f, bl, bh);
        const double mag = prim ? fmax(fabs(B.lo), fabs(B.hi)) + kk : 0.0;
        // prefix scans over the lanes (inclusive): min hi, sum and max of the magnitudes,
        // sum of k / 4, count of primitives
        double u_ex = prim ? B.hi : kHuge, m_in = mag, x_in = mag, ka_in = 0.25 * kk, c_in = prim ? 1.0 : 0.0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double uo = __shfl_up(u_ex, d, 64), mo = __shfl_up(m_in, d, 64);
            const double xo = __shfl_up(x_in, d, 64), ko = __shfl_up(ka_in, d, 64), co = __shfl_up(c_in, d, 64);
            if (ln >= (uint32_t)d) {
                u_ex = fmin(u_ex, uo);
                m_in += mo;
                x_in = fmax(x_in, xo);
                ka_in += ko;
                c_in += co;
            }
        }
        const double u_in = u_ex;  // inclusive
        u_ex = __shfl_up(u_in, 1, 64);
        if (ln == 0) u_ex = kHuge;
        u_ex = fmin(u_ex, carry_u);
        m_in += carry_mag;
        x_in = fmax(x_in, carry_max);
        ka_in += carry_ka;
        c_in += carry_cnt;
        const double slack = kMargin * (m_in + c_in * (x_in + ka_in));
        const double tiny = 1e-12 * (fabs(u_ex) + fabs(B.lo) + fabs(B.hi) + kk);
        const bool drop = prim && q >= 1 && B.known && u_ex < kHuge && B.lo >= u_ex + kk + tiny + slack;
        // an operator (even q >= 2) is kept with its primitive q - 1
        const uint64_t pk = __ballot(prim && !drop);
        const bool prev_prim_kept = ln == 0 ? prev_kept : ((pk >> (ln - 1)) & 1ull) != 0;
        const bool kept = valid && (prim ? !drop : prev_prim_kept);
        const uint64_t km = __ballot(kept);
        if (kept) out[nkept + (uint32_t)__builtin_popcountll(km & ((1ull << ln) - 1ull))] = (uint16_t)q;
        nkept += (uint32_t)__builtin_popcountll(km);
        // one-sign bounds over the kept primitives
        double lo_k = (prim && !drop) ? B.lo : kHuge;
        double k4 = (prim && !drop) ? 0.25 * kk : 0.0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            lo_k = fmin(lo_k, __shfl_xor(lo_k, d, 64));
            k4 += __shfl_xor(k4, d, 64);
        }
        carry_lo = fmin(carry_lo, lo_k);
        carry_k4 += k4;
        carry_u = fmin(carry_u, __shfl(u_in, 63, 64));
        carry_mag = __shfl(m_in, 63, 64);
        carry_max = __shfl(x_in, 63, 64);
        carry_ka = __shfl(ka_in, 63, 64);
        carry_cnt = __shfl(c_in, 63, 64);
        // the chunk's last primitive: at the highest odd q of the chunk (q = c0 + 63 when
        // the chunk is full; 64 is even, so lane 63 is odd)
        const uint32_t last = min(nops - c0, 64u) - 1u;
        const uint32_t lastp = (c0 + last) & 1u ? last : (last ? last - 1u : 0u);
        prev_kept = ((pk >> lastp) & 1ull) != 0;
    }
    const double slack = kMargin * (carry_mag + carry_cnt * (carry_max + carry_ka));
    const bool one_sign = (carry_u + slack < 0.0) || (carry_lo > -kHuge && carry_lo - carry_k4 - slack > 0.0);
    if (ln == 0) P.super_lens[s] = one_sign ? kSingleSign : nkept;
}

constexpr uint32_t kSuperChunk = 16;  // ops of a lane's program staged in LDS at a time

__global__ __launch_bounds__(64) void k_sp_prune_super(SpPass P)
{
    extern __shared__ double prune_lds[];
    __shared__ double2 civ[kSuperChunk][64];   // primitive intervals
    __shared__ uint32_t ccode[kSuperChunk][64];
    __shared__ float ck[kSuperChunk][64];      // operator smoothing radius
    const uint32_t D = P.prune_depth, ln = thread_id() & 63u;
    const PruneStack S = PruneStack::at(prune_lds, D);
    const uint64_t s = (uint64_t)blockIdx.x * 64u + ln;
    if (s >= P.nsuper) return;
    const SpJob& J = P.jobs[job_of(P, s, true)];
    if (J.union_fold) return;  // k_sp_prune_fold
    const s2m_op* ops = J.ops;
    const uint32_t nops = J.nops;
    const double2* iv = P.super_iv + s * P.list_stride;
    uint16_t* out = P.super_lists + s * P.list_stride;
    PruneWalk w;
    for (uint32_t q0 = 0; q0 < nops; q0 += kSuperChunk) {
        // the lane's next ops and intervals: all loads in flight at once (each lane reads
        // back only its own column)
        const uint32_t m = min(kSuperChunk, nops - q0);
        for (uint32_t j = 0; j < m; ++j) {
            ccode[j][ln] = ops[q0 + j].code;
            ck[j][ln] = ops[q0 + j].f[0];
            civ[j][ln] = iv[q0 + j];  // operators' entries are unused
        }
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t q = q0 + j, code = ccode[j][ln];
            if (code < S2M_OP_UNION) {
                const double2 v = civ[j][ln];
                w.prim(q, q, Iv{v.x, v.y, v.x > -kHuge}, out, S, ln);  // known iff lo > -kHuge
            } else {
                w.op(q, code, ck[j][ln], out, S, ln);
            }
        }
    }
    P.super_lens[s] = w.one_sign() ? kSingleSign : w.len;
}

__global__ __launch_bounds__(64) void k_sp_prune(SpPass P)
{
    extern __shared__ double prune_lds[];
    __shared__ StagedTape T;          // the super-tile's list (at most kStageOps ops)
    __shared__ uint16_t idx[kStageOps];  // op index of each staged op
    __shared__ uint32_t sums[kStageOps];  // its flops | kFoldBit for primitives other than the Mandelbulb
    const uint32_t D = P.prune_depth, ln = thread_id() & 63u;
    const PruneStack S = PruneStack::at(prune_lds, D);
    const uint64_t s = blockIdx.x;
    const uint32_t jid = __builtin_amdgcn_readfirstlane(job_of(P, s, true));
    const SpJob& J = P.jobs[jid];
    const uint32_t sl = __builtin_amdgcn_readfirstlane(P.super_lens[s]);
    uint32_t t0[3], t1[3];
    super_tiles(J, s, t0, t1);
    const uint32_t tx = t0[0] + (ln & 3u), ty = t0[1] + ((ln >> 2) & 3u), tz = t0[2] + (ln >> 4);
    const bool valid = tx <= t1[0] && ty <= t1[1] && tz <= t1[2];
    const uint64_t tile = J.tile_base + ((uint64_t)(tz - J.tz0) * J.nty + ty) * J.ntx + tx;
    const s2m_op* ops = J.ops;
    const uint16_t* in = P.super_lists + s * P.list_stride;
    const bool staged = sl != kSingleSign && sl <= kStageOps;
    auto op_sum = [&](uint32_t i) {
        const uint32_t c = ops[i].code;
        return J.op_flops[i] | (c < S2M_OP_UNION && c != S2M_OP_MANDELBULB ? kFoldBit : 0u);
    };
    if (staged) {
        stage_tape(T, ops, in, sl);
        for (uint32_t k = ln; k < sl; k += 64u) {
            idx[k] = in[k];
            sums[k] = op_sum(in[k]);
        }
    }
    __syncthreads();
    bool surface = false;
    uint32_t lens = 0;
    if (valid && sl == kSingleSign) {  // no surface anywhere in the super-tile
        P.lens[tile] = 0;
        P.flops[tile] = 0;
    } else if (valid) {
        const uint32_t tt[3] = {tx, ty, tz};
        double bl[3], bh[3];
        tiles_box(J, tt, tt, bl, bh);
        uint16_t* L = P.lists + (size_t)tile * P.list_stride;
        auto walk = [&](auto& w) {
            if (staged) {  // every lane walks the same list: wave-uniform dispatch
                for (uint32_t q = 0; q < sl; ++q) {
                    const uint32_t code = uniform_code(T, q), i = idx[q], sum = sums[q];
                    if (code < S2M_OP_UNION) w.prim(q, i, iv_primitive(code, T.f[q], bl, bh), L, S, ln, sum);
                    else w.op(i, code, T.f[q][0], L, S, ln, sum);
                }
            } else {
                for (uint32_t q = 0; q < sl; ++q) {
                    const uint32_t i = __builtin_amdgcn_readfirstlane(in[q]);
                    const uint32_t code = ops[i].code, sum = op_sum(i);
                    if (code < S2M_OP_UNION) w.prim(q, i, iv_primitive(code, ops[i].f, bl, bh), L, S, ln, sum);
                    else w.op(i, code, ops[i].f[0], L, S, ln, sum);
                }
            }
            surface = !w.one_sign();
            // the pruned program is one subtree: its flops and whether it is a left fold
            const bool fold = (w.top_sum & kFoldBit) && w.len <= kStageOps;
            lens = w.len | (surface && fold ? kLinearFlag : 0u);
            P.lens[tile] = lens;
            P.flops[tile] = w.top_sum & ~kFoldBit;
        };
        if (J.left_fold) {  // the super-tile's list is a left fold too
            FoldWalk w;
            walk(w);
        } else {
            PruneWalk w;
            walk(w);
        }
    }
    // slots for the surviving tiles: one counter update per wave
    const uint64_t m = __ballot(surface);
    uint64_t wbase = 0;
    if (ln == 0 && m) wbase = atomicAdd((unsigned long long*)&P.counters[kSlotsUsed], (unsigned long long)__builtin_popcountll(m));
    wbase = __shfl(wbase, 0, 64);
    int32_t slot = -1;
    if (surface) {
        const uint64_t sidx = wbase + (uint64_t)__builtin_popcountll(m & ((1ull << ln) - 1ull));
        if (sidx < P.slot_cap) {
            slot = (int32_t)sidx;
            P.tile_of_slot[sidx] = (uint32_t)tile;
            P.job_of_slot[sidx] = jid;
        }
    }
    if (valid) P.slot_of_tile[tile] = slot;
    const uint64_t act = __ballot(slot >= 0);  // bit l: tile l of the super-tile has a slot
    if (ln == 0) P.super_act[s] = act;

}

// Per-axis corner tables of a tile's 8 cells: box-relative lattice index and coordinate
// of each cell's lo / hi corner.
struct TileAxes {
    uint8_t lo[3][8], hi[3][8];  // box indices (< 16)
    float clo[3][8], chi[3][8];
};

// Entry e (< 24) of A: axis e / 8, cell e % 8.  Cells past the end of the grid repeat the
// last cell's entries (their cells are never contoured).
__device__ __forceinline__ void load_tile_axes(TileAxes& A, const SpJob& J, uint32_t tx, uint32_t ty,
                                               uint32_t tz, uint32_t e)
{
    const uint32_t ax = e >> 3, c = e & 7u;
    const uint32_t t = ax == 0 ? tx : ax == 1 ? ty : tz;
    const uint32_t r = ax == 0 ? J.rx : ax == 1 ? J.ry : J.z0 + J.nz;  // cells on the axis
    const uint32_t* lo = ax == 0 ? J.lox : ax == 1 ? J.loy : J.loz;
    const uint32_t* hi = ax == 0 ? J.hix : ax == 1 ? J.hiy : J.hiz;
    const float* co = ax == 0 ? J.cx : ax == 1 ? J.cy : J.cz;
    const uint32_t org = J.ibox[2 * (ax == 0 ? 0 : ax == 1 ? J.ntx : J.ntx + J.nty) + t];
    const uint32_t g = t * 8u + c;   // cell index along the axis (z: of the job)
    const uint32_t gi = g < r ? g : r - 1;
    A.lo[ax][c] = lo[gi] - org;
    A.hi[ax][c] = hi[gi] - org;
    A.clo[ax][c] = co[lo[gi]];
    A.chi[ax][c] = co[hi[gi]];
}

// Per slot, written by k_sp_bricks: what k_sp_tiles<1> stages for the slot, in its LDS
// layout, so that one read of the record (made while the block contours the slot before)
// replaces the dependent loads of header, job, pruned list, ops, axis tables and
// coordinates.  Header and live bricks are written for every slot, the rest for slots
// whose program is a left fold.
struct alignas(16) TileRec {
    SlotHdr H;
    uint64_t live;              // bricks (2x2x2 cells) that may hold an active cell
    uint16_t xb[4];             // per x-brick: the box x indices its cells use (bit mask)
    TileAxes A;
    uint8_t bm[2][16];          // per box y (z) index: the y (z) bricks whose cells use it
    float pc[3][16];            // lattice coordinates of the box per axis
    StagedFold T;               // the pruned program (the used rows are read)
};
static_assert(sizeof(TileRec) == kTileRecBytes, "record size shared with the host");
constexpr uint32_t kRecVec = kTileRecBytes / 16;   // 16-byte words per record
constexpr uint32_t kRecPre = 64;                    // words read ahead: header .. codes, ops 0..22
constexpr uint32_t kRecRowsAt = (uint32_t)offsetof(TileRec, T) + (uint32_t)offsetof(StagedFold, f);
static_assert(kRecRowsAt + 32u * 12u <= 16u * kRecPre, "ops 0..22 are read ahead");
struct TileGen {  // k_sp_tiles<0>: the same parts, staged per slot
    StagedTape T;
    TileAxes A;
    float pc[3][16];
};

// ---------------------------------------------------------------------------------
// n / d for n, d < 2^16 with m = ceil(2^32 / d): the error term n * (m*d - 2^32) / 2^32 / d
// stays below 1/d, so it never reaches the next integer.
__device__ __forceinline__ uint32_t magic16(uint32_t d)
{
    return d <= 1u ? 0u : (uint32_t)((0x100000000ull + d - 1u) / d);
}
__device__ __forceinline__ uint32_t div16(uint32_t n, uint32_t d, uint32_t m)
{
    return d <= 1u ? n : __umulhi(n, m);
}

// ---------------------------------------------------------------------------------
// Bricks: the 4x4x4 blocks of 2x2x2 cells of a tile, brick (bx, by, bz) = bit
// (bz * 4 + by) * 4 + bx of the slot's `slot_live`.  Per slot (one 64-lane block
// iteration, lane = brick): interval bounds of the tile's program on the brick's lattice
// points; a brick whose values provably keep one sign holds no active cell, so its cells
// are not contoured and its points are not evaluated.  Left-fold programs only (other
// programs keep every brick).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_sp_bricks(SpPass P)
{
    __shared__ StagedFold T;
    __shared__ TileAxes A;
    const uint32_t nslots = active_slots(P);
    // The slots' tile, job and list length are read ahead (tile and job two slots ahead,
    // the length one slot ahead) with per-lane loads, which barriers do not wait for.
    const uint32_t z = thread_id() & 0u;  // a per-lane zero: vector loads
    const uint32_t G = gridDim.x;
    uint32_t t1 = 0, j1 = 0, l1 = 0, t2 = 0, j2 = 0;
    if (blockIdx.x < nslots) {
        t1 = P.tile_of_slot[blockIdx.x + z];
        j1 = P.job_of_slot[blockIdx.x + z];
        l1 = P.lens[__builtin_amdgcn_readfirstlane(t1) + z];
    }
    if (blockIdx.x + G < nslots) {
        t2 = P.tile_of_slot[blockIdx.x + G + z];
        j2 = P.job_of_slot[blockIdx.x + G + z];
    }
    for (uint32_t slot = blockIdx.x; slot < nslots; slot += G) {
        const uint32_t b = thread_id();
        const uint64_t tile = __builtin_amdgcn_readfirstlane(t1);
        const uint32_t lens = __builtin_amdgcn_readfirstlane(l1);
        const uint32_t n = lens & ~kLinearFlag;
        const uint32_t jid = __builtin_amdgcn_readfirstlane(j1);
        t1 = t2;
        j1 = j2;
        if (slot + G < nslots) l1 = P.lens[__builtin_amdgcn_readfirstlane(t1) + z];
        if (slot + 2u * G < nslots) {
            t2 = P.tile_of_slot[slot + 2u * G + z];
            j2 = P.job_of_slot[slot + 2u * G + z];
        }
        const SpJob& J = P.jobs[jid];
        uint32_t tx, ty, tzl;
        tile_xyz(J, tile, tx, ty, tzl);
        TileRec& RG = *reinterpret_cast<TileRec*>(P.slot_rec + (size_t)slot * kRecVec);
        if (b < 48) {  // the slot's header (lane a < 3: axis a's box origin and size) and coordinates
            const uint32_t a = b >> 4, i = b & 15u, t = a == 0 ? tx : a == 1 ? ty : J.tz0 + tzl;
            const uint32_t nt = a == 0 ? J.ntx : a == 1 ? J.nty : J.ntz;
            const uint32_t* ib = J.ibox + 2 * (a == 0 ? 0 : a == 1 ? J.ntx : J.ntx + J.nty);
            const uint32_t o = ib[t], e = ib[nt + t];
            const bool linear = (lens & kLinearFlag) && n <= kStageOps;
            if (linear && i <= e - o) RG.pc[a][i] = (a == 0 ? J.cx : a == 1 ? J.cy : J.cz)[o + i];
            if (i == 0) {
                const uint32_t box = o | ((e - o + 1) << 24);
                P.slot_hdr[slot].box[a] = box;
                RG.H.box[a] = box;
            }
            if (b == 0) {
                SlotHdr H;
                H.tile = (uint32_t)tile;
                H.job = jid;
                H.lens = lens;
                H.txy = tx | (ty << 16);
                H.tzl = tzl;
                for (SlotHdr* d : {&P.slot_hdr[slot], &RG.H}) {
                    d->tile = H.tile;
                    d->job = H.job;
                    d->lens = H.lens;
                    d->txy = H.txy;
                    d->tzl = H.tzl;
                }
            }
        }
        if (b >= 48 && b < 56) {  // the slot's look-back neighbours (nb_row): (tx - dx, ty - dy, tzl - dz)
            const uint32_t i = b - 48u, dx = i & 1u, dy = (i >> 1) & 1u, dz = i >> 2;
            int32_t ns = -1;
            if (tx >= dx && ty >= dy && tzl >= dz)
                ns = P.slot_of_tile[J.tile_base + ((uint64_t)(tzl - dz) * J.nty + (ty - dy)) * J.ntx + (tx - dx)];
            P.slot_nb[(size_t)slot * 8u + i] = ns;
        }
        if (!(lens & kLinearFlag) || n > kStageOps) {  // every brick is kept
            if (b == 0) P.slot_live[slot] = RG.live = ~0ull;
            continue;
        }
        __syncthreads();  // the previous slot's readers of T and A are done
        stage_tape(T, J.ops, P.lists + (size_t)tile * P.list_stride, n);
        if (b < 24) load_tile_axes(A, J, tx, ty, J.tz0 + tzl, b);
        __syncthreads();
        {  // the record's axes and program (codes and the rows in use)
            const uint32_t words = (uint32_t)(offsetof(StagedFold, f) + 32u * ((n + 1u) >> 1)) / 4u;
            for (uint32_t k = b; k < words; k += 64u)
                reinterpret_cast<uint32_t*>(&RG.T)[k] = reinterpret_cast<const uint32_t*>(&T)[k];
            for (uint32_t k = b; k < (uint32_t)sizeof(TileAxes) / 4u; k += 64u)
                reinterpret_cast<uint32_t*>(&RG.A)[k] = reinterpret_cast<const uint32_t*>(&A)[k];
            if (b < 4) {
                const uint32_t c = 2u * b;
                RG.xb[b] = (uint16_t)((1u << A.lo[0][c]) | (1u << A.hi[0][c]) | (1u << A.lo[0][c + 1]) |
                                      (1u << A.hi[0][c + 1]));
            } else if (b >= 32) {  // lane 32 + 16 (a - 1) + i: axis a (y, z), box index i
                const uint32_t a = 1u + ((b - 32u) >> 4), i = b & 15u;
                uint32_t m = 0;
                for (uint32_t c = 0; c < 8u; ++c)
                    if (A.lo[a][c] == i || A.hi[a][c] == i) m |= 1u << (c >> 1);
                RG.bm[a - 1u][i] = (uint8_t)m;
            }
        }
        // brick b: cells c, c + 1 per axis; the first cell's lo corner, the second's hi corner
        const uint32_t c[3] = {(b & 3u) * 2u, ((b >> 2) & 3u) * 2u, (b >> 4) * 2u};
        double bl[3], bh[3];
        for (int a = 0; a < 3; ++a) {
            bl[a] = (double)A.clo[a][c[a]];
            bh[a] = (double)A.chi[a][c[a] + 1];
        }
        Iv acc = iv_primitive(uniform_code(T, 0), T.f_at(0), bl, bh);
        for (uint32_t k = 1; k + 1 < n; k += 2) {
            const Iv d = iv_primitive(uniform_code(T, k), T.f_at(k), bl, bh);
            int keep;
            acc = iv_op(uniform_code(T, k + 1), acc, d, T.f_at(k + 1)[0], keep);
        }
        const bool may_cross = !(acc.known && (acc.lo > 0.0 || acc.hi < 0.0));
        const uint64_t m = __ballot(may_cross);
        if (b == 0) P.slot_live[slot] = RG.live = m;
    }
}

// ---------------------------------------------------------------------------------
// The cells of one tile (all threads of the block; the tile's box values in LDS), in
// three steps: the cells whose 8 corners change sign are listed (cells of bricks outside
// `live` are skipped: their corner values keep one sign and were not evaluated); each
// listed cell gets its vertex (dualcontour.wgsl:72-131) and sign bits (:57-69); the
// cells with a vertex are staged compactly for k_sp_emit with their rank in their row
// segment, their bits go to the slot's `cellbits` and each row segment's count to the
// slot's kRowVtx row.  Contains barriers.
// ---------------------------------------------------------------------------------
// Staged cells are allocated in chunks of kStageChunk per block (one counter update per
// chunk).  Each slot reserves 8 entries per live brick (its cells) when its tile starts,
// so a slot's cells are contiguous; a chunk with fewer entries left than a slot reserves
// is abandoned.  Hence at most 2 x (reserved entries) + blocks x kStageChunk are used; the
// host grows the buffer from the device count if that exceeds it.
constexpr uint32_t kStageChunk = kStageChunkHost;
constexpr uint32_t kTileThreads = 256;  // threads of a k_sp_tiles block (one tile at a time)
constexpr int kLinearPoints = 2;        // points per lane sharing each op's decode (left folds)
static_assert(kRecVec - kRecPre <= kTileThreads, "one thread per record word past the read-ahead");

struct ContourLds {  // block-shared state of contour_tile (and `cand`, 512 entries: the cells
                     // with a sign change among their corners)
    uint32_t rows[64];   // row czl * 8 + cyl: bit cxl = the cell has a vertex
    uint32_t ncand, nact;
    uint32_t base, left, slot_base;  // staging chunk of the block
};

// Reset for the next tile (thread_id() < 64; a barrier must follow before contour_tile).
__device__ __forceinline__ void contour_reset(ContourLds& L)
{
    if (thread_id() < 64) L.rows[thread_id()] = 0;
    if (thread_id() == 0) L.ncand = L.nact = 0;
}

// Thread 0: staging space for the slot's cells with a vertex, at most `cells` (8 per live
// brick), from the block's chunk; published by the barriers before contour_tile.
__device__ __forceinline__ void contour_reserve(const SpPass& P, ContourLds& L, uint32_t cells)
{
    if (L.left < cells) {
        L.base = (uint32_t)atomicAdd((unsigned long long*)&P.counters[kStageUsed], (unsigned long long)kStageChunk);
        L.left = kStageChunk;
    }
    L.slot_base = L.base + (kStageChunk - L.left);
    L.left -= cells;
}

__device__ __forceinline__ void tile_corners(const float* box, const TileAxes& A, uint32_t bx, uint32_t by,
                                             uint32_t c, float cv[8])
{
    const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
    const uint32_t xi[2] = {A.lo[0][cxl], A.hi[0][cxl]};
    const uint32_t yi[2] = {A.lo[1][cyl], A.hi[1][cyl]};
    const uint32_t zi[2] = {A.lo[2][czl], A.hi[2][czl]};
#pragma unroll
    for (int k = 0; k < 8; ++k) cv[k] = box[(zi[(k >> 2) & 1] * by + yi[(k >> 1) & 1]) * bx + xi[k & 1]];
}

// The job's extents as wave-uniform values, read once per tile: the job table sits in
// global memory that the kernel also writes, so the compiler re-reads fields of a
// reference to it inside every loop.
struct TileGrid {
    uint32_t rx, ry, nz;  // cells per axis of the job's grid (nz: of this pass)
    uint32_t bx, by;      // box strides
};

__device__ __forceinline__ TileGrid tile_grid(const SpJob& J)
{
    TileGrid G;
    G.rx = __builtin_amdgcn_readfirstlane(J.rx);
    G.ry = __builtin_amdgcn_readfirstlane(J.ry);
    G.nz = __builtin_amdgcn_readfirstlane(J.nz);
    G.bx = __builtin_amdgcn_readfirstlane(J.bx);
    G.by = __builtin_amdgcn_readfirstlane(J.by);
    return G;
}

__device__ __forceinline__ void contour_tile(const SpPass& P, const TileGrid& G, uint32_t slot, uint32_t tx,
                                             uint32_t ty, uint32_t tzl, const float* box, const TileAxes& A,
                                             uint32_t nlive, const uint8_t* lb, uint16_t* cand, ContourLds& L)
{
    // 1. cells with a sign change: the 8 cells of each live brick (lb: the live bricks,
    // nlive of them), one per thread
    for (uint32_t k0 = 0; k0 < nlive * 8u; k0 += kTileThreads) {
        const uint32_t k = k0 + thread_id();
        uint32_t c = 0;
        bool cand_c = false;
        if (k < nlive * 8u) {
            const uint32_t b = lb[k >> 3], j = k & 7u;  // brick (x fastest), cell within it
            const uint32_t cxl = (b & 3u) * 2u + (j & 1u), cyl = ((b >> 2) & 3u) * 2u + ((j >> 1) & 1u);
            const uint32_t czl = (b >> 4) * 2u + (j >> 2);
            c = (czl * 8u + cyl) * 8u + cxl;
            if (tx * 8u + cxl < G.rx && ty * 8u + cyl < G.ry && tzl * 8u + czl < G.nz) {
                float cv[8];
                tile_corners(box, A, G.bx, G.by, c, cv);
                uint32_t pm = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i) pm |= (uint32_t)(cv[i] > 0.0f) << i;
                cand_c = pm != 0u && pm != 0xffu;
            }
        }
        const uint64_t w = __ballot(cand_c);
        const uint32_t lead = (uint32_t)__builtin_ctzll(w | (1ull << 63));
        uint32_t wbase = 0;
        if (w && lane() == lead) wbase = atomicAdd(&L.ncand, (uint32_t)__builtin_popcountll(w));
        wbase = (uint32_t)__builtin_amdgcn_readlane((int)wbase, (int)lead);
        if (cand_c) cand[wbase + (uint32_t)__builtin_popcountll(w & ((1ull << lane()) - 1ull))] = (uint16_t)c;
    }
    __syncthreads();
    // 2. vertex and sign bits of each listed cell
    const uint32_t ncand = L.ncand;
    bool act[2] = {false, false};
    float pos[2][3];
    uint32_t cell[2] = {0, 0}, sg[2] = {0, 0}, k_at[2] = {0, 0};
#pragma unroll
    for (uint32_t r = 0; r < 2; ++r) {
        const uint32_t k = r * kTileThreads + thread_id();
        if (k < ncand) {
            const uint32_t c = cand[k];
            const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
            float cv[8];
            tile_corners(box, A, G.bx, G.by, c, cv);
            const float lo[3] = {A.clo[0][cxl], A.clo[1][cyl], A.clo[2][czl]};
            const float hi[3] = {A.chi[0][cxl], A.chi[1][cyl], A.chi[2][czl]};
            act[r] = contour_cell_sp(cv, lo, hi, pos[r]);
            cell[r] = c;
            // cell_sign_changes (dualcontour.wgsl:57-69)
            sg[r] = (uint32_t)(cv[1] > 0.0f) | ((uint32_t)(cv[2] > 0.0f) << 1) |
                    ((uint32_t)(cv[4] > 0.0f) << 2) | ((uint32_t)(cv[0] > 0.0f) << 3);
        }
        const uint64_t w = __ballot(act[r]);
        const uint32_t lead = (uint32_t)__builtin_ctzll(w | (1ull << 63));
        uint32_t wbase = 0;
        if (w && lane() == lead) wbase = atomicAdd(&L.nact, (uint32_t)__builtin_popcountll(w));
        wbase = (uint32_t)__builtin_amdgcn_readlane((int)wbase, (int)lead);
        if (act[r]) {
            k_at[r] = wbase + (uint32_t)__builtin_popcountll(w & ((1ull << lane()) - 1ull));
            atomicOr(&L.rows[cell[r] >> 3], 1u << (cell[r] & 7u));
        }
    }
    __syncthreads();
    // 3. the staged cells (at the space reserved by contour_reserve), row segment counts
    // and the slot's cell bits
    if (thread_id() == 0) {
        P.slot_stage[2 * slot] = L.slot_base;
        P.slot_stage[2 * slot + 1] = L.nact;
    }
    if (thread_id() < 64) {
        const uint32_t row = thread_id();
        slot_row(P, kRowVtx, slot)[row] = (uint32_t)__builtin_popcount(L.rows[row]);  // 0 outside the grid
        if (row < 16)
            P.cellbits[(size_t)slot * 16u + row] = L.rows[4 * row] | (L.rows[4 * row + 1] << 8) |
                                                   (L.rows[4 * row + 2] << 16) | (L.rows[4 * row + 3] << 24);
    }
    const uint32_t stage_base = L.slot_base;
#pragma unroll
    for (uint32_t r = 0; r < 2; ++r) {
        if (!act[r]) continue;
        const uint32_t c = cell[r];
        const uint32_t rank = (uint32_t)__builtin_popcount(L.rows[c >> 3] & ((1u << (c & 7u)) - 1u));
        const uint64_t at = (uint64_t)stage_base + k_at[r];
        if (at < P.stage_cap)
            P.stage[at] = make_float4(pos[r][0], pos[r][1], pos[r][2], __uint_as_float(c | (rank << 9) | (sg[r] << 12)));
    }
}

constexpr uint32_t kList = 1024;   // needed points held before they are evaluated
constexpr uint32_t kBoxAxis = 16;  // most lattice indices of a tile box per axis
static_assert(kList >= 8 * 8 * 8, "the point list doubles as the contour's candidate cells");
static_assert(kTileThreads >= kBoxAxis * kBoxAxis, "one box row per thread");

// k_sp_tiles: per slot (persistent, slots drawn from a work queue), the tile's pruned
// program, staged in
// LDS, at the lattice points of its closed box that the tile's cells need (all points for
// programs that are not left folds), into an LDS box; then the tile's cells are contoured
// from it (contour_tile).  LINEAR = 1 runs the slots whose program is a left fold (two
// points per lane), LINEAR = 0 the others; both launches walk all slots and skip the
// other kind.
template <int LINEAR>
__device__ __forceinline__ void tiles_body(const SpPass& P)
{
    extern __shared__ float box[];  // P.box_stride floats
    // left folds: the slot's record (k_sp_bricks); other programs: staged here
    __shared__ std::conditional_t<LINEAR != 0, TileRec, TileGen> R;
    // points to evaluate; then the contour's candidate cells (contour_tile's `cand`)
    __shared__ uint16_t list[LINEAR ? kList : 8 * 8 * 8];
    __shared__ uint32_t nlist;
    __shared__ ContourLds L;
    __shared__ uint32_t sh_slot, sh_next;
    __shared__ uint8_t lb[64];  // the live bricks of the slot (contour_tile)
    const uint32_t nslots = active_slots(P);
    constexpr uint32_t kQueue = LINEAR ? 0u : 1u;
    if (thread_id() == 0) L.left = 0;  // published by the first slot's barriers
    uint64_t tiles_done = 0, points_done = 0, flops_done = 0;  // thread 0: this block's share
    uint32_t next = thread_id() == 0 ? queue_take(P, kQueue, blockIdx.x, gridDim.x) : 0u;
    // Left folds: threads < kRecPre hold the first kRecPre x 16 bytes of the next slot's
    // record, read while the current slot is contoured and stored to R at the loop's top.
    uint4 pre = make_uint4(0u, 0u, 0u, 0u);
    auto prefetch = [&]() {  // all threads; `next` published in sh_next before a barrier
        const uint32_t s = sh_next;
        if (thread_id() < kRecPre && s < nslots) pre = P.slot_rec[(size_t)s * kRecVec + thread_id()];
    };
    if constexpr (LINEAR != 0) {
        if (thread_id() == 0) nlist = 0;
        if (thread_id() == 0) sh_next = next;
        __syncthreads();
        prefetch();
    }
    for (;;) {
        __syncthreads();  // the previous slot's readers of sh_slot, R, list, box, L are done
        if (thread_id() == 0) sh_slot = next;
        if (LINEAR && thread_id() < kRecPre) reinterpret_cast<uint4*>(&R)[thread_id()] = pre;
        __syncthreads();
        const uint32_t slot = sh_slot;
        if (slot >= nslots) break;
        if (thread_id() == 0) next = queue_take(P, kQueue, blockIdx.x, gridDim.x);  // prefetched
        const SlotHdr& H = [&]() -> const SlotHdr& {
            if constexpr (LINEAR != 0) return R.H;
            else return P.slot_hdr[__builtin_amdgcn_readfirstlane(slot)];
        }();
        const uint32_t lens = __builtin_amdgcn_readfirstlane(H.lens);
        auto skip = [&]() {  // to the next slot (left folds: its record is read first)
            if (LINEAR) {
                if (thread_id() == 0) sh_next = next;
                __syncthreads();
                prefetch();
            }
        };
        if (((lens & kLinearFlag) != 0) != (LINEAR != 0)) {  // same for the whole block
            skip();
            continue;
        }
        const uint64_t tile = __builtin_amdgcn_readfirstlane(H.tile);
        // thread 0's flop count for the block statistics, read now so that the load has
        // landed when the tile ends
        const uint32_t tile_flops = thread_id() == 0 ? P.flops[tile] : 0u;
        const uint32_t n = lens & ~kLinearFlag;
        const SpJob& J = P.jobs[__builtin_amdgcn_readfirstlane(H.job)];
        const uint32_t txy = __builtin_amdgcn_readfirstlane(H.txy);
        const uint32_t tx = txy & 0xffffu, ty = txy >> 16, tzl = __builtin_amdgcn_readfirstlane(H.tzl);
        const uint32_t bxh = __builtin_amdgcn_readfirstlane(H.box[0]), byh = __builtin_amdgcn_readfirstlane(H.box[1]);
        const uint32_t bzh = __builtin_amdgcn_readfirstlane(H.box[2]);
        const uint32_t nx = bxh >> 24, ny = byh >> 24, nz = bzh >> 24;
        const TileGrid G = tile_grid(J);
        const uint32_t bx = G.bx, by = G.by;
        const uint16_t* Lst = P.lists + (size_t)tile * P.list_stride;
        uint64_t lv = ~0ull;
        uint32_t evaluated = 0;
        contour_reset(L);
        if constexpr (LINEAR != 0) {
            lv = R.live;
            if (lv == 0) {  // no brick can hold an active cell: empty rows, no cell bits
                if (thread_id() < 64) slot_row(P, kRowVtx, slot)[thread_id()] = 0;
                if (thread_id() < 16) P.cellbits[(size_t)slot * 16u + thread_id()] = 0;
                if (thread_id() == 0) {
                    P.slot_stage[2 * slot] = 0;
                    P.slot_stage[2 * slot + 1] = 0;
                    tiles_done += 1;
                }
                skip();
                continue;
            }
            // the rest of a long program's record (rows past the prefetched part): loaded
            // now, stored to R once the point list is built (nothing before reads it)
            const uint32_t need = (kRecRowsAt + 32u * ((n + 1u) >> 1) + 15u) >> 4;
            const uint32_t kr = kRecPre + thread_id();
            uint4 rest = make_uint4(0u, 0u, 0u, 0u);
            if (kr < need) rest = P.slot_rec[(size_t)slot * kRecVec + kr];
            // The points that some live brick uses.  Per box row (iz, iy): the x-bricks of the
            // live bricks whose cells use it (rm), hence the row's needed x positions (the
            // union of those x-bricks' positions, xb).  One row per thread; each wave reserves
            // its rows' points in `list` with one atomic; entries are iz << 10 | iy << 5 | ix.
            // The list is evaluated window by window (one window unless more than kList points).
            if (thread_id() == 0) contour_reserve(P, L, 8u * (uint32_t)__builtin_popcountll(lv));
            const uint32_t q = thread_id();
            uint32_t M = 0, code = 0;
            if (q < ny * nz) {
                const uint32_t iz = div16(q, ny, magic16(ny)), iy = q - iz * ny;
                // rm = OR over the z-bricks bz and y-bricks by that use the row of the live
                // bits of x-bricks (bz, by, 0..3): the used z layers of lv ORed (16 bits,
                // by * 4 + bx), then the used y rows of that
                const uint32_t mz = R.bm[1][iz], my = R.bm[0][iy];
                uint32_t layers = 0, rm = 0;
#pragma unroll
                for (uint32_t bz = 0; bz < 4; ++bz)
                    if ((mz >> bz) & 1u) layers |= (uint32_t)(lv >> (16u * bz)) & 0xffffu;
#pragma unroll
                for (uint32_t by = 0; by < 4; ++by)
                    if ((my >> by) & 1u) rm |= (layers >> (4u * by)) & 0xfu;
                M = ((rm & 1u) ? R.xb[0] : 0u) | ((rm & 2u) ? R.xb[1] : 0u) | ((rm & 4u) ? R.xb[2] : 0u) |
                    ((rm & 8u) ? R.xb[3] : 0u);
                code = (iz << 10) | (iy << 5);
            }
            // the exclusive prefix sum of cnt (<= 16, 5 bits) over the wave's lanes, bit by
            // bit: lanes below with the bit set (mbcnt over its ballot) times its weight
            const uint32_t cnt = (uint32_t)__builtin_popcount(M);
            uint32_t below = 0, wtotal = 0;
#pragma unroll
            for (uint32_t b = 0; b < 5; ++b) {
                const uint64_t m = __ballot((int)((cnt >> b) & 1u));
                below += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
                wtotal += (uint32_t)__builtin_popcountll(m) << b;
            }
            uint32_t wbase = 0;
            if (lane() == 0 && wtotal) wbase = atomicAdd(&nlist, wtotal);
            const uint32_t off = (uint32_t)__builtin_amdgcn_readlane((int)wbase, 0) + below;
            if (kr < need) reinterpret_cast<uint4*>(&R)[kr] = rest;
            // window 0 is filled at once (every wave has its offsets); the total is read
            // after the barrier that completes it
            uint32_t total = 0;
            for (uint32_t w0 = 0;; w0 += kList) {
                if (w0) __syncthreads();  // the previous window's readers of the list are done
                if (off < w0 + kList && off + cnt > w0) {
                    // the row's k-th needed x position at off + k, over the box's x positions
                    // (a loop that is the same for every lane)
                    const uint32_t at0 = off - w0;  // wraps below w0: skipped by the bound test
                    for (uint32_t b = 0; b < nx; ++b) {
                        const uint32_t at = at0 + (uint32_t)__builtin_popcount(M & ((1u << b) - 1u));
                        if (((M >> b) & 1u) && at < kList) list[at] = (uint16_t)(code | b);
                    }
                }
                __syncthreads();
                if (w0 == 0) total = nlist;  // reset for the next slot after the box is complete
                const uint32_t m = min(kList, total - w0);
                for (uint32_t i0 = thread_id(); i0 < m; i0 += kLinearPoints * kTileThreads) {
                    float x[kLinearPoints], y[kLinearPoints], z[kLinearPoints], v[kLinearPoints];
                    uint32_t at[kLinearPoints];
#pragma unroll
                    for (int p = 0; p < kLinearPoints; ++p) {
                        const uint32_t i = i0 + p * kTileThreads;
                        const uint32_t e = list[i < m ? i : i0];
                        const uint32_t ix = e & 31u, iy = (e >> 5) & 31u, iz = e >> 10;
                        x[p] = R.pc[0][ix];
                        y[p] = R.pc[1][iy];
                        z[p] = R.pc[2][iz];
                        at[p] = (iz * by + iy) * bx + ix;
                    }
                    run_linear<kLinearPoints>(R.T, n, x, y, z, v);
#pragma unroll
                    for (int p = 0; p < kLinearPoints; ++p)
                        if (i0 + p * kTileThreads < m) box[at[p]] = v[p];
                }
                if (w0 + kList >= total) break;
            }
            evaluated = total;
            if (thread_id() == 0) sh_next = next;
        } else {
            const uint32_t x0 = bxh & 0xffffffu, y0 = byh & 0xffffffu, z0 = bzh & 0xffffffu;
            const uint32_t nxy = nx * ny, npts = nxy * nz;
            const uint32_t mx = magic16(nx), mxy = magic16(nxy);
            const bool staged = n <= kStageOps;
            if (thread_id() == 0) contour_reserve(P, L, 8u * 8u * 8u);
            if (staged) stage_tape(R.T, J.ops, Lst, n);
            if (thread_id() < 24) load_tile_axes(R.A, J, tx, ty, J.tz0 + tzl, thread_id());
            if (thread_id() < 48) {
                const uint32_t a = thread_id() >> 4, i = thread_id() & 15u;
                const float* c = a == 0 ? J.cx + x0 : a == 1 ? J.cy + y0 : J.cz + z0;
                if (i < (a == 0 ? nx : a == 1 ? ny : nz)) R.pc[a][i] = c[i];
            }
            __syncthreads();
            for (uint32_t q = thread_id(); q < npts; q += kTileThreads) {
                const uint32_t iz = div16(q, nxy, mxy), r = q - iz * nxy;
                const uint32_t iy = div16(r, nx, mx), ix = r - iy * nx;
                const float x = R.pc[0][ix], y = R.pc[1][iy], z = R.pc[2][iz];
                box[(iz * by + iy) * bx + ix] =
                    staged ? run_general(R.T, n, x, y, z) : tape_eval_list(J.ops, Lst, n, x, y, z);
            }
            evaluated = npts;
        }
        if (thread_id() < 64 && ((lv >> thread_id()) & 1ull))  // the live bricks, in order
            lb[__builtin_popcountll(lv & ((1ull << thread_id()) - 1ull))] = (uint8_t)thread_id();
        __syncthreads();  // the box is complete
        if (LINEAR) {
            if (thread_id() == 0) nlist = 0;  // every thread has read the total
            prefetch();  // the next slot's record, while this one is contoured
        }
        contour_tile(P, G, slot, tx, ty, tzl, box, R.A, (uint32_t)__builtin_popcountll(lv), lb, list, L);
        if (thread_id() == 0) {
            tiles_done += 1;
            points_done += evaluated;
            flops_done += (uint64_t)evaluated * tile_flops;
        }
    }
    if (thread_id() == 0 && tiles_done) {  // the block's own entry (no shared address): k_sp_job_end sums them
        uint64_t* st = P.block_stats + 3u * blockIdx.x;
        st[0] += tiles_done;
        st[1] += points_done;
        st[2] += flops_done;
    }
}

template <int LINEAR>
__global__ void k_sp_tiles(SpPass P);
template <>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_sp_tiles<1>(SpPass P)
{
    tiles_body<1>(P);
}
template <>
__global__ __launch_bounds__(kTileThreads) void k_sp_tiles<0>(SpPass P)
{
    tiles_body<0>(P);
}

// ---------------------------------------------------------------------------------
// Quads of a cell with a vertex (mesh.rs:267-324) from its sign bits s and its look-back
// neighbours (exists(dx, dy, dz), dx, dy, dz in {-1, 0}; x, y, z: the cell in the job):
// the mask of valid quads (bit 0: x-edge, 1: y-edge, 2: z-edge).  `invalid` counts the
// quads with a missing vertex, and with the readback lag the z = 0 ones that the lagged
// reference looks up (label z - 1) and misses.
// ---------------------------------------------------------------------------------
template <typename E>
__device__ __forceinline__ uint32_t cell_quads(uint32_t s, uint32_t x, uint32_t y, uint32_t z, bool lag,
                                               const E& exists, uint32_t& invalid)
{
    const bool c0 = s & 1u, c1 = (s >> 1) & 1u, c2 = (s >> 2) & 1u, c3 = (s >> 3) & 1u;
    uint32_t nq = 0;
    if (c0 != c3 && y > 0) {  // x-edge: v(x,y-1,z-1), v(x,y,z-1), v, v(x,y-1,z)
        if (z > 0) {
            if (exists(0, -1, -1) && exists(0, 0, -1) && exists(0, -1, 0)) nq |= 1u;
            else ++invalid;
        } else if (lag) {
            ++invalid;
        }
    }
    if (c1 != c3 && x > 0) {  // y-edge: v(x-1,y,z-1), v(x,y,z-1), v, v(x-1,y,z)
        if (z > 0) {
            if (exists(-1, 0, -1) && exists(0, 0, -1) && exists(-1, 0, 0)) nq |= 2u;
            else ++invalid;
        } else if (lag) {
            ++invalid;
        }
    }
    if (c2 != c3 && x > 0 && y > 0) {  // z-edge: v(x-1,y-1,z), v(x,y-1,z), v, v(x-1,y,z)
        if (exists(-1, -1, 0) && exists(0, -1, 0) && exists(-1, 0, 0)) nq |= 4u;
        else ++invalid;
    }
    return nq;
}

// ---------------------------------------------------------------------------------
// k_sp_quadcount: per slot (one 64-lane block iteration), the valid quads of the tile's
// cells with a vertex (their staged entries), two triangles each, summed per row segment
// into the slot's kRowTri row; the invalid ones are added to the pass's and the job's
// counts.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sp_quadcount(SpPass P, MeshBufs b)
{
    __shared__ uint32_t ex[81];    // row (k, j): bit i = cell (8tx-1+i, 8ty-1+j, 8tz-1+k) has a vertex
    __shared__ uint32_t rowt[64];  // triangles of each row segment of the tile
    const uint32_t nslots = active_slots(P);
    for (uint32_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        const SlotHdr& H = P.slot_hdr[slot];
        const SpJob& J = P.jobs[__builtin_amdgcn_readfirstlane(H.job)];
        const uint32_t base = __builtin_amdgcn_readfirstlane(P.slot_stage[2 * slot]);
        const uint32_t n = __builtin_amdgcn_readfirstlane(P.slot_stage[2 * slot + 1]);
        const bool lag = (J.flags & S2M_COMPAT_DROP_LAST_SLICE) != 0;
        const uint32_t txy = __builtin_amdgcn_readfirstlane(H.txy);
        const uint32_t tx = txy & 0xffffu, ty = txy >> 16, tzl = __builtin_amdgcn_readfirstlane(H.tzl);
        __syncthreads();  // the previous slot's readers of ex / rowt are done
        int32_t ns[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ns[i] = P.slot_nb[(size_t)slot * 8u + i];
        for (uint32_t t = thread_id(); t < 81u; t += 64u)
            ex[t] = nb_row<false>(P, J, tx, ty, J.tz0 + tzl, ns, (int)(ty * 8u + t % 9u) - 1,
                                  (int)(J.z0 + tzl * 8u + t / 9u) - 1, nullptr);
        rowt[thread_id()] = 0;
        __syncthreads();
        uint32_t invalid = 0;
        for (uint32_t k = thread_id(); k < n; k += 64u) {
            if ((uint64_t)base + k >= P.stage_cap) break;
            const uint32_t meta = __float_as_uint(P.stage[(uint64_t)base + k].w);
            const uint32_t c = meta & 511u, sg = meta >> 12;
            const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
            auto exists = [&](int dx, int dy, int dz) {
                return ((ex[(czl + 1 + dz) * 9 + (cyl + 1 + dy)] >> (cxl + 1 + dx)) & 1u) != 0;
            };
            const uint32_t nq = cell_quads(sg, tx * 8u + cxl, ty * 8u + cyl, J.z0 + tzl * 8u + czl, lag, exists,
                                           invalid);
            if (nq) atomicAdd(&rowt[c >> 3], 2u * (uint32_t)__builtin_popcount(nq));
        }
        __syncthreads();
        {
            const uint32_t row = thread_id();
            slot_row(P, kRowTri, slot)[row] = rowt[row];  // 0 outside the grid
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) invalid += __shfl_xor(invalid, d, 64);
        if (thread_id() == 0 && invalid) {
            atomicAdd((unsigned long long*)&b.counters[kInvalid], (unsigned long long)invalid);
            atomicAdd((unsigned long long*)&P.job_out[(size_t)J.out_index * kJobOut + 2], (unsigned long long)invalid);
        }
    }
}

// ---------------------------------------------------------------------------------
// k_sp_emit: per slot (one 64-lane block iteration), the staged vertices at their ordered
// index (row-segment offset + rank) with key, sign bits and the 4-tap normal of the
// tile's pruned program (sdf3d_normal.wgsl:4-9, dualcontour.wgsl:171), then the
// triangles of their valid quads at the row segment's triangle offset plus the
// triangles of the cells before them in the segment (make_triangles, mesh.rs:250-265).
// Triangle vertex indices are relative to the job's first vertex.  LINEAR = 1 handles the
// slots whose program is a left fold, LINEAR = 0 the others (launched only when a program
// of the pass is not a left fold; its stack interpreter needs more registers).
// ---------------------------------------------------------------------------------
template <int LINEAR>
__global__ __launch_bounds__(64) void k_sp_emit(SpPass P, MeshBufs b)
{
    __shared__ StagedTape T;      // programs that are not left folds (left folds: read in place)
    __shared__ uint32_t nb[729];  // vertex of cell (8tx-1+i, 8ty-1+j, 8tz-1+k) at nb[(k*9+j)*9+i]
    __shared__ uint32_t tc[128];  // triangles of cell c in byte c % 4 of tc[c / 4]
    const uint32_t nslots = active_slots(P);
    for (uint32_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        const SlotHdr& H = P.slot_hdr[slot];
        const uint64_t tile = __builtin_amdgcn_readfirstlane(H.tile);
        const SpJob& J = P.jobs[__builtin_amdgcn_readfirstlane(H.job)];
        const uint32_t base = __builtin_amdgcn_readfirstlane(P.slot_stage[2 * slot]);
        const uint32_t n = __builtin_amdgcn_readfirstlane(P.slot_stage[2 * slot + 1]);
        const uint32_t txy = __builtin_amdgcn_readfirstlane(H.txy);
        const uint32_t tx = txy & 0xffffu, ty = txy >> 16, tzl = __builtin_amdgcn_readfirstlane(H.tzl);
        const bool normals = !(J.flags & S2M_NO_NORMALS);
        const bool lag = (J.flags & S2M_COMPAT_DROP_LAST_SLICE) != 0;
        const uint32_t vbase = (uint32_t)P.job_out[(size_t)J.out_index * kJobOut + 0];
        const uint16_t* Lst = P.lists + (size_t)tile * P.list_stride;
        const uint32_t lens = __builtin_amdgcn_readfirstlane(H.lens);
        const uint32_t len = lens & ~kLinearFlag;
        const bool linear = (lens & kLinearFlag) != 0, staged = len <= kStageOps;
        if (linear != (LINEAR != 0)) continue;  // the other launch's slot
        __syncthreads();  // the previous slot's readers of T, nb, tc are done
        if (!LINEAR && normals && staged) stage_tape(T, J.ops, Lst, len);
        // left folds: the program as staged in the slot's record, read in place
        const ScalarFold TS =
            ScalarFold::at(reinterpret_cast<const char*>(P.slot_rec + (size_t)slot * kRecVec) + offsetof(TileRec, T));
        int32_t ns[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ns[i] = P.slot_nb[(size_t)slot * 8u + i];
        for (uint32_t t = thread_id(); t < 81u; t += 64u)
            (void)nb_row<true>(P, J, tx, ty, J.tz0 + tzl, ns, (int)(ty * 8u + t % 9u) - 1,
                               (int)(J.z0 + tzl * 8u + t / 9u) - 1, nb + 9u * t);
        for (uint32_t w = thread_id(); w < 128u; w += 64u) tc[w] = 0;
        __syncthreads();
        // vertices, and each cell's triangle count (the first round's cell, quads and row
        // triangle offset are kept for the triangle loop)
        uint32_t meta0 = 0, nq0 = 0;
        uint64_t trow0 = 0;
        for (uint32_t k = thread_id(); k < n; k += 64u) {
            if ((uint64_t)base + k >= P.stage_cap) break;
            const float4 e = P.stage[(uint64_t)base + k];
            const uint32_t meta = __float_as_uint(e.w);
            const uint32_t c = meta & 511u, sg = meta >> 12;
            const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
            const uint32_t at = ((czl + 1) * 9u + (cyl + 1)) * 9u + (cxl + 1);
            auto exists = [&](int dx, int dy, int dz) { return nb[(int)at + (dz * 9 + dy) * 9 + dx] != kNone; };
            uint32_t invalid = 0;
            const uint32_t nq =
                cell_quads(sg, tx * 8u + cxl, ty * 8u + cyl, J.z0 + tzl * 8u + czl, lag, exists, invalid);
            if (nq) atomicOr(&tc[c >> 2], (2u * (uint32_t)__builtin_popcount(nq)) << ((c & 3u) * 8u));
            if (k == thread_id()) {
                meta0 = meta;
                nq0 = nq;
                if (nq)
                    trow0 = (uint64_t)slot_row(P, kRowTriOff, slot)[c >> 3] +
                            P.row_off[P.nrows + row_index(J, tzl * 8u + czl, ty * 8u + cyl)];
            }
            const uint64_t vi = nb[at];
            if (vi >= b.vcap) continue;
            float* o = b.vtx + 6 * vi;
            o[0] = e.x;
            o[1] = e.y;
            o[2] = e.z;
            b.keys[vi] = (uint64_t)(tx * 8u + cxl) | ((uint64_t)(ty * 8u + cyl) << 16) |
                         ((uint64_t)(J.z0 + tzl * 8u + czl) << 32);
            b.signs[vi] = (uint8_t)sg;
            if (!normals) {
                o[3] = o[4] = o[5] = 0.0f;
                continue;
            }
            const float px = e.x, py = e.y, pz = e.z, eps = J.eps;
            // sdf3d_normal (sdf3d_normal.wgsl:4-9): taps p + v_i * eps, then normalize
            const float tx4[4] = {px + eps, px + (-eps), px + (-eps), px + eps};
            const float ty4[4] = {py + (-eps), py + (-eps), py + eps, py + eps};
            const float tz4[4] = {pz + (-eps), pz + eps, pz + (-eps), pz + eps};
            float d[4];
            if constexpr (LINEAR != 0) {
                run_linear<4>(TS, len, tx4, ty4, tz4, d);
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    d[t] = staged ? run_general(T, len, tx4[t], ty4[t], tz4[t])
                                  : tape_eval_list(J.ops, Lst, len, tx4[t], ty4[t], tz4[t]);
            }
            const float d1 = d[0], d2 = d[1], d3 = d[2], d4 = d[3];
            const float gx = ((d1 + (-d2)) + (-d3)) + d4;
            const float gy = (((-d1) + (-d2)) + d3) + d4;
            const float gz = (((-d1) + d2) + (-d3)) + d4;
            const float ln = sqrtf((gx * gx + gy * gy) + gz * gz);
            o[3] = canonical_nan(gx / ln);
            o[4] = canonical_nan(gy / ln);
            o[5] = canonical_nan(gz / ln);
        }
        __syncthreads();  // every cell's triangle count is in tc
        // triangles, in vertex order: each quad (q2, q1, q0), (q0, q3, q2)
        for (uint32_t k = thread_id(); k < n; k += 64u) {
            if ((uint64_t)base + k >= P.stage_cap) break;
            const bool first = k == thread_id();
            const uint32_t meta = first ? meta0 : __float_as_uint(P.stage[(uint64_t)base + k].w);
            const uint32_t c = meta & 511u, sg = meta >> 12;
            const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
            const uint32_t at = ((czl + 1) * 9u + (cyl + 1)) * 9u + (cxl + 1);
            auto nbv = [&](int dx, int dy, int dz) { return nb[(int)at + (dz * 9 + dy) * 9 + dx]; };
            auto exists = [&](int dx, int dy, int dz) { return nbv(dx, dy, dz) != kNone; };
            uint32_t invalid = 0;
            const uint32_t y = ty * 8u + cyl, zl = tzl * 8u + czl;
            const uint32_t nq = first ? nq0 : cell_quads(sg, tx * 8u + cxl, y, J.z0 + zl, lag, exists, invalid);
            if (!nq) continue;
            const uint64_t row2 = ((uint64_t)tc[(c >> 3) * 2u + 1] << 32) | tc[(c >> 3) * 2u];  // 8 counts
            uint64_t t = first ? trow0
                               : (uint64_t)slot_row(P, kRowTriOff, slot)[c >> 3] + P.row_off[P.nrows + row_index(J, zl, y)];
            for (uint32_t i = 0; i < cxl; ++i) t += (row2 >> (8u * i)) & 0xffu;
            const uint32_t v = nbv(0, 0, 0);
            const bool c0 = sg & 1u, c1 = (sg >> 1) & 1u, c2 = (sg >> 2) & 1u;
            for (int q = 0; q < 3; ++q) {
                if (!(nq & (1u << q))) continue;
                uint32_t f[4];
                bool rev;
                if (q == 0) {
                    f[0] = nbv(0, -1, -1), f[1] = nbv(0, 0, -1), f[3] = nbv(0, -1, 0), rev = c0;
                } else if (q == 1) {
                    f[0] = nbv(-1, 0, -1), f[1] = nbv(0, 0, -1), f[3] = nbv(-1, 0, 0), rev = !c1;
                } else {
                    f[0] = nbv(-1, -1, 0), f[1] = nbv(0, -1, 0), f[3] = nbv(-1, 0, 0), rev = c2;
                }
                f[2] = v;
                uint32_t qv[4];
                for (int i = 0; i < 4; ++i) qv[i] = (rev ? f[3 - i] : f[i]) - vbase;
                if (t + 1 < b.tcap) {
                    uint32_t* o = b.tri + 3 * t;
                    o[0] = qv[2];
                    o[1] = qv[1];
                    o[2] = qv[0];
                    o[3] = qv[0];
                    o[4] = qv[3];
                    o[5] = qv[2];
                }
                t += 2;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Offsets in the reference's VertexList order: job after job, then (z, y, tile x).
//  k_sp_rowsum    one wave per tile row (job, tile layer, ty), lane = row (z % 8, y % 8):
//                 the row segments of the tile row's active tiles (the super-tiles'
//                 activity words) in x order: each segment's offset within its row
//                 (the slot's kRowVtxOff / kRowTriOff rows) and the row's totals (row_sum)
//                 and each 4096-row chunk's sums (one atomic per wave and chunk)
//  k_sp_rowdown   exclusive scan of the rows' totals (row_off), block b adding the sums of
//                 the chunks before it
// A segment's first vertex (triangle) is row_off[row] + its offset within the row.
// ---------------------------------------------------------------------------------
constexpr uint32_t kScanPer = 16;                  // consecutive rows per thread
constexpr uint32_t kScanChunk = 256 * kScanPer;
static_assert(kScanChunk == kRowChunk, "rows per chunk of the offset scan");

__global__ __launch_bounds__(64) void k_sp_rowsum(SpPass P, uint32_t* __restrict__ bsum)
{
    const uint64_t tr = blockIdx.x;
    uint32_t lo = 0, hi = P.njobs - 1;  // last job whose first tile row <= tr
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (P.jobs[mid].trow_base <= tr) lo = mid;
        else hi = mid - 1;
    }
    const SpJob& J = P.jobs[__builtin_amdgcn_readfirstlane(lo)];
    const uint32_t local = (uint32_t)(tr - J.trow_base), ty = local % J.nty, tzl = local / J.nty;
    const uint32_t zl = tzl * 8u + (lane() >> 3), y = ty * 8u + (lane() & 7u);
    const bool row_ok = zl < J.nz && y < J.ry;
    const uint64_t srow = J.super_base + ((uint64_t)(tzl / kSuper) * J.nsy + ty / kSuper) * J.nsx;
    const uint32_t bit0 = ((tzl % kSuper) * kSuper + ty % kSuper) * kSuper;
    const uint64_t trow_tile = J.tile_base + ((uint64_t)tzl * J.nty + ty) * J.ntx;  // tile (0, ty, tzl)
    uint32_t sv = 0, st = 0;
    for (uint32_t sx0 = 0; sx0 < J.nsx; sx0 += 16u) {
        // active tiles 4 * sx0 .. 4 * sx0 + 63: lane l < 16 reads super-tile sx0 + l's word
        uint64_t m = 0;
        if (lane() < 16u && sx0 + lane() < J.nsx)
            m = ((P.super_act[srow + sx0 + lane()] >> bit0) & 0xfull) << (4u * lane());
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1) m |= __shfl_xor(m, d, 64);
        m = __shfl(m, 0, 64);
        // the active tiles' row counts (one 256-byte row of each slot), 8 tiles at a time
        while (m) {
            uint32_t slots[8], v[8], t[8];
            uint32_t nb = 0;
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                slots[k] = 0;
                if (m) {
                    const uint32_t tx = 4u * sx0 + (uint32_t)__builtin_ctzll(m);
                    m &= m - 1ull;
                    slots[k] = (uint32_t)__builtin_amdgcn_readfirstlane(P.slot_of_tile[trow_tile + tx]);
                    nb = k + 1;
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                v[k] = t[k] = 0;
                if (k < nb) {
                    v[k] = slot_row(P, kRowVtx, slots[k])[lane()];
                    t[k] = slot_row(P, kRowTri, slots[k])[lane()];
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                if (k < nb) {
                    slot_row(P, kRowVtxOff, slots[k])[lane()] = sv;
                    slot_row(P, kRowTriOff, slots[k])[lane()] = st;
                }
                sv += v[k];
                st += t[k];
            }
        }
    }
    const uint64_t row = row_ok ? row_index(J, zl, y) : 0;
    if (row_ok) {
        P.row_sum[row] = sv;
        P.row_sum[P.nrows + row] = st;
    }
    // the rows' totals into their chunks' sums (bsum: per chunk a vertex and a triangle sum,
    // kChunkSumStride words apart), one atomic per chunk the wave's rows fall in
    const uint32_t chunk = (uint32_t)(row / kScanChunk);
    uint64_t rem = __ballot(row_ok && (sv | st));
    while (rem) {
        const uint32_t c = __shfl(chunk, (int)__builtin_ctzll(rem), 64);
        const bool mine = ((rem >> lane()) & 1ull) && chunk == c;
        uint32_t v = mine ? sv : 0u, t = mine ? st : 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            v += __shfl_xor(v, d, 64);
            t += __shfl_xor(t, d, 64);
        }
        if (lane() == (uint32_t)__builtin_ctzll(rem)) {
            if (v) atomicAdd(&bsum[(2u * c) * kChunkSumStride], v);
            if (t) atomicAdd(&bsum[(2u * c + 1u) * kChunkSumStride], t);
        }
        rem &= ~__ballot(mine);
    }
}

// Exclusive offsets of the rows, block b = rows b * kScanChunk ...: the pass's first index
// (kVtxPrev / kTriPrev) plus the chunk sums of the blocks before it (bsum) plus the
// block-local scan.  The last block writes the totals after the pass (kVtxTotal / kTriTotal).
__global__ __launch_bounds__(256) void k_sp_rowdown(SpPass P, const uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t lds[2][4];
    __shared__ uint32_t pre[2][4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + thread_id() * kScanPer;
    uint32_t v[kScanPer], t[kScanPer], sv = 0, st = 0;
    for (uint32_t k = 0; k < kScanPer; ++k) {
        const bool in = base + k < P.nrows;
        v[k] = in ? P.row_sum[base + k] : 0u;
        t[k] = in ? P.row_sum[P.nrows + base + k] : 0u;
        sv += v[k];
        st += t[k];
    }
    uint32_t iv = sv, it = st;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t a = __shfl_up(iv, d, 64), b = __shfl_up(it, d, 64);
        if (lane() >= (uint32_t)d) {
            iv += a;
            it += b;
        }
    }
    if (lane() == 63) {
        lds[0][thread_id() >> 6] = iv;
        lds[1][thread_id() >> 6] = it;
    }
    {  // the chunks before this block
        uint32_t cv = 0, ct = 0;
        for (uint32_t c = thread_id(); c < blockIdx.x; c += blockDim.x) {
            cv += bsum[(2u * c) * kChunkSumStride];
            ct += bsum[(2u * c + 1u) * kChunkSumStride];
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            cv += __shfl_xor(cv, d, 64);
            ct += __shfl_xor(ct, d, 64);
        }
        if (lane() == 0) {
            pre[0][thread_id() >> 6] = cv;
            pre[1][thread_id() >> 6] = ct;
        }
    }
    __syncthreads();
    const uint32_t bv = (uint32_t)P.counters[kVtxPrev] + pre[0][0] + pre[0][1] + pre[0][2] + pre[0][3];
    const uint32_t bt = (uint32_t)P.counters[kTriPrev] + pre[1][0] + pre[1][1] + pre[1][2] + pre[1][3];
    if (blockIdx.x + 1 == gridDim.x && thread_id() == 0) {  // after the pass
        P.counters[kVtxTotal] = bv + lds[0][0] + lds[0][1] + lds[0][2] + lds[0][3];
        P.counters[kTriTotal] = bt + lds[1][0] + lds[1][1] + lds[1][2] + lds[1][3];
    }
    uint32_t rv = bv + iv - sv, rt = bt + it - st;
    for (uint32_t w = 0; w < (thread_id() >> 6); ++w) {
        rv += lds[0][w];
        rt += lds[1][w];
    }
    for (uint32_t k = 0; k < kScanPer && base + k < P.nrows; ++k) {
        P.row_off[base + k] = rv;
        P.row_off[P.nrows + base + k] = rt;
        rv += v[k];
        rt += t[k];
    }
    // the first vertex and triangle index of every job whose first slab this pass holds and
    // whose first row is this block's: the exclusive offsets of that row
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * kScanChunk, hi = lo + kScanChunk;
    for (uint32_t j = thread_id(); j < P.njobs; j += blockDim.x) {
        const SpJob& J = P.jobs[j];
        if (J.z0 != 0 || (uint64_t)J.nz * J.ry == 0 || J.row_base < lo || J.row_base >= hi) continue;
        uint64_t* o = P.job_out + (size_t)J.out_index * kJobOut;
        o[0] = (uint64_t)P.row_off[J.row_base];
        o[1] = (uint64_t)P.row_off[P.nrows + J.row_base];
    }
}

// Cell -> vertex map of the slab's last plane, for the next slab (single job).
__global__ __launch_bounds__(256) void k_sp_boundary(SpPass P)
{
    const SpJob& J = P.jobs[0];
    const uint32_t i = blockIdx.x * blockDim.x + thread_id();
    if (i >= J.rx * J.ry) return;
    J.boundary[i] = vertex_at(P, J, i % J.rx, i / J.rx, J.z0 + J.nz - 1);
}

// Per-pass reset of the slot counter; the largest count seen is kept for the host's
// capacity check (a kernel rather than a memset, so a pass can be graph-captured).
__global__ void k_sp_pass_begin(uint64_t* counters, uint32_t* queues, uint32_t* bsum, uint32_t nbsum)
{
    if (thread_id() == 0 && blockIdx.x == 0) {
        const uint64_t used = counters[kSlotsUsed];
        if (used > counters[kSlotsMax]) counters[kSlotsMax] = used;
        counters[kSlotsUsed] = 0;
        if (counters[kStageUsed] > counters[kStageMax]) counters[kStageMax] = counters[kStageUsed];
        counters[kStageUsed] = 0;
        counters[kVtxPrev] = counters[kVtxTotal];  // the next pass's first vertex and triangle
        counters[kTriPrev] = counters[kTriTotal];
    }
    for (uint32_t i = thread_id(); i < kQueues * kQueueGroups; i += blockDim.x) queues[i * 32u] = 0;
    for (uint32_t i = thread_id(); i < nbsum; i += blockDim.x) bsum[i * kChunkSumStride] = 0;
}

// After a pass (or a dense slab): the totals so far and the pass's use of the slot and
// staging capacities, written to `out` (host memory mapped for the device) so that the
// host can copy the pass's finished vertices and triangles while later passes run.
__global__ void k_snapshot(const uint64_t* counters, uint64_t* out)
{
    if (thread_id() == 0 && blockIdx.x == 0) {
        out[0] = counters[kVtxTotal];
        out[1] = counters[kTriTotal];
        out[2] = counters[kSlotsUsed];
        out[3] = counters[kStageUsed];
    }
}

__global__ __launch_bounds__(256) void k_sp_job_begin(uint64_t* counters, uint32_t n, uint32_t* queues,
                                                      uint32_t* bsum, uint32_t nbsum, uint64_t* stats,
                                                      uint32_t nstats, uint32_t* boundary, uint32_t nbnd)
{
    if (blockIdx.x == 0) {
        for (uint32_t i = thread_id(); i < n; i += blockDim.x) counters[i] = 0;
        for (uint32_t i = thread_id(); i < 3u * nstats; i += blockDim.x) stats[i] = 0;
        for (uint32_t i = thread_id(); i < kQueues * kQueueGroups; i += blockDim.x) queues[i * 32u] = 0;
        for (uint32_t i = thread_id(); i < nbsum; i += blockDim.x) bsum[i * kChunkSumStride] = 0;
    }
    // a multi-pass job's boundary plane: no vertex carried yet
    for (uint32_t i = blockIdx.x * blockDim.x + thread_id(); i < nbnd; i += gridDim.x * blockDim.x)
        boundary[i] = 0xffffffffu;
}

// After a job's last pass: the pass folded into the maxima as k_sp_pass_begin does, then
// counters[0 .. n) copied to `out` (host memory mapped for the device).
__global__ __launch_bounds__(256) void k_sp_job_end(uint64_t* counters, uint32_t n, uint64_t* out,
                                                    const uint64_t* stats, uint32_t nstats)
{
    __shared__ uint64_t part[4][3];
    uint64_t sum[3] = {0, 0, 0};  // the tile blocks' tiles, points and flops
    for (uint32_t i = thread_id(); i < nstats; i += blockDim.x)
        for (int k = 0; k < 3; ++k) sum[k] += stats[3u * i + k];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        for (int d = 32; d >= 1; d >>= 1) sum[k] += __shfl_xor(sum[k], d, 64);
    if (lane() == 0)
        for (int k = 0; k < 3; ++k) part[thread_id() >> 6][k] = sum[k];
    __syncthreads();
    if (thread_id() == 0) {
        if (counters[kSlotsUsed] > counters[kSlotsMax]) counters[kSlotsMax] = counters[kSlotsUsed];
        if (counters[kStageUsed] > counters[kStageMax]) counters[kStageMax] = counters[kStageUsed];
        counters[kTilesActive] = part[0][0] + part[1][0] + part[2][0] + part[3][0];
        counters[kEvalPoints] = part[0][1] + part[1][1] + part[2][1] + part[3][1];
        counters[kEvalFlops] = part[0][2] + part[1][2] + part[2][2] + part[3][2];
    }
    __syncthreads();
    for (uint32_t i = thread_id(); i < n; i += blockDim.x) out[i] = counters[i];
}

// ---------------------------------------------------------------------------------
hipError_t launch_sparse_job_end(hipStream_t s, uint64_t* counters, uint32_t n, uint64_t* out,
                                 const uint64_t* stats, uint32_t nstats)
{
    hipLaunchKernelGGL(k_sp_job_end, dim3(1), dim3(256), 0, s, counters, n, out, stats, nstats);
    return hipGetLastError();
}

hipError_t launch_snapshot(hipStream_t s, const uint64_t* counters, uint64_t* out)
{
    hipLaunchKernelGGL(k_snapshot, dim3(1), dim3(64), 0, s, counters, out);
    return hipGetLastError();
}

hipError_t launch_sparse_pass_begin(hipStream_t s, uint64_t* counters, uint32_t* queues, uint32_t* bsum,
                                    uint32_t nbsum)
{
    hipLaunchKernelGGL(k_sp_pass_begin, dim3(1), dim3(256), 0, s, counters, queues, bsum, nbsum);
    return hipGetLastError();
}

hipError_t launch_sparse_job_begin(hipStream_t s, uint64_t* counters, uint32_t n, uint32_t* queues,
                                   uint32_t* bsum, uint32_t nbsum, uint64_t* stats, uint32_t nstats,
                                   uint32_t* boundary, uint32_t nbnd)
{
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((nbnd + 255u) / 256u, 256u));
    hipLaunchKernelGGL(k_sp_job_begin, dim3(blocks), dim3(256), 0, s, counters, n, queues, bsum, nbsum, stats,
                       nstats, boundary, nbnd);
    return hipGetLastError();
}

// Blocks of a persistent launch: as many as the device holds at once (at most `cap`),
// so no block starts only when others end.  The occupancy query is slow (tens of
// microseconds), so its answer is cached per (kernel, block size, LDS bytes).
template <typename K>
static uint32_t resident_blocks(K kernel, int threads, size_t smem, int cus, uint32_t cap)
{
    struct Entry {
        const void* f;
        int threads;
        size_t smem;
        int per_cu;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    const void* f = reinterpret_cast<const void*>(kernel);
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (const Entry& e : cache)
            if (e.f == f && e.threads == threads && e.smem == smem) per_cu = e.per_cu;
    }
    if (per_cu == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, threads, smem) != hipSuccess || per_cu < 1)
            per_cu = 1;
        std::lock_guard<std::mutex> lock(mu);
        cache.push_back(Entry{f, threads, smem, per_cu});
    }
    return std::min<uint32_t>(cap, (uint32_t)cus * (uint32_t)per_cu);
}

hipError_t launch_sparse_pass(hipStream_t s, const SpPass& P, const MeshBufs& b, int cus,
                              uint32_t tile_blocks, bool general_tiles, bool fold_jobs, bool other_jobs,
                              int boundary_plane, hipEvent_t* marks)
{
    if (P.ntiles == 0) return hipSuccess;
    const uint64_t nsb64 = (P.nrows + kScanChunk - 1) / kScanChunk;
    if (nsb64 > kMaxScanBlocks) return hipErrorInvalidValue;
    const uint32_t nsb = (uint32_t)nsb64;
    const uint32_t big = 1u << 20;
    auto mark = [&](int i) {
        if (marks && marks[i]) (void)hipEventRecord(marks[i], s);
    };
    const size_t prune_smem = PruneStack::bytes(P.prune_depth);
    const uint32_t chunks = (P.list_stride + 63) / 64;
    if (P.nsuper * chunks > 0x7fffffffull) return hipErrorInvalidValue;
    mark(0);
    if (fold_jobs) hipLaunchKernelGGL(k_sp_prune_fold, dim3((uint32_t)P.nsuper), dim3(64), 0, s, P);
    if (other_jobs) {
        hipLaunchKernelGGL(k_sp_prune_prims, dim3((uint32_t)(P.nsuper * chunks)), dim3(64), 0, s, P, chunks);
        hipLaunchKernelGGL(k_sp_prune_super, dim3((uint32_t)((P.nsuper + 63) / 64)), dim3(64), prune_smem, s, P);
    }
    hipLaunchKernelGGL(k_sp_prune, dim3((uint32_t)P.nsuper), dim3(64), prune_smem, s, P);
    hipLaunchKernelGGL(k_sp_bricks, dim3(resident_blocks(k_sp_bricks, 64, 0, cus, big)), dim3(64), 0, s, P);
    mark(1);
    const size_t smem = (size_t)P.box_stride * sizeof(float);
    hipLaunchKernelGGL((k_sp_tiles<1>), dim3(resident_blocks(k_sp_tiles<1>, kTileThreads, smem, cus, tile_blocks)),
                       dim3(kTileThreads), smem, s, P);
    if (general_tiles)
        hipLaunchKernelGGL((k_sp_tiles<0>), dim3(resident_blocks(k_sp_tiles<0>, kTileThreads, smem, cus, tile_blocks)),
                           dim3(kTileThreads), smem, s, P);
    mark(2);
    hipLaunchKernelGGL(k_sp_quadcount, dim3(resident_blocks(k_sp_quadcount, 64, 0, cus, big)), dim3(64), 0, s,
                       P, b);
    mark(3);
    // b.bcount: the chunk sums (2 x nsb, cleared by the pass's begin kernel)
    if (P.ntrows) hipLaunchKernelGGL(k_sp_rowsum, dim3((uint32_t)P.ntrows), dim3(64), 0, s, P, b.bcount);
    if (nsb) hipLaunchKernelGGL(k_sp_rowdown, dim3(nsb), dim3(256), 0, s, P, b.bcount);
    mark(4);
    hipLaunchKernelGGL((k_sp_emit<1>), dim3(resident_blocks(k_sp_emit<1>, 64, 0, cus, big)), dim3(64), 0, s, P, b);
    if (general_tiles)
        hipLaunchKernelGGL((k_sp_emit<0>), dim3(resident_blocks(k_sp_emit<0>, 64, 0, cus, big)), dim3(64), 0, s, P, b);
    if (boundary_plane) {
        hipLaunchKernelGGL(k_sp_boundary, dim3((uint32_t)((P.boundary_cells + 255) / 256)), dim3(256), 0,
                           s, P);
    }
    mark(5);
    return hipGetLastError();
}

}  // namespace s2m
