// sparse.hip - the tile-sparse meshing pipeline, for one job or a batch of jobs.
//
// Programs that interval pruning applies to (every library primitive and operator except
// the Mandelbulb) are meshed tile by tile: cells are grouped in tiles of 8x8x8 and only
// tiles whose SDF can change sign are evaluated and contoured.  A pass covers the tiles
// of one slab of one job, or of every job of a batch (each job whole); the tiles of the
// pass are numbered job after job.  In stream order:
//   k_sp_prune      one thread per tile: interval bounds of the job's tape on the tile's
//                   padded box; single-signed tiles are dropped, the others get a slot in
//                   a pool shared by all jobs and an exactly pruned op list
//   k_sp_eval       per slot (grid-stride, persistent): the pruned tape at every lattice
//                   point of the tile's closed box, stored in the slot's compact box
//   k_sp_cells<0>   per slot: active cells per row segment (z, y, tile x)
//   reduce / k_scan / downsweep   exclusive scan of the row segments, job after job and
//                   (z, y, tile x) inside a job: the reference's VertexList order
//   k_sp_ranges<0>  each job's first vertex index
//   k_sp_cells<1>   vertices at their ordered index, the slot's cell -> vertex box, and
//                   the 4-tap normals of the tile's vertices with its pruned tape
//   k_sp_quads<0/1> the same count / scan / emit scheme for triangles (vertex order), with
//                   vertex indices relative to the job's first vertex
//   k_sp_ranges<1>  each job's first triangle index
//   k_sp_boundary   (multi-slab single job) cell -> vertex map of the slab's last plane
// Loops over slots read their bound on the device: no host round trip inside a pass.
#include "device_interval.h"
#include "device_sdf.h"
#include "internal.h"

namespace s2m {

namespace {

__device__ __forceinline__ uint32_t lane() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t active_slots(const SpPass& P)
{
    const uint64_t n = P.counters[kSlotsUsed];
    return (uint32_t)(n < P.slot_cap ? n : P.slot_cap);
}

// Row segment of row (zl, y) and tile column tx of job J.
__device__ __forceinline__ size_t seg_index(const SpJob& J, uint32_t zl, uint32_t y, uint32_t tx)
{
    return J.seg_base + ((size_t)zl * J.ry + y) * J.ntx + tx;
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
    return P.vbox[(size_t)slot * 512u + (((z & 7u) * 8u + (y & 7u)) * 8u + (x & 7u))];
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
    pos[0] = lo[0] + ((hi[0] - lo[0]) * ax) / cnt;
    pos[1] = lo[1] + ((hi[1] - lo[1]) * ay) / cnt;
    pos[2] = lo[2] + ((hi[2] - lo[2]) * az) / cnt;
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sp_prune(SpPass P)
{
    const uint64_t tile = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= P.ntiles) return;
    // job of the tile: jobs are sorted by tile_base
    uint32_t lo = 0, hi = P.njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (P.jobs[mid].tile_base <= tile) lo = mid;
        else hi = mid - 1;
    }
    const SpJob& J = P.jobs[lo];
    uint32_t tx, ty, tzl;
    tile_xyz(J, tile, tx, ty, tzl);
    const uint32_t tz = J.tz0 + tzl;
    const float* ex = J.ext;
    const float* ey = ex + 2 * J.ntx;
    const float* ez = ey + 2 * J.nty;
    const double bl[3] = {(double)ex[tx] - J.pad, (double)ey[ty] - J.pad, (double)ez[tz] - J.pad};
    const double bh[3] = {(double)ex[J.ntx + tx] + J.pad, (double)ey[J.nty + ty] + J.pad,
                          (double)ez[J.ntz + tz] + J.pad};
    const s2m_op* ops = J.ops;
    uint16_t* L = P.lists + (size_t)tile * P.list_stride;
    Iv st[S2M_MAX_STACK];
    uint32_t start[S2M_MAX_STACK];
    uint32_t sp = 0, len = 0;
    for (uint32_t i = 0; i < J.nops; ++i) {
        const uint32_t code = ops[i].code;
        if (code < S2M_OP_UNION) {
            st[sp] = iv_primitive(code, ops[i].f, bl, bh);
            start[sp] = len;
            ++sp;
            L[len++] = (uint16_t)i;
        } else {
            const Iv B = st[--sp];
            const uint32_t sB = start[sp];
            const Iv A = st[--sp];
            const uint32_t sA = start[sp];
            int keep;
            const Iv r = iv_op(code, A, B, ops[i].f[0], keep);
            if (keep == kKeepBoth) {
                L[len++] = (uint16_t)i;
                st[sp] = r;
            } else if (keep == kKeepA) {
                len = sB;  // drop B's subtree and the operator
                st[sp] = A;
            } else {       // drop A's subtree and the operator
                const uint32_t nb = len - sB;
                for (uint32_t j = 0; j < nb; ++j) L[sA + j] = L[sB + j];
                len = sA + nb;
                st[sp] = B;
            }
            start[sp] = sA;
            ++sp;
        }
    }
    const Iv root = st[0];
    const bool one_sign = root.known && (root.lo > 0.0 || root.hi < 0.0);
    uint32_t fl = 0;
    for (uint32_t j = 0; j < len; ++j) fl += J.op_flops[L[j]];
    P.lens[tile] = len;
    P.flops[tile] = fl;
    int32_t slot = -1;
    if (!one_sign) {
        const uint64_t s = atomicAdd((unsigned long long*)&P.counters[kSlotsUsed], 1ull);
        if (s < P.slot_cap) {
            slot = (int32_t)s;
            P.tile_of_slot[s] = (uint32_t)tile;
            P.job_of_slot[s] = lo;
        }
    }
    P.slot_of_tile[tile] = slot;
    if (slot < 0) {  // no vertex in this tile: zero its row segments for both scans
        const uint32_t z1 = min(tzl * 8u + 8u, J.nz), y1 = min(ty * 8u + 8u, J.ry);
        for (uint32_t zl = tzl * 8u; zl < z1; ++zl)
            for (uint32_t y = ty * 8u; y < y1; ++y) {
                P.seg_v[seg_index(J, zl, y, tx)] = 0;
                P.seg_t[seg_index(J, zl, y, tx)] = 0;
            }
    }
}

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sp_eval(SpPass P)
{
    const uint32_t nslots = active_slots(P);
    for (uint32_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        const uint64_t tile = P.tile_of_slot[slot];
        const SpJob& J = P.jobs[P.job_of_slot[slot]];
        uint32_t tx, ty, tzl;
        tile_xyz(J, tile, tx, ty, tzl);
        const uint32_t tz = J.tz0 + tzl;
        const uint32_t* bxr = J.ibox;
        const uint32_t* byr = bxr + 2 * J.ntx;
        const uint32_t* bzr = byr + 2 * J.nty;
        const uint32_t x0 = bxr[tx], nx = bxr[J.ntx + tx] - x0 + 1;
        const uint32_t y0 = byr[ty], ny = byr[J.nty + ty] - y0 + 1;
        const uint32_t z0 = bzr[tz], nz = bzr[J.ntz + tz] - z0 + 1;
        const uint32_t npts = nx * ny * nz;
        const uint16_t* L = P.lists + (size_t)tile * P.list_stride;
        const uint32_t n = P.lens[tile];
        float* box = P.boxes + (size_t)slot * P.box_stride;
        for (uint32_t q = threadIdx.x; q < npts; q += blockDim.x) {
            const uint32_t ix = q % nx, iy = (q / nx) % ny, iz = q / (nx * ny);
            box[(iz * J.by + iy) * J.bx + ix] =
                tape_eval_list(J.ops, L, n, J.cx[x0 + ix], J.cy[y0 + iy], J.cz[z0 + iz]);
        }
        if (threadIdx.x == 0) {
            atomicAdd((unsigned long long*)&P.counters[kTilesActive], 1ull);
            atomicAdd((unsigned long long*)&P.counters[kEvalPoints], (unsigned long long)npts);
            atomicAdd((unsigned long long*)&P.counters[kEvalFlops],
                      (unsigned long long)npts * P.flops[tile]);
        }
    }
}

// ---------------------------------------------------------------------------------
// Cells of one slot per block iteration; thread t owns cells t and t + 256 of the tile
// (x fastest), so the 8 cells of a row segment are 8 consecutive lanes.
// ---------------------------------------------------------------------------------
template <int EMIT>
__global__ __launch_bounds__(256) void k_sp_cells(SpPass P, MeshBufs b)
{
    extern __shared__ float box[];  // P.box_stride floats
    __shared__ float vpos[512 * 3];
    __shared__ uint32_t vidx_l[512];
    __shared__ uint32_t nact;
    const uint32_t nslots = active_slots(P);
    for (uint32_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        const uint64_t tile = P.tile_of_slot[slot];
        const SpJob& J = P.jobs[P.job_of_slot[slot]];
        uint32_t tx, ty, tzl;
        tile_xyz(J, tile, tx, ty, tzl);
        const uint32_t tz = J.tz0 + tzl;
        const uint32_t* bxr = J.ibox;
        const uint32_t* byr = bxr + 2 * J.ntx;
        const uint32_t* bzr = byr + 2 * J.nty;
        const uint32_t x0 = bxr[tx], y0 = byr[ty], z0l = bzr[tz];
        const float* gbox = P.boxes + (size_t)slot * P.box_stride;
        __syncthreads();  // the previous slot's readers of box / vpos are done
        for (uint32_t q = threadIdx.x; q < P.box_stride; q += blockDim.x) box[q] = gbox[q];
        if (threadIdx.x == 0) nact = 0;
        __syncthreads();
        for (uint32_t r = 0; r < 2; ++r) {
            const uint32_t c = r * 256u + threadIdx.x;
            const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
            const uint32_t x = tx * 8u + cxl, y = ty * 8u + cyl, zl = tzl * 8u + czl;
            const uint32_t z = J.z0 + zl;
            const bool inside = x < J.rx && y < J.ry && zl < J.nz;
            bool act = false;
            float pos[3] = {0.0f, 0.0f, 0.0f}, cv[8];
            uint32_t sg = 0;
            if (inside) {
                const uint32_t xi[2] = {J.lox[x] - x0, J.hix[x] - x0};
                const uint32_t yi[2] = {J.loy[y] - y0, J.hiy[y] - y0};
                const uint32_t zi[2] = {J.loz[z] - z0l, J.hiz[z] - z0l};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    cv[k] = box[(zi[(k >> 2) & 1] * J.by + yi[(k >> 1) & 1]) * J.bx + xi[k & 1]];
                const float lo[3] = {J.cx[J.lox[x]], J.cy[J.loy[y]], J.cz[J.loz[z]]};
                const float hi[3] = {J.cx[J.hix[x]], J.cy[J.hiy[y]], J.cz[J.hiz[z]]};
                act = contour_cell_sp(cv, lo, hi, pos);
                // cell_sign_changes (dualcontour.wgsl:57-69)
                sg = (uint32_t)(cv[1] > 0.0f) | ((uint32_t)(cv[2] > 0.0f) << 1) |
                     ((uint32_t)(cv[4] > 0.0f) << 2) | ((uint32_t)(cv[0] > 0.0f) << 3);
            }
            const uint64_t m = __ballot(act);
            const uint32_t g0 = lane() & ~7u;
            const uint32_t gmask = (uint32_t)((m >> g0) & 0xffull);
            const uint32_t below = (uint32_t)__builtin_popcount(gmask & ((1u << (lane() - g0)) - 1u));
            const bool row_ok = y < J.ry && zl < J.nz;
            if (!EMIT) {
                if (cxl == 0 && row_ok) P.seg_v[seg_index(J, zl, y, tx)] = (uint32_t)__builtin_popcount(gmask);
                continue;
            }
            uint32_t v = kNone;
            if (act) {
                const uint64_t vi = (uint64_t)P.seg_off_v[seg_index(J, zl, y, tx)] + below;
                v = (uint32_t)vi;
                if (vi < b.vcap) {
                    float* o = b.vtx + 6 * vi;
                    o[0] = pos[0];
                    o[1] = pos[1];
                    o[2] = pos[2];
                    o[3] = 0.0f;
                    o[4] = 0.0f;
                    o[5] = 0.0f;
                    b.keys[vi] = (uint64_t)x | ((uint64_t)y << 16) | ((uint64_t)z << 32);
                    b.signs[vi] = (uint8_t)sg;
                }
                const uint32_t k = atomicAdd(&nact, 1u);
                vpos[3 * k] = pos[0];
                vpos[3 * k + 1] = pos[1];
                vpos[3 * k + 2] = pos[2];
                vidx_l[k] = v;
            }
            P.vbox[(size_t)slot * 512u + c] = v;
        }
        if (EMIT && !(J.flags & S2M_NO_NORMALS)) {
            __syncthreads();
            const uint32_t n = nact;
            const uint16_t* L = P.lists + (size_t)tile * P.list_stride;
            const uint32_t len = P.lens[tile];
            const float eps = J.eps;
            for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
                const uint32_t v = vidx_l[k];
                if (v >= b.vcap) continue;
                const float px = vpos[3 * k], py = vpos[3 * k + 1], pz = vpos[3 * k + 2];
                // sdf3d_normal (sdf3d_normal.wgsl:4-9): taps p + v_i * eps, then normalize
                const float d1 = tape_eval_list(J.ops, L, len, px + eps, py + (-eps), pz + (-eps));
                const float d2 = tape_eval_list(J.ops, L, len, px + (-eps), py + (-eps), pz + eps);
                const float d3 = tape_eval_list(J.ops, L, len, px + (-eps), py + eps, pz + (-eps));
                const float d4 = tape_eval_list(J.ops, L, len, px + eps, py + eps, pz + eps);
                const float gx = ((d1 + (-d2)) + (-d3)) + d4;
                const float gy = (((-d1) + (-d2)) + d3) + d4;
                const float gz = (((-d1) + d2) + (-d3)) + d4;
                const float ln = sqrtf((gx * gx + gy * gy) + gz * gz);
                float* o = b.vtx + 6 * (uint64_t)v;
                o[3] = gx / ln;
                o[4] = gy / ln;
                o[5] = gz / ln;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Multi-block exclusive scan of a u32 array: per-block sums (kScanChunk entries per
// block), k_scan over the sums (adds and advances the running total), block-local scan
// plus the block offset.
// ---------------------------------------------------------------------------------
constexpr uint32_t kScanChunk = 4096;  // 256 threads x 16

__global__ __launch_bounds__(256) void k_sp_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                   uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t lds[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x * 16u;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) s += (base + k < n) ? in[base + k] : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane() == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

__global__ __launch_bounds__(256) void k_sp_downsweep(const uint32_t* __restrict__ in, uint64_t n,
                                                      const uint32_t* __restrict__ boff,
                                                      uint32_t* __restrict__ out)
{
    __shared__ uint32_t lds[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x * 16u;
    uint32_t v[16], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(inc, d, 64);
        if (lane() >= (uint32_t)d) inc += t;
    }
    if (lane() == 63) lds[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t run = boff[blockIdx.x] + inc - s;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) run += lds[w];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// First vertex (WHICH = 0) or triangle (WHICH = 1) index of every job whose first slab
// this pass holds: the exclusive offset of its first row segment.
template <int WHICH>
__global__ void k_sp_ranges(SpPass P)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.njobs) return;
    const SpJob& J = P.jobs[j];
    if (J.z0 != 0) return;
    const uint32_t* off = WHICH ? P.seg_off_t : P.seg_off_v;
    const uint64_t* total = P.counters + (WHICH ? kTriPrev : kVtxPrev);
    // a job with no row segment (empty grid) starts where the pass's output starts
    P.job_out[(size_t)J.out_index * kJobOut + (WHICH ? 1 : 0)] =
        (J.nz * J.ry * J.ntx) ? (uint64_t)off[J.seg_base] : *total;
}

// ---------------------------------------------------------------------------------
// Quads of one slot (mesh.rs:267-324): per cell with a vertex, the x-, y- and z-edge
// quads; invalid ones (a missing vertex) are counted in the EMIT = 0 pass.  Triangle
// vertex indices are relative to the job's first vertex.
// ---------------------------------------------------------------------------------
template <int EMIT>
__global__ __launch_bounds__(256) void k_sp_quads(SpPass P, MeshBufs b)
{
    const uint32_t nslots = active_slots(P);
    for (uint32_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        const uint64_t tile = P.tile_of_slot[slot];
        const uint32_t jid = P.job_of_slot[slot];
        const SpJob& J = P.jobs[jid];
        const bool lag = (J.flags & S2M_COMPAT_DROP_LAST_SLICE) != 0;
        const uint32_t vbase = EMIT ? (uint32_t)P.job_out[(size_t)J.out_index * kJobOut + 0] : 0u;
        uint32_t tx, ty, tzl;
        tile_xyz(J, tile, tx, ty, tzl);
        uint32_t invalid = 0;
        for (uint32_t r = 0; r < 2; ++r) {
            const uint32_t c = r * 256u + threadIdx.x;
            const uint32_t cxl = c & 7u, cyl = (c >> 3) & 7u, czl = c >> 6;
            const uint32_t x = tx * 8u + cxl, y = ty * 8u + cyl, zl = tzl * 8u + czl;
            const uint32_t z = J.z0 + zl;
            const uint32_t v = P.vbox[(size_t)slot * 512u + c];
            uint32_t q[3][4];
            uint32_t nq = 0;
            if (v != kNone) {  // implies the cell is inside the grid
                const uint32_t s = (v < b.vcap) ? b.signs[v] : 0u;
                const bool c0 = s & 1, c1 = (s >> 1) & 1, c2 = (s >> 2) & 1, c3 = (s >> 3) & 1;
                if (c0 != c3 && y > 0) {  // x-edge: Quad(v(x,y-1,z-1), v(x,y,z-1), v, v(x,y-1,z)).swap(c.0)
                    if (z > 0) {
                        const uint32_t t[4] = {vertex_at(P, J, x, y - 1, z - 1),
                                               vertex_at(P, J, x, y, z - 1), v,
                                               vertex_at(P, J, x, y - 1, z)};
                        for (int k = 0; k < 4; ++k) q[0][k] = c0 ? t[3 - k] : t[k];
                        if (t[0] != kNone && t[1] != kNone && t[3] != kNone) nq |= 1;
                        else ++invalid;
                    } else if (lag) {
                        ++invalid;  // the lagged reference looks up label z-1 and misses
                    }
                }
                if (c1 != c3 && x > 0) {  // y-edge: Quad(v(x-1,y,z-1), v(x,y,z-1), v, v(x-1,y,z)).swap(!c.1)
                    if (z > 0) {
                        const uint32_t t[4] = {vertex_at(P, J, x - 1, y, z - 1),
                                               vertex_at(P, J, x, y, z - 1), v,
                                               vertex_at(P, J, x - 1, y, z)};
                        for (int k = 0; k < 4; ++k) q[1][k] = !c1 ? t[3 - k] : t[k];
                        if (t[0] != kNone && t[1] != kNone && t[3] != kNone) nq |= 2;
                        else ++invalid;
                    } else if (lag) {
                        ++invalid;
                    }
                }
                if (c2 != c3 && x > 0 && y > 0) {  // z-edge: Quad(v(x-1,y-1,z), v(x,y-1,z), v, v(x-1,y,z)).swap(c.2)
                    const uint32_t t[4] = {vertex_at(P, J, x - 1, y - 1, z),
                                           vertex_at(P, J, x, y - 1, z), v,
                                           vertex_at(P, J, x - 1, y, z)};
                    for (int k = 0; k < 4; ++k) q[2][k] = c2 ? t[3 - k] : t[k];
                    if (t[0] != kNone && t[1] != kNone && t[3] != kNone) nq |= 4;
                    else ++invalid;
                }
            }
            const uint32_t ntri = 2u * (uint32_t)__builtin_popcount(nq);
            uint32_t inc = ntri;  // inclusive sum over the 8 lanes of the row segment
#pragma unroll
            for (int d = 1; d < 8; d <<= 1) {
                const uint32_t t = __shfl_up(inc, d, 8);
                if ((lane() & 7u) >= (uint32_t)d) inc += t;
            }
            const uint32_t seg_total = __shfl(inc, 7, 8);
            const bool row_ok = y < J.ry && zl < J.nz;
            if (!EMIT) {
                if (cxl == 0 && row_ok) P.seg_t[seg_index(J, zl, y, tx)] = seg_total;
            } else if (nq) {
                uint64_t t = (uint64_t)P.seg_off_t[seg_index(J, zl, y, tx)] + (inc - ntri);
                for (int e = 0; e < 3; ++e) {
                    if (!(nq & (1u << e))) continue;
                    if (t + 1 < b.tcap) {  // make_triangles: (q2, q1, q0), (q0, q3, q2)
                        uint32_t* o = b.tri + 3 * t;
                        o[0] = q[e][2] - vbase;
                        o[1] = q[e][1] - vbase;
                        o[2] = q[e][0] - vbase;
                        o[3] = q[e][0] - vbase;
                        o[4] = q[e][3] - vbase;
                        o[5] = q[e][2] - vbase;
                    }
                    t += 2;
                }
            }
        }
        if (!EMIT && invalid) {
            atomicAdd((unsigned long long*)&b.counters[kInvalid], (unsigned long long)invalid);
            atomicAdd((unsigned long long*)&P.job_out[(size_t)J.out_index * kJobOut + 2],
                      (unsigned long long)invalid);
        }
    }
}

// Cell -> vertex map of the slab's last plane, for the next slab (single job).
__global__ __launch_bounds__(256) void k_sp_boundary(SpPass P)
{
    const SpJob& J = P.jobs[0];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= J.rx * J.ry) return;
    J.boundary[i] = vertex_at(P, J, i % J.rx, i / J.rx, J.z0 + J.nz - 1);
}

// Per-pass reset of the slot counter; the largest count seen is kept for the host's
// capacity check (a kernel rather than a memset, so a pass can be graph-captured).
__global__ void k_sp_pass_begin(uint64_t* counters)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint64_t used = counters[kSlotsUsed];
        if (used > counters[kSlotsMax]) counters[kSlotsMax] = used;
        counters[kSlotsUsed] = 0;
    }
}

// ---------------------------------------------------------------------------------
hipError_t launch_sparse_pass_begin(hipStream_t s, uint64_t* counters)
{
    hipLaunchKernelGGL(k_sp_pass_begin, dim3(1), dim3(64), 0, s, counters);
    return hipGetLastError();
}

hipError_t launch_sparse_pass(hipStream_t s, const SpPass& P, const MeshBufs& b, int persistent_blocks,
                              int boundary_plane, hipEvent_t* marks)
{
    if (P.ntiles == 0) return hipSuccess;
    const uint64_t nsb64 = (P.nseg + kScanChunk - 1) / kScanChunk;
    if (nsb64 > kMaxScanBlocks) return hipErrorInvalidValue;
    const uint32_t nsb = (uint32_t)nsb64;
    const uint32_t pb = (uint32_t)persistent_blocks;
    const size_t smem = (size_t)P.box_stride * sizeof(float);
    const uint32_t jb = (P.njobs + 63) / 64;
    auto mark = [&](int i) {
        if (marks && marks[i]) (void)hipEventRecord(marks[i], s);
    };
    mark(0);
    hipLaunchKernelGGL(k_sp_prune, dim3((uint32_t)((P.ntiles + 63) / 64)), dim3(64), 0, s, P);
    mark(1);
    hipLaunchKernelGGL(k_sp_eval, dim3(pb), dim3(256), 0, s, P);
    mark(2);
    hipLaunchKernelGGL((k_sp_cells<0>), dim3(pb), dim3(256), smem, s, P, b);
    if (nsb) {
        hipLaunchKernelGGL(k_sp_reduce, dim3(nsb), dim3(256), 0, s, P.seg_v, P.nseg, b.bcount);
        (void)launch_scan(s, b, nsb, 0);
        hipLaunchKernelGGL(k_sp_downsweep, dim3(nsb), dim3(256), 0, s, P.seg_v, P.nseg, b.boffset,
                           P.seg_off_v);
    }
    hipLaunchKernelGGL((k_sp_ranges<0>), dim3(jb), dim3(64), 0, s, P);
    mark(3);
    hipLaunchKernelGGL((k_sp_cells<1>), dim3(pb), dim3(256), smem, s, P, b);
    mark(4);
    hipLaunchKernelGGL((k_sp_quads<0>), dim3(pb), dim3(256), 0, s, P, b);
    if (nsb) {
        hipLaunchKernelGGL(k_sp_reduce, dim3(nsb), dim3(256), 0, s, P.seg_t, P.nseg, b.bcount);
        (void)launch_scan(s, b, nsb, 1);
        hipLaunchKernelGGL(k_sp_downsweep, dim3(nsb), dim3(256), 0, s, P.seg_t, P.nseg, b.boffset,
                           P.seg_off_t);
    }
    hipLaunchKernelGGL((k_sp_ranges<1>), dim3(jb), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_sp_quads<1>), dim3(pb), dim3(256), 0, s, P, b);
    if (boundary_plane) {
        hipLaunchKernelGGL(k_sp_boundary, dim3((uint32_t)((P.boundary_cells + 255) / 256)), dim3(256), 0,
                           s, P);
    }
    mark(5);
    return hipGetLastError();
}

}  // namespace s2m
