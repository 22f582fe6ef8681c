// internal.h - declarations shared by the host runtime (api.cpp) and the kernels
// (kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sdf2mesh.h"

namespace s2m {

constexpr uint32_t kBlock = 256;             // threads per block of the cell kernels
constexpr uint32_t kRounds = 4;              // cells per thread (strided by kBlock)
constexpr uint32_t kCellsPerBlock = kBlock * kRounds;
constexpr uint32_t kMaxScanBlocks = 65536;   // most block counts one dense slab scans
constexpr uint32_t kDenseScanChunk = 1024;        // block counts per workgroup of the dense scan
constexpr uint32_t kDenseScanParts = kMaxScanBlocks / kDenseScanChunk;  // its chunk sums, after the counters
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
    kTilesActive = 5,  // tiles whose cells were evaluated (tiled path)
    kEvalPoints = 6,   // lattice points evaluated (tiled path)
    kEvalFlops = 7,    // algorithmic flops executed by the pruned tapes (tiled path)
    kSlotsUsed = 8,    // active tiles of the current slab (sparse path)
    kSlotsMax = 9,     // largest kSlotsUsed of the job so far
    kStageUsed = 10,   // staged active cells of the current pass (sparse path)
    kStageMax = 12,    // largest kStageUsed of the job so far
    kNumCounters = 16
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

constexpr uint32_t kTile = 8;  // cells per tile edge (sparse pipeline)
constexpr uint32_t kSuper = 4;  // tiles per super-tile edge (first pruning level)
constexpr uint32_t kSingleSign = 0xffffffffu;  // super_lens: the SDF keeps one sign there
constexpr uint32_t kStageChunkHost = 1024;  // staged-cell allocation chunk (sparse.hip)
constexpr uint32_t kTileBlocksPerCU = 8;    // most k_sp_tiles blocks per compute unit
constexpr uint32_t kTileRecBytes = 2720;    // per slot: k_sp_tiles' staged record (sparse.hip)
constexpr uint32_t kRowChunk = 4096;        // rows per chunk of the row offset scan (sparse.hip)
constexpr uint32_t kChunkSumStride = 32;    // words between chunk sums (one 128-byte line each)
constexpr uint32_t kQueues = 2;        // work queues of a pass (k_sp_tiles<1>, k_sp_tiles<0>)
constexpr uint32_t kQueueGroups = 64;  // counters per queue, 32 u32 (128 bytes) apart
// Profiling marks of one sparse pass: prune (tiles, bricks) | eval+contour | quad count |
// scan | emission
constexpr int kSparseMarks = 6;

// Tile-sparse pipeline (sparse.hip).  One job of a pass: a slab of a single job, or a
// whole job of a batch.  Pointers are device addresses.
constexpr uint32_t kJobOut = 4;  // per job: first vertex, first triangle, invalid quads, spare

struct SpJob {
    uint32_t rx, ry;           // cells per row / rows per plane
    uint32_t nz, z0;           // cell planes of this pass
    uint32_t ntx, nty, ntz;    // tiles of the whole job per axis
    uint32_t tz0;              // tile layer of plane z0
    uint32_t nops, flags;
    uint32_t union_fold;       // the program is a left fold of unions / smooth unions (k > 0)
    uint32_t left_fold;        // the program is a left fold of library primitives
    uint32_t bx, by;           // lattice box dims of the job's tiles (slot box indexing)
    uint32_t out_index;        // row of the pass's job_out table
    float eps;
    double pad;                // tile box padding (normal taps, rounding)
    uint64_t tile_base;        // first tile of the job in the pass's numbering
    uint64_t row_base;         // first row (zl, y) of the job
    uint64_t trow_base;        // first tile row (tile layer, ty) of the job
    uint64_t super_base;       // first super-tile of the job
    uint32_t nsx, nsy, nsz;    // super-tiles of the job in this pass per axis
    const s2m_op* ops;
    const float* cx;
    const float* cy;
    const float* cz;
    const uint32_t* lox;
    const uint32_t* hix;
    const uint32_t* loy;
    const uint32_t* hiy;
    const uint32_t* loz;
    const uint32_t* hiz;
    const float* ext;          // tile coordinate extents: xmin[ntx] xmax[ntx] ymin.. zmax[ntz]
    const uint32_t* ibox;      // tile lattice index boxes, same layout
    const uint32_t* op_flops;
    uint32_t* boundary;        // [ry][rx] vertex of cell plane z0 - 1 (multi-slab job)
};

// Per slot, written by k_sp_bricks for the kernels after it.
struct SlotHdr {
    uint32_t tile, job, lens;
    uint32_t txy;    // tile x | tile y << 16
    uint32_t tzl;    // tile layer within the pass
    uint32_t box[3]; // per axis: first lattice index of the tile's box | points << 24
};

struct SpPass {
    const SpJob* jobs;
    uint32_t njobs;
    uint32_t list_stride;      // >= nops of every job
    uint32_t prune_depth;      // >= 1, >= the deepest operand stack of any job's program - 1
    uint64_t ntiles;           // tiles of the pass
    uint64_t nrows;            // rows (z, y) of the pass
    uint64_t ntrows;           // tile rows (tile layer, ty) of the pass
    uint64_t boundary_cells;   // rx * ry of jobs[0] when a boundary plane is written
    uint64_t nsuper;           // super-tiles of the pass
    double2* super_iv;         // [super-tile][list_stride] primitive intervals (lo, hi)
    uint16_t* super_lists;     // [super-tile][list_stride] pruned op indices
    uint32_t* super_lens;      // [super-tile] list length or kSingleSign
    uint64_t* super_act;       // [super-tile] bit (z * 4 + y) * 4 + x: that tile has a slot
    uint16_t* lists;           // [tile][list_stride] pruned op indices
    uint32_t* lens;            // [tile]
    uint32_t* flops;           // [tile] flops per point of the pruned tape
    int32_t* slot_of_tile;     // [tile], -1: no active cell
    uint32_t* tile_of_slot;    // [slot]
    uint32_t* job_of_slot;     // [slot]
    uint32_t slot_cap;
    uint32_t box_stride;       // floats of a tile's lattice box (LDS)
    uint32_t* cellbits;        // [slot][16] bit c: cell c of the tile has a vertex (row r = c / 8
                               // in byte r % 4 of word r / 4)
    float4* stage;             // [stage_cap] active cells: position, cell | rank << 9 | signs << 12
    uint64_t stage_cap;
    uint32_t* slot_stage;      // [slot][2] first staged cell, staged cells
    uint64_t* slot_live;       // [slot] bricks (2x2x2 cells) that may hold an active cell
    SlotHdr* slot_hdr;         // [slot]
    int32_t* slot_nb;          // [slot][8] slots of the tile (tx - dx, ty - dy, tz - dz) at (dz * 2 + dy) * 2 + dx
    uint4* slot_rec;           // [slot] kTileRecBytes: k_sp_tiles<1>'s staged record
    uint64_t* block_stats;     // [k_sp_tiles block][3]: tiles, points, flops of the job so far
    uint32_t* slot_rows;       // [4][slot_cap][64] per row of each slot's tile: vertices,
                               // triangles, first vertex / triangle within the row
    uint32_t* row_sum;         // [2][nrows] vertices, triangles per row
    uint32_t* row_off;         // [2][nrows] exclusive offsets: global vertex / triangle index
    uint64_t* counters;
    uint32_t* queues;          // [kQueues][kQueueGroups][32] work queue counters (word 0)
    uint64_t* job_out;         // [job][kJobOut]
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
    uint64_t* actbits;     // dense slab: bit c % 64 of word c / 64 = cell c has a vertex (zero between slabs)
};

// ---- launchers (kernels.hip) ----
hipError_t launch_eval_lattice(hipStream_t s, const s2m_op* ops, uint32_t nops, const float* cx,
                               const float* cy, const float* cz, uint32_t nx, uint32_t ny,
                               uint32_t lz0, uint32_t nzl, uint32_t pitch, float* lat);
// The lattice of a tape that is one Mandelbulb op (k_eval_mandelbulb: equal to
// launch_eval_lattice's values).
hipError_t launch_eval_mandelbulb(hipStream_t s, const s2m_op* op, const float* cx, const float* cy,
                                  const float* cz, uint32_t nx, uint32_t ny, uint32_t lz0, uint32_t nzl,
                                  uint32_t pitch, float* lat);
hipError_t launch_cells(hipStream_t s, const SlabArgs& a, const MeshBufs& b, int emit,
                        int write_zero_normals);
hipError_t launch_scan(hipStream_t s, const MeshBufs& b, uint32_t nblocks, uint32_t which);
// cnt[0..4) (zeroed by the caller): k_sincos_check's mismatch counts.
hipError_t launch_sincos_check(hipStream_t s, uint64_t* cnt);
// cnt[0..2) (zeroed by the caller): k_bailout_check's comparisons and disagreements.
hipError_t launch_bailout_check(hipStream_t s, uint64_t* cnt);
// A tape just copied to the device: every Mandelbulb op's f[3] set to the bailout
// threshold of its f[1] (sd_mandelbulb reads it there instead of computing it per point).
hipError_t launch_tape_prepare(hipStream_t s, s2m_op* ops, uint32_t n);
hipError_t launch_normals(hipStream_t s, const s2m_op* ops, uint32_t nops, float eps,
                          const MeshBufs& b, bool mandel = false);  // mandel: ops is one Mandelbulb op
// After a slab's quad pass: its activity bits zeroed (the words of `ncells` cells).
hipError_t launch_clear_bits(hipStream_t s, uint64_t* bits, uint32_t ncells);
hipError_t launch_quads(hipStream_t s, const SlabArgs& a, const MeshBufs& b, int emit);
// Between passes of a job: slot and staging counters folded into their maxima, the
// previous pass's totals made the next pass's first indices, work queues and the
// offset scan's nbsum chunk sums (bsum, kChunkSumStride words apart) cleared.
hipError_t launch_sparse_pass_begin(hipStream_t s, uint64_t* counters, uint32_t* queues, uint32_t* bsum,
                                    uint32_t nbsum);
// After a job's last pass: slot and staging counters folded into their maxima, then
// counters[0 .. n) copied to `out` (a device address of mapped host memory).
// The tile blocks' statistics (nstats entries of tiles, points, flops) summed into
// kTilesActive, kEvalPoints and kEvalFlops first.
hipError_t launch_sparse_job_end(hipStream_t s, uint64_t* counters, uint32_t n, uint64_t* out,
                                 const uint64_t* stats, uint32_t nstats);
// A job's first pass: counters[0 .. n) (the counters and the job output rows), the work
// queues and the nbsum chunk sums (bsum, kChunkSumStride words apart) cleared, and the
// nbnd entries of a multi-pass job's boundary plane set to "no vertex".
hipError_t launch_sparse_job_begin(hipStream_t s, uint64_t* counters, uint32_t n, uint32_t* queues,
                                   uint32_t* bsum, uint32_t nbsum, uint64_t* stats, uint32_t nstats,
                                   uint32_t* boundary, uint32_t nbnd);
// cus: compute units of the device; tile_blocks: most blocks of a k_sp_tiles launch (each
// holds a staging chunk); general_tiles: some tile's pruned program may not be a left fold
// (false when every job's whole program is one: its pruned programs are then left folds).
// fold_jobs / other_jobs: some job's program is / is not a left fold of unions (pruned at
// the super-tile level without a sequential walk).
hipError_t launch_sparse_pass(hipStream_t s, const SpPass& P, const MeshBufs& b, int cus,
                              uint32_t tile_blocks, bool general_tiles, bool fold_jobs, bool other_jobs,
                              int boundary_plane, hipEvent_t* marks);
// out[0..4) (a device address of mapped host memory) = the vertex and triangle totals so
// far and the current pass's slot and staging use, after the kernels queued before it.
hipError_t launch_snapshot(hipStream_t s, const uint64_t* counters, uint64_t* out);
// Contour corner values given per cell (test hook): cells of the slab read their 8
// corners from `corners` (8 floats per cell, cells in (z,y,x) order from plane z0).
hipError_t launch_cells_from_corners(hipStream_t s, const SlabArgs& a, const float* corners,
                                     const MeshBufs& b, int emit);

// ---- no-allocation scratch (kernels.hip) ----
constexpr size_t kNoAllocScratchBytes = (size_t)S2M_NOALLOC_SCRATCH_BYTES;
constexpr uint32_t kMaxDevices = 64;
// Device address of the scratch array on `device`, which must be the current device.
// Looked up once per device (hipGetSymbolAddress); later calls return the kept address
// and the first lookup's status.
hipError_t noalloc_scratch(int device, void** ptr);

// ---- host memory for results (hostmem.cpp) ----
// Host memory for a result array: page-locked and recycled when large, malloc otherwise.
// Returned to the pool (or freed) by host_release; any other pointer is passed to free().
void* host_alloc(size_t bytes);
// A page-locked block shared by `refs` results: each points somewhere inside it and is
// released with host_release; the block returns to the pool after the last release.
void* host_alloc_shared(size_t bytes, uint32_t refs);
void host_release(void* p);

// ---- autotuner table (tune.cpp) ----
// Shapes the meshing calls use instead of their rules (0 = no choice for that knob).
struct TuneChoice {
    uint32_t stream_passes = 0;  // z-passes of a single tile-sparse job streamed to host memory
    uint32_t dense_slab = 0;     // cell planes per slab of the dense path
    uint64_t batch_cells = 0;    // most cells per pass of a batch
};
// The calling thread's override while s2m_tune times a candidate, else the table's entry.
bool tune_lookup(const std::string& key, TuneChoice* out);
// The table's entry only, with the time it was chosen for.
bool tune_cached(const std::string& key, TuneChoice* out, double* ms);
void tune_store(const std::string& key, const TuneChoice& c, double ms);
void tune_override(const TuneChoice* c);  // NULL: back to the table
// Index of the chosen candidate: the first one, replaced by any later one faster by more
// than 2 % (so a choice depends only on the timings and their order).
uint32_t tune_pick(const double* ms, uint32_t n);

// ---- host helpers (api.cpp) for the server ----
// A program from a gfx950 code object holding s2m_c_eval_lattice and s2m_c_normals.
int program_from_code_object(s2m_ctx* ctx, const void* code, size_t size, s2m_program** out);
s2m_program* program_view(int device, const s2m_op* host_ops, uint32_t n, const s2m_op* d_ops);
hipStream_t context_stream(s2m_ctx* ctx);
bool tape_valid(const s2m_op* ops, uint32_t n);
int job_geometry(const s2m_params* p, uint32_t res[3], uint32_t* nz_cells);
int set_error(int code, const char* msg);
// Whether a request of the program (or tape `ops`) meshes alone on the dense path, by the
// rule s2m_mesh_batch uses (compiled, S2M_DENSE_EVAL, or a Mandelbulb above 128^3).
bool dense_request(const s2m_program* prog, const s2m_op* ops, uint32_t n, const s2m_params* p);
// Estimated device bytes of a dense job (working buffers of its slabs + outputs).
uint64_t dense_request_bytes(const s2m_params* p);
int program_device(const s2m_program* prog);

}  // namespace s2m
