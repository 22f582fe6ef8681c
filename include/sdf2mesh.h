/*
 * sdf2mesh.h - C ABI of the MI355X meshing library (libsdf2mesh.so).
 *
 * The library turns a signed distance function (SDF) into a triangle mesh with the
 * reference's mass-point dual contouring, computed by HIP kernels for gfx950.
 *
 * Plain C types only (no HIP or torch types), so a Rust `-sys` crate, ctypes or any
 * other FFI can bind it.  Every entry point is thread-compatible: use one context per
 * host thread.  Errors: every `int` return is an s2m_status; the message of the last
 * failure on the calling thread is returned by s2m_last_error().
 *
 * Reference interfaces replaced (paths relative to the sdf2mesh repository):
 *   s2m_create / s2m_destroy       wgpu adapter/device/queue set-up, main.rs:180-196
 *   s2m_program_from_tape          Sdf3DShader::from_path + create_shader_module
 *                                  (shader.rs:44-67,220-225) for programs built from
 *                                  the library primitives and operators
 *                                  (sdf3d_primitives.wgsl, sdf_op.wgsl)
 *   s2m_program_from_source        Sdf3DShader::from_path / from_glsl_fragment_shader
 *                                  + create_shader_module (shader.rs:44-104,220-225)
 *   s2m_translate                  Sdf3DShader::write_to_file for --debug-wgsl
 *                                  (shader.rs:206-209, main.rs:223-225)
 *   s2m_mesh                       the per-slice loop and host scan, main.rs:294-357,
 *                                  plus TriangleMesh::from(VertexList), mesh.rs:334-341
 *   s2m_write_stl / s2m_write_ply  TriangleMesh::write_to_file, mesh.rs:182-210
 *   s2m_params                     AppState + From<&Arguments>, main.rs:28-33,139-175
 *   s2m_mesh_batch(_device)        no reference counterpart: many jobs per device pass
 *   s2m_server_*                   no reference counterpart: continuous-batching request loop
 */
#ifndef SDF2MESH_H
#define SDF2MESH_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------
 * Status codes.  The first six mirror ShaderProcessingError (shadertoy.rs:70-80):
 * RequestError, ShaderError, ParseErrors, WgslError, ValidationError, MissingSdf.
 * ---------------------------------------------------------------------------------- */
enum s2m_status {
    S2M_OK = 0,
    S2M_E_REQUEST = 1,     /* network request (ShaderToy); never available offline  */
    S2M_E_SHADER = 2,      /* shader text error (e.g. function not found)            */
    S2M_E_PARSE = 3,       /* parse error in WGSL / GLSL input                       */
    S2M_E_WGSL = 4,        /* lowering / code generation error                       */
    S2M_E_VALIDATION = 5,  /* type or program validation error                       */
    S2M_E_MISSING_SDF = 6, /* no sdf3d (or named) function, shader.rs:96-98          */
    S2M_E_HIP = 7,         /* HIP runtime or hiprtc failure                          */
    S2M_E_OOM = 8,         /* device or host allocation failed                       */
    S2M_E_BUDGET = 9,      /* request exceeds the configured HBM budget (serving)    */
    S2M_E_ARG = 10,        /* invalid argument                                       */
    S2M_E_IO = 11,         /* file I/O error                                         */
    S2M_E_BUSY = 12        /* no result ready yet (batch/server polling)             */
};

/* ------------------------------------------------------------------------------------
 * SDF tape: a postfix program over the reference's primitive and operator library.
 * Each primitive pushes one distance, each operator pops d2 (top) and d1 (below) and
 * pushes op(d1, d2).  A valid tape leaves exactly one value on the stack.
 * Semantics follow the WGSL library literally (fp32, operation order as written):
 *   SPHERE    f[0..2]=c, f[3]=s            sdf3d_sphere(p - c, s)      primitives.wgsl:27-30
 *   BOX       f[0..2]=c, f[3..5]=b         sdf3d_box(p - c, b)         primitives.wgsl:7-11
 *   TORUS     f[0..2]=c, f[3..4]=t         sdf3d_torus(p - c, t)       primitives.wgsl:32-36
 *   CYLINDER  f[0..2]=c, f[3]=h, f[4]=r    sdf3d_cylinder(p - c, h, r) primitives.wgsl:13-17
 *   CAPSULE   f[0..2]=a, f[3..5]=b, f[6]=r sdf3d_capsule(p, a, b, r)   primitives.wgsl:19-25
 *   MANDELBULB f[0]=power, f[1]=bailout, f[2]=offset, iarg = iterations | (swap_yz<<16)
 *             | (trig<<17)
 *             the distance estimator of examples/mandelmesh.frag:3-28 (f[3] is not read
 *             from the caller's tape: the device copy holds the bailout's r^2 threshold there)
 *   UNION min(d1,d2)  INTERSECTION max(d1,d2)  SUBTRACTION max(-d1,d2)
 *   SMOOTH_UNION / SMOOTH_INTERSECTION / SMOOTH_SUBTRACTION  f[0]=k,  sdf_op.wgsl:7-23
 * `p - c` is a per-component fp32 subtraction, so c = 0 leaves p unchanged.
 * ---------------------------------------------------------------------------------- */
enum s2m_opcode {
    S2M_OP_SPHERE = 1,
    S2M_OP_BOX = 2,
    S2M_OP_TORUS = 3,
    S2M_OP_CYLINDER = 4,
    S2M_OP_CAPSULE = 5,
    S2M_OP_MANDELBULB = 6,
    S2M_OP_UNION = 16,
    S2M_OP_INTERSECTION = 17,
    S2M_OP_SUBTRACTION = 18,
    S2M_OP_SMOOTH_UNION = 19,
    S2M_OP_SMOOTH_INTERSECTION = 20,
    S2M_OP_SMOOTH_SUBTRACTION = 21
};

typedef struct s2m_op {
    uint32_t code; /* s2m_opcode */
    uint32_t iarg; /* integer operand (MANDELBULB iterations and flags) */
    float f[10];   /* float operands */
} s2m_op;          /* 48 bytes */

#define S2M_MAX_TAPE 4096u  /* longest accepted tape                  */
#define S2M_MAX_STACK 16u   /* deepest accepted operand stack         */
#define S2M_MANDEL_SWAP_YZ (1u << 16) /* apply p.xyz = p.xzy first (mandelmesh.frag:4) */
#define S2M_MANDEL_TRIG (1u << 17)    /* evaluate the escape step as written, with atan, asin,
                                         pow, sin and cos; by default a power-8 step takes
                                         the same values by complex squaring, without
                                         transcendentals (fewer roundings, several times
                                         fewer instructions) */

/* ------------------------------------------------------------------------------------
 * Meshing parameters: the AppState uniform (main.rs:28-33) as plain fields.
 * ---------------------------------------------------------------------------------- */
#define S2M_COMPAT_POW2 0x1u            /* round res up as `2 << ilog2(res)` unless a
                                           power of two (main.rs:141-148)            */
#define S2M_COMPAT_DROP_LAST_SLICE 0x2u /* drop cells of the top z-slice, as the
                                           reference's one-slice readback lag does
                                           (main.rs:298-356: staging buffers are mapped
                                           before the slice's submit)                  */
#define S2M_CORNERS_EXACT 0x4u          /* evaluate the SDF at the reference's per-cell
                                           corner coordinates cmin=min+h*i, cmax=cmin+h
                                           (dualcontour.wgsl:22-43) instead of a shared
                                           lattice min+h*i                             */
#define S2M_NO_NORMALS 0x8u             /* skip the 4-tap normal (STL does not use it) */
#define S2M_DENSE_EVAL 0x10u            /* evaluate every lattice point with the full
                                           program; by default tiles of 8^3 cells whose
                                           SDF provably keeps one sign are skipped and
                                           each tile runs a pruned program that gives
                                           identical values on the tile                 */
#define S2M_REFERENCE_COMPAT (S2M_COMPAT_POW2 | S2M_COMPAT_DROP_LAST_SLICE | S2M_CORNERS_EXACT)

typedef struct s2m_params {
    float bb_min[3];  /* AppState.bb_min.xyz                                  */
    float bb_max[3];  /* AppState.bb_max.xyz                                  */
    uint32_t res[3];  /* AppState.dims.xyz (cells per axis)                   */
    float eps;        /* AppState.bb_max.w: normal tap distance (1e-4)        */
    uint32_t flags;   /* S2M_* flags above                                    */
    uint32_t slab;    /* cell planes per device slab, 0 = automatic           */
} s2m_params;

/* Fill p the way From<&Arguments> does (main.rs:139-175): cube of edge `bounds`
 * centred at the origin (lib.rs:115-118), res on all three axes, eps 1e-4,
 * flags = S2M_REFERENCE_COMPAT. */
void s2m_params_reference(s2m_params* p, uint32_t resolution, float bounds);

/* Resolution after the reference's rounding rule (main.rs:141-148). */
uint32_t s2m_round_resolution(uint32_t res);

/* ------------------------------------------------------------------------------------
 * Mesh results.  Vertices are in ascending cell order (z, then y, then x), as the
 * reference's VertexList (mesh.rs:213-248); triangles in VertexList order with the
 * x-, y- then z-edge quad of each cell (mesh.rs:267-324, lib.rs:188-211).
 * ---------------------------------------------------------------------------------- */
typedef struct s2m_mesh_out {
    float* vtx;             /* 6 floats per vertex: pos.xyz, normal.xyz            */
    uint64_t nvtx;
    uint32_t* tri;          /* 3 vertex indices per triangle                       */
    uint64_t ntri;
    uint64_t invalid_quads; /* quads skipped for a missing vertex (mesh.rs:270-278) */
    uint64_t* cell_keys;    /* per vertex x | y<<16 | z<<32 (mesh.rs:224-226)      */
    uint8_t* signs;         /* per vertex bits [c100>0]|[c010>0]<<1|[c001>0]<<2|
                               [c000>0]<<3 (dualcontour.wgsl:57-69)                */
} s2m_mesh_out;

/* Device-resident result of s2m_mesh_device: pointers into context-owned HBM, valid
 * until the next meshing call on the same context. */
typedef struct s2m_device_mesh {
    const float* vtx;        /* 6 floats per vertex (device memory)   */
    const uint32_t* tri;     /* 3 u32 per triangle (device memory)    */
    const uint64_t* cell_keys;
    const uint8_t* signs;
    uint64_t nvtx;
    uint64_t ntri;
    uint64_t invalid_quads;
} s2m_device_mesh;

/* Per-stage device time of the last meshing call, from HIP events recorded on the
 * context's stream (milliseconds).  Filled only when profiling is on: level 1 records the
 * whole call and the evaluation launches (total_ms, eval_ms; the pruned path's other
 * stage fields stay 0), level 2 every stage (each event adds a marker between kernels,
 * so level 2 costs a little device time).  Counts are filled at every level. */
typedef struct s2m_timing {
    double total_ms;    /* first kernel start to last kernel end                   */
    double eval_ms;     /* lattice evaluation kernels                               */
    double contour_ms;  /* active-cell count, scan (and vertex emission when dense);
                           pruned path: the vertex and triangle offset scans        */
    double normal_ms;   /* 4-tap normals (pruned path: vertex, normal and triangle
                           emission)                                                */
    double quad_ms;     /* quad count, scan, triangle emission (pruned path: the
                           quad count)                                              */
    double d2h_ms;      /* device-to-host copies (s2m_mesh only)                    */
    uint64_t eval_launches;
    uint64_t lattice_points; /* SDF evaluations done on the lattice                 */
    uint64_t slabs;
    uint64_t tiles_total;    /* tiles of 8^3 cells (pruned path; 0 when dense)       */
    uint64_t tiles_active;   /* tiles whose lattice points were evaluated            */
    uint64_t eval_flops;     /* algorithmic flops executed on the lattice            */
    double prune_ms;         /* interval pruning of tiles and bricks (pruned path)   */
    uint64_t graph_launches; /* 1 when the job's kernels ran as a captured graph      */
} s2m_timing;

typedef struct s2m_ctx s2m_ctx;
typedef struct s2m_program s2m_program;

/* ---- context -------------------------------------------------------------------- */
s2m_ctx* s2m_create(int device);
void s2m_destroy(s2m_ctx* ctx);
const char* s2m_last_error(void);
int s2m_set_profiling(s2m_ctx* ctx, int level);  /* 0 off, 1 total + evaluation, 2 every stage */
int s2m_last_timing(const s2m_ctx* ctx, s2m_timing* out);

/* ---- programs ------------------------------------------------------------------- */
/* Validate a tape (stack discipline, opcodes, sizes) and upload it to the device. */
int s2m_program_from_tape(s2m_ctx* ctx, const s2m_op* ops, uint32_t n_ops, s2m_program** out);
void s2m_program_free(s2m_program* prog);

/* ------------------------------------------------------------------------------------
 * Programs from SDF source (Sdf3DShader::from_path, shader.rs:44-67, and
 * Sdf3DShader::from_glsl_fragment_shader, shader.rs:73-104; main.rs:198-227).
 *   S2M_LANG_SDF3D  WGSL text whose `use <module>;` lines import the library modules
 *                   (sdf3d::*, sdf3d::primitives, sdf3d::normal, sdf::op, sdf::*) and whose
 *                   `include "file";` lines splice files in (shader.rs:159-203); it defines
 *                   `fn sdf3d(p: vec3f) -> f32`.
 *   S2M_LANG_GLSL   a GLSL fragment shader with `void main()` and `float <sdf_name>(vec3)`
 *                   (sdf_name NULL = "sdf", the --glsl-sdf default, main.rs:218); a
 *                   function named sdf3d, if present, is the SDF (shader.rs:88-95).
 * A program built from library primitives at p - c and library operators with constant
 * operands becomes a tape (pruned sparse path, batchable); any other program is compiled
 * for gfx950 at run time with hiprtc (dense path).  S2M_FORCE_COMPILED compiles every
 * program.  Errors: S2M_E_IO (unreadable file), S2M_E_PARSE, S2M_E_VALIDATION (names,
 * types, recursion, sdf3d_normal not visible), S2M_E_MISSING_SDF (GLSL sdf function
 * absent), S2M_E_SHADER (GLSL without main), S2M_E_WGSL (code generation), S2M_E_HIP.
 * ---------------------------------------------------------------------------------- */
#define S2M_LANG_SDF3D 0u
#define S2M_LANG_GLSL 1u
#define S2M_SOURCE_IS_PATH 0x100u /* `source` is a file path, not the text            */
#define S2M_FORCE_COMPILED 0x200u /* compile even a program that lowers to a tape      */

int s2m_program_from_source(s2m_ctx* ctx, const char* source, uint32_t lang_flags,
                            const char* sdf_name, s2m_program** out);

#define S2M_PROGRAM_TAPE 0
#define S2M_PROGRAM_COMPILED 1
/* S2M_PROGRAM_TAPE or S2M_PROGRAM_COMPILED; -1 for NULL. */
int s2m_program_kind(const s2m_program* prog);
/* Length of a tape program's tape (0 for a compiled program); copies min(length, cap)
 * ops to `ops` when it is not NULL. */
uint32_t s2m_program_tape(const s2m_program* prog, s2m_op* ops, uint32_t cap);

/* Front-end stages without a device: writes min(size, cap) bytes of the artifact to buf
 * and returns its size, or 0 with *status (may be NULL) and s2m_last_error() set.
 *   S2M_STAGE_SOURCE  the source after `use` / `include` expansion, before validation
 *                     (the --debug-wgsl output, main.rs:223-225)
 *   S2M_STAGE_TAPE    the tape (s2m_op records); S2M_E_WGSL if the program is not one
 *   S2M_STAGE_HIP     the generated HIP C++
 *   S2M_STAGE_CODE    the gfx950 code object hiprtc builds from it */
#define S2M_STAGE_SOURCE 0u
#define S2M_STAGE_TAPE 1u
#define S2M_STAGE_HIP 2u
#define S2M_STAGE_CODE 3u
uint64_t s2m_translate(const char* source, uint32_t lang_flags, const char* sdf_name,
                       uint32_t stage, void* buf, uint64_t cap, int* status);
/* Host-side tape validation only (no device needed): returns S2M_OK or S2M_E_VALIDATION. */
int s2m_tape_validate(const s2m_op* ops, uint32_t n_ops);
/* Algorithmic fp32 operation count of one evaluation of the tape: + - * / min max abs
 * sqrt and each transcendental count 1, clamp 2, mix 3, length(vec3) 6, length(vec2) 4,
 * dot(vec3) 5, constant subexpressions folded (the static count of the WGSL text). */
uint64_t s2m_tape_flops(const s2m_op* ops, uint32_t n_ops);

/* ---- meshing -------------------------------------------------------------------- */
/* Mesh `prog` with `params`; results are copied into library-allocated host memory
 * that s2m_mesh_free releases.  cell_keys/signs are filled only if want_keys != 0. */
int s2m_mesh(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
             int want_keys, s2m_mesh_out* out);
/* Same computation, result left in device memory owned by ctx (no PCIe transfer). */
int s2m_mesh_device(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
                    s2m_device_mesh* out);
void s2m_mesh_free(s2m_mesh_out* out);

/* ---- meshing without device allocation ------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (serving loops that must not call hipMalloc, harnesses that forbid it): the mesh of
 * s2m_mesh_device, computed by the dense-path kernels, written into the caller's device
 * buffers.  Every working buffer (coordinate tables, lattice slab, cell -> vertex planes,
 * block counts, activity bits, counters) is carved from one static device array of
 * S2M_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process
 * ends; slabs shrink until they fit it.  The call performs no hipMalloc, hipFree or
 * other allocating HIP call.  Errors: S2M_E_ARG when one cell plane per slab does not fit
 * the scratch; S2M_E_OOM when the mesh exceeds vcap / tcap, with out->nvtx / out->ntri set
 * to the sizes it needs (the triangle count is a lower bound when the vertices
 * overflowed).  All calls on one GPU share the array: calls that overlapped would race,
 * so the library runs them one at a time per device (a lock held until the call has
 * synchronised the context's stream).  Not for use under stream capture: it
 * synchronises the stream. */
#define S2M_NOALLOC_SCRATCH_BYTES (64u << 20)
typedef struct s2m_mesh_buffers {
    float* vtx;          /* device: 6 floats per vertex, vcap vertices                    */
    uint64_t* cell_keys; /* device: vcap keys (working storage of the quad pass too)      */
    uint8_t* signs;      /* device: vcap sign bytes (read by the quad pass)               */
    uint64_t vcap;
    uint32_t* tri;       /* device: 3 u32 per triangle, tcap triangles                    */
    uint64_t tcap;
} s2m_mesh_buffers;
int s2m_mesh_device_noalloc(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
                            const s2m_mesh_buffers* bufs, s2m_device_mesh* out);

/* ---- batching ------------------------------------------------------------------- */
/* Mesh n independent jobs together: the tiles of all jobs go through one pass of the
 * tile-sparse kernels (one set of launches for the batch), so small jobs share the
 * device instead of each paying its own launch and synchronisation latency.  Each
 * job's result is identical to meshing it alone with s2m_mesh.  outs[i] receives job
 * i's mesh (free each with s2m_mesh_free).  Batches too large for one pass are split
 * into consecutive passes.  Compiled programs (no tape) are meshed one at a time. */
int s2m_mesh_batch(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs,
                   const s2m_params* params, int want_keys, s2m_mesh_out* outs);
/* One pass only, results left in device buffers the context owns (valid until its next
 * meshing call): outs[i] points at job i's vertices and triangles.  Jobs for the dense
 * path (compiled programs, the Mandelbulb above 128^3 cells, S2M_DENSE_EVAL) are meshed
 * one at a time before the pass, each into buffers of its own. */
int s2m_mesh_batch_device(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs,
                          const s2m_params* params, s2m_device_mesh* outs);

/* ---- autotuning ---------------------------------------------------------------- */
/* No reference counterpart (the reference dispatches one z-slice at a time, main.rs:298-356).
 * The library picks three shapes by rule unless the autotuner has chosen one for the job's
 * key (device, program class, grid, flags; batches: device, job count, total cells):
 *   stream_passes  z-passes of a single tile-sparse job meshed by s2m_mesh (the copy of one
 *                  pass to host memory overlaps the next pass's work)
 *   dense_slab     cell planes per slab of the dense path (Mandelbulb, compiled programs)
 *   batch_cells    most cells per tile-sparse pass of s2m_mesh_batch
 * The tile edge is fixed at 8 cells: the tile kernels' LDS layouts are built for it.
 * s2m_tune times each candidate of the knob that applies to (prog, params) with s2m_mesh,
 * largest device buffers first (one untimed run, then the median of `reps` runs, at least
 * 3; regrowing a buffer frees device memory, which the driver clears on a DMA engine while
 * copies to host memory slow down, so buffers grow once), keeps the first candidate
 * unless a later one is more than 2 % faster (s2m_tune_pick), stores the choice in a
 * process-wide table that every later meshing call of that key uses, and reports it.
 * A key already in the table is returned without timing (measured = 0).  s2m_params.slab
 * != 0 overrides any choice. */
typedef struct s2m_tuning {
    uint32_t stream_passes; /* 0: not this job's knob */
    uint32_t dense_slab;
    uint64_t batch_cells;
    double ms;              /* median time of the chosen shape when it was timed */
    uint32_t measured;      /* candidates timed by this call */
} s2m_tuning;
int s2m_tune(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params, uint32_t reps,
             s2m_tuning* out);
int s2m_tune_batch(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs,
                   const s2m_params* params, uint32_t reps, s2m_tuning* out);
/* The table as a JSON file ({"entries": [{"key": ..., "stream_passes": ..., "dense_slab": ...,
 * "batch_cells": ..., "ms": ...}, ...]}): save writes every entry, load merges a file's
 * entries into the table (S2M_E_IO, S2M_E_PARSE on failure), clear empties it. */
int s2m_tuning_save(const char* path);
int s2m_tuning_load(const char* path);
void s2m_tuning_clear(void);
/* The choice rule on its own: index of the chosen one of n timings (0 for n = 0). */
uint32_t s2m_tune_pick(const double* ms, uint32_t n);

/* ---- serving ------------------------------------------------------------------- */
/* A request loop with continuous batching.  `workers` + `dense_workers` threads each own
 * a context (HIP streams + resident device buffers) on `device`.  A submitted request
 * waits in one of two FIFO lanes: the dense lane (compiled programs, S2M_DENSE_EVAL, and
 * Mandelbulbs above 128^3 cells, meshed alone on the dense path with each slab's mesh
 * copied to host memory while the next one computes) or the pass lane (everything else).
 * An idle worker admits requests from the head of its lane (from the other lane when its
 * own is empty) while their estimated device memory fits in what is left of
 * `hbm_budget`: up to `max_batch` pass-lane requests meshed in one batched pass, or one
 * dense request; then it returns their memory to the budget.  Requests arriving while
 * others are on the device are admitted as soon as a worker and memory free up, and the
 * passes and dense jobs of different workers overlap on the GPU, so a large dense job
 * does not hold back the small requests behind it.  Every result is identical to meshing
 * the request alone with s2m_mesh. */
typedef struct s2m_server s2m_server;

typedef struct s2m_server_config {
    int device;
    uint64_t hbm_budget; /* bytes of device memory admitted requests may use (0: 16 GiB) */
    uint32_t max_batch;  /* requests per pass (0: 64)                                  */
    uint32_t workers;    /* workers preferring the pass lane (0: 2)                    */
    int want_keys;       /* also return cell keys and sign bits                         */
    uint32_t dense_workers; /* workers preferring the dense lane (0: 1)                 */
} s2m_server_config;

typedef struct s2m_request_stats {
    double queue_ms;     /* submit -> admitted to a pass                                */
    double latency_ms;   /* submit -> result ready                                       */
    uint32_t batch_size; /* requests in the pass that meshed it                          */
    uint32_t worker;
} s2m_request_stats;

s2m_server* s2m_server_create(const s2m_server_config* cfg);
/* Queue a request (the tape is copied).  Fails with S2M_E_BUDGET when its estimated
 * memory exceeds the whole budget, S2M_E_VALIDATION for an invalid tape. */
int s2m_server_submit(s2m_server* srv, const s2m_op* ops, uint32_t n_ops, const s2m_params* params,
                      uint64_t* ticket);
/* Queue a request for any program: a tape program's tape is copied as s2m_server_submit
 * does; a compiled program is meshed as it is and must stay valid until its request has
 * been taken.  S2M_E_ARG for a program of another device. */
int s2m_server_submit_program(s2m_server* srv, const s2m_program* prog, const s2m_params* params,
                              uint64_t* ticket);
/* Take a finished request's mesh: S2M_E_BUSY while it is queued or running; with
 * wait != 0 blocks until it is done.  Each ticket can be taken once. */
int s2m_server_take(s2m_server* srv, uint64_t ticket, int wait, s2m_mesh_out* out,
                    s2m_request_stats* stats);
/* Estimated device bytes of one request (the admission measure). */
uint64_t s2m_server_request_bytes(const s2m_op* ops, uint32_t n_ops, const s2m_params* params);
/* Finish queued work, stop the workers and free everything. */
void s2m_server_destroy(s2m_server* srv);

/* ---- output --------------------------------------------------------------------- */
/* ASCII STL exactly as STLWriter (mesh.rs:8-48,155-180), floats in Rust `{}` form. */
int s2m_write_stl(const s2m_mesh_out* mesh, const char* path);
/* ASCII PLY exactly as PLYWriter (mesh.rs:50-141,198-210). */
int s2m_write_ply(const s2m_mesh_out* mesh, const char* path);
/* Dispatch on the extension, case-insensitive (mesh.rs:182-196).  Unknown
 * extensions write nothing and return S2M_OK, as the reference logs and continues. */
int s2m_write_to_file(const s2m_mesh_out* mesh, const char* path);
/* Format one f32 as Rust's `{}` Display does (shortest round-trip digits, never an
 * exponent).  Returns the string length; buf needs at least 64 bytes. */
int s2m_format_f32(float x, char* buf);

/* ---- test hooks ----------------------------------------------------------------- */
/* The page-locked host memory results are copied into: stats[0] blocks page-locked so far
 * (hipHostMalloc calls), [1] page-locked bytes held, [2] idle blocks kept for reuse,
 * [3] their bytes.  A repeated meshing call of the same size adds no block. */
void s2m_host_pool_stats(uint64_t stats[4]);
/* No reference counterpart: returns every idle page-locked block of that pool to the
 * system (results still held by callers are not touched). */
void s2m_host_pool_trim(void);
/* Per-axis coordinate tables used for corner evaluation.  coord receives the
 * distinct coordinates (at most 2*res+1), lo/hi (res entries each) the index of each
 * cell's min and max corner.  Returns the number of coordinates, or 0 on error. */
uint32_t s2m_axis_table(float bmin, float bmax, uint32_t res, int exact_corners,
                        float* coord, uint32_t* lo, uint32_t* hi);
/* Evaluate the program on the corner lattice of params (x fastest), copy it to
 * `out` (capacity cap floats) and report the lattice extent in dims[3]. */
int s2m_eval_lattice(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
                     float* out, uint64_t cap, uint32_t dims[3]);
/* Evaluate the program on the lattice planes that hold the corners of cell planes
 * [z0, z1) (after the resolution rule of params): lattice planes lz0 .. lz0+dims[2]-1 of
 * the corner lattice, written to `out` x fastest.  dims receives the window's extent and
 * *lz0 its first plane.  Lets a test contour one z-window of a grid too large to copy
 * whole.  No reference counterpart (the reference never exposes corner values). */
int s2m_eval_lattice_range(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
                           uint32_t z0, uint32_t z1, float* out, uint64_t cap, uint32_t dims[3],
                           uint32_t* lz0);
/* Contour given per-cell corner values (8 per cell in the order 000,100,010,110,001,
 * 101,011,111; cells in (z,y,x) order) with cell bounds from per-axis tables
 * (cell_lo/cell_hi hold res_x, then res_y, then res_z floats).  No SDF is involved,
 * so normals are zero.  flags: S2M_COMPAT_DROP_LAST_SLICE only. */
int s2m_contour_cells(s2m_ctx* ctx, const float* corners, const uint32_t res[3],
                      const float* cell_lo, const float* cell_hi, uint32_t flags,
                      s2m_mesh_out* out);
/* No reference counterpart: checks a device math identity the Mandelbulb kernels rely on.  The
 * Mandelbulb's escape step takes the sine and cosine of an angle with one sincosf,
 * where the reference's shader calls sin and cos.  Over every float bit pattern,
 * counts[0] / counts[2] receive the sine / cosine results of sincosf that differ from
 * sinf / cosf, counts[1] / counts[3] those of them where both are NaN.  A fraction of a
 * second on one GPU. */
int s2m_check_sincos(s2m_ctx* ctx, uint64_t counts[4]);

/* No reference counterpart: checks a device identity the Mandelbulb kernels rely on.  The power-8
 * Mandelbulb step tests the bailout (`if(r > 2.0) continue;`, examples/mandelmesh.frag:13-14)
 * as r^2 > t, t the largest float whose correctly rounded square root does not exceed
 * the bailout.  Over 2^20 bailouts spread across every bit pattern and the squared
 * lengths around each bailout's square, counts[0] receives the comparisons made and
 * counts[1] those where r^2 > t and sqrtf(r^2) > bailout disagree. */
int s2m_check_bailout(s2m_ctx* ctx, uint64_t counts[2]);

#ifdef __cplusplus
}
#endif
#endif /* SDF2MESH_H */
