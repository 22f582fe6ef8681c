// api.cpp - host runtime behind the C ABI (include/sdf2mesh.h).
//
// One context = one device + one non-blocking HIP stream + device buffers that are
// grown on demand and reused across jobs.  A job is streamed through the device in
// z-slabs of cells; the only host synchronisation per job is the final read of the
// device counters (vertex / triangle totals), after all slabs are queued.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

using namespace s2m;

namespace {

thread_local std::string g_err;
std::mutex g_noalloc_mu[s2m::kMaxDevices];  // s2m_mesh_device_noalloc: one call per device at a time

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(S2M_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                        __FILE__, __LINE__);                                                \
    } while (0)

// Device buffer that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        size_t alloc = std::max<size_t>(need, 256);
        hipError_t e = hipMalloc(&p, alloc);
        if (e == hipSuccess) bytes = alloc;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Page-locked host memory (hipHostMalloc), grown on demand: copies from and to it are
// DMA transfers that do not stage through a driver buffer.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipHostFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        size_t alloc = std::max<size_t>(need, 4096);
        hipError_t e = hipHostMalloc(&p, alloc, hipHostMallocDefault);
        if (e == hipSuccess) bytes = alloc;
        return e;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Per-axis corner coordinate table (host side).
struct AxisTable {
    std::vector<float> coord;
    std::vector<uint32_t> lo, hi;
};

}  // namespace

struct s2m_program {
    int device = 0;
    s2m_op* d_ops = nullptr;
    uint32_t n = 0;
    std::vector<s2m_op> ops;
    bool owns_device_copy = true;  // false: d_ops points into a caller-owned arena
    // compiled programs (hiprtc code object): dense-path kernels instead of a tape
    hipModule_t module = nullptr;
    hipFunction_t k_eval = nullptr, k_normals = nullptr;
};

struct s2m_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int profiling = 0;  // 0 off, 1 total + evaluation kernels, 2 every stage
    s2m_timing timing{};
    DevBuf lattice, vidx, bcount, boffset, actbits, counters, tables, check;
    DevBuf lattice2, actbits2, bcount2;  // the dense path's second slab buffers (evaluation and
                                         // count pass one slab ahead)
    // s2m_mesh_batch_device: the meshes of its dense-path jobs (vertices, keys, signs,
    // triangles: four buffers per job), kept beside the shared pass's output buffers
    std::vector<DevBuf> batch_dense;
    // s2m_mesh_batch: two sets of what a pass leaves for the host (its output buffers, its
    // tables image and the mapped block its counters land in).  Pass p + 1 is queued into
    // one set before pass p, in the other, is read back, and pass p's copy to host memory
    // (on copy_stream) runs beside pass p + 1.  Per set: "pass done" and "copied out".
    struct PassSet {
        DevBuf vtx, keys, signs, tri, tile_tab;
        PinnedBuf host_img, host_back;
        std::vector<uint8_t> dev_img;
        const void* dev_img_at = nullptr;
        hipEvent_t done = nullptr, copied = nullptr;
    };
    PassSet pset[2];
    DevBuf vtx, keys, signs, tri, corners;
    DevBuf tile_tab, tile_lists, tile_lens, tile_flops;
    DevBuf sp_slot_of_tile, sp_tile_of_slot, sp_job_of_slot, sp_cellbits, sp_slot_rows, sp_boundary;
    DevBuf sp_stage, sp_slot_nb, sp_slot_stage, sp_slot_live, sp_slot_hdr, sp_slot_rec, sp_block_stats;
    DevBuf sp_queues, sp_super_iv, sp_super_lists, sp_super_lens, sp_super_act, sp_rows;
    uint64_t slot_cap = 0;
    uint64_t stage_cap = 0;  // staged-cell entries of sp_stage
    int cus = 256;
    std::string arch = "gfx";  // the device's architecture name (autotuner keys)
    PinnedBuf host_img;   // staging of the last pass's tables (kept until its sync)
    std::vector<uint8_t> dev_img;  // the tables image in tile_tab (an identical image is not re-sent)
    const void* dev_img_at = nullptr;  // ... at this device address
    PinnedBuf host_back;  // device counters and per-job ranges read back after a pass
    uint64_t vcap = 0, tcap = 0;
    uint64_t last_nvtx = 0, last_ntri = 0, last_invalid = 0;
    std::vector<hipEvent_t> events;
    size_t events_used = 0;
    // streamed readback (s2m_mesh): device-to-host copies of finished passes run on
    // copy_stream while later passes compute on `stream`
    hipStream_t copy_stream = nullptr;
    // dense slabs: the lattice of slab k + 1 is evaluated on eval_stream while slab k is
    // contoured on `stream`
    hipStream_t eval_stream = nullptr;
    PinnedBuf host_snap;                 // per pass: k_snapshot's totals and capacity use
    // captured graphs of repeated one-pass jobs, most recent first, and the keys of jobs
    // seen once (run_graph)
    static constexpr size_t kGraphs = 8, kSeen = 16;
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
    std::vector<std::string> graph_seen;
    std::vector<hipEvent_t> pass_done;   // per pass: recorded after its snapshot

    hipEvent_t pass_event(size_t k)
    {
        while (pass_done.size() <= k) {
            hipEvent_t e;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
            pass_done.push_back(e);
        }
        return pass_done[k];
    }

    hipEvent_t next_event()
    {
        if (events_used == events.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            events.push_back(e);
        }
        return events[events_used++];
    }
};

// ---------------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------------
namespace s2m {
FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{d, 1, 0};
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    // m = floor(2^32 * (2^l - d) / d) + 1
    const uint64_t num = ((uint64_t(1) << l) - d) << 32;
    f.m = (uint32_t)(num / d + 1);
    f.l = l;
    return f;
}
}  // namespace s2m

namespace {

// cell_bounds (dualcontour.wgsl:22-27) along one axis: size = (max-min)/f32(res-1),
// cmin_i = min + size*f32(i), cmax_i = cmin_i + size.  exact != 0 keeps both values and
// merges those that coincide; otherwise cell i uses lattice points i and i+1 at
// min + size*f32(j).  coord lists distinct coordinates in lattice order; lo/hi give
// each cell's corner indices.
void build_axis(float bmin, float bmax, uint32_t res, bool exact, AxisTable& t)
{
    t.coord.clear();
    t.lo.assign(res, 0);
    t.hi.assign(res, 0);
    if (res == 0) return;
    const float v = (float)(uint32_t)(res - 1u);
    const float size = (bmax - bmin) / v;
    if (!exact) {
        t.coord.resize((size_t)res + 1);
        for (uint32_t j = 0; j <= res; ++j) t.coord[j] = bmin + size * (float)j;
        for (uint32_t i = 0; i < res; ++i) {
            t.lo[i] = i;
            t.hi[i] = i + 1;
        }
        return;
    }
    // lattice position j carries L_j = min + size*j (j < res) and H_j = L_{j-1} + size
    // (j >= 1); both are kept when they differ (NaN never compares equal, so a NaN
    // coordinate is kept twice, like two distinct values).
    t.coord.reserve(2 * (size_t)res + 1);
    float prev_lo = 0.0f;
    for (uint32_t j = 0; j <= res; ++j) {
        const bool has_lo = j < res, has_hi = j >= 1;
        const float L = has_lo ? bmin + size * (float)j : 0.0f;
        const float H = has_hi ? prev_lo + size : 0.0f;
        if (has_hi) {
            t.hi[j - 1] = (uint32_t)t.coord.size();
            t.coord.push_back(H);
        }
        if (has_lo) {
            if (has_hi && L == H && !std::signbit(L) == !std::signbit(H)) {
                t.lo[j] = t.hi[j - 1];
            } else {
                t.lo[j] = (uint32_t)t.coord.size();
                t.coord.push_back(L);
            }
            prev_lo = L;
        }
    }
}

uint32_t round_res(uint32_t res)
{
    // main.rs:141-148: res = 2 << res.ilog2() when res has more than one bit set
    if (__builtin_popcount(res) > 1) {
        uint32_t lg = 31u - (uint32_t)__builtin_clz(res);
        return lg >= 31 ? 0u : (2u << lg);
    }
    return res;
}

bool tape_ok(const s2m_op* ops, uint32_t n)
{
    if (!ops || n == 0 || n > S2M_MAX_TAPE) return false;
    int sp = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = ops[i].code;
        if (c >= S2M_OP_SPHERE && c <= S2M_OP_MANDELBULB) {
            if (++sp > (int)S2M_MAX_STACK) return false;
            if (c == S2M_OP_MANDELBULB && (ops[i].iarg & 0xffffu) > 1024u) return false;
        } else if (c >= S2M_OP_UNION && c <= S2M_OP_SMOOTH_SUBTRACTION) {
            if (sp < 2) return false;
            --sp;
        } else {
            return false;
        }
    }
    return sp == 1;
}

// Algorithmic flops of one op (the static count of the WGSL library text; a translation
// p - c counts 3 unless c = 0, where the WGSL form is the untranslated call).
uint32_t op_flops(const s2m_op& o)
{
    const bool tr = !(o.f[0] == 0.0f && o.f[1] == 0.0f && o.f[2] == 0.0f && !std::signbit(o.f[0]) &&
                      !std::signbit(o.f[1]) && !std::signbit(o.f[2]));
    switch (o.code) {
    case S2M_OP_SPHERE: return 7 + (tr ? 3 : 0);     // length3, -s
    case S2M_OP_BOX: return 19 + (tr ? 3 : 0);       // abs3, -3, max3, length3, max2, min, +
    case S2M_OP_TORUS: return 10 + (tr ? 3 : 0);     // length2, -, length2, -
    case S2M_OP_CYLINDER: return 17 + (tr ? 3 : 0);  // length2, abs2, -2, max, min, +, max2, length2
    case S2M_OP_CAPSULE: return 24;                  // pa, dot, /, clamp, ba*h, -, length3, -r
    case S2M_OP_MANDELBULB: return (o.iarg & 0xffffu) * 30u + 5u;
    case S2M_OP_UNION: case S2M_OP_INTERSECTION: return 1;
    case S2M_OP_SUBTRACTION: return 2;
    default: return 13;                              // smooth operators
    }
}

bool tape_prunable(const std::vector<s2m_op>& ops)
{
    for (const s2m_op& o : ops)
        if (o.code == S2M_OP_MANDELBULB) return false;
    return true;
}

struct Job {
    uint32_t res[3];
    uint32_t nz_cells;  // cell planes processed (res_z, minus the top one under the lag)
    AxisTable ax[3];
    bool exact;
    bool tiled = false;
};

int prepare_job(const s2m_params* p, Job& j)
{
    if (!p) return fail(S2M_E_ARG, "params is NULL");
    for (int a = 0; a < 3; ++a) {
        uint32_t r = p->res[a];
        if (p->flags & S2M_COMPAT_POW2) r = round_res(r);
        if (r > 65535u)
            return fail(S2M_E_ARG, "resolution %u exceeds 65535 (u16 cell keys, mesh.rs:214)", r);
        j.res[a] = r;
    }
    j.exact = (p->flags & S2M_CORNERS_EXACT) != 0;
    for (int a = 0; a < 3; ++a) build_axis(p->bb_min[a], p->bb_max[a], j.res[a], j.exact, j.ax[a]);
    j.nz_cells = j.res[2];
    if ((p->flags & S2M_COMPAT_DROP_LAST_SLICE) && j.nz_cells > 0) j.nz_cells -= 1;
    return S2M_OK;
}

uint32_t pick_slab(const s2m_params* p, const Job& j)
{
    const uint64_t plane = (uint64_t)j.res[0] * j.res[1];
    if (plane == 0) return 1;
    uint64_t t = p->slab;
    // ~32M cells per slab: at 1024^2 cells per plane, 32 planes (fewer, fuller launches;
    // measured 66 ms per Mandelbulb job against 75 ms at 8 planes, no gain beyond 32)
    if (t == 0) t = std::max<uint64_t>(1, (32ull << 20) / plane);
    // keep the per-slab block count within one scan block
    const uint64_t max_t = std::max<uint64_t>(1, ((uint64_t)kMaxScanBlocks * kCellsPerBlock) / plane);
    t = std::min<uint64_t>(t, max_t);
    t = std::min<uint64_t>(t, std::max<uint32_t>(j.nz_cells, 1));
    return (uint32_t)t;
}

// Bytes of the three axis tables as upload_tables lays them out.
size_t tables_bytes(const Job& j)
{
    size_t nf = 0, nu = 0;
    for (int a = 0; a < 3; ++a) {
        nf += j.ax[a].coord.size();
        nu += 2 * j.ax[a].lo.size();
    }
    return nf * 4 + nu * 4 + 64;
}

// Upload the three axis tables: [cx cy cz | lox hix loy hiy loz hiz] as one buffer, into
// `dst` (tables_bytes(j) bytes of device memory) or, when dst is NULL, the context's
// table buffer.
int upload_tables(s2m_ctx* ctx, const Job& j, const float* (&c)[3], const uint32_t* (&lo)[3],
                  const uint32_t* (&hi)[3], void* dst = nullptr)
{
    size_t nf = 0;
    for (int a = 0; a < 3; ++a) nf += j.ax[a].coord.size();
    std::vector<uint8_t> host(tables_bytes(j));
    float* fp = reinterpret_cast<float*>(host.data());
    uint32_t* up = reinterpret_cast<uint32_t*>(host.data() + nf * 4);
    size_t fo[3], lo_o[3], hi_o[3], f = 0, u = 0;
    for (int a = 0; a < 3; ++a) {
        fo[a] = f;
        std::copy(j.ax[a].coord.begin(), j.ax[a].coord.end(), fp + f);
        f += j.ax[a].coord.size();
        lo_o[a] = u;
        std::copy(j.ax[a].lo.begin(), j.ax[a].lo.end(), up + u);
        u += j.ax[a].lo.size();
        hi_o[a] = u;
        std::copy(j.ax[a].hi.begin(), j.ax[a].hi.end(), up + u);
        u += j.ax[a].hi.size();
    }
    if (!dst) {
        HIP_TRY(ctx->tables.ensure(host.size()));
        dst = ctx->tables.p;
    }
    HIP_TRY(hipMemcpyAsync(dst, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    // the host vector must outlive the async copy
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const float* dfp = static_cast<const float*>(dst);
    const uint32_t* dup = reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(dst) + nf * 4);
    for (int a = 0; a < 3; ++a) {
        c[a] = dfp + fo[a];
        lo[a] = dup + lo_o[a];
        hi[a] = dup + hi_o[a];
    }
    return S2M_OK;
}

// ---------------------------------------------------------------------------------
// Tile-sparse pipeline (sparse.hip): host side.
// ---------------------------------------------------------------------------------
// Device tables of one job, packed into one host byte image: axis coordinate tables,
// cell -> coordinate index tables, tile extents and index boxes, per-op flops.
struct SpTables {
    std::vector<uint8_t> bytes;
    size_t off_c[3], off_lo[3], off_hi[3], off_ext, off_ibox, off_flops;
    uint32_t nt[3];
    uint32_t bmax[3];
    double pad;
};

void build_sp_tables(const Job& j, const s2m_program* prog, float eps, SpTables& T)
{
    std::vector<uint8_t>& B = T.bytes;
    B.clear();
    auto put = [&](const void* src, size_t n) {
        const size_t off = (B.size() + 255) & ~size_t(255);
        B.resize(off + n);
        if (n) std::memcpy(B.data() + off, src, n);
        return off;
    };
    for (int a = 0; a < 3; ++a) {
        T.off_c[a] = put(j.ax[a].coord.data(), j.ax[a].coord.size() * 4);
        T.off_lo[a] = put(j.ax[a].lo.data(), j.ax[a].lo.size() * 4);
        T.off_hi[a] = put(j.ax[a].hi.data(), j.ax[a].hi.size() * 4);
    }
    const uint32_t ncell[3] = {j.res[0], j.res[1], j.nz_cells};
    for (int a = 0; a < 3; ++a) {
        T.nt[a] = (ncell[a] + kTile - 1) / kTile;
        T.bmax[a] = 1;
    }
    const size_t next = 2 * ((size_t)T.nt[0] + T.nt[1] + T.nt[2]);
    std::vector<float> ext(next);
    std::vector<uint32_t> ibox(next);
    double maxabs = 0;
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
        const AxisTable& ax = j.ax[a];
        for (uint32_t ti = 0; ti < T.nt[a]; ++ti) {
            const uint32_t c0 = ti * kTile, c1 = std::min(c0 + kTile, ncell[a]) - 1;
            const uint32_t i0 = ax.lo[c0], i1 = ax.hi[c1];
            float mn = ax.coord[i0], mx = ax.coord[i0];
            for (uint32_t i = i0; i <= i1; ++i) {
                mn = std::fmin(mn, ax.coord[i]);
                mx = std::fmax(mx, ax.coord[i]);
                maxabs = std::fmax(maxabs, std::fabs((double)ax.coord[i]));
            }
            ext[off + ti] = mn;
            ext[off + T.nt[a] + ti] = mx;
            ibox[off + ti] = i0;
            ibox[off + T.nt[a] + ti] = i1;
            T.bmax[a] = std::max(T.bmax[a], i1 - i0 + 1);
        }
        off += 2 * (size_t)T.nt[a];
    }
    T.off_ext = put(ext.data(), next * 4);
    T.off_ibox = put(ibox.data(), next * 4);
    std::vector<uint32_t> fl(prog->n);
    for (uint32_t i = 0; i < prog->n; ++i) fl[i] = op_flops(prog->ops[i]);
    T.off_flops = put(fl.data(), fl.size() * 4);
    // tile boxes must also hold the normal taps p +- eps of vertices in their cells and
    // absorb the fp32 rounding of p + eps and of the vertex position
    T.pad = 2.0 * std::fabs((double)eps) + 1e-6 * (1.0 + maxabs);
}

// Result range of one job of a sparse run, in the context's output buffers.
struct SpRange {
    uint64_t v0, nv, t0, nt, invalid;
};

// Host arrays one job's mesh is copied into while the job runs (s2m_mesh).  After each
// pass of the sparse pipeline (or slab of the dense one) the compute stream records a
// snapshot of the totals (k_snapshot, into mapped host memory) and an event;
// readback_pass waits for a pass's event and queues the copy of that pass's vertices and
// triangles on the copy stream.  The host calls it for pass k right after queueing pass
// k + 1, so the copy of pass k starts as soon as pass k ends and overlaps the computation
// of pass k + 1 (waiting until every pass is queued would delay the first copy by the
// launches of all later passes).  Vertices and triangles are appended in pass order and
// never rewritten, so a finished pass's range is final.
struct Readback {
    int want_keys = 0;
    float* vtx = nullptr;
    uint32_t* tri = nullptr;
    uint64_t* keys = nullptr;
    uint8_t* signs = nullptr;
    uint64_t vcap = 0, tcap = 0;      // entries of the host arrays
    uint64_t v_done = 0, t_done = 0;  // copies queued so far
    uint32_t passes_done = 0;         // passes whose copies are queued (or that stopped them)
    bool stopped = false;             // a pass overflowed a capacity: the attempt repeats
    hipEvent_t first = nullptr, last = nullptr;  // profiling: around the copies

    void release()
    {
        host_release(vtx);
        host_release(tri);
        host_release(keys);
        host_release(signs);
        vtx = nullptr;
        tri = nullptr;
        keys = nullptr;
        signs = nullptr;
        vcap = tcap = 0;
    }
};

// Before an attempt of the job: host arrays of at least vcap vertices and tcap triangles
// (kept from an earlier call when large enough), no copy queued, no earlier attempt's copy
// still reading the device buffers the attempt will write.
int readback_begin(s2m_ctx* ctx, Readback& rb, uint64_t vcap, uint64_t tcap)
{
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    if (rb.vcap < vcap || !rb.vtx) {
        host_release(rb.vtx);
        host_release(rb.keys);
        host_release(rb.signs);
        rb.keys = nullptr;
        rb.signs = nullptr;
        rb.vtx = static_cast<float*>(host_alloc(std::max<uint64_t>(vcap, 1) * 6 * sizeof(float)));
        if (rb.want_keys) {
            rb.keys = static_cast<uint64_t*>(host_alloc(std::max<uint64_t>(vcap, 1) * sizeof(uint64_t)));
            rb.signs = static_cast<uint8_t*>(host_alloc(std::max<uint64_t>(vcap, 1)));
        }
        rb.vcap = vcap;
    }
    if (rb.tcap < tcap || !rb.tri) {
        host_release(rb.tri);
        rb.tri = static_cast<uint32_t*>(host_alloc(std::max<uint64_t>(tcap, 1) * 3 * sizeof(uint32_t)));
        rb.tcap = tcap;
    }
    if (!rb.vtx || !rb.tri || (rb.want_keys && (!rb.keys || !rb.signs)))
        return fail(S2M_E_OOM, "host allocation of %llu vertices / %llu triangles failed",
                    (unsigned long long)vcap, (unsigned long long)tcap);
    rb.v_done = rb.t_done = 0;
    rb.passes_done = 0;
    rb.stopped = false;
    rb.first = rb.last = nullptr;
    return S2M_OK;
}

// After pass (or slab) k's kernels: its snapshot and event.
int readback_mark(s2m_ctx* ctx, const uint64_t* counters, uint32_t k)
{
    void* dev = nullptr;
    HIP_TRY(hipHostGetDevicePointer(&dev, ctx->host_snap.p, 0));
    HIP_TRY(launch_snapshot(ctx->stream, counters, static_cast<uint64_t*>(dev) + 4 * (size_t)k));
    hipEvent_t e = ctx->pass_event(k);
    if (!e) return fail(S2M_E_HIP, "event creation failed");
    HIP_TRY(hipEventRecord(e, ctx->stream));
    return S2M_OK;
}

// Queue the copies of [rb.v_done, nv) vertices and [rb.t_done, nt) triangles.
int readback_copy(s2m_ctx* ctx, Readback& rb, uint64_t nv, uint64_t nt)
{
    hipStream_t s = ctx->copy_stream;
    if (ctx->profiling && !rb.first) {
        rb.first = ctx->next_event();
        if (rb.first) HIP_TRY(hipEventRecord(rb.first, s));
    }
    if (nv > rb.v_done) {
        const uint64_t v0 = rb.v_done, m = nv - v0;
        HIP_TRY(hipMemcpyAsync(rb.vtx + 6 * v0, ctx->vtx.as<float>() + 6 * v0, m * 6 * sizeof(float),
                               hipMemcpyDeviceToHost, s));
        if (rb.want_keys) {
            HIP_TRY(hipMemcpyAsync(rb.keys + v0, ctx->keys.as<uint64_t>() + v0, m * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(rb.signs + v0, ctx->signs.as<uint8_t>() + v0, m, hipMemcpyDeviceToHost, s));
        }
        rb.v_done = nv;
    }
    if (nt > rb.t_done) {
        const uint64_t t0 = rb.t_done;
        HIP_TRY(hipMemcpyAsync(rb.tri + 3 * t0, ctx->tri.as<uint32_t>() + 3 * t0, (nt - t0) * 3 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
        rb.t_done = nt;
    }
    return S2M_OK;
}

// Pass rb.passes_done is queued and followed by readback_mark: once it finishes, queue
// the copies of its vertices and triangles.  A pass whose totals or capacity use exceed
// the attempt's capacities stops the copies (the attempt will repeat).
int readback_pass(s2m_ctx* ctx, Readback& rb, uint64_t vcap, uint64_t tcap, uint64_t slot_cap, uint64_t stage_cap)
{
    const uint32_t k = rb.passes_done++;
    if (rb.stopped) return S2M_OK;
    const volatile uint64_t* snap = ctx->host_snap.as<uint64_t>();
    HIP_TRY(hipEventSynchronize(ctx->pass_done[k]));
    const uint64_t nv = snap[4 * k], nt = snap[4 * k + 1], slots = snap[4 * k + 2], staged = snap[4 * k + 3];
    if (nv > vcap || nt > tcap || nv > rb.vcap || nt > rb.tcap || slots > slot_cap || staged > stage_cap ||
        nv < rb.v_done || nt < rb.t_done) {
        rb.stopped = true;
        return S2M_OK;
    }
    HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->pass_done[k], 0));
    return readback_copy(ctx, rb, nv, nt);
}

// Every pass up to npass-1 whose copies are not queued yet, in order.
int readback_passes(s2m_ctx* ctx, Readback& rb, uint32_t npass, uint64_t vcap, uint64_t tcap, uint64_t slot_cap,
                    uint64_t stage_cap)
{
    while (rb.passes_done < npass) {
        const int rc = readback_pass(ctx, rb, vcap, tcap, slot_cap, stage_cap);
        if (rc) return rc;
    }
    return S2M_OK;
}

// After a successful attempt with nv vertices and nt triangles: the rest copied (when a
// pass stopped the streaming) and every copy finished.
int readback_finish(s2m_ctx* ctx, Readback& rb, uint64_t nv, uint64_t nt)
{
    if (!rb.vtx || nv > rb.vcap || nt > rb.tcap) {  // not streamed, or larger than its arrays
        HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
        rb.release();
        const int brc = readback_begin(ctx, rb, std::max<uint64_t>(nv, 1), std::max<uint64_t>(nt, 1));
        if (brc) return brc;
    }
    int rc = readback_copy(ctx, rb, nv, nt);
    if (rc) return rc;
    if (rb.first) {
        rb.last = ctx->next_event();
        if (rb.last) HIP_TRY(hipEventRecord(rb.last, ctx->copy_stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    if (rb.first && rb.last) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, rb.first, rb.last));
        ctx->timing.d2h_ms = ms;
    }
    return S2M_OK;
}

// Passes of a single sparse job whose mesh is streamed to host memory: enough that the
// copy of one pass overlaps the next pass's work (copying a CSG-64 512^3 mesh over PCIe
// takes about twice its device time), few enough that every pass keeps the device busy.
uint32_t stream_passes(uint32_t tile_layers)
{
    return std::max<uint32_t>(1, std::min<uint32_t>(4, tile_layers / 16));
}

// Is the program a left fold of library primitives (P0 P1 op1 P2 op2 ...) short enough to
// be staged (sparse.hip, kStageOps = 128)?  Every pruned sub-program of one is one too:
// pruning keeps either the accumulated value or the right operand, a single primitive.
// Is the program a left fold of unions and smooth unions (k > 0) of library primitives
// other than the Mandelbulb (k_sp_prune_fold)?
bool union_fold_program(const std::vector<s2m_op>& ops)
{
    const size_t n = ops.size();
    if (n == 0 || (n & 1u) == 0) return false;
    for (size_t k = 0; k < n; ++k) {
        const uint32_t c = ops[k].code;
        if (k == 0 || (k & 1u)) {
            if (c >= S2M_OP_UNION || c == S2M_OP_MANDELBULB) return false;
        } else if (!(c == S2M_OP_UNION || (c == S2M_OP_SMOOTH_UNION && ops[k].f[0] > 0.0f))) {
            return false;
        }
    }
    return true;
}

bool left_fold_program(const std::vector<s2m_op>& ops)
{
    const size_t n = ops.size();
    if (n == 0 || n > 128 || (n & 1u) == 0) return false;
    for (size_t k = 0; k < n; ++k) {
        const bool prim = ops[k].code < S2M_OP_UNION;
        if (prim != (k == 0 || (k & 1u)) || ops[k].code == S2M_OP_MANDELBULB) return false;
    }
    return true;
}

// Run `enqueue` (which queues a job's kernels on the context's stream) as a graph: replay
// the context's graph captured for `key`; capture it and replay it when `key` was seen
// before (a repeated job); else queue the kernels as usual and remember `key`.  The
// context keeps the kGraphs graphs and kSeen keys used last.  A graph holds kernels only
// (every clear of the path is a kernel).
template <class F>
int run_graph(s2m_ctx* ctx, const std::string& key, F enqueue)
{
    auto& G = ctx->graphs;
    for (size_t i = 0; i < G.size(); ++i)
        if (G[i].first == key) {
            std::rotate(G.begin(), G.begin() + i, G.begin() + i + 1);  // most recent first
            HIP_TRY(hipGraphLaunch(G[0].second, ctx->stream));
            ctx->timing.graph_launches = 1;
            return S2M_OK;
        }
    auto& S = ctx->graph_seen;
    auto it = std::find(S.begin(), S.end(), key);
    if (it == S.end()) {
        S.insert(S.begin(), key);
        if (S.size() > s2m_ctx::kSeen) S.pop_back();
        return enqueue();
    }
    S.erase(it);
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(S2M_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return fail(S2M_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
    if (G.size() == s2m_ctx::kGraphs) {
        (void)hipGraphExecDestroy(G.back().second);
        G.pop_back();
    }
    G.insert(G.begin(), {key, x});
    HIP_TRY(hipGraphLaunch(x, ctx->stream));
    ctx->timing.graph_launches = 1;
    return S2M_OK;
}

// Autotuner key of a single job: device, program class, grid and the flags that change
// the work (corner mode, lag, normals, dense evaluation).
std::string tune_key(const s2m_ctx* ctx, const s2m_program* prog, const s2m_params* p, const Job& j)
{
    const char* cls = prog->module ? "compiled"
                      : ((p->flags & S2M_DENSE_EVAL) || !tape_prunable(prog->ops)) ? "dense"
                      : union_fold_program(prog->ops)                             ? "fold"
                      : left_fold_program(prog->ops)                              ? "linear"
                                                                                  : "tape";
    char buf[192];
    std::snprintf(buf, sizeof buf, "%s:%d|%s|%ux%ux%u|%u", ctx->arch.c_str(), ctx->cus, cls, j.res[0], j.res[1],
                  j.res[2],
                  p->flags & (S2M_CORNERS_EXACT | S2M_COMPAT_DROP_LAST_SLICE | S2M_NO_NORMALS | S2M_DENSE_EVAL));
    return buf;
}

// Autotuner key of a batch: device, job count and total cells (to a power of two).
std::string tune_batch_key(const s2m_ctx* ctx, uint32_t n, uint64_t cells)
{
    uint32_t lg = 0;
    while (lg < 63 && (1ull << (lg + 1)) <= cells) ++lg;
    char buf[128];
    std::snprintf(buf, sizeof buf, "%s:%d|batch|%u|2^%u", ctx->arch.c_str(), ctx->cus, n, lg);
    return buf;
}

// Capacities a sparse pass was queued with (slots, staged cells, vertices, triangles).
struct SparseCaps {
    uint64_t want_v = 0, want_t = 0, want_s = 0, stage_cap = 0;
    uint64_t tiles_pass = 0, tile_blocks = 0;
};

// A batch pass queued by run_sparse without waiting for it (s2m_mesh_batch reads it back
// with sparse_read once its event has passed).
struct SparsePending {
    bool active = false;
    uint32_t n = 0, npass = 0;
    uint64_t tiles_total = 0, graph_launches = 0;
    SparseCaps caps;
};

// Exchange the context's pass outputs, tables image and read-back block with batch set k
// (twice restores them): run_sparse then writes set k.
void swap_pass_set(s2m_ctx* ctx, int k)
{
    s2m_ctx::PassSet& S = ctx->pset[k];
    std::swap(ctx->vtx, S.vtx);
    std::swap(ctx->keys, S.keys);
    std::swap(ctx->signs, S.signs);
    std::swap(ctx->tri, S.tri);
    std::swap(ctx->tile_tab, S.tile_tab);
    std::swap(ctx->host_img, S.host_img);
    std::swap(ctx->host_back, S.host_back);
    std::swap(ctx->dev_img, S.dev_img);
    std::swap(ctx->dev_img_at, S.dev_img_at);
}

// A finished sparse pass: the counters and per-job rows its last kernel wrote to `back`
// (mapped host memory).  Sets the context's timing and last-mesh fields; *fits is true
// with every job's range in `out`, false when a capacity overflowed, with `c` grown for
// the next attempt.
int sparse_read(s2m_ctx* ctx, const uint64_t* back, uint32_t n, uint32_t npass, uint64_t tiles_total,
                SparseCaps& c, std::vector<SpRange>& out, bool* fits)
{
    const uint64_t* cnt = back;
    const uint64_t* jo = cnt + kNumCounters;
    *fits = false;
    ctx->timing.eval_launches = npass;
    ctx->timing.slabs = npass;
    ctx->timing.lattice_points = cnt[kEvalPoints];
    ctx->timing.tiles_total = tiles_total;
    ctx->timing.tiles_active = cnt[kTilesActive];
    ctx->timing.eval_flops = cnt[kEvalFlops];
    if (cnt[kVtxTotal] >= 0xffffffffull || cnt[kTriTotal] >= 0xffffffffull)
        return fail(S2M_E_ARG, "output of %llu vertices / %llu triangles exceeds u32 indices",
                    (unsigned long long)cnt[kVtxTotal], (unsigned long long)cnt[kTriTotal]);
    const bool ok_v = cnt[kVtxTotal] <= c.want_v, ok_t = cnt[kTriTotal] <= c.want_t;
    const bool ok_s = cnt[kSlotsMax] <= c.want_s;
    const bool ok_stage = cnt[kStageMax] <= c.stage_cap;
    if (ok_v && ok_t && ok_s && ok_stage) {
        out.assign(n, SpRange{});
        for (uint32_t i = 0; i < n; ++i) {
            SpRange& r = out[i];
            r.v0 = jo[(size_t)i * kJobOut + 0];
            r.t0 = jo[(size_t)i * kJobOut + 1];
            r.invalid = jo[(size_t)i * kJobOut + 2];
            const uint64_t v1 = (i + 1 < n) ? jo[(size_t)(i + 1) * kJobOut + 0] : cnt[kVtxTotal];
            const uint64_t t1 = (i + 1 < n) ? jo[(size_t)(i + 1) * kJobOut + 1] : cnt[kTriTotal];
            r.nv = v1 - r.v0;
            r.nt = t1 - r.t0;
        }
        ctx->last_nvtx = cnt[kVtxTotal];
        ctx->last_ntri = cnt[kTriTotal];
        ctx->last_invalid = cnt[kInvalid];
        *fits = true;
        return S2M_OK;
    }
    if (!ok_s) c.want_s = std::min<uint64_t>(c.tiles_pass, cnt[kSlotsMax] + cnt[kSlotsMax] / 8 + 64);
    if (!ok_stage) c.stage_cap = cnt[kStageMax] + cnt[kStageMax] / 8 + c.tile_blocks * kStageChunkHost;
    if (!ok_v) c.want_v = cnt[kVtxTotal] + cnt[kVtxTotal] / 8 + 1024;
    if (!ok_t) c.want_t = cnt[kTriTotal] + cnt[kTriTotal] / 8 + 2048;
    return S2M_OK;
}

// Mesh `n` jobs with the sparse pipeline.  n == 1 may be cut into z-slabs; a batch runs
// as one pass with every job whole.  Capacities (slots, vertices, triangles) start from
// estimates and grow if the device counters report more.  With `pend` (a batch, no
// profiling) the pass is only queued: the caller reads it back with sparse_read once it
// has finished, and queues it again (without `pend`) if a capacity overflowed.
int run_sparse(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs, const s2m_params* const* ps,
               const Job* jobs, std::vector<SpRange>& out, Readback* rb = nullptr, SparsePending* pend = nullptr)
{
    if (pend) pend->active = false;
    out.assign(n, SpRange{});
    std::vector<SpTables> tabs(n);
    std::vector<size_t> tab_off(n);
    size_t tab_bytes = 0;
    uint32_t list_stride = 1, box_stride = 1, prune_depth = 1;
    bool general_tiles = false;  // a tile's pruned program may not be a left fold
    bool fold_jobs = false, other_jobs = false;  // programs that are / are not union folds
    uint64_t ntiles_all = 0, nrows_all = 0, plane_max = 0, cells_all = 0, planes_sum = 0;
    for (uint32_t i = 0; i < n; ++i) {
        build_sp_tables(jobs[i], progs[i], ps[i]->eps, tabs[i]);
        tab_off[i] = tab_bytes;
        tab_bytes += (tabs[i].bytes.size() + 255) & ~size_t(255);
        list_stride = std::max(list_stride, progs[i]->n);
        if (!left_fold_program(progs[i]->ops)) general_tiles = true;
        (union_fold_program(progs[i]->ops) ? fold_jobs : other_jobs) = true;
        box_stride = std::max(box_stride, tabs[i].bmax[0] * tabs[i].bmax[1] * tabs[i].bmax[2]);
        uint32_t depth = 0;  // operand stack entries below the top (k_sp_prune's LDS stack)
        for (const s2m_op& o : progs[i]->ops) {
            if (o.code < S2M_OP_UNION) prune_depth = std::max(prune_depth, depth++);
            else --depth;
        }
        ntiles_all += (uint64_t)tabs[i].nt[0] * tabs[i].nt[1] * tabs[i].nt[2];
        nrows_all += (uint64_t)jobs[i].nz_cells * jobs[i].res[1];
        plane_max = std::max<uint64_t>(plane_max, (uint64_t)jobs[i].res[0] * jobs[i].res[1]);
        cells_all += (uint64_t)jobs[i].res[0] * jobs[i].res[1] * jobs[i].nz_cells;
        planes_sum += (uint64_t)jobs[i].res[0] * jobs[i].res[1];
    }
    if (box_stride * sizeof(float) > 40u * 1024u)
        return fail(S2M_E_ARG, "tile box of %u points exceeds the LDS staging limit", box_stride);
    // a tile's box spans lattice positions c0 .. c0 + 8: L of the first, H of the last and at
    // most both of each of the 7 between, so at most 16 coordinates per axis (k_sp_tiles
    // gives each box row (z, y) a thread)
    for (uint32_t i = 0; i < n; ++i)
        if (tabs[i].bmax[0] > 16 || tabs[i].bmax[1] > 16 || tabs[i].bmax[2] > 16)
            return fail(S2M_E_ARG, "tile box of %u x %u x %u lattice points", tabs[i].bmax[0], tabs[i].bmax[1],
                        tabs[i].bmax[2]);
    // slabs (single job only): whole tile layers, per-pass buffers bounded to ~1 GiB
    uint32_t layers = tabs[0].nt[2];
    if (n == 1) {
        if (ps[0]->slab) layers = std::max<uint32_t>(1, (ps[0]->slab + kTile - 1) / kTile);
        else if (rb) {  // streamed to host memory: the autotuner's passes, else the rule's
            TuneChoice tc;
            uint32_t passes = stream_passes(tabs[0].nt[2]);
            if (tune_lookup(tune_key(ctx, progs[0], ps[0], jobs[0]), &tc) && tc.stream_passes)
                passes = tc.stream_passes;
            passes = std::min(passes, std::max<uint32_t>(tabs[0].nt[2], 1));
            layers = (tabs[0].nt[2] + passes - 1) / passes;
        }
        layers = std::min(layers, std::max<uint32_t>(tabs[0].nt[2], 1));
        auto pass_bytes = [&](uint64_t L) {
            const uint64_t t = (uint64_t)tabs[0].nt[0] * tabs[0].nt[1] * L;
            return t * (list_stride * 2 + 12) + 16ull * L * kTile * jobs[0].res[1] * tabs[0].nt[0];
        };
        while (layers > 1 && pass_bytes(layers) > (1ull << 30)) layers = (layers + 1) / 2;
    }
    const uint32_t npass = n == 1 ? (tabs[0].nt[2] + layers - 1) / std::max(layers, 1u) : 1;
    if (npass == 1) rb = nullptr;  // nothing to overlap: the caller copies the finished mesh
    const uint64_t tiles_pass = n == 1 ? (uint64_t)tabs[0].nt[0] * tabs[0].nt[1] * layers : ntiles_all;
    const uint64_t rows_pass = n == 1 ? (uint64_t)layers * kTile * jobs[0].res[1] : nrows_all;
    if (tiles_pass > 0x7fffffffull || rows_pass > 0xffffffffull || plane_max > 0xffffffffull ||
        (rows_pass + 4095) / 4096 > kMaxScanBlocks)
        return fail(S2M_E_ARG, "grid too large for one pass; mesh the jobs separately");

    // device tables + job descriptors (one per pass for a single job's slabs, else per job)
    const size_t njob_desc = n == 1 ? npass : n;
    const size_t desc_off = (tab_bytes + 255) & ~size_t(255);
    const size_t img_bytes = desc_off + njob_desc * sizeof(SpJob);
    HIP_TRY(ctx->host_img.ensure(img_bytes));
    uint8_t* img = ctx->host_img.as<uint8_t>();
    std::memset(img, 0, img_bytes);
    for (uint32_t i = 0; i < n; ++i)
        std::memcpy(img + tab_off[i], tabs[i].bytes.data(), tabs[i].bytes.size());
    HIP_TRY(ctx->tile_tab.ensure(img_bytes));
    uint8_t* dimg = ctx->tile_tab.as<uint8_t>();
    std::vector<SpJob> desc(njob_desc);
    uint64_t tile_base = 0, super_base = 0, nsuper_pass = 0, row_base = 0, trow_base = 0;
    uint64_t nrows_pass = 0, ntrows_pass = 0;
    for (size_t d = 0; d < njob_desc; ++d) {
        const uint32_t i = n == 1 ? 0 : (uint32_t)d;
        const Job& j = jobs[i];
        const SpTables& T = tabs[i];
        uint8_t* base = dimg + tab_off[i];
        SpJob J{};
        J.rx = j.res[0];
        J.ry = j.res[1];
        if (n == 1) {
            J.z0 = (uint32_t)d * layers * kTile;
            J.nz = std::min(j.nz_cells - std::min(j.nz_cells, J.z0), layers * kTile);
            J.tz0 = (uint32_t)d * layers;
            J.tile_base = 0;
            J.row_base = 0;
            J.trow_base = 0;
        } else {
            J.z0 = 0;
            J.nz = j.nz_cells;
            J.tz0 = 0;
            J.tile_base = tile_base;
            J.row_base = row_base;
            J.trow_base = trow_base;
            tile_base += (uint64_t)T.nt[0] * T.nt[1] * T.nt[2];
            row_base += (uint64_t)j.nz_cells * j.res[1];
            trow_base += (uint64_t)T.nt[2] * T.nt[1];
        }
        {   // rows and tile rows of this job's pass (per pass for a single job's slabs)
            const uint64_t r = (uint64_t)J.nz * J.ry, tr = (uint64_t)((J.nz + kTile - 1) / kTile) * T.nt[1];
            nrows_pass = n == 1 ? std::max(nrows_pass, r) : row_base;
            ntrows_pass = n == 1 ? std::max(ntrows_pass, tr) : trow_base;
        }
        J.ntx = T.nt[0];
        J.nty = T.nt[1];
        J.ntz = T.nt[2];
        // super-tiles: kSuper^3 tiles, over the tile layers this pass holds
        J.nsx = (J.ntx + kSuper - 1) / kSuper;
        J.nsy = (J.nty + kSuper - 1) / kSuper;
        J.nsz = ((J.nz + kTile - 1) / kTile + kSuper - 1) / kSuper;
        const uint64_t nsup = (uint64_t)J.nsx * J.nsy * J.nsz;
        J.super_base = n == 1 ? 0 : super_base;
        super_base += nsup;
        nsuper_pass = n == 1 ? std::max(nsuper_pass, nsup) : super_base;
        J.nops = progs[i]->n;
        J.union_fold = union_fold_program(progs[i]->ops) ? 1u : 0u;
        J.left_fold = left_fold_program(progs[i]->ops) ? 1u : 0u;
        J.flags = ps[i]->flags;
        J.bx = T.bmax[0];
        J.by = T.bmax[1];
        J.out_index = i;
        J.eps = ps[i]->eps;
        J.pad = T.pad;
        J.ops = progs[i]->d_ops;
        J.cx = reinterpret_cast<const float*>(base + T.off_c[0]);
        J.cy = reinterpret_cast<const float*>(base + T.off_c[1]);
        J.cz = reinterpret_cast<const float*>(base + T.off_c[2]);
        J.lox = reinterpret_cast<const uint32_t*>(base + T.off_lo[0]);
        J.hix = reinterpret_cast<const uint32_t*>(base + T.off_hi[0]);
        J.loy = reinterpret_cast<const uint32_t*>(base + T.off_lo[1]);
        J.hiy = reinterpret_cast<const uint32_t*>(base + T.off_hi[1]);
        J.loz = reinterpret_cast<const uint32_t*>(base + T.off_lo[2]);
        J.hiz = reinterpret_cast<const uint32_t*>(base + T.off_hi[2]);
        J.ext = reinterpret_cast<const float*>(base + T.off_ext);
        J.ibox = reinterpret_cast<const uint32_t*>(base + T.off_ibox);
        J.op_flops = reinterpret_cast<const uint32_t*>(base + T.off_flops);
        J.boundary = ctx->sp_boundary.as<uint32_t>();
        desc[d] = J;
    }
    HIP_TRY(ctx->sp_boundary.ensure(std::max<uint64_t>(plane_max, 1) * sizeof(uint32_t)));
    for (auto& J : desc) J.boundary = ctx->sp_boundary.as<uint32_t>();
    std::memcpy(img + desc_off, desc.data(), desc.size() * sizeof(SpJob));
    // the tables of consecutive calls with the same jobs are identical: send a changed image only
    if (ctx->dev_img_at != dimg || ctx->dev_img.size() != img_bytes ||
        std::memcmp(ctx->dev_img.data(), img, img_bytes) != 0) {
        ctx->dev_img_at = nullptr;  // until the copy is queued
        HIP_TRY(hipMemcpyAsync(dimg, img, img_bytes, hipMemcpyHostToDevice, ctx->stream));
        ctx->dev_img.assign(img, img + img_bytes);
        ctx->dev_img_at = dimg;
    }
    const SpJob* ddesc = reinterpret_cast<const SpJob*>(dimg + desc_off);

    HIP_TRY(ctx->tile_lists.ensure(tiles_pass * list_stride * sizeof(uint16_t)));
    HIP_TRY(ctx->tile_lens.ensure(tiles_pass * sizeof(uint32_t)));
    HIP_TRY(ctx->tile_flops.ensure(tiles_pass * sizeof(uint32_t)));
    HIP_TRY(ctx->sp_slot_of_tile.ensure(tiles_pass * sizeof(int32_t)));
    HIP_TRY(ctx->sp_rows.ensure(std::max<uint64_t>(nrows_pass, 1) * 4 * sizeof(uint32_t)));
    const uint32_t nsb_max = (uint32_t)((nrows_pass + kRowChunk - 1) / kRowChunk) + 1;  // vertex and triangle sums
    HIP_TRY(ctx->bcount.ensure((size_t)2 * nsb_max * kChunkSumStride * sizeof(uint32_t)));
    // the counters, then per job kJobOut words (first vertex, first triangle, invalid quads)
    HIP_TRY(ctx->counters.ensure((kNumCounters + (size_t)n * kJobOut) * sizeof(uint64_t)));
    HIP_TRY(ctx->sp_queues.ensure((size_t)kQueues * kQueueGroups * 32 * sizeof(uint32_t)));
    const uint64_t nsup_alloc = std::max<uint64_t>(nsuper_pass, 1);
    HIP_TRY(ctx->sp_super_iv.ensure(nsup_alloc * list_stride * 2 * sizeof(double)));
    HIP_TRY(ctx->sp_super_lists.ensure(nsup_alloc * list_stride * sizeof(uint16_t)));
    HIP_TRY(ctx->sp_super_lens.ensure(nsup_alloc * sizeof(uint32_t)));
    HIP_TRY(ctx->sp_super_act.ensure(nsup_alloc * sizeof(uint64_t)));

    // a closed surface crosses O(res^2) cells: start from 6 per row-plane cell
    const uint64_t est_v = std::min<uint64_t>(cells_all, std::max<uint64_t>(65536, 6 * planes_sum));
    uint64_t want_v = std::max<uint64_t>(est_v, ctx->vcap);
    uint64_t want_t = std::max<uint64_t>(3 * want_v, ctx->tcap);
    uint64_t want_s = std::min<uint64_t>(tiles_pass, std::max<uint64_t>({4096, tiles_pass / 4, ctx->slot_cap}));
    const uint64_t tile_blocks = (uint64_t)std::max(1, ctx->cus) * kTileBlocksPerCU;
    uint64_t stage_cap = ctx->stage_cap;
    if (rb) HIP_TRY(ctx->host_snap.ensure((size_t)npass * 4 * sizeof(uint64_t)));

    for (int attempt = 0; attempt < 4; ++attempt) {
        HIP_TRY(ctx->vtx.ensure(want_v * 6 * sizeof(float)));
        HIP_TRY(ctx->keys.ensure(want_v * sizeof(uint64_t)));
        HIP_TRY(ctx->signs.ensure(want_v));
        HIP_TRY(ctx->tri.ensure(want_t * 3 * sizeof(uint32_t)));
        HIP_TRY(ctx->sp_tile_of_slot.ensure(want_s * sizeof(uint32_t)));
        HIP_TRY(ctx->sp_job_of_slot.ensure(want_s * sizeof(uint32_t)));
        HIP_TRY(ctx->sp_cellbits.ensure(want_s * 16 * sizeof(uint32_t)));
        HIP_TRY(ctx->sp_slot_rows.ensure(want_s * 4 * 64 * sizeof(uint32_t)));
        // staged cells are reserved 8 per live brick of a slot, in per-block chunks: about
        // 2 x vertices + blocks x chunk (two launches); the device count is checked below
        stage_cap = std::max(stage_cap, 2 * want_v + 2 * tile_blocks * kStageChunkHost);
        HIP_TRY(ctx->sp_stage.ensure(stage_cap * sizeof(float4)));
        ctx->stage_cap = stage_cap;
        HIP_TRY(ctx->sp_slot_stage.ensure(want_s * 2 * sizeof(uint32_t)));
        HIP_TRY(ctx->sp_slot_live.ensure(want_s * sizeof(uint64_t)));
        HIP_TRY(ctx->sp_slot_hdr.ensure(want_s * sizeof(SlotHdr)));
        HIP_TRY(ctx->sp_slot_nb.ensure(want_s * 8 * sizeof(int32_t)));
        HIP_TRY(ctx->sp_slot_rec.ensure(want_s * kTileRecBytes));
        ctx->vcap = std::max(ctx->vcap, want_v);
        ctx->tcap = std::max(ctx->tcap, want_t);
        ctx->slot_cap = std::max(ctx->slot_cap, want_s);
        if (rb) {  // host arrays for this job's own estimate, not the context's largest job
            const int rrc = readback_begin(ctx, *rb, std::min(want_v, est_v), std::min(want_t, 3 * est_v));
            if (rrc) return rrc;
        }

        MeshBufs b{};
        b.vtx = ctx->vtx.as<float>();
        b.keys = ctx->keys.as<uint64_t>();
        b.signs = ctx->signs.as<uint8_t>();
        b.tri = ctx->tri.as<uint32_t>();
        b.vcap = want_v;
        b.tcap = want_t;
        b.counters = ctx->counters.as<uint64_t>();
        b.bcount = ctx->bcount.as<uint32_t>();  // the offset scan's chunk sums

        uint64_t* job_out = b.counters + kNumCounters;
        HIP_TRY(ctx->sp_block_stats.ensure(tile_blocks * 3 * sizeof(uint64_t)));
        // the last pass folded and the counters and job rows written to mapped host memory
        const uint32_t nback = kNumCounters + n * kJobOut;
        HIP_TRY(ctx->host_back.ensure(nback * sizeof(uint64_t)));
        void* back_dev = nullptr;
        HIP_TRY(hipHostGetDevicePointer(&back_dev, ctx->host_back.p, 0));

        ctx->events_used = 0;
        std::vector<hipEvent_t> ev;
        hipEvent_t ev_start = nullptr, ev_end = nullptr;
        uint64_t tiles_total = 0;
        auto enqueue = [&]() -> int {  // the job's kernels, in stream order
        HIP_TRY(launch_sparse_job_begin(ctx->stream, b.counters, kNumCounters + n * kJobOut,
                                        ctx->sp_queues.as<uint32_t>(), b.bcount, 2 * nsb_max,
                                        ctx->sp_block_stats.as<uint64_t>(), (uint32_t)tile_blocks,
                                        ctx->sp_boundary.as<uint32_t>(),
                                        n == 1 && npass > 1 ? (uint32_t)plane_max : 0u));
        if (ctx->profiling) {
            ev_start = ctx->next_event();
            HIP_TRY(hipEventRecord(ev_start, ctx->stream));
        }
        for (uint32_t pass = 0; pass < npass; ++pass) {
            SpPass P{};
            P.jobs = n == 1 ? ddesc + pass : ddesc;
            P.njobs = n == 1 ? 1 : n;
            P.list_stride = list_stride;
            P.prune_depth = prune_depth;
            if (n == 1) {
                const SpJob& J = desc[pass];
                P.ntiles = (uint64_t)J.ntx * J.nty * ((J.nz + kTile - 1) / kTile);
                P.nrows = (uint64_t)J.nz * J.ry;
                P.ntrows = (uint64_t)((J.nz + kTile - 1) / kTile) * J.nty;
                P.boundary_cells = (uint64_t)J.rx * J.ry;
            } else {
                P.ntiles = ntiles_all;
                P.nrows = nrows_pass;
                P.ntrows = ntrows_pass;
            }
            if (n == 1) {
                const SpJob& J = desc[pass];
                P.nsuper = (uint64_t)J.nsx * J.nsy * J.nsz;
            } else {
                P.nsuper = nsuper_pass;
            }
            P.lists = ctx->tile_lists.as<uint16_t>();
            P.lens = ctx->tile_lens.as<uint32_t>();
            P.flops = ctx->tile_flops.as<uint32_t>();
            P.slot_of_tile = ctx->sp_slot_of_tile.as<int32_t>();
            P.tile_of_slot = ctx->sp_tile_of_slot.as<uint32_t>();
            P.job_of_slot = ctx->sp_job_of_slot.as<uint32_t>();
            P.slot_cap = (uint32_t)want_s;
            P.box_stride = box_stride;
            P.cellbits = ctx->sp_cellbits.as<uint32_t>();
            P.stage = ctx->sp_stage.as<float4>();
            P.stage_cap = stage_cap;
            P.slot_stage = ctx->sp_slot_stage.as<uint32_t>();
            P.slot_live = ctx->sp_slot_live.as<uint64_t>();
            P.slot_hdr = ctx->sp_slot_hdr.as<SlotHdr>();
            P.slot_nb = ctx->sp_slot_nb.as<int32_t>();
            P.slot_rec = ctx->sp_slot_rec.as<uint4>();
            P.block_stats = ctx->sp_block_stats.as<uint64_t>();
            P.slot_rows = ctx->sp_slot_rows.as<uint32_t>();
            P.row_sum = ctx->sp_rows.as<uint32_t>();
            P.row_off = P.row_sum + 2 * P.nrows;
            P.counters = b.counters;
            P.queues = ctx->sp_queues.as<uint32_t>();
            P.super_iv = ctx->sp_super_iv.as<double2>();
            P.super_lists = ctx->sp_super_lists.as<uint16_t>();
            P.super_lens = ctx->sp_super_lens.as<uint32_t>();
            P.super_act = ctx->sp_super_act.as<uint64_t>();
            P.job_out = job_out;
            hipEvent_t m[kSparseMarks] = {};
            if (ctx->profiling >= 2)
                for (auto& e : m) e = ctx->next_event();
            else if (ctx->profiling)  // the evaluation launches only: fewer markers between kernels
                m[1] = ctx->next_event(), m[2] = ctx->next_event();
            if (pass > 0)
                HIP_TRY(launch_sparse_pass_begin(ctx->stream, b.counters, ctx->sp_queues.as<uint32_t>(), b.bcount,
                                                 2 * nsb_max));
            const int boundary = (n == 1 && pass + 1 < npass) ? 1 : 0;
            HIP_TRY(launch_sparse_pass(ctx->stream, P, b, ctx->cus, (uint32_t)tile_blocks, general_tiles, fold_jobs,
                                       other_jobs, boundary,
                                         ctx->profiling ? m : nullptr));
            if (ctx->profiling) ev.insert(ev.end(), m, m + kSparseMarks);
            tiles_total += P.ntiles;
            if (rb) {  // this pass's snapshot; the previous pass's copies once it ends
                int rrc = readback_mark(ctx, b.counters, pass);
                if (!rrc && pass > 0) rrc = readback_pass(ctx, *rb, want_v, want_t, want_s, stage_cap);
                if (rrc) return rrc;
            }
        }
        HIP_TRY(launch_sparse_job_end(ctx->stream, b.counters, nback, static_cast<uint64_t*>(back_dev),
                                      ctx->sp_block_stats.as<uint64_t>(), (uint32_t)tile_blocks));
        if (ctx->profiling) {
            ev_end = ctx->next_event();
            HIP_TRY(hipEventRecord(ev_end, ctx->stream));
        }
        return S2M_OK;
        };
        // A one-pass job without streamed readback or profiling events (its kernels and
        // their arguments fixed by the tables image, the buffers and the capacities) is
        // captured into a graph the second time the same job comes, and the graph replayed
        // from then on: one launch instead of a dozen.
        ctx->timing.graph_launches = 0;
        if (!rb && !ctx->profiling && npass == 1) {
            std::string key(reinterpret_cast<const char*>(img), img_bytes);
            auto put = [&key](const void* q, size_t m) { key.append(static_cast<const char*>(q), m); };
            for (const DevBuf* d : {&ctx->vtx, &ctx->keys, &ctx->signs, &ctx->tri, &ctx->counters, &ctx->bcount,
                                    &ctx->sp_queues, &ctx->sp_block_stats, &ctx->tile_tab, &ctx->tile_lists,
                                    &ctx->tile_lens, &ctx->tile_flops, &ctx->sp_slot_of_tile, &ctx->sp_tile_of_slot,
                                    &ctx->sp_job_of_slot, &ctx->sp_cellbits, &ctx->sp_stage, &ctx->sp_slot_stage,
                                    &ctx->sp_slot_live, &ctx->sp_slot_hdr, &ctx->sp_slot_nb, &ctx->sp_slot_rec,
                                    &ctx->sp_slot_rows, &ctx->sp_rows, &ctx->sp_super_iv, &ctx->sp_super_lists,
                                    &ctx->sp_super_lens, &ctx->sp_super_act, &ctx->sp_boundary})
                put(&d->p, sizeof d->p);
            const uint64_t scal[] = {want_v, want_t, want_s, stage_cap, tile_blocks, nsb_max, list_stride,
                                     prune_depth, box_stride, (uint64_t)general_tiles, (uint64_t)fold_jobs,
                                     (uint64_t)other_jobs, n, (uint64_t)(uintptr_t)back_dev};
            put(scal, sizeof scal);
            const int grc = run_graph(ctx, key, enqueue);
            if (grc) return grc;
            tiles_total = n == 1 ? (uint64_t)desc[0].ntx * desc[0].nty * ((desc[0].nz + kTile - 1) / kTile)
                                 : ntiles_all;
        } else {
            const int erc = enqueue();
            if (erc) return erc;
        }
        SparseCaps caps;
        caps.want_v = want_v;
        caps.want_t = want_t;
        caps.want_s = want_s;
        caps.stage_cap = stage_cap;
        caps.tiles_pass = tiles_pass;
        caps.tile_blocks = tile_blocks;
        if (pend && !rb && !ctx->profiling && attempt == 0) {  // queued only: read back by the caller
            pend->active = true;
            pend->n = n;
            pend->npass = npass;
            pend->tiles_total = tiles_total;
            pend->graph_launches = ctx->timing.graph_launches;
            pend->caps = caps;
            return S2M_OK;
        }
        if (rb) {  // copies of each pass's mesh while the later passes run
            const int rrc = readback_passes(ctx, *rb, npass, want_v, want_t, want_s, stage_cap);
            if (rrc) return rrc;
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (ctx->profiling) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_end));
            ctx->timing.total_ms = ms;
            for (size_t q = 0; q + kSparseMarks - 1 < ev.size(); q += kSparseMarks) {
                float t[kSparseMarks - 1] = {};
                for (int k = 0; k + 1 < kSparseMarks; ++k)
                    if (ev[q + k] && ev[q + k + 1]) HIP_TRY(hipEventElapsedTime(&t[k], ev[q + k], ev[q + k + 1]));
                ctx->timing.prune_ms += t[0];
                ctx->timing.eval_ms += t[1];     // evaluation + contouring
                ctx->timing.quad_ms += t[2];     // quad count
                ctx->timing.contour_ms += t[3];  // vertex and triangle offset scans
                ctx->timing.normal_ms += t[4];   // vertex, normal and triangle emission
            }
        }
        bool fits = false;
        const int src = sparse_read(ctx, ctx->host_back.as<uint64_t>(), n, npass, tiles_total, caps, out, &fits);
        if (src || fits) return src;
        want_v = caps.want_v;
        want_t = caps.want_t;
        want_s = caps.want_s;
        stage_cap = caps.stage_cap;
    }
    return fail(S2M_E_OOM, "output buffers kept overflowing");
}

// Lattice evaluation and vertex normals of a program: the tape interpreter's kernels,
// or the kernels of a compiled program (same bodies, device_dense.h).
hipError_t eval_lattice_prog(hipStream_t s, const s2m_program* prog, const float* cx, const float* cy,
                             const float* cz, uint32_t nx, uint32_t ny, uint32_t lz0, uint32_t nzl,
                             uint32_t pitch, float* lat)
{
    if (!prog->module && prog->n == 1 && prog->ops[0].code == S2M_OP_MANDELBULB)
        return launch_eval_mandelbulb(s, prog->d_ops, cx, cy, cz, nx, ny, lz0, nzl, pitch, lat);
    if (!prog->module)
        return launch_eval_lattice(s, prog->d_ops, prog->n, cx, cy, cz, nx, ny, lz0, nzl, pitch, lat);
    if (nx == 0 || ny == 0 || nzl == 0) return hipSuccess;
    void* args[] = {&cx, &cy, &cz, &nx, &ny, &lz0, &nzl, &pitch, &lat};
    return hipModuleLaunchKernel(prog->k_eval, (nx + 15) / 16, (ny + 3) / 4, (nzl + 3) / 4, 256, 1, 1, 0, s,
                                 args, nullptr);
}

hipError_t normals_prog(hipStream_t s, const s2m_program* prog, float eps, const MeshBufs& b)
{
    if (!prog->module)
        return launch_normals(s, prog->d_ops, prog->n, eps, b, prog->n == 1 && prog->ops[0].code == S2M_OP_MANDELBULB);
    float* vtx = b.vtx;
    const uint64_t* counters = b.counters;
    uint64_t vcap = b.vcap;
    void* args[] = {&eps, &vtx, &counters, &vcap};
    return hipModuleLaunchKernel(prog->k_normals, 1024, 1, 1, 256, 1, 1, 0, s, args, nullptr);
}

// Buffers of a dense job that must not allocate (s2m_mesh_device_noalloc): every working
// buffer is carved from `scratch` (device memory of `scratch_bytes`), the mesh goes to
// the caller's buffers.
struct FixedBufs {
    uint8_t* scratch;
    size_t scratch_bytes;
    float* vtx;
    uint64_t* keys;
    uint8_t* signs;
    uint32_t* tri;
    uint64_t vcap, tcap;
};

// Working-buffer sizes of a dense job streamed in slabs of T cell planes.
struct DenseLayout {
    uint32_t T, nx, ny, pitch, max_nzl, max_blocks;
    size_t tables, lattice, vidx, blocks, actbits, counters;
    size_t carved() const  // bytes when carved from one region, each part 256-byte aligned
    {
        size_t b = 0;
        for (size_t n : {tables, lattice, vidx, blocks, blocks, actbits, counters}) b += (n + 255) & ~size_t(255);
        return b;
    }
};

DenseLayout dense_layout(const Job& j, uint32_t T, bool lattice)
{
    DenseLayout L{};
    const uint64_t plane = (uint64_t)j.res[0] * j.res[1];
    L.T = T;
    L.nx = (uint32_t)j.ax[0].coord.size();
    L.ny = (uint32_t)j.ax[1].coord.size();
    L.pitch = (L.nx + 63u) & ~63u;
    for (uint32_t z0 = 0; z0 < j.nz_cells; z0 += T) {  // largest lattice z extent of any slab
        const uint32_t z1 = std::min(z0 + T, j.nz_cells);
        L.max_nzl = std::max(L.max_nzl, j.ax[2].hi[z1 - 1] - j.ax[2].lo[z0] + 1);
    }
    L.max_blocks = (uint32_t)((plane * T + kCellsPerBlock - 1) / kCellsPerBlock);
    L.tables = tables_bytes(j);
    L.lattice = lattice ? (size_t)L.pitch * L.ny * L.max_nzl * sizeof(float) : 0;
    L.vidx = (size_t)(T + 1) * plane * sizeof(uint32_t);
    L.blocks = (size_t)L.max_blocks * sizeof(uint32_t);
    L.actbits = (size_t)L.max_blocks * (kCellsPerBlock / 64) * sizeof(uint64_t);
    L.counters = (kNumCounters + kDenseScanParts) * sizeof(uint64_t);  // the scan's chunk sums after them
    return L;
}

// Mesh one job on the device.  corners != NULL: contour the given per-cell corner
// dump (test hook) instead of evaluating a program.  fx != NULL: allocate nothing (the
// working buffers come from fx->scratch, the slab shrinks until they fit, and a mesh
// larger than the caller's buffers fails with S2M_E_OOM and its size in last_nvtx /
// last_ntri instead of being repeated with larger buffers).
int run_job(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* p, const Job& j,
            const float* h_corners, const FixedBufs* fx = nullptr, Readback* rb = nullptr)
{
    int rc = S2M_OK;
    ctx->timing = s2m_timing{};
    ctx->events_used = 0;
    const uint64_t plane = (uint64_t)j.res[0] * j.res[1];
    const uint64_t total_cells = plane * j.nz_cells;
    ctx->last_nvtx = ctx->last_ntri = ctx->last_invalid = 0;
    if (total_cells == 0) return S2M_OK;
    if (total_cells > 0xffffffffull * 4ull)
        return fail(S2M_E_ARG, "grid of %llu cells is too large", (unsigned long long)total_cells);
    if (fx && h_corners) return fail(S2M_E_ARG, "corner dumps are not contoured without allocation");
    const bool prof = ctx->profiling && !fx;  // events are created on first use

    uint32_t slab = pick_slab(p, j);
    if (prog && !fx && p->slab == 0) {  // the autotuner's slab for this job, if it chose one
        TuneChoice tc;
        if (tune_lookup(tune_key(ctx, prog, p, j), &tc) && tc.dense_slab) {
            s2m_params q = *p;
            q.slab = tc.dense_slab;
            slab = pick_slab(&q, j);
        }
    }
    DenseLayout L = dense_layout(j, slab, !h_corners);
    if (fx) {  // the largest slab (at most the usual one) whose buffers fit the scratch
        while (L.T > 1 && L.carved() > fx->scratch_bytes) L = dense_layout(j, L.T - 1, true);
        if (L.carved() > fx->scratch_bytes)
            return fail(S2M_E_ARG, "a %u x %u grid needs %llu bytes of working memory per cell plane; "
                                   "the no-allocation scratch holds %llu",
                        j.res[0], j.res[1], (unsigned long long)L.carved(),
                        (unsigned long long)fx->scratch_bytes);
    }
    const uint32_t T = L.T, nx = L.nx, ny = L.ny, pitch = L.pitch, max_blocks = L.max_blocks;

    float* lat = nullptr;
    uint32_t *vidx = nullptr, *bcount = nullptr, *boffset = nullptr;
    uint64_t *actbits = nullptr, *counters = nullptr;
    void* tables_at = nullptr;
    if (fx) {
        uint8_t* at = fx->scratch;
        auto take = [&at](size_t n) {
            uint8_t* q = at;
            at += (n + 255) & ~size_t(255);
            return q;
        };
        tables_at = take(L.tables);
        lat = reinterpret_cast<float*>(take(L.lattice));
        vidx = reinterpret_cast<uint32_t*>(take(L.vidx));
        bcount = reinterpret_cast<uint32_t*>(take(L.blocks));
        boffset = reinterpret_cast<uint32_t*>(take(L.blocks));
        actbits = reinterpret_cast<uint64_t*>(take(L.actbits));
        counters = reinterpret_cast<uint64_t*>(take(L.counters));
    } else {
        if (!h_corners) HIP_TRY(ctx->lattice.ensure(L.lattice));
        HIP_TRY(ctx->vidx.ensure(L.vidx));
        HIP_TRY(ctx->bcount.ensure(L.blocks));
        HIP_TRY(ctx->boffset.ensure(L.blocks));
        HIP_TRY(ctx->actbits.ensure(L.actbits));
        HIP_TRY(ctx->counters.ensure(L.counters));
        lat = ctx->lattice.as<float>();
        vidx = ctx->vidx.as<uint32_t>();
        bcount = ctx->bcount.as<uint32_t>();
        boffset = ctx->boffset.as<uint32_t>();
        actbits = ctx->actbits.as<uint64_t>();
        counters = ctx->counters.as<uint64_t>();
    }
    const float* dc[3];
    const uint32_t* dlo[3];
    const uint32_t* dhi[3];
    rc = upload_tables(ctx, j, dc, dlo, dhi, tables_at);
    if (rc) return rc;
    if (h_corners) {
        HIP_TRY(ctx->corners.ensure((size_t)total_cells * 8 * sizeof(float)));
        HIP_TRY(hipMemcpyAsync(ctx->corners.p, h_corners, (size_t)total_cells * 8 * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
    }

    // Output capacity: a closed surface crosses O(res^2) cells; start from 12 per
    // row-plane cell and grow once if the counters report more.
    const uint64_t est_v = std::min<uint64_t>(total_cells, std::max<uint64_t>(65536, 12 * plane));
    uint64_t want_v = std::max<uint64_t>(est_v, ctx->vcap);
    uint64_t want_t = std::max<uint64_t>(3 * want_v, ctx->tcap);
    if (fx) {
        want_v = fx->vcap;
        want_t = fx->tcap;
    }
    const bool normals = prog && !(p->flags & S2M_NO_NORMALS);
    const uint32_t nslabs = (j.nz_cells + T - 1) / T;
    if (nslabs == 1) rb = nullptr;  // nothing to overlap: the caller copies the finished mesh
    if (rb) HIP_TRY(ctx->host_snap.ensure((size_t)nslabs * 4 * sizeof(uint64_t)));
    // Several slabs: slab k + 1's lattice is evaluated (VALU-bound) on eval_stream, into
    // the other of two sets of slab buffers, while slab k is counted (the activity bits and
    // block counts: latency-bound gathers) and gets its map, vertices, normals and
    // triangles on the context's stream; the two streams then carry about equal work per
    // slab.  Events order it: a slab's count waits for its evaluation, an evaluation waits
    // until the slab two before has finished with its buffers (its last kernel), and the
    // first evaluation waits for the tables and clears queued before it.
    const bool piped = !fx && !h_corners && nslabs > 1;
    hipStream_t es = piped ? ctx->eval_stream : ctx->stream;
    float* lat_buf[2] = {lat, lat};
    uint64_t* act_buf[2] = {actbits, actbits};
    uint32_t* cnt_buf[2] = {bcount, bcount};
    if (piped) {
        HIP_TRY(ctx->lattice2.ensure(L.lattice));
        HIP_TRY(ctx->actbits2.ensure(L.actbits));
        HIP_TRY(ctx->bcount2.ensure(L.blocks));
        lat_buf[1] = ctx->lattice2.as<float>();
        act_buf[1] = ctx->actbits2.as<uint64_t>();
        cnt_buf[1] = ctx->bcount2.as<uint32_t>();
    }

    for (int attempt = 0; attempt < (fx ? 1 : 3); ++attempt) {
        MeshBufs b{};
        if (fx) {
            b.vtx = fx->vtx;
            b.keys = fx->keys;
            b.signs = fx->signs;
            b.tri = fx->tri;
        } else {
            HIP_TRY(ctx->vtx.ensure(want_v * 6 * sizeof(float)));
            HIP_TRY(ctx->keys.ensure(want_v * sizeof(uint64_t)));
            HIP_TRY(ctx->signs.ensure(want_v));
            HIP_TRY(ctx->tri.ensure(want_t * 3 * sizeof(uint32_t)));
            ctx->vcap = want_v;
            ctx->tcap = want_t;
            b.vtx = ctx->vtx.as<float>();
            b.keys = ctx->keys.as<uint64_t>();
            b.signs = ctx->signs.as<uint8_t>();
            b.tri = ctx->tri.as<uint32_t>();
        }
        b.vcap = want_v;
        b.tcap = want_t;
        b.counters = counters;
        b.bcount = bcount;
        b.boffset = boffset;
        b.actbits = actbits;
        if (rb) {  // host arrays for this job's own estimate, not the context's largest job
            const int rrc = readback_begin(ctx, *rb, std::min(want_v, est_v), std::min(want_t, 3 * est_v));
            if (rrc) return rrc;
        }

        HIP_TRY(hipMemsetAsync(b.counters, 0, kNumCounters * sizeof(uint64_t), ctx->stream));
        HIP_TRY(hipMemsetAsync(vidx, 0xff, (size_t)plane * sizeof(uint32_t), ctx->stream));
        // the count pass ORs into zeroed words; each slab's words are cleared after its quad pass
        for (uint64_t* words : {act_buf[0], act_buf[1]})
            HIP_TRY(hipMemsetAsync(words, 0, (size_t)max_blocks * (kCellsPerBlock / 64) * sizeof(uint64_t),
                                   ctx->stream));

        ctx->events_used = 0;
        // per slab 6 marks: evaluation start and end (on the evaluation's stream), then on
        // the context's stream the ends of cells, normals and quads, and the cells' start
        std::vector<hipEvent_t> ev;
        hipEvent_t ev_start = nullptr, ev_end = nullptr;
        if (prof) {
            ev_start = ctx->next_event();
            HIP_TRY(hipEventRecord(ev_start, ctx->stream));
        }
        auto mark = [&](hipStream_t st) -> hipEvent_t {  // a new event recorded on st, or null
            hipEvent_t e = ctx->next_event();
            if (e && hipEventRecord(e, st) != hipSuccess) e = nullptr;
            return e;
        };
        hipEvent_t lat_free[2] = {nullptr, nullptr};  // per buffer set: its slab's last kernel done
        if (piped) {
            hipEvent_t ready = mark(ctx->stream);
            if (!ready) return fail(S2M_E_HIP, "event record failed");
            HIP_TRY(hipStreamWaitEvent(es, ready, 0));
        }
        uint64_t lattice_points = 0, slabs = 0;
        for (uint32_t z0 = 0; z0 < j.nz_cells; z0 += T) {
            const uint32_t z1 = std::min(z0 + T, j.nz_cells);
            const uint32_t nzc = z1 - z0;
            const uint32_t lz0 = j.ax[2].lo[z0];
            const uint32_t nzl = j.ax[2].hi[z1 - 1] - lz0 + 1;

            float* const lat_s = lat_buf[slabs & 1];
            SlabArgs a{};
            a.lat = lat_s;
            a.pitch = pitch;
            a.lat_ny = ny;
            a.lz0 = lz0;
            a.cx = dc[0];
            a.cy = dc[1];
            a.cz = dc[2];
            a.lox = dlo[0];
            a.hix = dhi[0];
            a.loy = dlo[1];
            a.hiy = dhi[1];
            a.loz = dlo[2];
            a.hiz = dhi[2];
            a.rx = j.res[0];
            a.ry = j.res[1];
            a.z0 = z0;
            a.ncells = (uint32_t)(plane * nzc);
            a.div_rx = make_fastdiv(j.res[0]);
            a.div_rxy = make_fastdiv((uint32_t)plane);
            a.vidx = vidx;
            a.flags = p->flags;
            const uint32_t nb = (uint32_t)(((uint64_t)a.ncells + kCellsPerBlock - 1) / kCellsPerBlock);

            hipEvent_t m[6] = {};
            if (prof)
                for (auto& e : m) e = ctx->next_event();
            if (piped && lat_free[slabs & 1]) HIP_TRY(hipStreamWaitEvent(es, lat_free[slabs & 1], 0));
            if (m[0]) HIP_TRY(hipEventRecord(m[0], es));
            if (!h_corners) {
                HIP_TRY(eval_lattice_prog(es, prog, dc[0], dc[1], dc[2], nx, ny, lz0, nzl, pitch, lat_s));
                lattice_points += (uint64_t)nx * ny * nzl;
            }
            if (m[1]) HIP_TRY(hipEventRecord(m[1], es));
            if (m[5]) HIP_TRY(hipEventRecord(m[5], es));
            MeshBufs bs = b;  // this slab's activity bits and block counts
            bs.actbits = act_buf[slabs & 1];
            bs.bcount = cnt_buf[slabs & 1];
            if (piped) {
                hipEvent_t evaluated = mark(es);
                if (!evaluated) return fail(S2M_E_HIP, "event record failed");
                HIP_TRY(hipStreamWaitEvent(ctx->stream, evaluated, 0));
            }
            if (!h_corners) HIP_TRY(launch_cells(ctx->stream, a, bs, 0, 0));  // the count pass
            if (h_corners) {
                const float* src = ctx->corners.as<float>() + (size_t)z0 * plane * 8;
                HIP_TRY(launch_cells_from_corners(ctx->stream, a, src, b, 0));
                HIP_TRY(launch_scan(ctx->stream, b, nb, 0));
                HIP_TRY(launch_cells_from_corners(ctx->stream, a, src, b, 1));
            } else {
                HIP_TRY(launch_scan(ctx->stream, bs, nb, 0));
                HIP_TRY(launch_cells(ctx->stream, a, bs, 1, normals ? 0 : 1));
            }
            if (m[2]) HIP_TRY(hipEventRecord(m[2], ctx->stream));
            if (normals) HIP_TRY(normals_prog(ctx->stream, prog, p->eps, bs));
            if (m[3]) HIP_TRY(hipEventRecord(m[3], ctx->stream));
            HIP_TRY(launch_quads(ctx->stream, a, bs, 0));
            HIP_TRY(launch_scan(ctx->stream, bs, nb, 1));
            HIP_TRY(launch_quads(ctx->stream, a, bs, 1));
            HIP_TRY(launch_clear_bits(ctx->stream, bs.actbits, a.ncells));  // for the count two slabs on
            if (m[4]) HIP_TRY(hipEventRecord(m[4], ctx->stream));
            if (piped) {  // the slab's lattice, bits and counts are free for the slab two later
                lat_free[slabs & 1] = mark(ctx->stream);
                if (!lat_free[slabs & 1]) return fail(S2M_E_HIP, "event record failed");
            }
            if (rb) {  // the slab's vertices and triangles are final; the copies of the slab two
                       // before (waiting for it leaves this slab and the next evaluation queued)
                int rrc = readback_mark(ctx, b.counters, (uint32_t)slabs);
                if (!rrc && slabs > 1) rrc = readback_pass(ctx, *rb, want_v, want_t, ~0ull, ~0ull);
                if (rrc) return rrc;
            }
            if (z1 < j.nz_cells) {  // last map plane becomes the previous plane
                HIP_TRY(hipMemcpyAsync(vidx, vidx + (size_t)nzc * plane, plane * sizeof(uint32_t),
                                       hipMemcpyDeviceToDevice, ctx->stream));
            }
            if (prof) ev.insert(ev.end(), m, m + 6);
            ++slabs;
        }
        if (prof) {
            ev_end = ctx->next_event();
            HIP_TRY(hipEventRecord(ev_end, ctx->stream));
        }
        if (rb) {  // copies of each slab's mesh while the later slabs run
            const int rrc = readback_passes(ctx, *rb, (uint32_t)slabs, want_v, want_t, ~0ull, ~0ull);
            if (rrc) return rrc;
        }
        uint64_t cnt[kNumCounters];
        HIP_TRY(hipMemcpyAsync(cnt, b.counters, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (prof) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_end));
            ctx->timing.total_ms = ms;
            for (size_t s = 0; s + 5 < ev.size(); s += 6) {
                float t0, t1, t2, t3;
                HIP_TRY(hipEventElapsedTime(&t0, ev[s], ev[s + 1]));
                HIP_TRY(hipEventElapsedTime(&t1, ev[s + 5], ev[s + 2]));
                HIP_TRY(hipEventElapsedTime(&t2, ev[s + 2], ev[s + 3]));
                HIP_TRY(hipEventElapsedTime(&t3, ev[s + 3], ev[s + 4]));
                ctx->timing.eval_ms += t0;
                ctx->timing.contour_ms += t1;
                ctx->timing.normal_ms += t2;
                ctx->timing.quad_ms += t3;
            }
        }
        ctx->timing.eval_launches = h_corners ? 0 : slabs;
        ctx->timing.slabs = slabs;
        ctx->timing.lattice_points = lattice_points;
        if (prog) {
            uint64_t f = 0;
            for (const s2m_op& o : prog->ops) f += op_flops(o);
            ctx->timing.eval_flops = f * lattice_points;
        }
        if (cnt[kVtxTotal] >= 0xffffffffull)
            return fail(S2M_E_ARG, "mesh has %llu vertices; indices are u32",
                        (unsigned long long)cnt[kVtxTotal]);
        if (cnt[kVtxTotal] <= want_v && cnt[kTriTotal] <= want_t) {
            ctx->last_nvtx = cnt[kVtxTotal];
            ctx->last_ntri = cnt[kTriTotal];
            ctx->last_invalid = cnt[kInvalid];
            return S2M_OK;
        }
        if (fx) {  // no second attempt: report the sizes the mesh needs
            ctx->last_nvtx = cnt[kVtxTotal];
            ctx->last_ntri = cnt[kTriTotal];
            ctx->last_invalid = cnt[kInvalid];
            return fail(S2M_E_OOM, "the mesh needs %llu vertices and %llu%s triangles; the buffers hold %llu and %llu",
                        (unsigned long long)cnt[kVtxTotal], (unsigned long long)cnt[kTriTotal],
                        cnt[kVtxTotal] > want_v ? " or more" : "", (unsigned long long)want_v,
                        (unsigned long long)want_t);
        }
        // vertices past the capacity form no quads, so after a vertex overflow the triangle
        // count is low: size triangles from the vertices as at the start
        if (cnt[kVtxTotal] > want_v) want_t = std::max<uint64_t>(want_t, 3 * (cnt[kVtxTotal] + cnt[kVtxTotal] / 8));
        want_v = std::max<uint64_t>(want_v, cnt[kVtxTotal] + cnt[kVtxTotal] / 8 + 1024);
        want_t = std::max<uint64_t>(want_t, cnt[kTriTotal] + cnt[kTriTotal] / 8 + 2048);
    }
    return fail(S2M_E_OOM, "output buffers kept overflowing");
}

int copy_out(s2m_ctx* ctx, int want_keys, s2m_mesh_out* out, double* d2h_ms)
{
    std::memset(out, 0, sizeof *out);
    const uint64_t nv = ctx->last_nvtx, nt = ctx->last_ntri;
    out->nvtx = nv;
    out->ntri = nt;
    out->invalid_quads = ctx->last_invalid;
    out->vtx = (float*)host_alloc(std::max<uint64_t>(nv, 1) * 6 * sizeof(float));
    out->tri = (uint32_t*)host_alloc(std::max<uint64_t>(nt, 1) * 3 * sizeof(uint32_t));
    if (want_keys) {
        out->cell_keys = (uint64_t*)host_alloc(std::max<uint64_t>(nv, 1) * sizeof(uint64_t));
        out->signs = (uint8_t*)host_alloc(std::max<uint64_t>(nv, 1));
    }
    if (!out->vtx || !out->tri || (want_keys && (!out->cell_keys || !out->signs))) {
        s2m_mesh_free(out);
        return fail(S2M_E_OOM, "host allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profiling) {
        e0 = ctx->next_event();
        e1 = ctx->next_event();
        HIP_TRY(hipEventRecord(e0, ctx->stream));
    }
    if (nv) HIP_TRY(hipMemcpyAsync(out->vtx, ctx->vtx.p, nv * 6 * sizeof(float),
                                   hipMemcpyDeviceToHost, ctx->stream));
    if (nt) HIP_TRY(hipMemcpyAsync(out->tri, ctx->tri.p, nt * 3 * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
    if (want_keys && nv) {
        HIP_TRY(hipMemcpyAsync(out->cell_keys, ctx->keys.p, nv * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(out->signs, ctx->signs.p, nv, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (e1) HIP_TRY(hipEventRecord(e1, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (e1) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        *d2h_ms = ms;
    }
    return S2M_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------
extern "C" {

const char* s2m_last_error(void) { return g_err.c_str(); }

uint32_t s2m_round_resolution(uint32_t res) { return round_res(res); }

void s2m_params_reference(s2m_params* p, uint32_t resolution, float bounds)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    // Bounds3D::cube(a, zero) (lib.rs:115-118): min = 0 - a*0.5, max = 0 + a*0.5
    const float half = bounds * 0.5f;
    for (int a = 0; a < 3; ++a) {
        p->bb_min[a] = 0.0f - half;
        p->bb_max[a] = 0.0f + half;
        p->res[a] = resolution;
    }
    p->eps = 0.0001f;
    p->flags = S2M_REFERENCE_COMPAT;
}

s2m_ctx* s2m_create(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        fail(S2M_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
        return nullptr;
    }
    if (device < 0 || device >= n) {
        fail(S2M_E_ARG, "device %d out of range (%d devices)", device, n);
        return nullptr;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        fail(S2M_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
        return nullptr;
    }
    s2m_ctx* ctx = new s2m_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        if (prop.multiProcessorCount > 0) ctx->cus = prop.multiProcessorCount;
        const char* colon = std::strchr(prop.gcnArchName, ':');
        ctx->arch.assign(prop.gcnArchName, colon ? (size_t)(colon - prop.gcnArchName) : std::strlen(prop.gcnArchName));
    }
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->eval_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fail(S2M_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
        if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return nullptr;
    }
    // set-up of the no-allocation entry point: the scratch array's address on this
    // device, looked up once per device while it is current (a failure is kept and
    // reported by s2m_mesh_device_noalloc)
    void* scratch_at = nullptr;
    (void)noalloc_scratch(device, &scratch_at);
    return ctx;
}

void s2m_destroy(s2m_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->copy_stream);
    (void)hipStreamSynchronize(ctx->eval_stream);
    for (DevBuf* b : {&ctx->lattice, &ctx->lattice2, &ctx->actbits2, &ctx->bcount2, &ctx->vidx, &ctx->bcount, &ctx->boffset, &ctx->actbits, &ctx->check, &ctx->counters,
                      &ctx->tables, &ctx->vtx, &ctx->keys, &ctx->signs, &ctx->tri, &ctx->corners,
                      &ctx->tile_tab, &ctx->tile_lists, &ctx->tile_lens,
                      &ctx->tile_flops, &ctx->sp_slot_of_tile, &ctx->sp_tile_of_slot,
                      &ctx->sp_job_of_slot, &ctx->sp_cellbits, &ctx->sp_slot_rows,
                      &ctx->sp_boundary, &ctx->sp_queues, &ctx->sp_super_iv,
                      &ctx->sp_super_lists, &ctx->sp_super_lens, &ctx->sp_super_act, &ctx->sp_rows,
                      &ctx->sp_stage, &ctx->sp_slot_nb, &ctx->sp_slot_stage, &ctx->sp_slot_live, &ctx->sp_slot_hdr, &ctx->sp_slot_rec, &ctx->sp_block_stats})
        b->release();
    for (DevBuf& b : ctx->batch_dense) b.release();
    for (s2m_ctx::PassSet& S : ctx->pset) {
        for (DevBuf* b : {&S.vtx, &S.keys, &S.signs, &S.tri, &S.tile_tab}) b->release();
        S.host_img.release();
        S.host_back.release();
        for (hipEvent_t e : {S.done, S.copied})
            if (e) (void)hipEventDestroy(e);
    }
    ctx->host_img.release();
    ctx->host_back.release();
    ctx->host_snap.release();
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->pass_done) (void)hipEventDestroy(e);
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    (void)hipStreamDestroy(ctx->eval_stream);
    (void)hipStreamDestroy(ctx->copy_stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int s2m_set_profiling(s2m_ctx* ctx, int on)
{
    if (!ctx) return fail(S2M_E_ARG, "ctx is NULL");
    ctx->profiling = on < 0 ? 0 : on > 2 ? 2 : on;
    return S2M_OK;
}

int s2m_last_timing(const s2m_ctx* ctx, s2m_timing* out)
{
    if (!ctx || !out) return fail(S2M_E_ARG, "NULL argument");
    *out = ctx->timing;
    return S2M_OK;
}

int s2m_tape_validate(const s2m_op* ops, uint32_t n_ops)
{
    return tape_ok(ops, n_ops) ? S2M_OK
                               : fail(S2M_E_VALIDATION, "invalid tape (opcode, operand count or "
                                                        "stack depth)");
}

uint64_t s2m_tape_flops(const s2m_op* ops, uint32_t n_ops)
{
    if (!tape_ok(ops, n_ops)) return 0;
    uint64_t f = 0;
    for (uint32_t i = 0; i < n_ops; ++i) f += op_flops(ops[i]);
    return f;
}

int s2m_program_from_tape(s2m_ctx* ctx, const s2m_op* ops, uint32_t n_ops, s2m_program** out)
{
    if (!ctx || !out) return fail(S2M_E_ARG, "NULL argument");
    *out = nullptr;
    int rc = s2m_tape_validate(ops, n_ops);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    s2m_program* p = new s2m_program();
    p->device = ctx->device;
    p->n = n_ops;
    p->ops.assign(ops, ops + n_ops);
    hipError_t e = hipMalloc(&p->d_ops, n_ops * sizeof(s2m_op));
    if (e == hipSuccess)
        e = hipMemcpy(p->d_ops, ops, n_ops * sizeof(s2m_op), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_tape_prepare(ctx->stream, p->d_ops, n_ops);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        if (p->d_ops) (void)hipFree(p->d_ops);
        delete p;
        return fail(S2M_E_HIP, "tape upload: %s", hipGetErrorString(e));
    }
    *out = p;
    return S2M_OK;
}

void s2m_program_free(s2m_program* prog)
{
    if (!prog) return;
    if (prog->d_ops && prog->owns_device_copy) {
        (void)hipSetDevice(prog->device);
        (void)hipFree(prog->d_ops);
    }
    if (prog->module) {
        (void)hipSetDevice(prog->device);
        (void)hipModuleUnload(prog->module);
    }
    delete prog;
}

int s2m_program_kind(const s2m_program* prog)
{
    if (!prog) {
        fail(S2M_E_ARG, "NULL argument");
        return -1;
    }
    return prog->module ? S2M_PROGRAM_COMPILED : S2M_PROGRAM_TAPE;
}

uint32_t s2m_program_tape(const s2m_program* prog, s2m_op* ops, uint32_t cap)
{
    if (!prog || prog->module) return 0;
    if (ops) std::copy(prog->ops.begin(), prog->ops.begin() + std::min(cap, prog->n), ops);
    return prog->n;
}

}  // extern "C"

namespace {

// Mesh one job: the tile-sparse pipeline for prunable tapes, the dense slabs otherwise.
// With rb, the mesh is also streamed into rb's host arrays pass by pass.
int mesh_one(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params, Readback* rb)
{
    if (!ctx || !prog || !params) return fail(S2M_E_ARG, "NULL argument");
    if (prog->device != ctx->device) return fail(S2M_E_ARG, "program built for another device");
    HIP_TRY(hipSetDevice(ctx->device));
    Job j;
    int rc = prepare_job(params, j);
    if (rc) return rc;
    j.tiled = !prog->module && !(params->flags & S2M_DENSE_EVAL) && tape_prunable(prog->ops);
    ctx->timing = s2m_timing{};
    ctx->last_nvtx = ctx->last_ntri = ctx->last_invalid = 0;
    const uint64_t cells = (uint64_t)j.res[0] * j.res[1] * j.nz_cells;
    if (j.tiled && cells > 0) {
        std::vector<SpRange> r;
        return run_sparse(ctx, 1, &prog, &params, &j, r, rb);
    }
    return run_job(ctx, prog, params, j, nullptr, nullptr, rb);
}

}  // namespace

extern "C" {

int s2m_mesh_device(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
                    s2m_device_mesh* out)
{
    if (!out) return fail(S2M_E_ARG, "NULL argument");
    int rc = mesh_one(ctx, prog, params, nullptr);
    if (rc) return rc;
    out->vtx = ctx->vtx.as<float>();
    out->tri = ctx->tri.as<uint32_t>();
    out->cell_keys = ctx->keys.as<uint64_t>();
    out->signs = ctx->signs.as<uint8_t>();
    out->nvtx = ctx->last_nvtx;
    out->ntri = ctx->last_ntri;
    out->invalid_quads = ctx->last_invalid;
    return S2M_OK;
}

int s2m_mesh(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params, int want_keys,
             s2m_mesh_out* out)
{
    if (!out) return fail(S2M_E_ARG, "out is NULL");
    std::memset(out, 0, sizeof *out);
    // the mesh is copied to host memory pass by pass while the job runs
    Readback rb;
    rb.want_keys = want_keys;
    int rc = mesh_one(ctx, prog, params, &rb);
    if (rc == S2M_OK) rc = readback_finish(ctx, rb, ctx->last_nvtx, ctx->last_ntri);
    if (rc) {
        (void)hipStreamSynchronize(ctx->copy_stream);
        rb.release();
        return rc;
    }
    out->vtx = rb.vtx;
    out->tri = rb.tri;
    out->cell_keys = rb.keys;
    out->signs = rb.signs;
    out->nvtx = ctx->last_nvtx;
    out->ntri = ctx->last_ntri;
    out->invalid_quads = ctx->last_invalid;
    return S2M_OK;
}

namespace {

// Copy the job ranges of a batch pass's output buffers to fresh host meshes.
int copy_ranges(const s2m_ctx::PassSet& src, const SpRange* r, uint32_t n, int want_keys, s2m_mesh_out* outs,
                hipStream_t s)
{
    // One page-locked block per array for the whole pass (one DMA copy each, from the
    // pass's output buffers in `src`, queued on `s` and unfinished on return): every job's
    // array points inside it, and the block is recycled after the last job's mesh is freed.
    // Jobs without vertices (or triangles) get arrays of their own.
    uint64_t v_lo = ~0ull, v_hi = 0, t_lo = ~0ull, t_hi = 0;
    uint32_t v_refs = 0, t_refs = 0;
    for (uint32_t i = 0; i < n; ++i) {
        std::memset(&outs[i], 0, sizeof outs[i]);
        if (r[i].nv) {
            v_lo = std::min(v_lo, r[i].v0);
            v_hi = std::max(v_hi, r[i].v0 + r[i].nv);
            ++v_refs;
        }
        if (r[i].nt) {
            t_lo = std::min(t_lo, r[i].t0);
            t_hi = std::max(t_hi, r[i].t0 + r[i].nt);
            ++t_refs;
        }
    }
    float* V = nullptr;
    uint32_t* Tr = nullptr;
    uint64_t* K = nullptr;
    uint8_t* S = nullptr;
    if (v_refs) {
        V = static_cast<float*>(host_alloc_shared((v_hi - v_lo) * 6 * sizeof(float), v_refs));
        if (want_keys) {
            K = static_cast<uint64_t*>(host_alloc_shared((v_hi - v_lo) * sizeof(uint64_t), v_refs));
            S = static_cast<uint8_t*>(host_alloc_shared(v_hi - v_lo, v_refs));
        }
    }
    if (t_refs) Tr = static_cast<uint32_t*>(host_alloc_shared((t_hi - t_lo) * 3 * sizeof(uint32_t), t_refs));
    bool ok = (!v_refs || (V && (!want_keys || (K && S)))) && (!t_refs || Tr);
    uint32_t v_given = 0, t_given = 0;  // references of the blocks handed to jobs
    for (uint32_t i = 0; i < n && ok; ++i) {
        s2m_mesh_out* o = &outs[i];
        v_given += r[i].nv ? 1u : 0u;
        t_given += r[i].nt ? 1u : 0u;
        o->nvtx = r[i].nv;
        o->ntri = r[i].nt;
        o->invalid_quads = r[i].invalid;
        o->vtx = r[i].nv ? V + 6 * (r[i].v0 - v_lo) : static_cast<float*>(host_alloc(6 * sizeof(float)));
        o->tri = r[i].nt ? Tr + 3 * (r[i].t0 - t_lo) : static_cast<uint32_t*>(host_alloc(3 * sizeof(uint32_t)));
        if (want_keys) {
            o->cell_keys = r[i].nv ? K + (r[i].v0 - v_lo) : static_cast<uint64_t*>(host_alloc(sizeof(uint64_t)));
            o->signs = r[i].nv ? S + (r[i].v0 - v_lo) : static_cast<uint8_t*>(host_alloc(1));
        }
        ok = o->vtx && o->tri && (!want_keys || (o->cell_keys && o->signs));
    }
    if (!ok) {  // release every reference the jobs hold, then the blocks' unused ones
        for (uint32_t i = 0; i < n; ++i) s2m_mesh_free(&outs[i]);
        for (; v_given < v_refs; ++v_given) {
            host_release(V);
            host_release(K);
            host_release(S);
        }
        for (; t_given < t_refs; ++t_given) host_release(Tr);
        return fail(S2M_E_OOM, "host allocation failed");
    }
    const uint64_t nv = v_refs ? v_hi - v_lo : 0, nt = t_refs ? t_hi - t_lo : 0;
    if (v_refs) {
        HIP_TRY(hipMemcpyAsync(V, src.vtx.as<float>() + 6 * v_lo, nv * 6 * sizeof(float), hipMemcpyDeviceToHost, s));
        if (want_keys) {
            HIP_TRY(hipMemcpyAsync(K, src.keys.as<uint64_t>() + v_lo, nv * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                   s));
            HIP_TRY(hipMemcpyAsync(S, src.signs.as<uint8_t>() + v_lo, nv, hipMemcpyDeviceToHost, s));
        }
    }
    if (t_refs)
        HIP_TRY(hipMemcpyAsync(Tr, src.tri.as<uint32_t>() + 3 * t_lo, nt * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               s));
    return S2M_OK;
}

// A job of a batch meshed on its own with the dense slabs rather than in the batch's tile
// pass: compiled programs (no tape for the pass), S2M_DENSE_EVAL, and programs interval
// pruning cannot cut (the Mandelbulb) on grids above 128^3 cells, where the dense slabs
// evaluate each lattice point once instead of every point of every tile's box.
bool dense_rule(bool compiled, bool prunable, uint32_t flags, const Job& j)
{
    if (compiled || (flags & S2M_DENSE_EVAL)) return true;
    const uint64_t cells = (uint64_t)j.res[0] * j.res[1] * j.nz_cells;
    return !prunable && cells > (1ull << 21);
}

bool dense_job(const s2m_program* prog, const s2m_params& p, const Job& j)
{
    return dense_rule(prog->module != nullptr, tape_prunable(prog->ops), p.flags, j);
}

int prepare_batch(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs, const s2m_params* params,
                  std::vector<Job>& jobs)
{
    if (!ctx || (n && (!progs || !params))) return fail(S2M_E_ARG, "NULL argument");
    jobs.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (!progs[i]) return fail(S2M_E_ARG, "program %u is NULL", i);
        if (progs[i]->device != ctx->device) return fail(S2M_E_ARG, "program built for another device");
        int rc = prepare_job(&params[i], jobs[i]);
        if (rc) return rc;
    }
    return S2M_OK;
}

}  // namespace

int s2m_mesh_batch(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs,
                   const s2m_params* params, int want_keys, s2m_mesh_out* outs)
{
    if (!outs && n) return fail(S2M_E_ARG, "outs is NULL");
    for (uint32_t i = 0; i < n; ++i) std::memset(&outs[i], 0, sizeof outs[i]);
    std::vector<Job> jobs;
    int rc = prepare_batch(ctx, n, progs, params, jobs);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->timing = s2m_timing{};
    // consecutive passes of at most pass_cells cells and 2M tiles (empty grids cost
    // nothing); jobs for the dense path are meshed one at a time (dense_job)
    uint32_t i = 0;
    s2m_timing sum{};
    auto add_timing = [&sum](const s2m_timing& t) {
        sum.total_ms += t.total_ms;
        sum.eval_ms += t.eval_ms;
        sum.prune_ms += t.prune_ms;
        sum.contour_ms += t.contour_ms;
        sum.normal_ms += t.normal_ms;
        sum.quad_ms += t.quad_ms;
        sum.d2h_ms += t.d2h_ms;
        sum.eval_launches += t.eval_launches;
        sum.lattice_points += t.lattice_points;
        sum.slabs += t.slabs;
        sum.tiles_total += t.tiles_total;
        sum.tiles_active += t.tiles_active;
        sum.eval_flops += t.eval_flops;
        sum.graph_launches += t.graph_launches;
    };
    uint64_t pass_cells = 64ull << 20;  // the rule; the autotuner may have chosen for this batch
    {
        uint64_t all = 0;
        for (uint32_t q = 0; q < n; ++q)
            if (!dense_job(progs[q], params[q], jobs[q]))
                all += (uint64_t)jobs[q].res[0] * jobs[q].res[1] * jobs[q].nz_cells;
        TuneChoice tc;
        if (tune_lookup(tune_batch_key(ctx, n, all), &tc) && tc.batch_cells) pass_cells = tc.batch_cells;
    }
    for (uint32_t q = 0; q < n; ++q)
        if (dense_job(progs[q], params[q], jobs[q])) {
            rc = s2m_mesh(ctx, progs[q], &params[q], want_keys, &outs[q]);
            if (rc) return rc;
            add_timing(ctx->timing);
        }
    // The passes: consecutive jobs up to pass_cells cells and 2M tiles; empty grids and
    // the dense jobs meshed above belong to no pass.
    struct PassJobs {
        std::vector<const s2m_program*> pp;
        std::vector<const s2m_params*> qq;
        std::vector<Job> jj;
        std::vector<uint32_t> idx;
    };
    std::vector<PassJobs> passes;
    while (i < n) {
        uint32_t k = i;
        uint64_t cells = 0, tiles = 0;
        PassJobs P;
        while (k < n) {
            const Job& j = jobs[k];
            const uint64_t c = (uint64_t)j.res[0] * j.res[1] * j.nz_cells;
            const uint64_t t = ((j.res[0] + 7) / 8ull) * ((j.res[1] + 7) / 8ull) * ((j.nz_cells + 7) / 8ull);
            const bool batched = c && !dense_job(progs[k], params[k], j);  // else: meshed above
            if (batched && !P.idx.empty() && (cells + c > pass_cells || tiles + t > (2ull << 20))) break;
            if (batched) {
                P.pp.push_back(progs[k]);
                P.qq.push_back(&params[k]);
                P.jj.push_back(j);
                P.idx.push_back(k);
                cells += c;
                tiles += t;
            }
            ++k;
        }
        if (!P.idx.empty()) passes.push_back(std::move(P));
        i = k;
    }
    for (uint32_t q = 0; q < n; ++q)  // empty grids (not meshed on the dense path above): empty meshes
        if (!outs[q].vtx && (uint64_t)jobs[q].res[0] * jobs[q].res[1] * jobs[q].nz_cells == 0) {
            outs[q].vtx = (float*)std::malloc(6 * sizeof(float));
            outs[q].tri = (uint32_t*)std::malloc(3 * sizeof(uint32_t));
            if (want_keys) {
                outs[q].cell_keys = (uint64_t*)std::malloc(sizeof(uint64_t));
                outs[q].signs = (uint8_t*)std::malloc(1);
            }
        }
    // Pass p runs in set p % 2.  Pass p + 1 is queued before pass p is read back, so the
    // device goes from one pass to the next while the host reads the finished pass's
    // counters and queues its copies to host memory on copy_stream, which run beside the
    // next pass; pass p + 2 waits for those copies before writing the set again.  With
    // profiling on, every pass runs to completion in turn.
    for (s2m_ctx::PassSet& S : ctx->pset)
        for (hipEvent_t* e : {&S.done, &S.copied})
            if (!*e) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    SparsePending pend[2];
    std::vector<SpRange> ranges[2];
    auto drain = [ctx](int code) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->copy_stream);
        return code;
    };
    auto enqueue_pass = [&](size_t p) -> int {
        const int k = (int)(p & 1u);
        const PassJobs& P = passes[p];
        if (p >= 2) HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->pset[k].copied, 0));
        swap_pass_set(ctx, k);
        ctx->timing = s2m_timing{};
        const int erc = run_sparse(ctx, (uint32_t)P.idx.size(), P.pp.data(), P.qq.data(), P.jj.data(), ranges[k],
                                   nullptr, &pend[k]);
        swap_pass_set(ctx, k);
        if (erc) return erc;
        if (!pend[k].active) add_timing(ctx->timing);  // ran to completion
        HIP_TRY(hipEventRecord(ctx->pset[k].done, ctx->stream));
        return S2M_OK;
    };
    auto collect_pass = [&](size_t p) -> int {
        const int k = (int)(p & 1u);
        const PassJobs& P = passes[p];
        s2m_ctx::PassSet& S = ctx->pset[k];
        if (pend[k].active) {
            HIP_TRY(hipEventSynchronize(S.done));
            bool fits = false;
            ctx->timing = s2m_timing{};
            const int crc = sparse_read(ctx, S.host_back.as<uint64_t>(), pend[k].n, pend[k].npass,
                                        pend[k].tiles_total, pend[k].caps, ranges[k], &fits);
            if (crc) return crc;
            ctx->timing.graph_launches = pend[k].graph_launches;
            add_timing(ctx->timing);
            if (!fits) {  // a capacity overflowed: the pass again, to completion, with the grown capacities
                ctx->vcap = std::max(ctx->vcap, pend[k].caps.want_v);
                ctx->tcap = std::max(ctx->tcap, pend[k].caps.want_t);
                ctx->slot_cap = std::max(ctx->slot_cap, pend[k].caps.want_s);
                ctx->stage_cap = std::max(ctx->stage_cap, pend[k].caps.stage_cap);
                swap_pass_set(ctx, k);
                ctx->timing = s2m_timing{};
                const int rrc = run_sparse(ctx, (uint32_t)P.idx.size(), P.pp.data(), P.qq.data(), P.jj.data(),
                                           ranges[k]);
                swap_pass_set(ctx, k);
                if (rrc) return rrc;
                add_timing(ctx->timing);
                HIP_TRY(hipEventRecord(S.done, ctx->stream));
            }
            pend[k].active = false;
        }
        std::vector<s2m_mesh_out> tmp(P.idx.size());
        HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, S.done, 0));
        const int crc = copy_ranges(S, ranges[k].data(), (uint32_t)P.idx.size(), want_keys, tmp.data(),
                                    ctx->copy_stream);
        for (size_t q = 0; q < P.idx.size(); ++q) outs[P.idx[q]] = tmp[q];
        if (crc) return crc;
        HIP_TRY(hipEventRecord(S.copied, ctx->copy_stream));
        return S2M_OK;
    };
    for (size_t p = 0; p < passes.size(); ++p) {
        rc = enqueue_pass(p);
        if (!rc && p > 0) rc = collect_pass(p - 1);
        if (rc) return drain(rc);
    }
    if (!passes.empty()) {
        rc = collect_pass(passes.size() - 1);
        if (rc) return drain(rc);
    }
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    ctx->timing = sum;
    return S2M_OK;
}

int s2m_mesh_batch_device(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs,
                          const s2m_params* params, s2m_device_mesh* outs)
{
    if (!outs && n) return fail(S2M_E_ARG, "outs is NULL");
    std::vector<Job> jobs;
    int rc = prepare_batch(ctx, n, progs, params, jobs);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->timing = s2m_timing{};
    std::vector<const s2m_program*> pp;
    std::vector<const s2m_params*> qq;
    std::vector<Job> jj;
    std::vector<uint32_t> idx;
    for (uint32_t i = 0; i < n; ++i) std::memset(&outs[i], 0, sizeof outs[i]);
    // jobs for the dense path (compiled programs, the Mandelbulb above 128^3, S2M_DENSE_EVAL)
    // are meshed one at a time first; each mesh moves to buffers of its own, since the
    // shared pass below writes the context's output buffers
    uint32_t ndense = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Job& j = jobs[i];
        if ((uint64_t)j.res[0] * j.res[1] * j.nz_cells == 0 || !dense_job(progs[i], params[i], j)) continue;
        rc = mesh_one(ctx, progs[i], &params[i], nullptr);
        if (rc) return rc;
        const uint64_t nv = ctx->last_nvtx, nt = ctx->last_ntri;
        if (ctx->batch_dense.size() < 4 * (size_t)(ndense + 1)) ctx->batch_dense.resize(4 * (size_t)(ndense + 1));
        DevBuf* d = &ctx->batch_dense[4 * (size_t)ndense++];
        HIP_TRY(d[0].ensure(std::max<uint64_t>(nv, 1) * 6 * sizeof(float)));
        HIP_TRY(d[1].ensure(std::max<uint64_t>(nv, 1) * sizeof(uint64_t)));
        HIP_TRY(d[2].ensure(std::max<uint64_t>(nv, 1)));
        HIP_TRY(d[3].ensure(std::max<uint64_t>(nt, 1) * 3 * sizeof(uint32_t)));
        HIP_TRY(hipMemcpyAsync(d[0].p, ctx->vtx.p, nv * 6 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(d[1].p, ctx->keys.p, nv * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(d[2].p, ctx->signs.p, nv, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(d[3].p, ctx->tri.p, nt * 3 * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
        s2m_device_mesh& o = outs[i];
        o.vtx = d[0].as<float>();
        o.cell_keys = d[1].as<uint64_t>();
        o.signs = d[2].as<uint8_t>();
        o.tri = d[3].as<uint32_t>();
        o.nvtx = nv;
        o.ntri = nt;
        o.invalid_quads = ctx->last_invalid;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const Job& j = jobs[i];
        if ((uint64_t)j.res[0] * j.res[1] * j.nz_cells == 0 || dense_job(progs[i], params[i], j)) continue;
        pp.push_back(progs[i]);
        qq.push_back(&params[i]);
        jj.push_back(jobs[i]);
        idx.push_back(i);
    }
    if (idx.empty()) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));  // the dense meshes' copies
        return S2M_OK;
    }
    std::vector<SpRange> r;
    rc = run_sparse(ctx, (uint32_t)idx.size(), pp.data(), qq.data(), jj.data(), r);
    if (rc) return rc;
    for (size_t q = 0; q < idx.size(); ++q) {
        s2m_device_mesh& o = outs[idx[q]];
        o.vtx = ctx->vtx.as<float>() + 6 * r[q].v0;
        o.tri = ctx->tri.as<uint32_t>() + 3 * r[q].t0;
        o.cell_keys = ctx->keys.as<uint64_t>() + r[q].v0;
        o.signs = ctx->signs.as<uint8_t>() + r[q].v0;
        o.nvtx = r[q].nv;
        o.ntri = r[q].nt;
        o.invalid_quads = r[q].invalid;
    }
    return S2M_OK;
}

int s2m_mesh_device_noalloc(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params,
                            const s2m_mesh_buffers* bufs, s2m_device_mesh* out)
{
    if (!ctx || !prog || !params || !bufs || !out) return fail(S2M_E_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    if (prog->device != ctx->device) return fail(S2M_E_ARG, "program built for another device");
    if (!bufs->vtx || !bufs->cell_keys || !bufs->signs || !bufs->tri)
        return fail(S2M_E_ARG, "an output buffer is NULL");
    if (ctx->device < 0 || ctx->device >= (int)kMaxDevices)
        return fail(S2M_E_ARG, "device %d has no scratch slot", ctx->device);
    HIP_TRY(hipSetDevice(ctx->device));
    void* scratch_at = nullptr;
    HIP_TRY(noalloc_scratch(ctx->device, &scratch_at));
    Job j;
    int rc = prepare_job(params, j);
    if (rc) return rc;
    const FixedBufs fx{static_cast<uint8_t*>(scratch_at), kNoAllocScratchBytes, bufs->vtx, bufs->cell_keys,
                       bufs->signs, bufs->tri, bufs->vcap, bufs->tcap};
    {
        // one call at a time per device: every call on the device uses the same array
        std::lock_guard<std::mutex> g(g_noalloc_mu[ctx->device]);
        rc = run_job(ctx, prog, params, j, nullptr, &fx);
    }
    out->vtx = bufs->vtx;
    out->tri = bufs->tri;
    out->cell_keys = bufs->cell_keys;
    out->signs = bufs->signs;
    out->nvtx = ctx->last_nvtx;
    out->ntri = ctx->last_ntri;
    out->invalid_quads = ctx->last_invalid;
    return rc;
}

void s2m_mesh_free(s2m_mesh_out* out)
{
    if (!out) return;
    host_release(out->vtx);
    host_release(out->tri);
    host_release(out->cell_keys);
    host_release(out->signs);
    std::memset(out, 0, sizeof *out);
}

}  // extern "C"

namespace {

using TuneClock = std::chrono::steady_clock;

// Median wall time (ms) of `reps` runs of `run` after one untimed run; run returns a status.
template <class F>
int time_median(uint32_t reps, F run, double* ms)
{
    int rc = run();
    std::vector<double> t;
    for (uint32_t r = 0; r < reps && rc == S2M_OK; ++r) {
        const TuneClock::time_point t0 = TuneClock::now();
        rc = run();
        t.push_back(std::chrono::duration<double, std::milli>(TuneClock::now() - t0).count());
    }
    if (rc) return rc;
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return S2M_OK;
}

// Time each candidate with `run` under the calling thread's override, pick one with
// tune_pick, store it under `key`.
template <class F>
int tune_candidates(const std::string& key, const std::vector<TuneChoice>& cand, uint32_t reps, F run,
                    s2m_tuning* out)
{
    std::vector<double> med(cand.size(), 0.0);
    int rc = S2M_OK;
    for (size_t k = 0; k < cand.size() && rc == S2M_OK; ++k) {
        tune_override(&cand[k]);
        rc = time_median(reps, run, &med[k]);
        tune_override(nullptr);
    }
    if (rc) return rc;
    const uint32_t best = tune_pick(med.data(), (uint32_t)med.size());
    tune_store(key, cand[best], med[best]);
    out->stream_passes = cand[best].stream_passes;
    out->dense_slab = cand[best].dense_slab;
    out->batch_cells = cand[best].batch_cells;
    out->ms = med[best];
    out->measured = (uint32_t)cand.size();
    return S2M_OK;
}

}  // namespace

extern "C" {

int s2m_tune(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params, uint32_t reps, s2m_tuning* out)
{
    if (!ctx || !prog || !params || !out) return fail(S2M_E_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    Job j;
    int rc = prepare_job(params, j);
    if (rc) return rc;
    s2m_params p = *params;
    p.slab = 0;  // an explicit slab would bypass the choice being timed
    const std::string key = tune_key(ctx, prog, &p, j);
    TuneChoice tc;
    double ms = 0;
    if (tune_cached(key, &tc, &ms)) {
        out->stream_passes = tc.stream_passes;
        out->dense_slab = tc.dense_slab;
        out->batch_cells = tc.batch_cells;
        out->ms = ms;
        return S2M_OK;
    }
    // Candidates in decreasing order of the device buffers they need, so that buffers grow
    // once, in the first candidate's untimed run: device memory freed when a buffer is
    // regrown is cleared by the driver on a DMA engine, and copies to host memory run at
    // half rate until that is done, which would bias the candidates measured meanwhile.
    std::vector<TuneChoice> cand;
    const bool tiled = !prog->module && !(p.flags & S2M_DENSE_EVAL) && tape_prunable(prog->ops);
    if (tiled) {  // z-passes of the streamed readback (one pass holds the whole job)
        const uint32_t layers = std::max<uint32_t>(1, (j.nz_cells + kTile - 1) / kTile);
        for (uint32_t c : {1u, 2u, 3u, 4u, 6u, 8u})
            if (c <= layers) cand.push_back(TuneChoice{c, 0, 0});
    } else {  // cell planes per dense slab
        for (uint32_t c : {128u, 64u, 32u, 16u, 8u, 4u})
            if (c <= std::max<uint32_t>(j.nz_cells, 4)) cand.push_back(TuneChoice{0, c, 0});
    }
    auto run = [&]() {
        s2m_mesh_out m;
        const int r = s2m_mesh(ctx, prog, &p, 0, &m);
        s2m_mesh_free(&m);
        return r;
    };
    return tune_candidates(key, cand, std::max<uint32_t>(reps, 3), run, out);
}

int s2m_tune_batch(s2m_ctx* ctx, uint32_t n, const s2m_program* const* progs, const s2m_params* params,
                   uint32_t reps, s2m_tuning* out)
{
    if (!out) return fail(S2M_E_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    std::vector<Job> jobs;
    int rc = prepare_batch(ctx, n, progs, params, jobs);
    if (rc) return rc;
    uint64_t all = 0;
    for (uint32_t q = 0; q < n; ++q)
        if (!dense_job(progs[q], params[q], jobs[q])) all += (uint64_t)jobs[q].res[0] * jobs[q].res[1] * jobs[q].nz_cells;
    const std::string key = tune_batch_key(ctx, n, all);
    TuneChoice tc;
    double ms = 0;
    if (tune_cached(key, &tc, &ms)) {
        out->batch_cells = tc.batch_cells;
        out->ms = ms;
        return S2M_OK;
    }
    std::vector<TuneChoice> cand;  // largest passes first (buffers grow once, as in s2m_tune)
    for (uint64_t c : {512ull << 20, 384ull << 20, 256ull << 20, 128ull << 20, 64ull << 20, 32ull << 20, 16ull << 20})
        cand.push_back(TuneChoice{0, 0, c});
    std::vector<s2m_mesh_out> outs(n);
    auto run = [&]() {
        const int r = s2m_mesh_batch(ctx, n, progs, params, 0, outs.data());
        for (s2m_mesh_out& m : outs) s2m_mesh_free(&m);
        return r;
    };
    return tune_candidates(key, cand, std::max<uint32_t>(reps, 3), run, out);
}

uint32_t s2m_axis_table(float bmin, float bmax, uint32_t res, int exact_corners, float* coord,
                        uint32_t* lo, uint32_t* hi)
{
    AxisTable t;
    build_axis(bmin, bmax, res, exact_corners != 0, t);
    if (coord) std::copy(t.coord.begin(), t.coord.end(), coord);
    if (lo) std::copy(t.lo.begin(), t.lo.end(), lo);
    if (hi) std::copy(t.hi.begin(), t.hi.end(), hi);
    return (uint32_t)t.coord.size();
}

int s2m_check_sincos(s2m_ctx* ctx, uint64_t counts[4])
{
    if (!ctx || !counts) return fail(S2M_E_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ctx->check.ensure(4 * sizeof(uint64_t)));
    HIP_TRY(hipMemsetAsync(ctx->check.p, 0, 4 * sizeof(uint64_t), ctx->stream));
    HIP_TRY(launch_sincos_check(ctx->stream, ctx->check.as<uint64_t>()));
    HIP_TRY(hipMemcpyAsync(counts, ctx->check.p, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return S2M_OK;
}

int s2m_check_bailout(s2m_ctx* ctx, uint64_t counts[2])
{
    if (!ctx || !counts) return fail(S2M_E_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ctx->check.ensure(4 * sizeof(uint64_t)));
    HIP_TRY(hipMemsetAsync(ctx->check.p, 0, 2 * sizeof(uint64_t), ctx->stream));
    HIP_TRY(launch_bailout_check(ctx->stream, ctx->check.as<uint64_t>()));
    HIP_TRY(hipMemcpyAsync(counts, ctx->check.p, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return S2M_OK;
}

int s2m_eval_lattice_range(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params, uint32_t z0,
                           uint32_t z1, float* out, uint64_t cap, uint32_t dims[3], uint32_t* lz0)
{
    if (!ctx || !prog || !params || !dims || !lz0) return fail(S2M_E_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    Job j;
    int rc = prepare_job(params, j);
    if (rc) return rc;
    dims[0] = (uint32_t)j.ax[0].coord.size();
    dims[1] = (uint32_t)j.ax[1].coord.size();
    dims[2] = 0;
    *lz0 = 0;
    if (z1 > j.res[2] || z0 > z1)
        return fail(S2M_E_ARG, "cell planes [%u, %u) outside the grid of %u planes", z0, z1, j.res[2]);
    if (z0 == z1 || dims[0] == 0 || dims[1] == 0) return S2M_OK;
    // the lattice planes holding the corners of cell planes z0 .. z1-1
    *lz0 = j.ax[2].lo[z0];
    dims[2] = j.ax[2].hi[z1 - 1] - *lz0 + 1;
    const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    if (!out || cap < n) return fail(S2M_E_ARG, "output holds %llu floats, lattice has %llu",
                                     (unsigned long long)cap, (unsigned long long)n);
    const float* dc[3];
    const uint32_t* dlo[3];
    const uint32_t* dhi[3];
    rc = upload_tables(ctx, j, dc, dlo, dhi);
    if (rc) return rc;
    HIP_TRY(ctx->lattice.ensure((size_t)n * sizeof(float)));
    HIP_TRY(eval_lattice_prog(ctx->stream, prog, dc[0], dc[1], dc[2], dims[0], dims[1], *lz0, dims[2], dims[0],
                              ctx->lattice.as<float>()));
    HIP_TRY(hipMemcpyAsync(out, ctx->lattice.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return S2M_OK;
}

int s2m_eval_lattice(s2m_ctx* ctx, const s2m_program* prog, const s2m_params* params, float* out,
                     uint64_t cap, uint32_t dims[3])
{
    if (!ctx || !prog || !params || !dims) return fail(S2M_E_ARG, "NULL argument");
    Job j;
    int rc = prepare_job(params, j);
    if (rc) return rc;
    // every cell plane: lattice planes 0 .. hi[res-1], the whole table
    uint32_t lz0 = 0;
    rc = s2m_eval_lattice_range(ctx, prog, params, 0, j.res[2], out, cap, dims, &lz0);
    if (rc == S2M_OK && j.res[2] == 0) dims[2] = (uint32_t)j.ax[2].coord.size();
    return rc;
}

int s2m_contour_cells(s2m_ctx* ctx, const float* corners, const uint32_t res[3],
                      const float* cell_lo, const float* cell_hi, uint32_t flags,
                      s2m_mesh_out* out)
{
    if (!ctx || !corners || !res || !cell_lo || !cell_hi || !out)
        return fail(S2M_E_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    HIP_TRY(hipSetDevice(ctx->device));
    for (int a = 0; a < 3; ++a)
        if (res[a] > 65535u) return fail(S2M_E_ARG, "resolution exceeds 65535");
    // The given cell bounds become coordinate tables coord = [lo..., hi...] with
    // lo[i] = i and hi[i] = res + i, so the slab loop reads them unchanged.
    s2m_params p{};
    p.flags = (flags & S2M_COMPAT_DROP_LAST_SLICE) | S2M_NO_NORMALS;
    Job j;
    j.exact = true;
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
        p.res[a] = res[a];
        j.res[a] = res[a];
        AxisTable& t = j.ax[a];
        t.coord.assign(cell_lo + off, cell_lo + off + res[a]);
        t.coord.insert(t.coord.end(), cell_hi + off, cell_hi + off + res[a]);
        t.lo.resize(res[a]);
        t.hi.resize(res[a]);
        for (uint32_t i = 0; i < res[a]; ++i) {
            t.lo[i] = i;
            t.hi[i] = res[a] + i;
        }
        off += res[a];
    }
    j.nz_cells = res[2];
    if ((p.flags & S2M_COMPAT_DROP_LAST_SLICE) && j.nz_cells > 0) j.nz_cells -= 1;
    int rc = run_job(ctx, nullptr, &p, j, corners);
    if (rc) return rc;
    double d2h = 0;
    return copy_out(ctx, 1, out, &d2h);
}

}  // extern "C"

namespace s2m {

int program_from_code_object(s2m_ctx* ctx, const void* code, size_t size, s2m_program** out)
{
    (void)size;
    HIP_TRY(hipSetDevice(ctx->device));
    s2m_program* p = new s2m_program();
    p->device = ctx->device;
    hipError_t e = hipModuleLoadData(&p->module, code);
    if (e == hipSuccess) e = hipModuleGetFunction(&p->k_eval, p->module, "s2m_c_eval_lattice");
    if (e == hipSuccess) e = hipModuleGetFunction(&p->k_normals, p->module, "s2m_c_normals");
    if (e != hipSuccess) {
        if (p->module) (void)hipModuleUnload(p->module);
        delete p;
        return fail(S2M_E_HIP, "loading the compiled program: %s", hipGetErrorString(e));
    }
    *out = p;
    return S2M_OK;
}

s2m_program* program_view(int device, const s2m_op* host_ops, uint32_t n, const s2m_op* d_ops)
{
    s2m_program* p = new s2m_program();
    p->device = device;
    p->n = n;
    p->ops.assign(host_ops, host_ops + n);
    p->d_ops = const_cast<s2m_op*>(d_ops);
    p->owns_device_copy = false;
    return p;
}

hipStream_t context_stream(s2m_ctx* ctx) { return ctx->stream; }

bool tape_valid(const s2m_op* ops, uint32_t n) { return tape_ok(ops, n); }

int job_geometry(const s2m_params* p, uint32_t res[3], uint32_t* nz_cells)
{
    Job j;
    int rc = prepare_job(p, j);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) res[a] = j.res[a];
    *nz_cells = j.nz_cells;
    return S2M_OK;
}

int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

bool dense_request(const s2m_program* prog, const s2m_op* ops, uint32_t n, const s2m_params* p)
{
    Job j;
    if (!p || prepare_job(p, j) != S2M_OK) return false;
    bool prunable = true;
    for (uint32_t i = 0; ops && i < n; ++i)
        if (ops[i].code == S2M_OP_MANDELBULB) prunable = false;
    return dense_rule(prog && prog->module, prunable, p->flags, j);
}

uint64_t dense_request_bytes(const s2m_params* p)
{
    Job j;
    if (!p || prepare_job(p, j) != S2M_OK) return 0;
    const uint64_t plane = (uint64_t)j.res[0] * j.res[1];
    const uint64_t est_v = std::min<uint64_t>(plane * j.nz_cells, std::max<uint64_t>(65536, 12 * plane));
    const DenseLayout L = dense_layout(j, pick_slab(p, j), true);
    return L.carved() + est_v * (24 + 8 + 1) + 3 * est_v * 12;
}

int program_device(const s2m_program* prog) { return prog->device; }

}  // namespace s2m
