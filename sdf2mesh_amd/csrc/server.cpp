// server.cpp - continuous-batching request loop (s2m_server_*).
//
// Worker threads, each with its own context (HIP stream, copy stream and device buffers
// that stay resident between passes) on one device.  A request goes at submission to
// one of two lanes:
//   * the pass lane: programs the tile-sparse pipeline meshes; an idle worker takes
//     requests from the head of the lane's FIFO while their estimated device memory fits
//     the remaining budget and meshes them together in one batched pass (s2m_mesh_batch);
//   * the dense lane: compiled programs and programs interval pruning cannot cut (the
//     Mandelbulb) on grids above 128^3; a worker takes one and meshes it alone on the
//     dense path, each slab's mesh copied to host memory while the next slab computes.
// Each worker prefers its own lane and takes from the other when its own is empty, so a
// large dense job occupies one worker while the others keep forming passes of the small
// requests that arrive meanwhile.  Requests arriving while passes run join the next pass
// a worker starts; the passes and dense jobs of different workers overlap on the GPU.
// Tapes are copied into a per-worker device arena, so the steady state performs no
// device allocation or free (a free would synchronise the device and stall the other
// workers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    std::vector<s2m_op> ops;             // the tape (empty for a caller's compiled program)
    const s2m_program* prog = nullptr;   // a caller's compiled program
    s2m_params params{};
    uint64_t bytes = 0;
    bool dense = false;                  // the dense lane
    Clock::time_point t_submit, t_admit, t_done;
    int status = -1;  // -1 queued or running, else s2m_status
    s2m_mesh_out out{};
    uint32_t batch_size = 0, worker = 0;
};

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// A worker's device copy of the tapes it meshes: grown on demand, never shrunk.
struct Arena {
    s2m_op* d = nullptr;
    size_t cap = 0;
    std::vector<s2m_op> staging;
    int ensure(size_t nops)
    {
        if (nops <= cap) return S2M_OK;
        if (d) (void)hipFree(d);  // growth only: rare after warm-up
        cap = std::max(nops, cap * 2);
        if (hipMalloc(&d, cap * sizeof(s2m_op)) != hipSuccess) {
            d = nullptr;
            cap = 0;
            return S2M_E_OOM;
        }
        return S2M_OK;
    }
};

}  // namespace

struct s2m_server {
    s2m_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> lane[2];  // 0: passes of tile-sparse requests, 1: dense requests
    std::unordered_map<uint64_t, Request*> all;
    uint64_t next_ticket = 1;
    uint64_t in_use = 0;
    bool stop = false;
    int start_error = 0;
    uint32_t started = 0;
    std::vector<std::thread> threads;

    bool fits(int l) const { return !lane[l].empty() && in_use + lane[l].front()->bytes <= cfg.hbm_budget; }
    void worker_main(uint32_t wid, int own);
    int enqueue(Request* r, uint64_t* ticket);
};

void s2m_server::worker_main(uint32_t wid, int own)
{
    s2m_ctx* ctx = s2m_create(cfg.device);
    {
        std::lock_guard<std::mutex> g(mu);
        if (!ctx) start_error = S2M_E_HIP;
        ++started;
    }
    cv_done.notify_all();
    if (!ctx) return;
    hipStream_t stream = s2m::context_stream(ctx);
    Arena arena;
    for (;;) {
        std::vector<Request*> batch;
        uint64_t bytes = 0;
        int l = own;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                if (stop && lane[0].empty() && lane[1].empty()) return true;
                return fits(0) || fits(1);
            });
            if (lane[0].empty() && lane[1].empty()) break;  // stopping
            if (!fits(l)) l = 1 - l;
            const Clock::time_point now = Clock::now();
            const size_t most = l == 1 ? 1 : cfg.max_batch;  // dense requests run alone
            while (!lane[l].empty() && batch.size() < most &&
                   in_use + bytes + lane[l].front()->bytes <= cfg.hbm_budget) {
                Request* r = lane[l].front();
                lane[l].pop_front();
                r->t_admit = now;
                bytes += r->bytes;
                batch.push_back(r);
            }
            in_use += bytes;
        }
        // the tapes go to the worker's device arena; compiled programs are used as they are
        size_t nops = 0;
        for (Request* r : batch) nops += r->ops.size();
        int rc = arena.ensure(nops);
        std::vector<s2m_program*> views;
        std::vector<const s2m_program*> progs;
        std::vector<s2m_params> params;
        std::vector<s2m_mesh_out> outs(batch.size());
        if (rc == S2M_OK) {
            arena.staging.clear();
            size_t off = 0;
            for (Request* r : batch) {
                if (r->prog) {
                    progs.push_back(r->prog);
                } else {
                    arena.staging.insert(arena.staging.end(), r->ops.begin(), r->ops.end());
                    views.push_back(s2m::program_view(cfg.device, r->ops.data(), (uint32_t)r->ops.size(),
                                                      arena.d + off));
                    progs.push_back(views.back());
                    off += r->ops.size();
                }
                params.push_back(r->params);
            }
            if (nops && (hipMemcpyAsync(arena.d, arena.staging.data(), nops * sizeof(s2m_op), hipMemcpyHostToDevice,
                                        stream) != hipSuccess ||
                         s2m::launch_tape_prepare(stream, arena.d, (uint32_t)nops) != hipSuccess))
                rc = S2M_E_HIP;
            if (rc == S2M_OK) {
                rc = batch.size() == 1 ? s2m_mesh(ctx, progs[0], &params[0], cfg.want_keys, &outs[0])
                                       : s2m_mesh_batch(ctx, (uint32_t)batch.size(), progs.data(), params.data(),
                                                        cfg.want_keys, outs.data());
            }
            for (s2m_program* p : views) s2m_program_free(p);
        }
        if (rc != S2M_OK)
            for (s2m_mesh_out& o : outs) s2m_mesh_free(&o);
        {
            std::lock_guard<std::mutex> g(mu);
            const Clock::time_point now = Clock::now();
            for (size_t i = 0; i < batch.size(); ++i) {
                Request* r = batch[i];
                r->status = rc;
                r->out = outs[i];
                r->t_done = now;
                r->batch_size = (uint32_t)batch.size();
                r->worker = wid;
            }
            in_use -= bytes;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
    if (arena.d) (void)hipFree(arena.d);
    s2m_destroy(ctx);
}

int s2m_server::enqueue(Request* r, uint64_t* ticket)
{
    if (r->bytes == 0) {
        delete r;
        return s2m::set_error(S2M_E_ARG, "invalid parameters");
    }
    if (r->bytes > cfg.hbm_budget) {
        delete r;
        return s2m::set_error(S2M_E_BUDGET, "request needs more device memory than the whole budget");
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(mu);
        r->ticket = next_ticket++;
        all[r->ticket] = r;
        lane[r->dense ? 1 : 0].push_back(r);
        *ticket = r->ticket;
    }
    cv_work.notify_all();
    return S2M_OK;
}

extern "C" {

uint64_t s2m_server_request_bytes(const s2m_op* ops, uint32_t n_ops, const s2m_params* params)
{
    uint32_t res[3], nz = 0;
    if (!params || s2m::job_geometry(params, res, &nz) != S2M_OK) return 0;
    if (ops && s2m::tape_valid(ops, n_ops) && s2m::dense_request(nullptr, ops, n_ops, params))
        return s2m::dense_request_bytes(params);
    const uint64_t plane = (uint64_t)res[0] * res[1];
    const uint64_t ntx = (res[0] + 7) / 8, nty = (res[1] + 7) / 8, ntz = (nz + 7) / 8;
    const uint64_t tiles = ntx * nty * ntz;
    const uint64_t verts = 6 * plane;                      // ~ cells a surface crosses
    uint64_t b = verts * (24 + 8 + 1) + 3 * verts * 12;     // vertices and triangles
    b += tiles * (2ull * n_ops + 16);                       // op lists, per-tile data
    b += 16ull * nz * res[1] * ntx;                         // row-segment counts and offsets
    b += (tiles / 4 + 1) * (13ull * 13 * 13 * 4 + 2048 + 8); // active-tile boxes
    return b;
}

s2m_server* s2m_server_create(const s2m_server_config* cfg)
{
    s2m_server* s = new s2m_server();
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget == 0) s->cfg.hbm_budget = 16ull << 30;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 64;
    if (s->cfg.workers == 0) s->cfg.workers = 2;
    if (s->cfg.dense_workers == 0) s->cfg.dense_workers = 1;
    const uint32_t total = s->cfg.workers + s->cfg.dense_workers;
    for (uint32_t w = 0; w < total; ++w)
        s->threads.emplace_back([s, w] { s->worker_main(w, w < s->cfg.workers ? 0 : 1); });
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->started == total; });
    if (s->start_error) {
        lk.unlock();
        s2m_server_destroy(s);
        s2m::set_error(S2M_E_HIP, "server: a worker could not create its device context");
        return nullptr;
    }
    return s;
}

int s2m_server_submit(s2m_server* srv, const s2m_op* ops, uint32_t n_ops, const s2m_params* params,
                      uint64_t* ticket)
{
    if (!srv || !params || !ticket) return s2m::set_error(S2M_E_ARG, "NULL argument");
    if (!s2m::tape_valid(ops, n_ops)) return s2m::set_error(S2M_E_VALIDATION, "invalid tape");
    Request* r = new Request();
    r->ops.assign(ops, ops + n_ops);
    r->params = *params;
    r->bytes = s2m_server_request_bytes(ops, n_ops, params);
    r->dense = s2m::dense_request(nullptr, ops, n_ops, params);
    return srv->enqueue(r, ticket);
}

int s2m_server_submit_program(s2m_server* srv, const s2m_program* prog, const s2m_params* params,
                              uint64_t* ticket)
{
    if (!srv || !prog || !params || !ticket) return s2m::set_error(S2M_E_ARG, "NULL argument");
    const uint32_t n = s2m_program_tape(prog, nullptr, 0);
    if (n) {  // a tape: copied, as s2m_server_submit does
        std::vector<s2m_op> ops(n);
        s2m_program_tape(prog, ops.data(), n);
        return s2m_server_submit(srv, ops.data(), n, params, ticket);
    }
    if (s2m::program_device(prog) != srv->cfg.device)
        return s2m::set_error(S2M_E_ARG, "program built for another device");
    Request* r = new Request();
    r->prog = prog;
    r->params = *params;
    r->bytes = s2m::dense_request_bytes(params);
    r->dense = true;
    return srv->enqueue(r, ticket);
}

int s2m_server_take(s2m_server* srv, uint64_t ticket, int wait, s2m_mesh_out* out,
                    s2m_request_stats* stats)
{
    if (!srv || !out) return s2m::set_error(S2M_E_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->all.find(ticket);
    if (it == srv->all.end()) return s2m::set_error(S2M_E_ARG, "unknown or already taken ticket");
    Request* r = it->second;
    if (r->status < 0) {
        if (!wait) return S2M_E_BUSY;
        srv->cv_done.wait(lk, [&] { return r->status >= 0; });
    }
    srv->all.erase(it);
    lk.unlock();
    *out = r->out;
    if (stats) {
        stats->queue_ms = ms_between(r->t_submit, r->t_admit);
        stats->latency_ms = ms_between(r->t_submit, r->t_done);
        stats->batch_size = r->batch_size;
        stats->worker = r->worker;
    }
    const int rc = r->status;
    delete r;
    return rc;
}

void s2m_server_destroy(s2m_server* srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (std::thread& t : srv->threads) t.join();
    for (auto& kv : srv->all) {
        s2m_mesh_free(&kv.second->out);
        delete kv.second;
    }
    delete srv;
}

}  // extern "C"
