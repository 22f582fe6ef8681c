// server.cpp - continuous-batching request loop (s2m_server_*).
//
// Worker threads, each with its own context (HIP stream and device buffers that stay
// resident between passes) on one device.  A worker that is idle takes requests from the
// head of the FIFO queue while their estimated device memory fits the remaining budget,
// meshes them in one batched pass (s2m_mesh_batch) and returns the memory to the budget.
// Requests that arrive while passes are running join the next pass a worker starts, so
// the device never waits for a full batch to form, and the passes of different workers
// overlap on the GPU.  Tapes are copied into a per-worker device arena, so the steady
// state performs no device allocation or free (a free would synchronise the device and
// stall the other workers).
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
    std::vector<s2m_op> ops;
    s2m_params params{};
    uint64_t bytes = 0;
    Clock::time_point t_submit, t_admit, t_done;
    int status = -1;  // -1 queued or running, else s2m_status
    bool taken = false;
    s2m_mesh_out out{};
    uint32_t batch_size = 0, worker = 0;
};

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct s2m_server {
    s2m_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> all;
    uint64_t next_ticket = 1;
    uint64_t in_use = 0;
    bool stop = false;
    int start_error = 0;
    uint32_t started = 0;
    std::vector<std::thread> threads;

    void worker_main(uint32_t wid);
};

void s2m_server::worker_main(uint32_t wid)
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
    s2m_op* arena = nullptr;
    size_t arena_cap = 0;
    std::vector<s2m_op> staging;
    for (;;) {
        std::vector<Request*> batch;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                if (stop && queue.empty()) return true;
                return !queue.empty() && in_use + queue.front()->bytes <= cfg.hbm_budget;
            });
            if (queue.empty()) break;  // stopping
            const Clock::time_point now = Clock::now();
            while (!queue.empty() && batch.size() < cfg.max_batch &&
                   in_use + bytes + queue.front()->bytes <= cfg.hbm_budget) {
                Request* r = queue.front();
                queue.pop_front();
                r->t_admit = now;
                bytes += r->bytes;
                batch.push_back(r);
            }
            in_use += bytes;
        }
        // pack the tapes into the worker's device arena
        size_t nops = 0;
        for (Request* r : batch) nops += r->ops.size();
        int rc = S2M_OK;
        if (nops > arena_cap) {
            if (arena) (void)hipFree(arena);  // growth only: rare after warm-up
            arena_cap = std::max(nops, arena_cap * 2);
            if (hipMalloc(&arena, arena_cap * sizeof(s2m_op)) != hipSuccess) {
                arena = nullptr;
                arena_cap = 0;
                rc = S2M_E_OOM;
            }
        }
        std::vector<s2m_program*> progs;
        std::vector<s2m_params> params;
        std::vector<s2m_mesh_out> outs(batch.size());
        if (rc == S2M_OK) {
            staging.clear();
            size_t off = 0;
            for (Request* r : batch) {
                staging.insert(staging.end(), r->ops.begin(), r->ops.end());
                progs.push_back(s2m::program_view(cfg.device, r->ops.data(), (uint32_t)r->ops.size(),
                                                  arena + off));
                params.push_back(r->params);
                off += r->ops.size();
            }
            if (hipMemcpyAsync(arena, staging.data(), nops * sizeof(s2m_op), hipMemcpyHostToDevice,
                               stream) != hipSuccess)
                rc = S2M_E_HIP;
            if (rc == S2M_OK)
                rc = s2m_mesh_batch(ctx, (uint32_t)batch.size(), progs.data(), params.data(),
                                    cfg.want_keys, outs.data());
            for (s2m_program* p : progs) s2m_program_free(p);
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
    if (arena) (void)hipFree(arena);
    s2m_destroy(ctx);
}

extern "C" {

uint64_t s2m_server_request_bytes(const s2m_op* ops, uint32_t n_ops, const s2m_params* params)
{
    (void)ops;
    uint32_t res[3], nz = 0;
    if (!params || s2m::job_geometry(params, res, &nz) != S2M_OK) return 0;
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
    for (uint32_t w = 0; w < s->cfg.workers; ++w)
        s->threads.emplace_back([s, w] { s->worker_main(w); });
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->started == s->cfg.workers; });
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
    const uint64_t bytes = s2m_server_request_bytes(ops, n_ops, params);
    if (bytes == 0) return s2m::set_error(S2M_E_ARG, "invalid parameters");
    if (bytes > srv->cfg.hbm_budget)
        return s2m::set_error(S2M_E_BUDGET, "request needs more device memory than the whole budget");
    Request* r = new Request();
    r->ops.assign(ops, ops + n_ops);
    r->params = *params;
    r->bytes = bytes;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(srv->mu);
        r->ticket = srv->next_ticket++;
        srv->all[r->ticket] = r;
        srv->queue.push_back(r);
        *ticket = r->ticket;
    }
    srv->cv_work.notify_one();
    return S2M_OK;
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
