// hostmem.cpp - page-locked host memory for mesh results, recycled across calls.
//
// s2m_mesh and the batch/serve entry points copy vertices and triangles from HBM into
// memory the library hands to the caller (freed with s2m_mesh_free).  Copies into
// page-locked memory run as direct DMA transfers at PCIe rate; copies into pageable
// memory stage through a driver buffer.  Page-locking is slow (milliseconds per 100 MB),
// so released blocks are kept and handed out again: in steady state a meshing call
// allocates nothing.  Blocks below kPinMin come from malloc (the copy is latency-bound
// at that size).  The pool keeps at most kKeepBytes of idle pinned memory; a block
// released above that limit is returned to the system.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace {

constexpr size_t kPinMin = 256u << 10;
constexpr size_t kKeepBytes = 2ull << 30;

struct Pool {
    std::mutex mu;
    std::unordered_map<void*, size_t> live;  // handed out: capacity
    std::multimap<size_t, void*> idle;       // released, by capacity
    size_t idle_bytes = 0;
};

// Never destroyed: results may be freed by static destructors of the caller after this
// translation unit's statics are gone.
Pool& pool()
{
    static Pool* p = new Pool();
    return *p;
}

// Capacity of a new block: `bytes` rounded up to a multiple of the largest power of two
// <= bytes / 16 (at most 1/16 more), so results of similar size reuse each other's blocks.
size_t block_size(size_t bytes)
{
    size_t step = 1;
    while ((step << 5) <= bytes) step <<= 1;
    return (bytes + step - 1) / step * step;
}

}  // namespace

namespace s2m {

void* host_alloc(size_t bytes)
{
    if (bytes < kPinMin) return std::malloc(bytes ? bytes : 1);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.idle.lower_bound(bytes);
        if (it != P.idle.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            P.idle_bytes -= it->first;
            P.live.emplace(p, it->first);
            P.idle.erase(it);
            return p;
        }
    }
    const size_t cap = block_size(bytes);
    void* p = nullptr;
    if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();  // leave no sticky error for the next call
        return std::malloc(bytes);
    }
    std::lock_guard<std::mutex> g(P.mu);
    P.live.emplace(p, cap);
    return p;
}

void host_release(void* p)
{
    if (!p) return;
    Pool& P = pool();
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.live.find(p);
        if (it == P.live.end()) {
            std::free(p);
            return;
        }
        P.idle.emplace(it->second, p);
        P.idle_bytes += it->second;
        P.live.erase(it);
        while (P.idle_bytes > kKeepBytes && !P.idle.empty()) {
            auto big = std::prev(P.idle.end());
            P.idle_bytes -= big->first;
            drop.push_back(big->second);
            P.idle.erase(big);
        }
    }
    for (void* q : drop) (void)hipHostFree(q);
}

}  // namespace s2m
