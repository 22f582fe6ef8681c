// hostmem.cpp - page-locked host memory for mesh results, recycled across calls.
//
// s2m_mesh and the batch/serve entry points copy vertices and triangles from HBM into
// memory the library hands to the caller (freed with s2m_mesh_free).  Copies into
// page-locked memory run as direct DMA transfers at PCIe rate; copies into pageable
// memory stage through a driver buffer.  Page-locking is slow (milliseconds per 100 MB),
// so released blocks are kept and handed out again: in steady state a meshing call
// allocates nothing.  A block may be shared by the results of several jobs (a batch's
// vertices arrive with one copy): every job's array points inside it, and the block
// returns to the pool when the last of them is released.  Unshared blocks below kPinMin
// come from malloc (the copy is latency-bound at that size).  The pool keeps at most
// kKeepBytes of idle pinned memory; a block released above that limit is returned to
// the system.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "internal.h"

namespace {

constexpr size_t kPinMin = 256u << 10;
constexpr size_t kKeepBytes = 2ull << 30;

struct Block {
    size_t cap;
    uint32_t refs;  // results still pointing into the block
    bool pinned;    // false: malloc fallback after a failed hipHostMalloc
};

struct Pool {
    std::mutex mu;
    std::map<uintptr_t, Block> live;     // handed out, by start address
    std::multimap<size_t, void*> idle;   // released pinned blocks, by capacity
    size_t idle_bytes = 0;
    uint64_t pinned_allocs = 0;          // hipHostMalloc calls made
    uint64_t pinned_bytes = 0;           // page-locked bytes held (live + idle)
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

void* host_alloc_shared(size_t bytes, uint32_t refs)
{
    if (bytes == 0) bytes = 1;
    if (refs == 0) refs = 1;
    if (refs == 1 && bytes < kPinMin) return std::malloc(bytes);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.idle.lower_bound(bytes);
        if (it != P.idle.end() && it->first <= 2 * bytes + kPinMin) {
            void* p = it->second;
            P.idle_bytes -= it->first;
            P.live.emplace(reinterpret_cast<uintptr_t>(p), Block{it->first, refs, true});
            P.idle.erase(it);
            return p;
        }
    }
    const size_t cap = block_size(bytes);
    void* p = nullptr;
    bool pinned = true;
    if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();  // leave no sticky error for the next call
        p = std::malloc(cap);
        pinned = false;
        if (!p) return nullptr;
    }
    std::lock_guard<std::mutex> g(P.mu);
    P.live.emplace(reinterpret_cast<uintptr_t>(p), Block{cap, refs, pinned});
    if (pinned) {
        P.pinned_allocs += 1;
        P.pinned_bytes += cap;
    }
    return p;
}

void* host_alloc(size_t bytes) { return host_alloc_shared(bytes, 1); }

void host_release(void* p)
{
    if (!p) return;
    Pool& P = pool();
    std::vector<void*> drop;
    void* unpinned = nullptr;
    {
        std::lock_guard<std::mutex> g(P.mu);
        const uintptr_t a = reinterpret_cast<uintptr_t>(p);
        auto it = P.live.upper_bound(a);  // the block starting at or below p, if p lies in it
        if (it == P.live.begin() || a >= std::prev(it)->first + std::prev(it)->second.cap) {
            std::free(p);  // not from the pool
            return;
        }
        --it;
        if (--it->second.refs > 0) return;
        void* start = reinterpret_cast<void*>(it->first);
        if (!it->second.pinned) {
            unpinned = start;
        } else {
            P.idle.emplace(it->second.cap, start);
            P.idle_bytes += it->second.cap;
            while (P.idle_bytes > kKeepBytes && !P.idle.empty()) {
                auto big = std::prev(P.idle.end());
                P.idle_bytes -= big->first;
                P.pinned_bytes -= big->first;
                drop.push_back(big->second);
                P.idle.erase(big);
            }
        }
        P.live.erase(it);
    }
    std::free(unpinned);
    for (void* q : drop) (void)hipHostFree(q);
}

}  // namespace s2m

extern "C" void s2m_host_pool_stats(uint64_t stats[4])
{
    if (!stats) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> g(P.mu);
    stats[0] = P.pinned_allocs;
    stats[1] = P.pinned_bytes;
    stats[2] = P.idle.size();
    stats[3] = P.idle_bytes;
}

extern "C" void s2m_host_pool_trim(void)
{
    Pool& P = pool();
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> g(P.mu);
        for (auto& kv : P.idle) {
            P.pinned_bytes -= kv.first;
            drop.push_back(kv.second);
        }
        P.idle.clear();
        P.idle_bytes = 0;
    }
    for (void* q : drop) (void)hipHostFree(q);
}
