// devmem.hip -- process-wide block cache behind devmem.hpp.
#include "devmem.hpp"

#include <mutex>
#include <unordered_map>
#include <vector>

namespace hfx_detail {
namespace {

constexpr size_t BLOCK_MIN = 256;
constexpr size_t BLOCK_MAX = size_t(256) << 20;   // larger blocks are not cached
constexpr size_t IDLE_MAX = size_t(4) << 30;      // idle bytes kept per cache
constexpr int MAX_DEV = 64;

size_t round_block(size_t n) {
    size_t b = BLOCK_MIN;
    while (b < n) b <<= 1;
    return b;
}

struct Cache {
    std::mutex m;
    std::unordered_map<void*, size_t> live;                // pointer -> block bytes (0: not cached)
    std::unordered_map<size_t, std::vector<void*>> idle;   // block bytes -> idle blocks
    size_t idle_bytes = 0;
};

Cache g_dev[MAX_DEV];
Cache g_host;

template <class Alloc, class Free>
hipError_t take(Cache& c, void** p, size_t bytes, Alloc alloc, Free release) {
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    const bool cached = bytes <= BLOCK_MAX;
    const size_t blk = cached ? round_block(bytes) : bytes;
    {
        std::lock_guard<std::mutex> lk(c.m);
        if (cached) {
            auto it = c.idle.find(blk);
            if (it != c.idle.end() && !it->second.empty()) {
                *p = it->second.back();
                it->second.pop_back();
                c.idle_bytes -= blk;
                c.live[*p] = blk;
                return hipSuccess;
            }
        }
    }
    hipError_t e = alloc(p, blk);
    if (e == hipErrorOutOfMemory) {
        // give the idle blocks back and retry once
        (void)hipGetLastError();
        std::vector<void*> drop;
        {
            std::lock_guard<std::mutex> lk(c.m);
            for (auto& kv : c.idle)
                for (void* q : kv.second) drop.push_back(q);
            c.idle.clear();
            c.idle_bytes = 0;
        }
        for (void* q : drop) release(q);
        e = alloc(p, blk);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> lk(c.m);
    c.live[*p] = cached ? blk : 0;
    return hipSuccess;
}

template <class Free>
void give(Cache& c, void* p, Free release) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(c.m);
        auto it = c.live.find(p);
        if (it != c.live.end()) {
            const size_t blk = it->second;
            c.live.erase(it);
            if (blk && c.idle_bytes + blk <= IDLE_MAX) {
                c.idle[blk].push_back(p);
                c.idle_bytes += blk;
                return;
            }
        }
    }
    release(p);
}

Cache& dev_cache() {
    int d = 0;
    (void)hipGetDevice(&d);
    return g_dev[d >= 0 && d < MAX_DEV ? d : 0];
}

}  // namespace

hipError_t dev_alloc(void** p, size_t bytes) {
    return take(dev_cache(), p, bytes, [](void** q, size_t n) { return hipMalloc(q, n); },
                [](void* q) { (void)hipFree(q); });
}

void dev_free(void* p) {
    give(dev_cache(), p, [](void* q) { (void)hipFree(q); });
}

hipError_t host_alloc(void** p, size_t bytes) {
    return take(g_host, p, bytes, [](void** q, size_t n) { return hipHostMalloc(q, n, hipHostMallocDefault); },
                [](void* q) { (void)hipHostFree(q); });
}

void host_free(void* p) {
    give(g_host, p, [](void* q) { (void)hipHostFree(q); });
}

}  // namespace hfx_detail
