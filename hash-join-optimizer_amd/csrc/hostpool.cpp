// hostpool.cpp - pinned host memory for result pages, recycled across results.
//
// Result pages go to host memory by DMA straight into their final blocks only if those blocks
// are pinned; pinning (and first-touching) fresh memory for every result costs more than the copy
// itself, so freed blocks are kept, by power-of-two size class, for the next results (up to
// kKeepBytes), and only blocks beyond that are returned to the system.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>

#include "hostpool.hpp"

namespace hj {
namespace {

constexpr uint64_t kMinClass = 64u << 10;
constexpr uint64_t kKeepBytes = 8ull << 30;

struct Pool {
    std::mutex mu;
    std::map<void*, uint64_t> owned;            // every pinned block -> its class size
    std::multimap<uint64_t, void*> free_blocks;  // class size -> free block
    uint64_t kept = 0;
};

Pool& pool() {
    static Pool* p = new Pool();  // never destroyed: blocks may be freed during exit
    return *p;
}

uint64_t size_class(uint64_t bytes) {
    uint64_t c = kMinClass;
    while (c < bytes) c <<= 1;
    return c;
}

}  // namespace

uint8_t* host_pages_alloc(uint64_t bytes) {
    const uint64_t c = size_class(bytes ? bytes : 1);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.free_blocks.find(c);
        if (it != P.free_blocks.end()) {
            void* b = it->second;
            P.free_blocks.erase(it);
            P.kept -= c;
            return (uint8_t*)b;
        }
    }
    void* b = nullptr;
    if (hipHostMalloc(&b, c, hipHostMallocDefault) != hipSuccess || !b) {
        (void)hipGetLastError();
        return (uint8_t*)std::malloc(bytes ? bytes : 1);
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.owned[b] = c;
    return (uint8_t*)b;
}

void host_pages_free(void* p) {
    if (!p) return;
    Pool& P = pool();
    std::unique_lock<std::mutex> lk(P.mu);
    auto it = P.owned.find(p);
    if (it == P.owned.end()) {
        lk.unlock();
        std::free(p);
        return;
    }
    const uint64_t c = it->second;
    if (P.kept + c <= kKeepBytes) {
        P.free_blocks.emplace(c, p);
        P.kept += c;
        return;
    }
    P.owned.erase(it);
    lk.unlock();
    (void)hipHostFree(p);
}

}  // namespace hj
