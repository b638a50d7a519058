// dircache.cpp - built pipeline tables kept in HBM across queries (hj_internal.hpp DirCache).
// No counterpart in the reference, whose execute() builds every join's table anew.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hj_internal.hpp"
#include "pipeline.hpp"

namespace hj {

DirCache::~DirCache() {
    for (auto& kv : map) {
        (void)hipFree(kv.second.table);
        (void)hipFree(kv.second.meta);
    }
}

const DirCache::Entry* DirCache::acquire(const std::string& key) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = map.find(key);
    if (it == map.end()) {
        ++misses;
        return nullptr;
    }
    ++hits;
    it->second.refs += 1;
    it->second.last_use = ++clock;
    return &it->second;
}

void DirCache::release(const std::string& key) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = map.find(key);
    if (it != map.end() && it->second.refs > 0) it->second.refs -= 1;
}

bool DirCache::insert(const std::string& key, const uint32_t* table, uint32_t k,
                      const uint32_t* meta, const uint32_t* bf, uint32_t tuples, hipStream_t s) {
    const uint64_t tbytes = (2ull * (1ull << k) + 8) * 4;
    const uint64_t mbytes = (4ull + kBloomWords) * 4;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (map.count(key)) return true;  // another query published it first
        if (tbytes + mbytes > budget) return false;
        // least recently used entries without a reference make room
        while (used + tbytes + mbytes > budget) {
            auto victim = map.end();
            for (auto it = map.begin(); it != map.end(); ++it)
                if (it->second.refs == 0 &&
                    (victim == map.end() || it->second.last_use < victim->second.last_use))
                    victim = it;
            if (victim == map.end()) return false;  // everything is in use
            (void)hipFree(victim->second.table);
            (void)hipFree(victim->second.meta);
            used -= victim->second.bytes;
            map.erase(victim);
            ++evictions;
        }
        used += tbytes + mbytes;  // reserved while the copy runs
    }
    Entry e;
    e.tuples = tuples;
    e.bytes = tbytes + mbytes;
    bool ok = hipMalloc(&e.table, tbytes) == hipSuccess && hipMalloc(&e.meta, mbytes) == hipSuccess;
    if (ok) {
        e.bf = e.meta + 4;
        ok = hipMemcpyAsync(e.table, table, tbytes, hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemcpyAsync(e.meta, meta, 16, hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemcpyAsync(e.bf, bf, kBloomWords * 4, hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!ok || map.count(key)) {
        if (e.table) (void)hipFree(e.table);
        if (e.meta) (void)hipFree(e.meta);
        used -= tbytes + mbytes;
        return ok;
    }
    e.last_use = ++clock;
    map.emplace(key, e);
    return true;
}

}  // namespace hj
