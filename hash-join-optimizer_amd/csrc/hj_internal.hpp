// hj_internal.hpp - host-side engine state: context, HBM arena, resident tables, device results.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hjoin.h"
#include "hj_device.hpp"

namespace hj {

// Error carried through the engine; the C ABI turns it into a return code + message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HJ_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::hj::Error(e_ == hipErrorOutOfMemory ? HJ_E_OOM : HJ_E_HIP,               \
                              std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

// One contiguous HBM region with two stacks: scratch grows up from the bottom (released by
// mark/reset after each operator), results grow down from the top (live until the query ends).
// This replaces the reference's three-level slab allocator (include/slab_allocator.h:15-107):
// one hipMalloc per context, no per-join allocation, everything released at query end.
struct Arena {
    uint8_t* base = nullptr;
    size_t cap = 0;
    size_t lo = 0;  // scratch top
    size_t hi = 0;  // results bottom (offset from base)
    size_t peak = 0;

    void init(uint8_t* b, size_t c) {
        base = b;
        cap = c;
        lo = 0;
        hi = c;
        peak = 0;
    }
    static size_t align(size_t n) { return (n + 255) & ~(size_t)255; }
    void* scratch(size_t n) {
        n = align(n ? n : 1);
        if (lo + n > hi)
            throw Error(HJ_E_OOM, "HBM arena exhausted (scratch " + std::to_string(n) +
                                      " B, free " + std::to_string(hi - lo) + " B)");
        void* p = base + lo;
        lo += n;
        if (lo + (cap - hi) > peak) peak = lo + (cap - hi);
        return p;
    }
    void* result(size_t n) {
        n = align(n ? n : 1);
        if (lo + n > hi)
            throw Error(HJ_E_OOM, "HBM arena exhausted (result " + std::to_string(n) +
                                      " B, free " + std::to_string(hi - lo) + " B)");
        hi -= n;
        if (lo + (cap - hi) > peak) peak = lo + (cap - hi);
        return base + hi;
    }
    size_t mark() const { return lo; }
    void reset(size_t m) { lo = m; }
    void clear() {
        lo = 0;
        hi = cap;
    }
    // everything above scratch mark m and every result
    void clear_to(size_t m) {
        lo = m;
        hi = cap;
    }
    size_t free_bytes() const { return hi - lo; }
};

struct DevColumn {
    int32_t type = HJ_INT32;
    uint32_t num_pages = 0;
    uint64_t num_rows = 0;
    uint64_t num_chars = 0;        // VARCHAR: total string bytes
    uint32_t max_len = 0;          // VARCHAR: longest string
    uint8_t* pages = nullptr;      // device
    uint32_t* row_start = nullptr; // device, num_pages + 1
    uint16_t* rank = nullptr;      // device
    uint32_t* group_start = nullptr;
    uint32_t* chunk_page = nullptr;
    uint32_t uniform_rows = 0;
    uint32_t all_dense = 1;
    // every page but the last holds >= 64 rows (and, VARCHAR, no long string): a query can index
    // the column in 64-row groups of row space (each group in at most two pages)
    bool regular = true;
    // VARCHAR: the first page past the reference's 18-bit StrRef page field that holds a string
    // (a non-NULL row or a long string's 0xffff page), ~0 when there is none
    uint64_t strref_page_over = ~0ull;
    bool has_strings = false;  // VARCHAR: some row is non-NULL (the reference makes a StrRef)
    std::vector<uint32_t> h_row_start;
    DCol view() const {
        DCol d;
        d.pages = pages;
        d.row_start = row_start;
        d.rank = rank;
        d.group_start = group_start;
        d.chunk_page = chunk_page;
        d.gidx = nullptr;
        d.flat = nullptr;
        d.vbits = nullptr;
        d.sref = nullptr;
        d.num_pages = num_pages;
        d.type = type;
        d.uniform_rows = uniform_rows;
        d.all_dense = all_dense;
        d.num_rows = num_rows;
        d.umagic = 0;
        if (type == HJ_INT32 && all_dense && uniform_rows >= 64 && num_pages) {
            // ceil(2^64 / uniform_rows) = floor((2^64 - 1) / uniform_rows) + 1
            d.umagic = ~0ull / uniform_rows + 1ull;
        }
        return d;
    }
};

}  // namespace hj

struct hj_dtable {
    uint64_t uid = 0;  // unique per upload in the process (a key of the directory cache)
    uint64_t num_rows = 0;
    std::vector<hj::DevColumn> cols;
    std::vector<void*> allocations;
};

// A device result: flat output columns in the arena's result stack.
struct hj_dresult {
    uint64_t num_rows = 0;
    struct Col {
        int32_t type;
        int32_t* values = nullptr;    // INT32
        uint64_t* offsets = nullptr;  // VARCHAR, n+1
        char* chars = nullptr;        // VARCHAR
        uint8_t* validity = nullptr;  // bitmap, bit i of byte i/8
        uint64_t num_chars = 0;
    };
    std::vector<Col> cols;
    bool live = false;
    bool batch_view = false;  // columns owned by an hj_batch (not freed with the result)
};

namespace hj {
// Built pipeline tables kept in HBM across queries (no counterpart in the reference, which builds
// every join's table in its execute()). An entry is one build side's bucketized table, its
// counters (tuple count, the INT32_MIN flag) and Bloom words, keyed by the build scan: the
// table's upload id, the key column and the scan's filters and predicate program. Entries in
// use by a running query hold a reference; the least recently used unreferenced ones are freed
// when a new entry would exceed the byte budget. Shared by the contexts of one device (a
// server's slots); every method takes the lock.
struct DirCache {
    struct Entry {
        uint32_t* table = nullptr;  // 2 * 2^k + 8 words
        uint32_t* meta = nullptr;   // 4 words, then kBloomWords Bloom words (one allocation)
        uint32_t* bf = nullptr;
        uint32_t tuples = 0;
        uint64_t bytes = 0;
        uint64_t last_use = 0;
        int refs = 0;
    };
    std::mutex mu;
    uint64_t budget = 0, used = 0, clock = 0;
    uint64_t hits = 0, misses = 0, evictions = 0;
    std::map<std::string, Entry> map;
    ~DirCache();
    // the entry for key with a reference taken, or nullptr
    const Entry* acquire(const std::string& key);
    void release(const std::string& key);
    // publish a copy of a built table (device pointers, complete in stream order on s): allocate,
    // copy on s, wait for s, then insert; false when it does not fit the budget
    bool insert(const std::string& key, const uint32_t* table, uint32_t k, const uint32_t* meta,
                const uint32_t* bf, uint32_t tuples, hipStream_t s);
};
}  // namespace hj

struct hj_context {
    int device = 0;
    std::shared_ptr<hj::DirCache> dcache;  // built tables kept across queries (nullptr: none)
    // set by a batch around one query it runs on its own: when that query's result comes
    // straight from the fused pass and every output is INT32, column c's values and validity
    // bitmap are written to emit_into[c] (the batch's buffers, sized for the source's rows)
    std::vector<std::pair<int32_t*, uint8_t*>> emit_into;
    hipStream_t stream = nullptr;
    hj::Arena arena;
    uint8_t* arena_mem = nullptr;
    std::string last_error;
    hj_dresult* live_result = nullptr;
    // pinned host staging: [0, 4 KiB) device->host scalars, [4 KiB, 64 KiB) kernel descriptors
    uint64_t* pinned = nullptr;
    static constexpr size_t kPinnedBytes = 64 * 1024;
    static constexpr size_t kDescOffset = 4096;
    bool use_pipeline = true;  // fused probe pipeline for left-deep chains (hj_set_option)
    bool part_build = true;    // pipeline builds partitioned in LDS (hj_set_option)
    bool gpu_pages = true;     // result pages written on the device (hj_set_option)
    bool xcd_slices = true;    // one-join pipelines over large tables: per-XCD table slices
    bool match_stash = true;   // join-at-a-time: the count pass keeps each row's lookup
    bool index_tables = true;  // fused probe: blocks index their source rows through LDS page
                               // tables (false: a page search per group)
    // pinned double buffer for result pages on their way to host memory (allocated on first use)
    uint8_t* page_stage = nullptr;
    hipEvent_t page_ev[2] = {nullptr, nullptr};
    static constexpr size_t kPageStageHalf = 32u << 20;
    int num_cus = 256;
    // kernel timing (hj_profile_enable)
    bool prof_on = false;
    std::string prof_prefix;
    struct ProfRec {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> event_pool;
};

namespace hj {
// brackets the launches issued during its lifetime with events when the name is profiled. With
// embed, the scope holds one launch that records the events itself (hipExtLaunchKernelGGL with
// a and b), so that timing adds no packets of its own to the stream.
struct Prof {
    hj_context* ctx;
    const char* name;
    bool embed;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    static hipEvent_t get(hj_context* c) {
        if (!c->event_pool.empty()) {
            hipEvent_t e = c->event_pool.back();
            c->event_pool.pop_back();
            return e;
        }
        // timing only: no system-scope fence when the event is recorded (the report
        // synchronizes the stream before reading the times), so a timed launch runs as an
        // untimed one would
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        return e;
    }
    Prof(hj_context* c, const char* n, bool embed_events = false, hipStream_t on = nullptr)
        : ctx(c), name(n), embed(embed_events), st(on ? on : c->stream) {
        if (!ctx->prof_on || std::strncmp(n, ctx->prof_prefix.c_str(), ctx->prof_prefix.size()))
            return;
        a = get(ctx);
        b = get(ctx);
        if (!a || !b) {
            a = b = nullptr;
            return;
        }
        if (!embed) (void)hipEventRecord(a, st);
    }
    ~Prof() {
        if (!a || !b) return;
        if (!embed) (void)hipEventRecord(b, st);
        ctx->prof.push_back({name, a, b});
    }
};
}  // namespace hj
