// wavecol.hpp - wave-level device helpers shared by the pipeline and the batched kernels
// (gfx950, wave64): bucketized open-addressing tables (lookup and compare-and-swap insert),
// 64-row-group column readers over the resident pages (buffer loads through the query's group
// index), and wave-mask utilities.
#pragma once
#include "pipeline.hpp"

namespace hj {
namespace {
// ---------------------------------------------------------------- tables
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // one 16-byte bucket

// home bucket of a key in a table of 2^k key slots (k >= kMinTableBits): 2^(k-2) buckets
__device__ inline uint32_t home_bucket(uint32_t key, uint32_t k) { return fib32(key) >> (34u - k); }

__device__ inline u32x4 load_bucket(const HJ_GLOBAL uint32_t* K, uint32_t b) {
    return *(const HJ_GLOBAL u32x4*)(K + 8u * b);
}

// slot of `key` (!= kEmptyKey), kDead when absent, given the keys v of its home bucket b;
// a key leaves its home bucket only when that bucket is full, so a bucket with a free slot ends
// the search, and at load factor <= 1/2 the home bucket nearly always does
__device__ inline uint32_t bucket_resolve(const HJ_GLOBAL uint32_t* K, uint32_t smask, uint32_t b,
                                          uint32_t key, u32x4 v) {
    for (uint32_t m = 0; m <= smask; ++m) {  // at most the whole segment
        if (v.x == key) return 4u * b;
        if (v.y == key) return 4u * b + 1u;
        if (v.z == key) return 4u * b + 2u;
        if (v.w == key) return 4u * b + 3u;
        if (v.x == kEmptyKey || v.y == kEmptyKey || v.z == kEmptyKey || v.w == kEmptyKey)
            return kDead;
        b = next_bucket(b, smask);
        v = load_bucket(K, b);
    }
    return kDead;
}

// ---------------------------------------------------------------- wave masks, buffer loads
// Liveness is kept as 64-bit wave masks (one per 64-row group, in scalar registers): validity
// words are lane masks already, a predicate becomes one with a ballot, and the lanes of a mask
// run a step with exactly that execution mask (inverse ballot). Columns are read with buffer
// loads: one per-lane offset serves every column and group of an iteration (the group is the
// instruction's immediate offset), and rows past the end of a column read 0 instead of faulting.
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, 32-bit data format (gfx9 family)
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRsrcWord3);
}

// lanes set in mask m below this lane
__device__ inline uint32_t mask_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ---------------------------------------------------------------- columns read by waves
// An INT32 column read in 64-row groups straight from its pages through the query's group
// index: a wave loads the group records it needs (lane u: the 16-byte record of group u), then,
// per group, each lane finds its row's value at A + 4 * (valid rows of the group before it), plus
// D for the rows in the next page (A, D and split from the record, wave-uniform), and loads it.
// The record also holds the group's validity bits. Branch-free: every column the pipeline
// streams has a group index.
struct IdxCol {
    __amdgpu_buffer_rsrc_t val;  // the pages
    __amdgpu_buffer_rsrc_t rec;  // the group index (empty for a dense column of uniform pages)
    uint64_t umagic;             // dense, uniform pages: ceil(2^64 / upr), else 0
    uint32_t upr;
};

__device__ inline IdxCol idx_col(const DCol& c, uint32_t rows) {
    return IdxCol{make_rsrc(c.pages, c.num_pages * HJ_PAGE_BYTES),
                  make_rsrc(c.gidx, c.gidx ? ((rows + 63u) >> 6) * 16u : 0u),
                  c.gidx ? 0ull : c.umagic, c.uniform_rows};
}

// the record of group grp (on the lane that loads it): validity bits in .x/.y, location in .z/.w
__device__ inline u32x4 idx_rec(const IdxCol& w, uint32_t grp) {
    return __builtin_amdgcn_raw_buffer_load_b128(w.rec, grp * 16u, 0, 0);
}

// idx_rec for a column that may be dense with pages of one row count (the build side): its
// record by arithmetic (every row valid; rows past the end are cut by the caller's range). The
// load is issued either way (it reads nothing from an empty resource) and the choice is a
// select, so no branch holds the load's wait.
__device__ inline u32x4 idx_rec_any(const IdxCol& w, uint32_t grp) {
    const u32x4 ld = idx_rec(w, grp);
    const uint32_t r = grp * 64u;
    const uint32_t p = dense_page(w.umagic, r);
    const uint32_t local0 = r - p * w.upr;
    const uint32_t split = w.upr - local0 < 64u ? w.upr - local0 : 64u;
    const uint64_t d = gint_pack(p * HJ_PAGE_BYTES + 4u + 4u * local0,
                                 HJ_PAGE_BYTES - 4u * (local0 + split), split);
    const bool dense = w.umagic != 0;
    return u32x4{dense ? ~0u : ld.x, dense ? ~0u : ld.y, dense ? (uint32_t)d : ld.z,
                 dense ? (uint32_t)(d >> 32) : ld.w};
}

// group u's validity bits and location (wave-uniform) from the records held by lanes 0 .. U-1
__device__ inline uint64_t rec_valid(const u32x4& r, int u) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)r.y, u) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)r.x, u);
}

// this lane's value of a group whose record lanes 0 .. U-1 hold (group u); its validity bits vw
__device__ inline int32_t idx_load(const IdxCol& w, const u32x4& r, int u, uint64_t vw,
                                   uint32_t lane) {
    const uint32_t A = (uint32_t)__builtin_amdgcn_readlane((int)r.z, u);
    const uint32_t E = (uint32_t)__builtin_amdgcn_readlane((int)r.w, u);
    // the whole offset in the per-lane operand, which the buffer's range check covers: a row
    // past the last page reads 0
    const uint32_t off = A + 4u * mask_rank(vw) + (lane >= (E >> 16) ? (E & 0xffffu) : 0u);
    return __builtin_amdgcn_raw_buffer_load_b32(w.val, off, 0, 0);
}

// A column of the query's row-space arrays (a predicate's results, a decoded column): the value
// of row r at r * 4, validity words beside it (none: every row valid)
struct FlatCol {
    __amdgpu_buffer_rsrc_t val;
    __amdgpu_buffer_rsrc_t vb;
    uint32_t all_valid;
};
__device__ inline FlatCol flat_col(const DCol& c, uint32_t rows) {
    return FlatCol{make_rsrc(c.flat, rows * 4u),
                   make_rsrc(c.vbits ? (const void*)c.vbits : (const void*)c.flat,
                             c.vbits ? ((rows + 63u) >> 6) * 8u : 0u),
                   c.vbits ? 0u : ~0u};
}

// the lanes of m run what the condition guards (m is wave-uniform)
__device__ inline bool in_mask(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// 64-bit word of lane u (wave-uniform)
__device__ inline uint64_t lane_word(uint32_t lo, uint32_t hi, int u) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, u) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)lo, u);
}

// lanes [0, rem) of a group (rem clamped to 64)
__device__ inline uint64_t range_mask(uint32_t rem) { return rem >= 64u ? ~0ull : (1ull << rem) - 1ull; }

// block-local append of the lanes of U (uniform) masks through an LDS counter: pos[u] is the
// lane's position among this block's appends for mask u (valid on the lanes of m[u])
template <int U>
__device__ inline void block_append_masks(uint32_t* lds_cnt, const uint64_t (&m)[U],
                                          uint32_t (&pos)[U]) {
    uint32_t tot = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) tot += (uint32_t)__popcll(m[u]);
    uint32_t base = 0;
    if (lane_id() == 0 && tot) base = atomicAdd(lds_cnt, tot);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        pos[u] = base + mask_rank(m[u]);
        base += (uint32_t)__popcll(m[u]);
    }
}

// one compare-and-swap of an insert at key slot sl: true once the key is placed (or found
// there: a duplicate build key, flagged in meta[1])
__device__ inline bool insert_at(HJ_GLOBAL uint32_t* K, uint32_t sl, uint32_t key, uint32_t row,
                                 HJ_GLOBAL uint32_t* meta) {
    const uint32_t old = atomicCAS((uint32_t*)&K[8u * (sl >> 2) + (sl & 3u)], kEmptyKey, key);
    if (old == kEmptyKey) {
        K[row_word(sl)] = row;
        return true;
    }
    if (old == key) {
        meta[1] = 1;
        return true;
    }
    return false;
}

// the rest of an insert whose first compare-and-swap (slot s of its home bucket) found another
// key: the other slots of the home bucket, then the following buckets of its segment whole (a
// segment that has no free slot left flags the build, as a repeated key does)
__device__ inline void insert_rest(HJ_GLOBAL uint32_t* K, uint32_t smask, uint32_t s,
                                   uint32_t key, uint32_t row, HJ_GLOBAL uint32_t* meta) {
    const uint32_t b0 = s >> 2;
    for (uint32_t i = 1; i < kBucketKeys; ++i)
        if (insert_at(K, 4u * b0 + ((s + i) & 3u), key, row, meta)) return;
    for (uint32_t b = next_bucket(b0, smask), m = 0; m < smask; b = next_bucket(b, smask), ++m)
        for (uint32_t j = 0; j < kBucketKeys; ++j)
            if (insert_at(K, 4u * b + j, key, row, meta)) return;
    meta[1] = 1;  // the segment is full
}

// insert: block x of a job inserts the tuples staged by block x of the filter pass. A key's

// key slot of kk from its home bucket's keys v: the slot, kDead (absent), or kDead - 1 (the
// home bucket is full: continue in the next buckets). The key kEmptyKey lives in slot 2^k when
// the build holds it (sp).
__device__ inline uint32_t resolve_home(u32x4 v, uint32_t kk, uint32_t hb, uint32_t k,
                                        uint32_t sp) {
    const uint32_t b4 = 4u * hb;
    uint32_t sl = v.w == kk ? b4 + 3u : kDead;
    sl = v.z == kk ? b4 + 2u : sl;
    sl = v.y == kk ? b4 + 1u : sl;
    sl = v.x == kk ? b4 : sl;
    const bool full = v.x != kEmptyKey && v.y != kEmptyKey && v.z != kEmptyKey && v.w != kEmptyKey;
    sl = (sl == kDead && full) ? kDead - 1u : sl;
    return kk == kEmptyKey ? (sp ? (1u << k) : kDead) : sl;
}

// slot of kk in table K (2^k key slots), kDead when absent; one 16-byte load settles nearly
// every lookup (the lanes whose home bucket is full continue through the following buckets)
__device__ inline uint32_t table_lookup(const HJ_GLOBAL uint32_t* K, uint32_t k, uint32_t sp,
                                        uint32_t kk) {
    const uint32_t hb = home_bucket(kk, k);
    uint32_t sl = resolve_home(load_bucket(K, hb), kk, hb, k, sp);
    if (sl == kDead - 1u) {
        const uint32_t smask = seg_mask(k);
        const uint32_t b = next_bucket(hb, smask);
        sl = bucket_resolve(K, smask, b, kk, load_bucket(K, b));
    }
    return sl;
}

// string bytes of one survivor (source row, build row per join) in each output VARCHAR column,

}  // namespace
}  // namespace hj
