// kernels.hpp - host launchers of the join kernels (all asynchronous on the given stream).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hj_device.hpp"

namespace hj {

constexpr int kMaxFilters = 4;
constexpr int kMaxSlots = 16;
constexpr uint32_t kLightMax = 64;       // a probe row with more matches goes to the heavy list
constexpr uint32_t kHeavyChunk = 65536;  // heavy rows are split into chunks of this many outputs

// A base-table scan: rows are the key column's rows; a row passes when every filter column is
// non-NULL and inside [lo, hi].
struct ScanSrc {
    DCol key;
    DCol fcol[kMaxFilters];
    int32_t flo[kMaxFilters];
    int32_t fhi[kMaxFilters];
    int32_t nf;
};

// An intermediate relation: row i maps to base row rid[i] of the key column's table.
struct RelSrc {
    DCol key;
    const uint32_t* rid;
    uint64_t n;
};

// Unchained directory over build tuples sorted by h = fib32(key).
//   dir[s] = (first tuple of slot s) << 16 | 16-bit Bloom filter, dir[D] = n << 16
//   sh[i]  = h of tuple i (ascending), srow[i] = its build-relation row
struct DirView {
    const uint64_t* dir;
    const uint32_t* sh;
    const uint32_t* srow;
    uint32_t k;
    uint32_t pad;
    uint64_t n;
};

// Where each output column of a join comes from: side 0 = probe relation, 1 = build relation;
// src = that relation's rid array of the slot (nullptr: the relation row itself).
struct OutSlots {
    uint32_t* dst[kMaxSlots];
    const uint32_t* src[kMaxSlots];
    int32_t side[kMaxSlots];
    int32_t n;
    // value mode (a root join whose outputs are INT32 columns): entry c is output column c,
    // dst[c] receives the value of row `row` of col[c] (0 for NULL) and valid8[c] its validity
    // byte, row being the build or probe row (side[c]) mapped through src[c]
    int32_t values;
    uint8_t* valid8[kMaxSlots];
    DCol col[kMaxSlots];
    // value mode, build-side columns: the values (and validity bytes, where valid8 is set) in
    // the directory's row order (bval[c][p] = column c at the build row of directory position
    // p), so the matches of a probe row read them consecutively; null: gathered per match
    const int32_t* bval[kMaxSlots];
    const uint8_t* bvalid[kMaxSlots];
};
// bval / bvalid of output column c of a value-mode join: one pass over the directory's rows
void dir_gather_values(const DirView& d, const OutSlots& o, int c, int32_t* bval, uint8_t* bvalid,
                       hipStream_t s);

struct HeavyEntry {
    uint32_t prow;
    uint32_t lo;
    uint32_t cnt;
    uint32_t pad;
    uint64_t out;
};

// ---- zero fill by kernel (safe to capture in a HIP graph, unlike hipMemsetAsync)
void zero_u32(uint32_t* p, uint64_t n, hipStream_t s);

// ---- resident-table page index
void launch_rank_index(const DCol& c, uint16_t* rank_out, hipStream_t s);
// INT32 pages -> flat values + row-space validity bitmap (vbits zeroed beforehand, or nullptr
// for a column without NULLs)
void launch_decode_int32(const DCol& c, int32_t* flat, uint64_t* vbits, hipStream_t s);
// VARCHAR pages -> sref[rows] (kNullRef / kLongRef | page / offset << 16 | length)
void launch_decode_varchar(const DCol& c, uint64_t* sref, hipStream_t s);
// group index of up to kIndexCols columns in one launch: out[c] receives 2
// words per 64-row group of col[c] (see hj_device.hpp); every page of col[c] but the last holds
// >= 64 rows
constexpr int kIndexCols = 16;
struct GroupIndexBatch {
    DCol col[kIndexCols];
    uint64_t* out[kIndexCols];
    uint32_t blk0[kIndexCols + 1];  // first block of each column, blk0[n] = the grid
    int32_t n;
};
void launch_group_index(const GroupIndexBatch& b, hipStream_t s);

// ---- generic exclusive scans (tile = 4096 elements per block)
size_t scan_partials_needed(uint64_t n);
// out[i] = sum of in[0..i); *total = sum of all (device pointer, may be null)
void scan_excl_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partial,
                       uint32_t* total, hipStream_t s);
void scan_excl_sum_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* partial,
                       uint64_t* total, hipStream_t s);
// dir[s] = (max(e[0..s)) << 16) | (bloom[s] & 0xffff) for s < D; dir[D] = n << 16
void scan_dir_words(const uint32_t* e, const uint32_t* bloom, uint64_t* dir, uint64_t D,
                    uint64_t n, uint32_t* partial, hipStream_t s);

// ---- radix sort of (key, value) u32 pairs by key, 4 passes of 8 bits; result in keys/vals.
size_t radix_sort_scratch_bytes(uint64_t n);
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                      uint64_t n, void* scratch, hipStream_t s);

// ---- directory over sorted hashes: e/bloom scratch of D u32 each
void build_directory(const uint32_t* sh, uint64_t n, uint32_t k, uint64_t* dir, uint32_t* e,
                     uint32_t* bloom, uint32_t* partial, hipStream_t s);

// ---- build-side key extraction (two passes: count per block, then write)
uint32_t extract_scan_blocks(const ScanSrc& src);
void extract_scan_count(const ScanSrc& src, uint32_t* counts, hipStream_t s);
void extract_scan_write(const ScanSrc& src, const uint32_t* offsets, uint32_t* out_h,
                        uint32_t* out_row, hipStream_t s);
uint32_t rel_blocks(uint64_t n);
void extract_rel_count(const RelSrc& src, uint32_t* counts, hipStream_t s);
void extract_rel_write(const RelSrc& src, const uint32_t* offsets, uint32_t* out_h,
                       uint32_t* out_row, hipStream_t s);

// ---- probe (two passes: per-block output counts, then writes). With a match array (one
// word per probe row: first directory position << 32 | match count, 0 for none) the count pass
// stores each row's lookup and the write pass reads it back instead of looking the key up again.
void probe_scan_count(const ScanSrc& src, const DirView& d, uint64_t* counts, uint64_t* match,
                      hipStream_t s);
void probe_scan_write(const ScanSrc& src, const DirView& d, const uint64_t* offsets,
                      const OutSlots& o, HeavyEntry* heavy, uint32_t* heavy_count,
                      uint64_t* match, hipStream_t s);
void probe_rel_count(const RelSrc& src, const DirView& d, uint64_t* counts, uint64_t* match,
                     hipStream_t s);
void probe_rel_write(const RelSrc& src, const DirView& d, const uint64_t* offsets,
                     const OutSlots& o, HeavyEntry* heavy, uint32_t* heavy_count, uint64_t* match,
                     hipStream_t s);
// validity bitmap (whole 64-bit words) of n rows from their validity bytes (nullptr: all set)
void pack_valid8(const uint8_t* valid8, uint64_t n, uint8_t* bits, hipStream_t s);
void expand_heavy(const DirView& d, const OutSlots& o, const HeavyEntry* heavy,
                  const uint32_t* heavy_count, hipStream_t s);

// ---- filtered scan without a join (root scan): selected row ids
void select_scan_count(const ScanSrc& src, uint32_t* counts, hipStream_t s);
void select_scan_write(const ScanSrc& src, const uint32_t* offsets, uint32_t* out_row,
                       hipStream_t s);


// ---- directory test surface: raw u32 keys
void hash_keys(const uint32_t* keys, uint64_t n, uint32_t* h, uint32_t* row, hipStream_t s);
void dir_probe_keys(const DirView& d, const uint32_t* keys, uint64_t m, uint64_t* counts,
                    uint64_t* first, hipStream_t s);
void dir_probe_pairs(const DirView& d, const uint32_t* keys, uint64_t m, const uint64_t* offsets,
                     uint64_t* pairs, hipStream_t s);

}  // namespace hj
