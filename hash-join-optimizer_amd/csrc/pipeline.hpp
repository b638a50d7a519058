// pipeline.hpp - the probe pipeline for left-deep join chains whose build sides are base scans
// with unique join keys (star and snowflake joins): every build side is built in the same three
// launches into an open-addressing table, and every join keyed by a column of the probe source
// runs in one fused streaming pass over the source; joins keyed by a column of an earlier
// build table run afterwards over the survivors.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace hj {

constexpr int kMaxStages = 8;
constexpr uint32_t kDead = 0xffffffffu;      // no match / a survivor that failed a later join
constexpr uint32_t kEmptyKey = 0x80000000u;  // a free key slot (INT32_MIN itself is kept aside)
constexpr uint32_t kBloomWords = 4096;       // 16 KiB: the LDS-resident filter of a small build side
constexpr uint32_t kBucketKeys = 4;          // one 16-byte load reads a whole bucket
constexpr uint32_t kMinTableBits = 6;
constexpr int kMaxOutStr = 16;                // output VARCHAR columns counted by the fused pass
constexpr uint32_t kSliceMaxK = 128;          // sliced probe: partition blocks per probe block
constexpr uint32_t kPartRows = 2048;          // sliced probe: source rows per partition block

// word index of slot s's row in a table
__host__ __device__ inline uint32_t row_word(uint32_t s) { return 8u * (s >> 2) + 4u + (s & 3u); }

// key slots for n build tuples: 2^bits >= 2n (load factor <= 1/2, at most two keys per 4-key
// bucket on average; a full bucket continues in the next one, which shares its 64-byte line half
// the time)
__host__ __device__ inline uint32_t table_bits(uint32_t n) {
    uint32_t k = 0;
    while ((1ull << k) < (uint64_t)n) ++k;
    return k + 1 < kMinTableBits ? kMinTableBits : k + 1;
}

// Linear probing runs over buckets within a segment of at most 2^kSegBits buckets (the next
// bucket after a segment's last is its first), so that one workgroup can build a segment in LDS.
constexpr uint32_t kSegBits = 9;
__host__ __device__ inline uint32_t seg_mask(uint32_t k) {
    return (k - 2 < kSegBits ? (1u << (k - 2)) : (1u << kSegBits)) - 1u;
}
__host__ __device__ inline uint32_t next_bucket(uint32_t b, uint32_t smask) {
    return (b & ~smask) | ((b + 1u) & smask);
}
// segment bits of a table with 2^k key slots (2^fine_bits(k) segments); a segment is the
// key slots whose hash shares its top fine_bits(k) bits
__host__ __device__ inline uint32_t fine_bits(uint32_t kmax) {
    return kmax - 2 > kSegBits ? kmax - 2 - kSegBits : 0u;
}
constexpr uint32_t kMaxFineBits = 12;   // partitioned builds: tables of at most 4096 segments ...
constexpr uint32_t kMaxPartBlocks = 512;  // ... and 512 filter blocks per job
// The filter pass orders each block's tuples by the top min(fine_bits(kmax), kPartBits) hash
// bits (kmax from the unfiltered rows). A segment of a table with more segments than partitions
// reads its partition's tuples and keeps its own (at most 4 segments per partition).
constexpr uint32_t kPartBits = 10;
// Device-scope atomics on one address serialize across the XCDs: the filter pass adds its
// blocks' tuple counts into kCountWays counters per job (block x into x % kCountWays)
constexpr uint32_t kCountWays = 8;

// rows per block of the build's filter and insert passes
constexpr uint32_t kBuildRows = 8192;

// One build side: the rows of a (filtered) base scan with a non-NULL key, in a bucketized
// open-addressing table (linear probing over buckets). A bucket is 32 bytes: kBucketKeys keys,
// then the build rows of those key slots, so an insert touches one cache line and a lookup
// reads the keys with one 16-byte load. Slot s is key (s & 3) of bucket s >> 2. A key equal to
// kEmptyKey is kept aside in slot 2^bits (its row at table_row(2^bits)), flagged by meta[2].
struct BuildJob {
    DCol key;
    DCol fcol[kMaxFilters];
    int32_t flo[kMaxFilters];
    int32_t fhi[kMaxFilters];
    int32_t nf;
    uint32_t rows;      // table rows (bound of the tuple count)
    uint32_t kmax;      // the table holds 2^kmax key slots
    uint32_t pbits;     // partitioned build: hash bits of the filter pass's staging order
    uint32_t blk0;      // first block of this job in the filter and insert grids ...
    uint32_t nblk;      // ... and its blocks, ceil(rows / kBuildRows) (at least 1)
    uint32_t* meta;     // [0] tuples, [1] a key occurs twice, [2] kEmptyKey present, [3] unused
    uint32_t* st_key;   // staged tuples, block x of the filter pass from x * kBuildRows on
    uint32_t* st_row;   // (capacity nblk * kBuildRows)
    uint32_t* blk;      // tuples staged per block of the filter pass
    uint32_t* table;    // 2 * 2^kmax + 8 words; 2^table_bits(tuples) key slots used
    uint16_t* poff;     // partitioned build: per filter block, the start of each of the
                        // 2^pbits partitions in its staged range, then its total
    uint32_t* bf;       // blocked Bloom filter, kBloomWords words (small build sides only)
    // pay_on: the table's row words hold the value of INT32 column `pay` (no NULL) at each
    // tuple's row instead of the row (the join's only build-side output, read by the probe
    // straight from the matching bucket)
    DCol pay;
    int32_t pay_on;
};

struct BuildBatch {
    BuildJob job[kMaxStages];
    uint32_t blk0[kMaxStages];  // job[j].blk0, ~0 for j >= njobs
    int32_t part;               // partitioned build (filter + segment passes) instead of
                                // filter + clear + insert
    uint32_t sgrid;             // blocks of the segment pass (a grid-stride loop)
    uint32_t* pcnt;             // partitioned build: tuples per job, kCountWays partial counts
                                // each (zeroed), summed by the segment pass into job[j].meta[0]
    int32_t njobs;
    uint32_t grid;      // blocks of the filter and insert passes (every job's nblk)
    uint32_t cgrid;     // blocks per job of the clear pass (grid-stride)
    // the probe side's group index, built by igrid extra blocks of the filter pass (the builds
    // do not read those columns; the probe launch comes after the build's last pass)
    GroupIndexBatch gib;
    uint32_t igrid;
};

struct PipeStage {
    int32_t job;        // build job of this join
    int32_t key_slot;   // 0: key is a column of the probe source; j+1: of stage j's build table
    DCol keycol;
};

// a join of the fused pass, in probe order (filled by the host, so that the pass reaches every
// pointer it needs in one load from the descriptor)
struct StreamJoin {
    DCol key;               // key column of the source (pages and group index) ...
    const uint32_t* table;  // the build's table ...
    const uint32_t* meta;   // ... its counters ([0] tuples, [2] kEmptyKey present) ...
    const uint32_t* bf;     // ... and Bloom words
};

constexpr int kProbeIdxCols = 4;  // source columns the fused probe's blocks index themselves

struct PipeDesc {
    // the source's columns whose group index the fused probe's blocks build, each block for
    // its own groups, as it starts
    DCol sidx_col[kProbeIdxCols];
    uint64_t* sidx_out[kProbeIdxCols];
    int32_t nsidx;
    uint32_t sidx_table;  // page-table entries a block may use (0: none, per-group page search)
    // sliced probe: ppay_on, the partition pass stores the source row's value of INT32 column
    // ppay (no NULL; the source's only output) in place of the row, and the survivors carry it
    DCol ppay;
    int32_t ppay_on;
    DCol fcol[kMaxFilters];       // filters of the source scan
    int32_t flo[kMaxFilters];
    int32_t fhi[kMaxFilters];
    int32_t nf;
    PipeStage st[kMaxStages];
    int32_t nst;
    int32_t nstream;              // stages keyed by the source, fused into one pass ...
    int32_t order[kMaxStages];    // ... in this order (smallest build first)
    int32_t qpos[kMaxStages];     // stage -> its position in order (-1: not streamed)
    int32_t ndep;                 // stages keyed by a build table (run over the survivors)
    uint32_t src_rows;
    int32_t nout;
    int32_t out_slot[kMaxSlots];  // 0: source row, j+1: build row of stage j
    uint32_t* final_out[kMaxSlots];
    uint32_t* stage[kMaxStages + 1];  // survivors of the fused pass, slot values (source row,
                                      // then build row per stage) at the start of each block's
                                      // range; compacted by a scan of the block counts
    uint32_t* blk_cnt;            // survivors per block of the fused pass
    uint32_t* blk_off;
    uint32_t probe_blocks;        // blocks of the fused pass ...
    uint32_t probe_gpb;           // ... and 64-row groups per block
    // sliced probe (xslices = 8: one streamed join over a large table, probed slice by slice,
    // one slice per XCD): the partition pass's (key, row) pairs (part_blocks * kPartRows each)
    // and slice counts (8 per partition block); probe block c * 8 + x takes slice x of the
    // partition blocks [c * part_k, (c + 1) * part_k) and stages from blk_origin[block] on
    uint32_t xslices;
    uint32_t part_blocks, part_k;
    uint32_t* ptk;
    uint32_t* ptr;
    uint32_t* pcnt;
    uint32_t* blk_origin;
    uint32_t* l_row;              // ndep > 0: survivors of the fused pass
    uint32_t* l_b[kMaxStages];    // build row per stage, kDead once a join failed
    uint32_t* tile_cnt;           // final compaction, per 4096 survivors
    uint32_t* tile_off;
    uint32_t* meta;               // [0] fall back (a build key occurs twice), [1] unused,
                                  // [2] survivors of the fused pass, [3] output rows; then the
                                  // counters of kMaxStages build jobs (4 words each)
    StreamJoin sj[kMaxStages];    // the fused joins in probe order
    // direct output: string bytes of the survivors of each block per output VARCHAR column
    DCol ostr[kMaxOutStr];        // the columns ...
    int32_t ostr_slot[kMaxOutStr];  // ... their row ids' stage array (0: source, 1 + q: join q)
    int32_t nostr;
    uint64_t* blk_bytes;          // per block, kMaxOutStr totals
    const BuildJob* jobs;
    uint32_t grid;                // blocks of the grid-stride kernels
};

// 64-row groups per wave and iteration of the fused pass over NS streamed joins
constexpr int probe_unroll(int ns) { return ns == 1 ? 4 : (ns == 2 ? 3 : (ns <= 4 ? 2 : 2)); }

// blocks of the fused pass over ns streamed joins that are resident at once on the device (the
// pass gives each a contiguous share of the source, so a second round of blocks would double
// its time)
uint32_t probe_resident_blocks(int ns, uint32_t num_cus);

// build: filter + stage, count + clear tables, insert
void pipe_build(const BuildBatch& hb, const BuildBatch* db, hipStream_t s);
// the fused pass over the source (one launch; ev_a/ev_b, when set, record its start and end) ...
void pipe_probe_fused(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s,
                      hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr);
// ... then the survivor compaction, the joins keyed by build tables and the final compaction
void pipe_probe_tail(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s);

}  // namespace hj
