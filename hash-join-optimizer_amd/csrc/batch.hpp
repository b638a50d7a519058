// batch.hpp - many independent join queries run together.
//
// No counterpart in the reference, whose execute() runs one plan at a time and spreads one
// join's probe over threads by fetch_add on a 4096-row chunk counter (src/execute.cpp:87-117).
// Here the work of a whole batch of queries is one list of items in HBM, in three kinds:
//   index   64-row groups of a column the batch reads through a group index (hj_device.hpp)
//   insert  build rows of one join: range filters, then a compare-and-swap into the join's
//           bucketized table (the pipeline's table layout, pipeline.hpp)
//   probe   64-row groups of one query's source: every join's key looked up in turn, survivors
//           staged in LDS, their output columns written at a range reserved with one atomic
// Each kind is one launch of resident workgroups that take its items in grid-stride order; the
// launches run in stream order (clear, index, insert, probe), so an item never waits for another.
// A build key that occurs twice, or a table segment without a free slot, flags its query, which
// the host then runs on its own.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace hj {

constexpr int kBatchJoins = 4;      // joins of one batched query (a chain keyed by the source)
constexpr int kBatchOut = 16;       // output columns of one batched query
constexpr uint32_t kBatchBlock = 256;
constexpr uint32_t kIndexItemGroups = 1024;  // 64-row groups per index item
constexpr uint32_t kStageWords = 6144;       // LDS survivor stage of a probe item (24 KiB)

enum : uint32_t { kItemIndex = 0, kItemInsert = 1, kItemProbe = 2 };

// one work item, 16 bytes
struct BatchItem {
    uint32_t kind_q;  // kind << 24 | query
    uint32_t sub;     // index: the column (BatchDesc::icol); insert: the join
    uint32_t lo, hi;  // index, probe: 64-row groups [lo, hi); insert: rows [lo, hi)
};

// Per-query counters (kQCtrWords words, zeroed by the clear pass):
//   [1] flag: a duplicate build key or a full segment  [2] output rows
//   [4 + j] join j's table holds the key kEmptyKey (INT32_MIN) in slot 2^k
constexpr uint32_t kQCtrWords = 8;

struct BatchJoin {
    DCol bkey;                   // build key column (of the build table)
    DCol bfcol[kMaxFilters];     // range filters of the build scan
    int32_t bflo[kMaxFilters];
    int32_t bfhi[kMaxFilters];
    int32_t nbf;
    uint32_t brows;              // build table rows
    uint32_t k;                  // the table holds 2^k key slots
    uint32_t pad;
    DCol pkey;                   // probe key: a column of the source table
    uint32_t* table;             // 2 * 2^k + 8 words, buckets of [4 keys | 4 rows]
};

struct BatchQuery {
    BatchJoin j[kBatchJoins];
    int32_t nj;
    int32_t nsf;                 // range filters of the source scan
    DCol sfcol[kMaxFilters];
    int32_t sflo[kMaxFilters];
    int32_t sfhi[kMaxFilters];
    uint32_t src_rows;
    int32_t nout;
    uint32_t onull;              // bit c: output column c may hold NULL (bitmap set by atomics)
    int32_t oslot[kBatchOut];    // 0: the source row, 1 + j: the build row of join j
    DCol ocol[kBatchOut];        // the column it is read from
    int32_t* out[kBatchOut];     // values (capacity src_rows)
    uint32_t* ovalid[kBatchOut]; // validity bitmaps (bit i of word i / 32)
    uint32_t* ctr;               // kQCtrWords counters
};

// one region the clear pass fills: words 32-bit words of value v
struct FillSeg {
    uint32_t* p;
    uint64_t words;
    uint32_t v;
    uint32_t pad;
};

struct BatchDesc {
    const BatchQuery* q;
    const BatchItem* items;
    uint32_t nitems;
    uint32_t nidx;               // index items (the first nidx items)
    const DCol* icol;            // columns indexed by the index items ...
    uint64_t* const* iout;       // ... and their group index arrays
    uint32_t* head;              // counters of the batch (unused words, zeroed by the clear pass)
    uint32_t probe0;             // first probe item (the insert items are [nidx, probe0))
    uint32_t wide0;              // first probe item of a query with three or four joins
    uint32_t radix_bits;         // insert items order their tuples by this many hash bits
    uint32_t coop;               // probe lookups: 0 one lane per row, 1 four lanes per row
    const FillSeg* fill;
    uint32_t nfill;
};

// blocks resident at once of each launch (probe: the variant for unroll x lane mapping)
struct BatchGrids {
    uint32_t index, insert, probe, wide;
};
BatchGrids batch_grids(uint32_t unroll, uint32_t coop, uint32_t num_cus);

struct BatchLaunch {
    uint32_t nfill;
    uint64_t max_fill_words;
    uint32_t nidx, nins, nprobe, nwide;  // items of each kind (probe: one or two joins, wide:
                                         // three or four)
    uint32_t unroll, coop;
    BatchGrids grids;
};
// the clear pass, then the index, insert and probe launches, on stream s (capturable in a HIP
// graph: no memset, no host synchronization)
// launches first .. last of the four (0 clear, 1 index, 2 insert, 3 probe)
void batch_launch(const BatchDesc* d_desc, const BatchLaunch& L, hipStream_t s, int first = 0,
                  int last = 3);

}  // namespace hj
