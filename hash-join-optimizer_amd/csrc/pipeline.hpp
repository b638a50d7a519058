// pipeline.hpp - the probe pipeline for left-deep join chains whose build sides are base scans:
// all build sides are built in one batch of launches (sort-free unchained directories); the probe
// source is streamed once for the first join, and the survivors run the remaining joins.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace hj {

constexpr int kMaxStages = 8;
constexpr uint32_t kMaxFusedSlot = 64;  // longest directory slot the pipeline accepts
constexpr uint32_t kDead = 0xffffffffu; // a survivor that failed a later join

// One build side: a (filtered) base scan turned into an unchained directory whose adjacency
// array is grouped by slot (order inside a slot is arbitrary; keys are expected unique).
struct BuildJob {
    DCol key;
    DCol fcol[kMaxFilters];
    int32_t flo[kMaxFilters];
    int32_t fhi[kMaxFilters];
    int32_t nf;
    uint32_t page_block0;     // first block of this job in page-parallel grids
    uint32_t kmax;            // slot arrays hold 2^kmax (+1) entries
    uint32_t* meta;           // [0] tuples, [1] unused, [2] longest slot
    uint32_t* blk;            // per key page: passing rows
    uint32_t* blk_off;        // per key page: exclusive prefix of blk
    uint32_t* cnt;            // per slot
    uint32_t* bloom;          // per slot
    uint64_t* dir;            // 2^kmax + 1 words: start << 32 | min(count, 0xffff) << 16 | Bloom
    uint32_t* part;           // slot-scan partials (one per 4096 slots)
    uint32_t* st_h;           // staged tuples at their page's first row (capacity: table rows)
    uint32_t* st_row;
    uint32_t* st_pos;         // position of the tuple inside its slot
    uint64_t* adj;            // adjacency array grouped by slot: row << 32 | hash
    uint32_t* bf;             // blocked Bloom filter, kBloomWords words (used for small builds)
    uint64_t rows;            // table rows (upper bound of the tuple count)
};

constexpr uint32_t kBloomWords = 4096;  // 16 KiB: the LDS-resident filter of a build side

struct BuildBatch {
    BuildJob job[kMaxStages];
    int32_t njobs;
    uint32_t page_blocks;
    uint32_t grid;            // blocks of the grid-stride kernels
};

struct PipeStage {
    int32_t job;       // build job of this join
    int32_t key_slot;  // 0: key is a column of the probe source; j+1: of stage j's build table
    DCol keycol;
};

struct PipeDesc {
    DCol driver;                  // source column holding stage 0's key; one block per page
    DCol fcol[kMaxFilters];       // filters of the source scan
    int32_t flo[kMaxFilters];
    int32_t fhi[kMaxFilters];
    int32_t nf;
    PipeStage st[kMaxStages];
    int32_t nst;
    int32_t nout;
    int32_t out_slot[kMaxSlots];  // 0: source row, j+1: build row of stage j
    uint32_t* final_out[kMaxSlots];
    uint32_t* s0_row;             // stage-0 survivors at their page's first row
    uint32_t* s0_b;
    uint32_t* page_cnt;           // per driver page
    uint32_t* page_off;
    uint32_t* l_row;              // dense survivor list after stage 0
    uint32_t* l_b[kMaxStages];    // build row per stage, kDead once a join failed
    uint32_t* tile_cnt;           // final compaction, per 4096 survivors
    uint32_t* tile_off;
    uint32_t* meta;               // [0] a key matched more than one build tuple,
                                  // [1] a directory slot exceeded kMaxFusedSlot,
                                  // [2] stage-0 survivors, [3] output rows
    const BuildJob* jobs;
    uint32_t grid;                // blocks of the grid-stride kernels
    // multi-join chains: joins keyed by the source run as streaming passes over their key
    // column in `order` (smallest build first), exchanging survivor bitmaps; the remaining
    // joins (keys from a build table) run over the compacted survivors afterwards
    int32_t nstream;
    int32_t order[kMaxStages];
    uint32_t* bm[2];              // survivor bitmaps, one bit per source row
    uint32_t bm_words;
};

// batch build: extract + stage, per-job page scan, clear, slot counts + Bloom bits,
// slot scan -> directory, scatter
void pipe_build(const BuildBatch& hb, const BuildBatch* db, hipStream_t s);
// stage 0 over the driver pages, survivor compaction, later stages, final compaction
void pipe_probe(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s);

}  // namespace hj
