// materialize.hpp - output columns of a join result gathered in one batch of launches.
//
// For up to kMatCols output columns per launch (descriptors passed by value):
//   pass a   per 256-row tile and column: INT32 values + validity; VARCHAR lengths (kept in
//            offsets[i+1]), validity and the tile's byte count
//   pass b   per VARCHAR column, one block: exclusive scan of the tile byte counts
//   pass c   per tile and VARCHAR column: offsets (tile base + in-tile scan) and the bytes
// The caller sizes each chars buffer from a bound (rows x the column's longest string) when it
// is small enough, so that no host round trip sits between the passes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hj_device.hpp"

namespace hj {

constexpr int kMatCols = 8;
constexpr int kMaxStageArrays = 9;  // stage arrays of the fused probe pass (source + 8 joins)
constexpr uint32_t kMatTile = 256;  // rows per block: one per thread

struct MatCol {
    DCol col;
    const uint32_t* rid;   // row ids into col (nullptr: row i)
    int32_t type;          // HJ_INT32 (0) or HJ_VARCHAR (3)
    int32_t pad;
    int32_t* values;       // INT32
    uint8_t* validity;     // bitmap, 64-bit words
    uint64_t* offsets;     // VARCHAR, n + 1
    char* chars;           // VARCHAR
    uint64_t* tile_sum;    // VARCHAR, tiles + 1 (the last entry receives the total)
    uint64_t* src;         // VARCHAR, n: the string reference (sref) of row i
};

struct MatBatch {
    MatCol c[kMatCols];
    int32_t ncol;
    uint32_t tiles;
    uint64_t n;
};

// Output of the fused probe pass (joins keyed by the source only) straight from its staged
// survivors: block b of the pass staged blk_cnt[b] rows (each output slot's row ids) from row
// b * rows_per_block of the stage arrays and counted their string bytes per output VARCHAR
// column (blk_bytes). Two launches, no host round trip:
//   emit       per `group` blocks of the pass: output row and string byte offsets from the
//              counts of the blocks before them, then INT32 values, string offsets and bytes,
//              validity bytes
//   emit bits  validity bitmaps from the validity bytes; totals are in `out`
// Output buffers hold cap_rows rows (an upper bound of the row count, known on the host), each
// chars buffer chars_cap bytes; a larger string total sets the overflow word instead of writing.
constexpr int kEmitCols = 16;
// k_emit sums the counts of every fused-pass block before its group with a fixed number of
// loads per thread: the pass may have at most this many blocks (emit_rows refuses more)
constexpr uint32_t kEmitMaxBlocks = 2048;

struct EmitCol {
    DCol col;
    int32_t slot;         // stage array of the row ids
    int32_t type;         // HJ_INT32 (0) or HJ_VARCHAR (3)
    int32_t sidx;         // VARCHAR: its entry in the pass's per-block string byte counts
    int32_t* values;      // INT32
    uint64_t* offsets;    // VARCHAR, cap_rows + 1
    char* chars;          // VARCHAR, chars_cap bytes
    uint64_t chars_cap;
    uint8_t* valid8;      // scratch: validity byte per output row (unused when nonull)
    uint8_t* validity;    // bitmap, 64-bit words
    int32_t nonull;       // no output row is NULL: the bitmap is set whole, no validity bytes
    int32_t staged;       // INT32: the stage array holds the values themselves (no gather)
};

struct EmitDesc {
    EmitCol c[kEmitCols];
    int32_t ncol;
    int32_t nstr;                // VARCHAR columns (blk_bytes entries 0 .. nstr-1)
    uint32_t blocks;             // blocks of the fused pass ...
    uint32_t group;              // ... per block of the emit pass (<= 16)
    uint32_t rows_per_block;     // stage rows per block of the fused pass
    uint32_t bytes_stride;       // entries per block in blk_bytes
    uint64_t cap_rows;
    const uint32_t* blk_cnt;     // survivors per block of the fused pass
    const uint32_t* blk_origin;  // first stage row of each block (null: block * rows_per_block)
    const uint64_t* blk_bytes;   // string bytes per block of the fused pass and VARCHAR column
    const uint32_t* stage[kMaxStageArrays];
    const uint32_t* meta;        // pipeline counters: [0] != 0 voids the pass (nothing written)
    uint64_t* out;               // [0] rows, [1 + c] string bytes of column c, [1 + kEmitCols]
                                 // overflow of a chars buffer (zeroed)
};

// false (nothing launched) when the pass has more than kEmitMaxBlocks blocks or a group is
// larger than 16
bool emit_rows(const EmitDesc& d, hipStream_t s);

// passes a and b; pass c separately, once chars buffers exist
void mat_gather(const MatBatch& b, hipStream_t s);
void mat_strings(const MatBatch& b, hipStream_t s);

}  // namespace hj
