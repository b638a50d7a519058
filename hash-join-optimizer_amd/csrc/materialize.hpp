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

// passes a and b; pass c separately, once chars buffers exist
void mat_gather(const MatBatch& b, hipStream_t s);
void mat_strings(const MatBatch& b, hipStream_t s);

}  // namespace hj
