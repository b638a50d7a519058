// predicate.hpp - a scan's load-time predicate program evaluated on the device.
//
// The reference evaluates Comparison / LogicalOperation trees column by column at load time
// (src/statement.cpp:46-230, include/inner_column.h:250-560) and keeps the rows whose result is
// set. Here the program (postfix, at most kMaxPredOps entries) runs once per row in one thread:
// comparisons push a bit, AND / OR / NOT combine the top of a 32-deep bit stack, and the result
// is written as 0 / 1 into an INT32 row-space array that the scan then filters on (= 1), so every
// scan path takes it as one more range filter.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hj_device.hpp"

namespace hj {

constexpr int kMaxPredOps = 32;
constexpr int kMaxPredCols = 8;   // distinct columns a program reads
constexpr int kMaxLikeTokens = 63;
constexpr int kMaxByteTokens = 31;  // byte automaton: bit kMaxByteTokens is its last state bit

struct PredOp {
    int32_t op;      // HJ_P_*
    int32_t col;     // comparisons: index into PredProg::col
    int32_t ival;    // INT32 literal; LIKE: 1 when the pattern is '%' and ASCII characters only
                     //   and at most kMaxByteTokens tokens (the automaton then runs over bytes)
    uint32_t loff;   // VARCHAR: literal bytes at lit + loff, llen of them; LIKE: tokens at
    uint32_t llen;   //   (const uint32_t*)(lit + loff), llen of them
    uint32_t pad;    // LIKE: the longest run of consecutive '%' tokens
    uint64_t pct;    // LIKE: bit j set iff token j - 1 is '%'
    uint64_t any;    // LIKE: bit j set iff token j - 1 is '_'
};

struct PredProg {
    PredOp op[kMaxPredOps];
    int32_t nops;
    int32_t ncol;
    DCol col[kMaxPredCols];
    const uint8_t* lit;  // device: literal bytes and LIKE tokens
    int32_t* out;        // device: 1 / 0 per row
    uint64_t n;
};

// LIKE tokens: a code point, or one of these
constexpr uint32_t kLikeAnyRun = 0xFFFFFFFFu;  // '%'
constexpr uint32_t kLikeAnyOne = 0xFFFFFFFEu;  // '_'

void eval_predicate(const PredProg& p, hipStream_t s);

}  // namespace hj
