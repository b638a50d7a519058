// pagewrite.hpp - the result's 8192-byte column pages written on the device.
//
// The reference writes its output ColumnarTable on the host, page by page, after the last join
// (table_to_columnar, src/execute.cpp:431-573, page rules include/plan.h:265-396). Here the flat
// device result columns are encoded in HBM and copied out page by page:
//   k_page_valid  counts each INT32 column's valid rows; a column without NULL is cut by
//                 arithmetic (every page but the last holds 1984 rows)
//   k_page_plan   one wave per column walks the rows and cuts pages exactly where the host
//                 writer would. A page's byte count is monotone in the rows it holds, so a wave
//                 finds the first row that does not fit among 64 candidates with one ballot:
//                 INT32 columns step 2048 rows per iteration (one validity word per lane),
//                 VARCHAR columns 512 rows (8 per lane); strings longer than 8185 bytes
//                 close the open page and take their own 0xffff / 0xfffe pages.
//   k_page_write  one block per page assembles the page in LDS (zeroed first, so every byte
//                 matches the host writer's zero-filled page) and stores it whole.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hj {

constexpr uint32_t kPageBytes = 8192;
constexpr uint32_t kLongStr = kPageBytes - 7;   // longer strings take pages of their own
constexpr uint32_t kLongPiece = kPageBytes - 4; // chars per 0xffff / 0xfffe page
constexpr int kPageCols = 16;
constexpr uint64_t kJumpMinRows = 65536;  // VARCHAR columns longer than this: planned by jumps

// one output page: rows [row0, row0 + rows) of the column, or piece `rows` of the long string
// in row row0 (kind 1: first piece, 0xffff page; kind 2: later pieces, 0xfffe pages)
struct PageRef {
    uint64_t row0;
    uint32_t rows;
    uint32_t kind;
};

struct PageCol {
    int32_t type;             // HJ_INT32 or HJ_VARCHAR
    uint32_t pad;
    uint64_t n;               // rows
    const uint8_t* validity;  // bit i of byte i/8, padded to whole 8-byte words
    const int32_t* values;    // INT32
    const uint64_t* offsets;  // VARCHAR, n + 1 entries into chars
    const char* chars;        // VARCHAR
    PageRef* refs;            // planner output
    uint64_t cap;             // refs capacity
    uint64_t first;           // writer: index of the column's first page in the page buffer
    // VARCHAR columns of more than kJumpMinRows rows: planned by jumps (scratch of the planner)
    uint32_t jump;            // 1: planned by jumps, 0: one wave walks the column
    uint32_t pad2;
    uint32_t* jlp;            // n: bytes within the row's 4096-row block, up to the row
    uint64_t* jbp;            // per 4096-row block: bytes of the rows before it
    uint32_t* jnx;            // n + 1: next page start after a page starting at the row
    uint32_t* ja;             // n + 1: next^32, then scratch
    uint32_t* jb;             // n + 1: next^1024
    uint32_t* jc;             // chain nodes at every 1024th page (n / 1024 + 3)
    uint32_t* jm;             // ... at every 32nd page (32 x as many)
    uint32_t* jnode;          // ... every page (1024 x as many)
    uint32_t* jcnt;           // nodes in jc
};

struct PageBatch {
    PageCol c[kPageCols];
    uint64_t jump_rows;  // the rows of the longest column planned by jumps (0: none)
    uint32_t ncol;
    uint32_t pad;
    uint64_t* counts;   // planner: pages per column
    uint64_t* valid_rows;  // planner: valid rows per INT32 column (zeroed before page_plan)
    uint8_t* pages;     // writer: all columns' pages, column after column
};

// plan every column (one wave each; INT32 columns without NULL by arithmetic); counts[c] =
// pages of column c (refs beyond cap unwritten); max_rows = the largest column's rows
void page_plan(const PageBatch& b, uint64_t max_rows, hipStream_t s);
// write pages p0 .. p0 + count - 1 of the batch (columns' pages in order, b.c[i].first set from
// the counts) to b.pages[0 .. count)
void page_write(const PageBatch& b, uint64_t p0, uint64_t count, hipStream_t s);

// refs capacity that holds every page of an INT32 column of n rows (a page takes at least 1984
// rows, any 1984 rows fitting in one page)
inline uint64_t int32_page_bound(uint64_t n) { return n / 1984 + 2; }

}  // namespace hj
