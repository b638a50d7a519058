// gindex.hpp - the query's group index of a column (see hj_device.hpp), built on the device:
// one thread per 64-row group of row space, kGroupIndexBlock threads per block. Run by its own
// launch (k_group_index) or by extra blocks of a pipeline build launch.
#pragma once
#include "kernels.hpp"

namespace hj {

constexpr uint32_t kGroupIndexBlock = 256;

// Group index of a query's columns: one thread per 64-row group of row space. The group's page
// comes from the chunk index and the page row starts, its validity bits from the bitmap of that
// page (from bit local0 on) and of the next one (from bit 0), its first value index from the
// page's group ranks. Reads the bitmaps once (1 bit per row) and writes 16 bytes per 64 rows.
// (One wave per page, with the first value indexes from a wave prefix sum, measured slower:
// a page holds 3-33 groups, so most lanes idle and the launch needs 90K short waves.)
// The record of group g (rows r = 64 g ..) given its page p (rows [rs, re), header hdr, first
// page-group gs) and the next page's header hq and row count nrq (used only where the group's
// rows run past re into a page p + 1 < num_pages).
__device__ inline void group_index_from(const DCol& c, uint64_t* gout, uint32_t g, uint32_t p,
                                        uint32_t rs, uint32_t re, uint32_t hdr, uint32_t gs,
                                        uint32_t hq, uint32_t nrq) {
    const uint32_t r = g * 64u;
    const uint32_t local0 = r - rs;
    const uint32_t split = re - r < 64u ? re - r : 64u;
    const uint64_t low_split = split >= 64u ? ~0ull : (1ull << split) - 1ull;
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
    uint64_t bits;
    uint32_t rank0;
    if (nn == nr) {
        bits = low_split;
        rank0 = local0;
    } else {
        const uint32_t nb = (nr + 7u) >> 3, gi = local0 >> 6, sh = local0 & 63u;
        const uint64_t w0 = group_bits(pg, nb, gi);
        // the next word lies past the bitmap only where the group's rows end before it
        const uint64_t w1 = sh ? group_bits(pg, nb, gi + 1u) : 0ull;
        bits = ((w0 >> sh) | (sh ? (w1 << (64u - sh)) : 0ull)) & low_split;
        rank0 = (uint32_t)gp(c.rank)[gs + gi] + (uint32_t)__popcll(w0 & ((1ull << sh) - 1ull));
    }
    if (split < 64u && p + 1u < c.num_pages) {
        const HJ_GLOBAL uint8_t* pq = page_ptr(c, p + 1u);
        const uint32_t want = 64u - split;
        const uint32_t avail = nrq < want ? nrq : want;
        const uint64_t low_avail = (1ull << avail) - 1ull;  // avail < 64
        const uint64_t bq = ((hq >> 16) == (hq & 0xffffu))
                                ? low_avail
                                : group_bits(pq, ((hq & 0xffffu) + 7u) >> 3, 0) & low_avail;
        bits |= bq << split;
    }
    HJ_GLOBAL uint64_t* out = gp(gout) + 2u * (uint64_t)g;
    out[0] = bits;
    if (c.type == 0) {  // INT32: byte offsets (the engine keeps such columns' pages < 4 GiB)
        const uint32_t in_first = (uint32_t)__popcll(bits & low_split);
        out[1] = gint_pack(p * HJ_PAGE_BYTES + 4u + 4u * rank0,
                           HJ_PAGE_BYTES - 4u * (rank0 + in_first), split);
    } else {
        out[1] = gloc_pack(p, rank0, split);
    }
}

// Group index of a query's columns: one thread per 64-row group of row space. The group's page
// comes from the chunk index and the page row starts, its validity bits from the bitmap of that
// page (from bit local0 on) and of the next one (from bit 0), its first value index from the
// page's group ranks. Reads the bitmaps once (1 bit per row) and writes 16 bytes per 64 rows.
// (One wave per page, with the first value indexes from a wave prefix sum, measured slower:
// a page holds 3-33 groups, so most lanes idle and the launch needs 90K short waves.)
__device__ inline void group_index_group(const DCol& c, uint64_t* gout, uint32_t g) {
    if ((uint64_t)g * 64u >= c.num_rows) return;
    const uint32_t p = find_page(c, g * 64u);
    const uint32_t rs = gp(c.row_start)[p], re = gp(c.row_start)[p + 1];
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)page_ptr(c, p);
    const bool nxt = p + 1u < c.num_pages;
    const uint32_t hq = nxt ? *(const HJ_GLOBAL uint32_t*)page_ptr(c, p + 1u) : 0u;
    const uint32_t nrq = nxt ? gp(c.row_start)[p + 2] - re : 0u;
    group_index_from(c, gout, g, p, rs, re, hdr, gp(c.group_start)[p], hq, nrq);
}

// The fused probe's prologue: block-local page tables. For a block's row range of a column,
// the row starts, headers and first page-groups of the pages p0 .. p0 + kPageTable - 1 (p0 the
// page of the range's first row), loaded by one wave into LDS (pt: 3 x (kPageTable + 1) words);
// then every group of the range finds its page in LDS and loads only its bitmap words and rank.
constexpr uint32_t kPageTable = 64;
constexpr uint32_t kPageTableWords = 3u * (kPageTable + 1u) + 2u;  // + p0, usable flag

// wave part: the table of rows [r_lo, r_hi) (r_lo < num_rows), usable when the page of row
// r_hi - 1 is at most entry limit - 2 (limit <= kPageTable - 1, so that a group's next page and
// its end are in the table; limit 0: never, every group as group_index_group)
__device__ inline void page_table_load(const DCol& c, uint32_t r_lo, uint32_t r_hi, uint32_t* pt,
                                       uint32_t lane, uint32_t limit) {
    const uint32_t p0 = find_page(c, r_lo);
    uint32_t* prs = pt;
    uint32_t* phdr = pt + (kPageTable + 1u);
    uint32_t* pgs = pt + 2u * (kPageTable + 1u);
    const uint32_t p = p0 + lane;  // lane < kPageTable
    const uint32_t rsv = p <= c.num_pages ? gp(c.row_start)[p] : 0xffffffffu;
    phdr[lane] = p < c.num_pages ? *(const HJ_GLOBAL uint32_t*)page_ptr(c, p) : 0u;
    pgs[lane] = p < c.num_pages ? gp(c.group_start)[p] : 0u;
    prs[lane] = rsv;
    if (lane == 0) {
        const uint32_t pe = p0 + kPageTable;
        prs[kPageTable] = pe <= c.num_pages ? gp(c.row_start)[pe] : 0xffffffffu;
        phdr[kPageTable] = 0u;
        pgs[kPageTable] = 0u;
    }
    // entries whose first row is at most r_hi - 1: the last of them holds row r_hi - 1
    const uint32_t k_last = (uint32_t)__popcll(__ballot(rsv <= r_hi - 1u)) - 1u;
    if (lane == 0) {
        pt[3u * (kPageTable + 1u)] = p0;
        const uint32_t lim = limit < kPageTable - 1u ? limit : kPageTable - 1u;
        pt[3u * (kPageTable + 1u) + 1u] = k_last + 2u <= lim ? 1u : 0u;
    }
}

// the aligned 8-byte word j (< 1024) of page pg's 8 KiB
__device__ inline uint64_t page_word(const HJ_GLOBAL uint8_t* pg, uint32_t j) {
    return ((const HJ_GLOBAL uint64_t*)pg)[j < 1023u ? j : 1023u];
}
__device__ inline uint64_t funnel64(uint64_t lo, uint64_t hi, uint32_t s) {
    return s ? (lo >> s) | (hi << (64u - s)) : lo;
}

// thread part: the records of the range's groups g = g0 + t + u * stride (u < kTablePer, those
// below the row count) of column c, from its page table; where the table is not usable, as
// group_index_group. Every load of the thread's groups is issued before any is used: the
// bitmap words as aligned words of the page (the bits a group uses lie in its page's bitmap;
// words beyond the page end are never loaded, and bits read past the bitmap are masked off),
// the page-group rank, the next page's first bitmap words.
constexpr int kTablePer = 2;
__device__ inline void page_table_groups(const DCol& c, uint64_t* gout, uint32_t g0,
                                         uint32_t ng, uint32_t t, uint32_t stride,
                                         const uint32_t* pt) {
    const uint32_t* prs = pt;
    const uint32_t* phdr = pt + (kPageTable + 1u);
    const uint32_t* pgs = pt + 2u * (kPageTable + 1u);
    const uint32_t p0 = pt[3u * (kPageTable + 1u)];
    const bool usable = pt[3u * (kPageTable + 1u) + 1u] != 0;
    for (uint32_t b0 = 0; b0 < ng; b0 += stride * kTablePer) {
        uint32_t g[kTablePer], p[kTablePer], rs[kTablePer], re[kTablePer], hdr[kTablePer];
        uint32_t hq[kTablePer], nrq[kTablePer], gi[kTablePer], s0[kTablePer], sq[kTablePer];
        uint64_t w0[kTablePer], w1[kTablePer], w2[kTablePer], q0[kTablePer], q1[kTablePer];
        uint32_t rk[kTablePer];
        bool have[kTablePer];
#pragma unroll
        for (int u = 0; u < kTablePer; ++u) {
            const uint32_t j = b0 + u * stride + t;
            g[u] = g0 + j;
            have[u] = j < ng && (uint64_t)g[u] * 64u < c.num_rows;
            if (have[u] && !usable) {
                group_index_group(c, gout, g[u]);
                have[u] = false;
            }
            const uint32_t r = g[u] * 64u;
            uint32_t lo = 0, hi = kPageTable;  // the last entry with a first row <= r
            while (have[u] && hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (prs[mid] <= r)
                    lo = mid;
                else
                    hi = mid;
            }
            p[u] = p0 + lo;
            rs[u] = prs[lo];
            re[u] = prs[lo + 1u];
            hdr[u] = phdr[lo];
            hq[u] = phdr[lo + 1u];
            nrq[u] = prs[lo + 2u] - prs[lo + 1u];
            const uint32_t nr = hdr[u] & 0xffffu, nb = (nr + 7u) >> 3;
            const uint32_t local0 = r - rs[u];
            gi[u] = local0 >> 6;
            const uint32_t pos = 8u * (HJ_PAGE_BYTES - nb) + 64u * gi[u];  // page-group gi's bit
            s0[u] = pos & 63u;
            const HJ_GLOBAL uint8_t* pg = page_ptr(c, have[u] ? p[u] : 0u);
            const bool dense = (hdr[u] >> 16) == nr;
            w0[u] = w1[u] = w2[u] = q0[u] = q1[u] = 0;
            rk[u] = 0;
            if (have[u] && !dense) {
                w0[u] = page_word(pg, pos >> 6);
                w1[u] = page_word(pg, (pos >> 6) + 1u);
                w2[u] = page_word(pg, (pos >> 6) + 2u);
                rk[u] = gp(c.rank)[pgs[lo] + gi[u]];
            }
            const uint32_t split = re[u] - r < 64u ? re[u] - r : 64u;
            const uint32_t nrq_h = hq[u] & 0xffffu;
            const uint32_t posq = 8u * (HJ_PAGE_BYTES - ((nrq_h + 7u) >> 3));
            sq[u] = posq & 63u;
            if (have[u] && split < 64u && p[u] + 1u < c.num_pages && (hq[u] >> 16) != nrq_h) {
                const HJ_GLOBAL uint8_t* pq = page_ptr(c, p[u] + 1u);
                q0[u] = page_word(pq, posq >> 6);
                q1[u] = page_word(pq, (posq >> 6) + 1u);
            }
        }
#pragma unroll
        for (int u = 0; u < kTablePer; ++u) {
            if (!have[u]) continue;
            const uint32_t r = g[u] * 64u, local0 = r - rs[u], sh = local0 & 63u;
            const uint32_t split = re[u] - r < 64u ? re[u] - r : 64u;
            const uint64_t low_split = split >= 64u ? ~0ull : (1ull << split) - 1ull;
            const uint32_t nr = hdr[u] & 0xffffu;
            uint64_t bits;
            uint32_t rank0;
            if ((hdr[u] >> 16) == nr) {
                bits = low_split;
                rank0 = local0;
            } else {
                const uint64_t a = funnel64(w0[u], w1[u], s0[u]);  // rows 64 gi .. 64 gi + 63
                const uint64_t b = funnel64(w1[u], w2[u], s0[u]);  // the next 64 rows
                bits = (sh ? (a >> sh) | (b << (64u - sh)) : a) & low_split;
                rank0 = rk[u] + (uint32_t)__popcll(a & ((1ull << sh) - 1ull));
            }
            if (split < 64u && p[u] + 1u < c.num_pages) {
                const uint32_t want = 64u - split;
                const uint32_t avail = nrq[u] < want ? nrq[u] : want;
                const uint64_t low_avail = (1ull << avail) - 1ull;  // avail < 64
                const uint64_t bq = ((hq[u] >> 16) == (hq[u] & 0xffffu))
                                        ? low_avail
                                        : funnel64(q0[u], q1[u], sq[u]) & low_avail;
                bits |= bq << split;
            }
            HJ_GLOBAL uint64_t* out = gp(gout) + 2u * (uint64_t)g[u];
            out[0] = bits;
            if (c.type == 0) {
                const uint32_t in_first = (uint32_t)__popcll(bits & low_split);
                out[1] = gint_pack(p[u] * HJ_PAGE_BYTES + 4u + 4u * rank0,
                                   HJ_PAGE_BYTES - 4u * (rank0 + in_first), split);
            } else {
                out[1] = gloc_pack(p[u], rank0, split);
            }
        }
    }
}

__device__ inline void group_index_thread(const GroupIndexBatch& b, uint32_t blk, uint32_t tid) {
    int ci = 0;
    for (int i = 1; i < b.n; ++i) ci += blk >= b.blk0[i] ? 1 : 0;
    group_index_group(b.col[ci], b.out[ci], (blk - b.blk0[ci]) * kGroupIndexBlock + tid);
}

}  // namespace hj
