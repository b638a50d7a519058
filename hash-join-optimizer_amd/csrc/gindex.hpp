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
__device__ inline void group_index_group(const DCol& c, uint64_t* gout, uint32_t g) {
    const uint64_t rows = c.num_rows;
    if ((uint64_t)g * 64u >= rows) return;
    const uint32_t r = g * 64u;
    const uint32_t p = find_page(c, r);
    const uint32_t rs = gp(c.row_start)[p], re = gp(c.row_start)[p + 1];
    const uint32_t local0 = r - rs;
    const uint32_t split = re - r < 64u ? re - r : 64u;
    const uint64_t low_split = split >= 64u ? ~0ull : (1ull << split) - 1ull;
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
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
        rank0 = (uint32_t)gp(c.rank)[gp(c.group_start)[p] + gi] +
                (uint32_t)__popcll(w0 & ((1ull << sh) - 1ull));
    }
    if (split < 64u && p + 1u < c.num_pages) {
        const uint32_t q = p + 1u;
        const HJ_GLOBAL uint8_t* pq = page_ptr(c, q);
        const uint32_t hq = *(const HJ_GLOBAL uint32_t*)pq;
        const uint32_t nrq = gp(c.row_start)[q + 1] - gp(c.row_start)[q];
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

__device__ inline void group_index_thread(const GroupIndexBatch& b, uint32_t blk, uint32_t tid) {
    int ci = 0;
    for (int i = 1; i < b.n; ++i) ci += blk >= b.blk0[i] ? 1 : 0;
    group_index_group(b.col[ci], b.out[ci], (blk - b.blk0[ci]) * kGroupIndexBlock + tid);
}

}  // namespace hj
