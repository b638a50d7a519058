// pagewrite.hip - device page encoder of the result columns (planner + writer, see pagewrite.hpp).
//
// Page rules restated from include/plan.h:265-396 and src/execute.cpp:447-569 (the host writer
// in pagecodec.cpp follows the same rules). Counting the row being added, row j of a page fits
//   INT32    while 4 + 4 * values + j / 8 + 1 <= 8192
//   VARCHAR  while 4 + 2 * strings + chars + j / 8 + 1 <= 8192
// Both left sides only grow with j, so the page ends before the first row whose cost exceeds
// 8192, and that row opens the next page.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/hjoin.h"
#include "hj_device.hpp"
#include "pagewrite.hpp"

namespace hj {
namespace {

constexpr int kWriteBlock = 256;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ inline void put_ref(const PageCol& c, uint64_t i, uint64_t row0, uint32_t rows,
                               uint32_t kind) {
    if (i < c.cap) {
        HJ_GLOBAL PageRef* r = gp(c.refs) + i;
        r->row0 = row0;
        r->rows = rows;
        r->kind = kind;
    }
}

// inclusive prefix sum over the wave with DPP moves: shifts by 1, 2, 4, 8 within each row of 16
// lanes, then the row totals broadcast from lane 15 and lane 31
__device__ inline uint32_t wave_scan(uint32_t x) {
    int v = (int)x;
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 -> rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 -> rows 2, 3
    return (uint32_t)v;
}

__device__ inline uint32_t lane_u32(uint32_t x, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)l);
}
__device__ inline uint64_t lane_u64(uint64_t x, uint32_t l) {
    return (uint64_t)lane_u32((uint32_t)x, l) | ((uint64_t)lane_u32((uint32_t)(x >> 32), l) << 32);
}
__device__ inline uint32_t first_lane(uint64_t m) {
    return (uint32_t)__builtin_ctzll(m);
}

// Planner state of one column: the open page starts at row s; g = its bytes (VARCHAR) or valid
// rows (INT32) in [s, b) for the window starting at row b (negative: minus those in [b, s) once
// s is inside the window); np = pages cut so far.
struct PlanState {
    uint64_t s = 0;
    int64_t g = 0;
    uint64_t np = 0;
};

// INT32 window of 2048 rows: lane i holds validity word i (rows b + 32i .. b + 32i + 31).
// Row positions are 32-bit and relative to b (the open page starts at most 65535 rows before
// the window), valid-row counts 32-bit.
__device__ inline void int32_window(const PageCol& c, PlanState& st, uint64_t b, uint32_t vword) {
    const uint32_t lane = threadIdx.x;
    const uint64_t rem = c.n - b;
    const int32_t nw = rem < 2048 ? (int32_t)rem : 2048;  // rows of the window (>= 1)
    const int32_t r0 = 32 * (int32_t)lane;
    uint32_t word = 0, nrow = 0;
    if (r0 < nw) {
        nrow = (uint32_t)(nw - r0 < 32 ? nw - r0 : 32);
        word = nrow < 32 ? vword & ((1u << nrow) - 1u) : vword;
    }
    const uint32_t cnt = (uint32_t)__popc(word);
    const uint32_t incl = wave_scan(cnt);
    const uint32_t excl = incl - cnt;
    const int32_t e = r0 + (int32_t)nrow;
    int32_t s = (int32_t)((int64_t)st.s - (int64_t)b);  // open page start
    int32_t g = (int32_t)st.g;
    for (;;) {
        // cost of the open page if it ended with this lane's last row
        const bool cand = nrow && e > s;
        const int32_t cost = 4 * (g + (int32_t)incl) + ((e - 1 - s) >> 3) + 5;
        const uint64_t m = __ballot(cand && cost > (int32_t)kPageBytes);
        if (!m) break;
        // the first row that does not fit lies in word L: lanes 0..31 test one row each
        const uint32_t L = first_lane(m);
        const uint32_t wl = lane_u32(word, L), xl = lane_u32(excl, L);
        const int32_t x = 32 * (int32_t)L + (int32_t)lane;
        const uint32_t upto = lane < 31 ? (2u << lane) - 1u : 0xffffffffu;
        const int32_t cx = 4 * (g + (int32_t)(xl + __popc(wl & upto))) + ((x - s) >> 3) + 5;
        const uint64_t m2 = __ballot(lane < 32 && x >= s && x < nw && cx > (int32_t)kPageBytes);
        const uint32_t t = first_lane(m2);
        const int32_t xb = 32 * (int32_t)L + (int32_t)t;
        if (lane == 0) put_ref(c, st.np, (uint64_t)((int64_t)b + s), (uint32_t)(xb - s), 0);
        ++st.np;
        g = -(int32_t)(xl + __popc(wl & ((1u << t) - 1u)));
        s = xb;
    }
    st.s = (uint64_t)((int64_t)b + s);
    st.g = (int64_t)g + (int64_t)lane_u32(incl, 63);
}

__device__ inline uint32_t pick8(const uint32_t (&v)[8], uint32_t k) {
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) x = (uint32_t)j == k ? v[j] : x;
    return x;
}

// VARCHAR window of 512 rows: lane i holds rows b + 8i .. b + 8i + 7 (their validity byte vbyte
// and offsets o[0..8]). A short string adds 2 + len bytes, a NULL row only its bit; a string
// longer than 8185 bytes closes the open page and takes pages of its own. Per lane, p[k] = bytes
// of its rows 0..k; over the wave, the lanes' totals are scanned. A page that overflows in lane
// L is cut at the first of L's rows whose cost exceeds 8192 (lanes 0..7 test one row each).
// Row positions inside the window are 32-bit and relative to b (the open page may have started
// before the window: a negative start, at least -65536 since a page holds at most 65535 rows),
// and byte counts are 32-bit (a window holds at most 512 * 8187 bytes of short strings).
// below counts the refs this call emits whose first row lies below cut_lim (the segmented
// planner's refs of its own segment)
__device__ inline void varchar_window(const PageCol& c, PlanState& st, uint64_t b, uint32_t vbyte,
                                      const uint64_t (&o)[9], uint64_t cut_lim, uint32_t& below) {
    const uint32_t lane = threadIdx.x;
    const uint64_t rem = c.n - b;
    const int32_t nw = rem < 512 ? (int32_t)rem : 512;  // rows of the window (>= 1)
    const int32_t r0 = 8 * (int32_t)lane;
    uint32_t p[8], lens[8], lng = 0, acc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const bool valid = r0 + k < nw && ((vbyte >> k) & 1u);
        const uint64_t len = valid ? o[k + 1] - o[k] : 0;
        const bool lg = valid && len > kLongStr;
        acc += (valid && !lg) ? 2u + (uint32_t)len : 0u;
        p[k] = acc;
        lens[k] = (uint32_t)len;
        lng |= (lg ? 1u : 0u) << k;
    }
    const uint32_t ain = wave_scan(acc), aex = ain - acc;
    int32_t s = (int32_t)((int64_t)st.s - (int64_t)b);  // open page start
    int32_t g = (int32_t)st.g;
    for (;;) {
        // the first long string at or after s
        const uint32_t ge = s <= r0 ? 0xffu : s >= r0 + 8 ? 0u : (0xffu << (uint32_t)(s - r0)) & 0xffu;
        const uint32_t lm = lng & ge;
        const uint64_t lb = __ballot(lm != 0);
        int32_t lim = 512;
        uint32_t LL = 0, kl = 0;
        if (lb) {
            LL = first_lane(lb);
            kl = (uint32_t)__builtin_ctz(lane_u32(lm, LL));
            lim = 8 * (int32_t)LL + (int32_t)kl;
        }
        // the open page's cost if it ended with this lane's last row before lim
        int32_t e = r0 + 7;
        e = e < lim - 1 ? e : lim - 1;
        e = e < nw - 1 ? e : nw - 1;
        const int32_t lo = s > r0 ? s : r0;
        const bool cand = r0 < lim && r0 < nw && e >= lo;
        const int32_t cost = 5 + g + (int32_t)aex + (int32_t)pick8(p, (uint32_t)(e - r0) & 7u) + ((e - s) >> 3);
        const uint64_t m = __ballot(cand && cost > (int32_t)kPageBytes);
        if (m) {
            const uint32_t L = first_lane(m);
            const int32_t rl = 8 * (int32_t)L, eL = (int32_t)lane_u32((uint32_t)e, L);
            const int32_t aexL = (int32_t)lane_u32(aex, L);
            uint32_t pj = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t v = lane_u32(p[j], L);
                pj = lane == (uint32_t)j ? v : pj;
            }
            const int32_t row = rl + (int32_t)lane;
            const int32_t cj = 5 + g + aexL + (int32_t)pj + ((row - s) >> 3);
            const uint64_t m2 = __ballot(lane < 8 && row >= s && row <= eL && cj > (int32_t)kPageBytes);
            const uint32_t t = first_lane(m2);
            const int32_t xb = rl + (int32_t)t;
            if (lane == 0) put_ref(c, st.np, (uint64_t)((int64_t)b + s), (uint32_t)(xb - s), 0);
            ++st.np;
            below += (uint64_t)((int64_t)b + s) < cut_lim ? 1u : 0u;
            // bytes of L's rows before row xb: p[t - 1] of lane L, which lane t - 1 holds in pj
            const uint32_t before = t ? lane_u32(pj, t - 1) : 0u;
            g = -(aexL + (int32_t)before);
            s = xb;
            continue;
        }
        if (!lb) break;
        // a long string at row lim: close the open page, then one page per piece of 8188 bytes
        if (lim > s) {
            if (lane == 0) put_ref(c, st.np, (uint64_t)((int64_t)b + s), (uint32_t)(lim - s), 0);
            ++st.np;
            below += (uint64_t)((int64_t)b + s) < cut_lim ? 1u : 0u;
        }
        const uint64_t ll = lane_u32(pick8(lens, kl), LL);
        const uint64_t pieces = (ll + kLongPiece - 1) / kLongPiece;
        const uint64_t rowl = b + (uint64_t)lim;
        for (uint64_t p0 = 0; p0 < pieces; p0 += 64) {
            const uint64_t q = p0 + lane;
            if (q < pieces) put_ref(c, st.np + q, rowl, (uint32_t)q, q == 0 ? 1u : 2u);
        }
        st.np += pieces;
        below += rowl < cut_lim ? (uint32_t)pieces : 0u;
        g = -(int32_t)(lane_u32(aex, LL) + lane_u32(pick8(p, kl), LL));
        s = lim + 1;
    }
    st.s = (uint64_t)((int64_t)b + s);
    st.g = (int64_t)g + (int64_t)lane_u32(ain, 63);
}

// The cut itself is a short chain of wave operations per window, so the loads run a group of
// windows ahead: the next group's data is requested before the current group is cut.
constexpr int kIntGroup = 16;  // INT32 windows per group (32768 rows)

__device__ uint64_t plan_int32(const PageCol& c) {
    const uint32_t lane = threadIdx.x;
    const uint64_t n = c.n;
    const HJ_GLOBAL uint32_t* vw = (const HJ_GLOBAL uint32_t*)gp(c.validity);
    constexpr uint64_t kRows = 2048ull * kIntGroup;
    uint32_t cur[kIntGroup], nxt[kIntGroup];
#pragma unroll
    for (int j = 0; j < kIntGroup; ++j) {
        const uint64_t r0 = 2048ull * j + 32ull * lane;
        nxt[j] = r0 < n ? vw[r0 >> 5] : 0u;
    }
    PlanState st;
    for (uint64_t B = 0; B < n; B += kRows) {
#pragma unroll
        for (int j = 0; j < kIntGroup; ++j) cur[j] = nxt[j];
#pragma unroll
        for (int j = 0; j < kIntGroup; ++j) {
            const uint64_t r0 = B + kRows + 2048ull * j + 32ull * lane;
            if (r0 < n) nxt[j] = vw[r0 >> 5];
        }
#pragma unroll
        for (int j = 0; j < kIntGroup; ++j)
            if (B + 2048ull * j < n) int32_window(c, st, B + 2048ull * j, cur[j]);
    }
    if (n > st.s) {
        if (lane == 0) put_ref(c, st.np, st.s, (uint32_t)(n - st.s), 0);
        ++st.np;
    }
    return st.np;
}

__device__ uint64_t plan_varchar(const PageCol& c) {
    const uint32_t lane = threadIdx.x;
    const uint64_t n = c.n;
    const HJ_GLOBAL uint8_t* vb = gp(c.validity);
    const HJ_GLOBAL uint64_t* off = gp(c.offsets);
    // window data one window ahead: validity byte and offsets of the lane's 8 rows
    uint32_t nb = 0;
    uint64_t no[9];
    auto load = [&](uint64_t b) {
        const uint64_t r0 = b + 8ull * lane;
        nb = r0 < n ? vb[r0 >> 3] : 0u;
#pragma unroll
        for (int k = 0; k < 9; ++k) no[k] = r0 + k <= n ? off[r0 + k] : 0ull;
    };
    load(0);
    PlanState st;
    for (uint64_t b = 0; b < n; b += 512) {
        const uint32_t cb = nb;
        uint64_t co[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) co[k] = no[k];
        if (b + 512 < n) load(b + 512);
        uint32_t below = 0;
        varchar_window(c, st, b, cb, co, 0, below);
    }
    if (n > st.s) {
        if (lane == 0) put_ref(c, st.np, st.s, (uint32_t)(n - st.s), 0);
        ++st.np;
    }
    return st.np;
}

// valid rows of each INT32 column (blockIdx.y), added into b.valid_rows (zeroed by the caller)
__global__ __launch_bounds__(256) void k_page_valid(PageBatch b) {
    __shared__ uint64_t red[4];
    const PageCol c = b.c[blockIdx.y];
    if (c.type != HJ_INT32) return;
    const HJ_GLOBAL uint32_t* vw = (const HJ_GLOBAL uint32_t*)gp(c.validity);
    const uint64_t words = (c.n + 31) / 32;
    uint64_t cnt = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = vw[w];
        const uint64_t r0 = 32 * w;
        if (c.n - r0 < 32) v &= (1u << (c.n - r0)) - 1u;
        cnt += (uint64_t)__popc(v);
    }
    cnt = block_sum<256>(cnt, red);
    if (threadIdx.x == 0 && cnt) atomicAdd((unsigned long long*)(b.valid_rows + blockIdx.y),
                                           (unsigned long long)cnt);
}

__global__ __launch_bounds__(64) void k_page_plan(PageBatch b) {
    const PageCol c = b.c[blockIdx.x];
    if (c.jump) return;  // planned by jumps (k_pj_*)
    uint64_t np;
    if (c.type == HJ_INT32 && gp(b.valid_rows)[blockIdx.x] == c.n) {
        // no NULL: every page but the last holds exactly 1984 rows
        np = (c.n + 1983) / 1984;
        for (uint64_t p = threadIdx.x; p < np; p += 64) {
            const uint64_t r0 = 1984 * p;
            put_ref(c, p, r0, (uint32_t)(c.n - r0 < 1984 ? c.n - r0 : 1984), 0);
        }
    } else {
        np = c.type == HJ_INT32 ? plan_int32(c) : plan_varchar(c);
    }
    if (threadIdx.x == 0) gp(b.counts)[blockIdx.x] = np;
}

// ---------------------------------------------------------------- VARCHAR planning by jumps
// The pages of a long VARCHAR column form a chain: where a page starts decides where it ends,
// and the next page starts there. One wave walks that chain window by window (plan_varchar);
// for a long column the chain is found in parallel instead, with the same cuts:
//   k_pj_cost  per row its page bytes (2 + length for a string of at most 8185 bytes, 0 for a
//              NULL row, kLongCost for a longer string, which no page of rows can hold), summed
//              within blocks of 4096 rows; k_pj_scan adds the blocks' totals up
//   k_pj_next  next(i) = the first row of the page after a page starting at row i: the first j
//              whose row does not fit, 5 + bytes(i .. j) + (j - i) / 8 > 8192 (a galloping then
//              binary search over the prefix sums); a long string at i is a page chain node of
//              its own, next(i) = i + 1
//   k_pj_sq    ten squarings: next^32 (D) and next^1024 (E), every row at once
//   k_pj_walk  one lane follows E from row 0 (a step per 1024 pages), k_pj_mid expands each of
//              those nodes by D into 32 nodes, k_pj_fine each of these by next into 32 pages
//   k_pj_refs  the pages in chain order as refs (a long string's pieces counted by a scan)
constexpr uint32_t kJBlk = 4096;          // rows per cost block
constexpr uint32_t kLongCost = 1u << 18;  // page bytes of a long string: more than a page

__device__ inline bool pj_valid(const PageCol& c, uint64_t r) {
    return (gp(c.validity)[r >> 3] >> (r & 7)) & 1u;
}
__device__ inline uint32_t pj_cost(const PageCol& c, uint64_t r) {
    if (!pj_valid(c, r)) return 0u;
    const uint64_t len = gp(c.offsets)[r + 1] - gp(c.offsets)[r];
    return len > kLongStr ? kLongCost : 2u + (uint32_t)len;
}
// bytes of rows [0, j)
__device__ inline uint64_t pj_prefix(const PageCol& c, uint64_t j) {
    return j ? gp(c.jbp)[(j - 1) / kJBlk] + gp(c.jlp)[j - 1] : 0ull;
}

__global__ __launch_bounds__(256) void k_pj_cost(PageBatch b) {
    __shared__ uint32_t lds[4];
    const PageCol& c = b.c[blockIdx.y];
    if (!c.jump) return;
    const uint64_t n = c.n;
    const uint64_t r0 = (uint64_t)blockIdx.x * kJBlk;
    if (r0 >= n) return;  // uniform for the block
    constexpr int K = (int)(kJBlk / 256);
    const uint64_t rt = r0 + (uint64_t)threadIdx.x * K;
    uint32_t v[K], acc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc += rt + k < n ? pj_cost(c, rt + k) : 0u;
        v[k] = acc;
    }
    uint32_t tot = 0;
    const uint32_t ex = block_excl_scan<256>(acc, lds, &tot);
    HJ_GLOBAL uint32_t* lp = gp(c.jlp);
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (rt + k < n) lp[rt + k] = ex + v[k];
    if (threadIdx.x == 0) gp(c.jbp)[blockIdx.x] = tot;  // the block's total, scanned next
}

// exclusive scan of the blocks' totals, one block per column
__global__ __launch_bounds__(1024) void k_pj_scan(PageBatch b) {
    __shared__ uint64_t lds[16];
    const PageCol& c = b.c[blockIdx.x];
    if (!c.jump) return;
    const uint64_t nb = (c.n + kJBlk - 1) / kJBlk;
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t i0 = (uint64_t)threadIdx.x * per;
    HJ_GLOBAL uint64_t* bp = gp(c.jbp);
    uint64_t acc = 0;
    for (uint64_t i = i0; i < i0 + per && i < nb; ++i) acc += bp[i];
    uint64_t tot = 0;
    uint64_t run = block_excl_scan<1024>(acc, lds, &tot);
    __syncthreads();  // every thread has read its totals before any is overwritten
    for (uint64_t i = i0; i < i0 + per && i < nb; ++i) {
        const uint64_t x = bp[i];
        bp[i] = run;
        run += x;
    }
}

__global__ __launch_bounds__(256) void k_pj_next(PageBatch b) {
    const PageCol& c = b.c[blockIdx.y];
    if (!c.jump) return;
    const uint64_t n = c.n;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    uint32_t nx = (uint32_t)n;
    if (i < n) {
        const uint64_t pi = pj_prefix(c, i);
        // does row j still fit in a page that starts at row i?
        auto fits = [&](uint64_t j) {
            return 5ull + (pj_prefix(c, j + 1) - pi) + ((j - i) >> 3) <= (uint64_t)kPageBytes;
        };
        if (!fits(i)) {
            nx = (uint32_t)(i + 1);  // a long string: its own pieces
        } else {
            uint64_t lo = i, hi = n;  // row lo fits; rows from hi on are past the page or the end
            for (uint64_t step = 64;; step <<= 1) {
                const uint64_t j = i + step;
                if (j >= hi) break;
                if (fits(j)) {
                    lo = j;
                } else {
                    hi = j;
                    break;
                }
            }
            while (hi - lo > 1) {
                const uint64_t mid = lo + (hi - lo) / 2;
                if (fits(mid))
                    lo = mid;
                else
                    hi = mid;
            }
            nx = (uint32_t)hi;
        }
    }
    gp(c.jnx)[i] = nx;
}

// squaring s of the jump tables: dst = src o src over rows [0, n] (src[n] = n)
__global__ __launch_bounds__(256) void k_pj_sq(PageBatch b, int step) {
    const PageCol& c = b.c[blockIdx.y];
    if (!c.jump) return;
    const uint64_t n = c.n;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    // steps 0-4: next -> next^32 in ja; steps 5-9: ja -> next^1024 in jb (jlp reused)
    const uint32_t* src;
    uint32_t* dst;
    switch (step) {
        case 0: src = c.jnx; dst = c.ja; break;
        case 1: src = c.ja; dst = c.jb; break;
        case 2: src = c.jb; dst = c.ja; break;
        case 3: src = c.ja; dst = c.jb; break;
        case 4: src = c.jb; dst = c.ja; break;
        case 5: src = c.ja; dst = c.jb; break;
        case 6: src = c.jb; dst = c.jlp; break;
        case 7: src = c.jlp; dst = c.jb; break;
        case 8: src = c.jb; dst = c.jlp; break;
        default: src = c.jlp; dst = c.jb; break;
    }
    const uint32_t x = gp(src)[i];
    gp(dst)[i] = gp(src)[x];
}

// the chain's nodes at every 1024th page (one lane per column)
__global__ __launch_bounds__(64) void k_pj_walk(PageBatch b) {
    const PageCol& c = b.c[blockIdx.x];
    if (!c.jump || threadIdx.x) return;
    const uint32_t n = (uint32_t)c.n;
    const HJ_GLOBAL uint32_t* E = gp(c.jb);
    HJ_GLOBAL uint32_t* C = gp(c.jc);
    uint32_t t = 0, x = 0;
    while (x < n) {
        C[t++] = x;
        x = E[x];
    }
    C[t] = n;
    gp(c.jcnt)[0] = t;
}

// each of those nodes expanded by next^32 into 32 nodes, then each of these by next into 32
__global__ __launch_bounds__(64) void k_pj_mid(PageBatch b) {
    const PageCol& c = b.c[blockIdx.y];
    if (!c.jump) return;
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= gp(c.jcnt)[0]) return;
    const uint32_t n = (uint32_t)c.n;
    const HJ_GLOBAL uint32_t* D = gp(c.ja);
    HJ_GLOBAL uint32_t* M = gp(c.jm) + (uint64_t)t * 32;
    uint32_t x = gp(c.jc)[t];
    for (int k = 0; k < 32; ++k) {
        M[k] = x;
        x = x < n ? D[x] : n;
    }
}

__global__ __launch_bounds__(256) void k_pj_fine(PageBatch b) {
    const PageCol& c = b.c[blockIdx.y];
    if (!c.jump) return;
    const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= (uint64_t)gp(c.jcnt)[0] * 32) return;
    const uint32_t n = (uint32_t)c.n;
    const HJ_GLOBAL uint32_t* X = gp(c.jnx);
    HJ_GLOBAL uint32_t* P = gp(c.jnode) + q * 32;
    uint32_t x = gp(c.jm)[q];
    for (int u = 0; u < 32; ++u) {
        P[u] = x;
        x = x < n ? X[x] : n;
    }
}

// the chain's pages in order as refs: one per node, a long string's pieces more (a scan of the
// nodes' ref counts), one block per column
__global__ __launch_bounds__(1024) void k_pj_refs(PageBatch b) {
    __shared__ uint64_t lds[16];
    const PageCol& c = b.c[blockIdx.x];
    if (!c.jump) return;
    const uint32_t n = (uint32_t)c.n;
    const uint64_t total_nodes = (uint64_t)gp(c.jcnt)[0] * 1024;
    const HJ_GLOBAL uint32_t* P = gp(c.jnode);
    const HJ_GLOBAL uint32_t* X = gp(c.jnx);
    auto refs_of = [&](uint32_t x, uint64_t* len) -> uint64_t {
        *len = 0;
        if (x >= n) return 0;
        if (pj_valid(c, x)) {
            const uint64_t l = gp(c.offsets)[(uint64_t)x + 1] - gp(c.offsets)[x];
            if (l > kLongStr) {
                *len = l;
                return (l + kLongPiece - 1) / kLongPiece;
            }
        }
        return 1;
    };
    uint64_t base = 0;
    for (uint64_t p0 = 0; p0 < total_nodes; p0 += 1024) {
        const uint64_t p = p0 + threadIdx.x;
        const uint32_t x = p < total_nodes ? P[p] : n;
        uint64_t len = 0;
        const uint64_t k = refs_of(x, &len);
        uint64_t tot = 0;
        const uint64_t ex = block_excl_scan<1024>(k, lds, &tot);
        const uint64_t at = base + ex;
        if (k == 1 && !len) {
            put_ref(c, at, x, X[x] - x, 0);
        } else if (k) {
            for (uint64_t q = 0; q < k; ++q) put_ref(c, at + q, x, (uint32_t)q, q == 0 ? 1u : 2u);
        }
        base += tot;
        if (!tot) break;  // past the chain's last node (uniform)
    }
    if (threadIdx.x == 0) gp(b.counts)[blockIdx.x] = base;
}

// 8 validity bits starting at bit q (bit i of byte i/8); nbytes = bytes holding the n rows
__device__ inline uint32_t valid8(const HJ_GLOBAL uint8_t* vb, uint64_t nbytes, uint64_t q) {
    const uint64_t k = q >> 3;
    const uint32_t sh = (uint32_t)(q & 7);
    uint32_t v = vb[k];
    if (sh && k + 1 < nbytes) v |= (uint32_t)vb[k + 1] << 8;
    return (v >> sh) & 0xffu;
}

// One block per page. Thread t handles rows f0 + 8t .. f0 + 8t + 7 of the page per 2048-row
// step: their validity is bitmap byte (f0 / 8 + t) of the page.
__global__ __launch_bounds__(kWriteBlock) void k_page_write(PageBatch b, uint64_t p0) {
    __shared__ uint32_t pg[kPageBytes / 4];
    __shared__ uint64_t s_src[kPageBytes / 2];  // VARCHAR: source offset of each string
    __shared__ uint64_t s_red[kWriteBlock / 64];
    uint8_t* pgb = (uint8_t*)pg;
    const uint32_t t = threadIdx.x;
    const uint64_t pid = p0 + blockIdx.x;
    uint32_t ci = 0;
    for (uint32_t i = 1; i < b.ncol; ++i) ci = pid >= b.c[i].first ? i : ci;
    const PageCol& c = b.c[ci];
    const HJ_GLOBAL PageRef* rp = gp(c.refs) + (pid - c.first);
    PageRef ref;
    ref.row0 = rp->row0;
    ref.rows = rp->rows;
    ref.kind = rp->kind;
    for (uint32_t i = t; i < kPageBytes / 4; i += kWriteBlock) pg[i] = 0;
    __syncthreads();
    const HJ_GLOBAL uint8_t* vb = gp(c.validity);
    const uint64_t nbytes = (c.n + 7) >> 3;
    if (ref.kind) {
        // piece ref.rows of the long string in row ref.row0
        const HJ_GLOBAL uint64_t* off = gp(c.offsets);
        const uint64_t st = off[ref.row0], len = off[ref.row0 + 1] - st;
        const uint64_t po = (uint64_t)ref.rows * kLongPiece;
        const uint32_t l = (uint32_t)(len - po < kLongPiece ? len - po : kLongPiece);
        if (t == 0) pg[0] = (ref.kind == 1 ? 0xffffu : 0xfffeu) | (l << 16);
        const HJ_GLOBAL char* src = gp(c.chars) + st + po;
        for (uint32_t x = t; x < l; x += kWriteBlock) pgb[4 + x] = (uint8_t)src[x];
    } else if (c.type == HJ_INT32) {
        const uint64_t r0 = ref.row0;
        const uint32_t R = ref.rows, nb = (R + 7) / 8;
        const HJ_GLOBAL int32_t* vals = gp(c.values);
        uint32_t nv = 0;
        for (uint32_t f0 = 0; f0 < R; f0 += 8 * kWriteBlock) {
            const uint32_t f = f0 + 8 * t;
            uint32_t byte = 0;
            if (f < R) {
                byte = valid8(vb, nbytes, r0 + f);
                if (R - f < 8) byte &= (1u << (R - f)) - 1u;
                pgb[kPageBytes - nb + (f >> 3)] = (uint8_t)byte;
            }
            uint64_t tot;
            uint32_t k = nv + (uint32_t)block_excl_scan<kWriteBlock>((uint64_t)__popc(byte), s_red, &tot);
            int32_t v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = ((byte >> i) & 1) ? vals[r0 + f + i] : 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if ((byte >> i) & 1) pg[1 + k++] = (uint32_t)v[i];
            nv += (uint32_t)tot;
        }
        if (t == 0) pg[0] = R | (nv << 16);
    } else {
        const uint64_t r0 = ref.row0;
        const uint32_t R = ref.rows, nb = (R + 7) / 8;
        const HJ_GLOBAL uint64_t* off = gp(c.offsets);
        // strings on the page first: the chars start after their end offsets
        uint32_t ns = 0;
        for (uint32_t f0 = 0; f0 < R; f0 += 8 * kWriteBlock) {
            const uint32_t f = f0 + 8 * t;
            uint32_t byte = 0;
            if (f < R) {
                byte = valid8(vb, nbytes, r0 + f);
                if (R - f < 8) byte &= (1u << (R - f)) - 1u;
                pgb[kPageBytes - nb + (f >> 3)] = (uint8_t)byte;
            }
            ns += (uint32_t)block_sum<kWriteBlock>((uint64_t)__popc(byte), s_red);
        }
        // end offsets (u16 at byte 4 + 2i) and source offsets of the strings, in row order
        uint32_t si = 0, cp = 0;
        for (uint32_t f0 = 0; f0 < R; f0 += 8 * kWriteBlock) {
            const uint32_t f = f0 + 8 * t;
            uint32_t byte = 0;
            uint64_t st[8];
            uint32_t ln[8];
            uint32_t sum = 0;
            if (f < R) {
                byte = valid8(vb, nbytes, r0 + f);
                if (R - f < 8) byte &= (1u << (R - f)) - 1u;
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                st[i] = 0;
                ln[i] = 0;
                if ((byte >> i) & 1) {
                    st[i] = off[r0 + f + i];
                    ln[i] = (uint32_t)(off[r0 + f + i + 1] - st[i]);
                }
                sum += ln[i];
            }
            uint64_t tot;
            const uint64_t ex = block_excl_scan<kWriteBlock>(
                ((uint64_t)sum << 32) | (uint64_t)__popc(byte), s_red, &tot);
            uint32_t k = si + (uint32_t)ex, e = cp + (uint32_t)(ex >> 32);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if ((byte >> i) & 1) {
                    e += ln[i];
                    *(uint16_t*)(pgb + 4 + 2 * k) = (uint16_t)e;
                    s_src[k] = st[i];
                    ++k;
                }
            }
            si += (uint32_t)tot;
            cp += (uint32_t)(tot >> 32);
        }
        if (t == 0) pg[0] = R | (ns << 16);
        __syncthreads();
        const uint16_t* ends = (const uint16_t*)(pgb + 4);
        const HJ_GLOBAL char* chars = gp(c.chars);
        uint8_t* dst = pgb + 4 + 2 * ns;
        for (uint32_t x = t; x < cp; x += kWriteBlock) {
            // string holding page char x: the first whose end offset exceeds x
            uint32_t lo = 0, hi = ns - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ends[mid] > x) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t s0 = lo ? ends[lo - 1] : 0u;
            dst[x] = (uint8_t)chars[s_src[lo] + (x - s0)];
        }
    }
    __syncthreads();
    HJ_GLOBAL u32x4* out = (HJ_GLOBAL u32x4*)gp(b.pages + (uint64_t)blockIdx.x * kPageBytes);
    const u32x4* src = (const u32x4*)pg;
    for (uint32_t i = t; i < kPageBytes / 16; i += kWriteBlock) out[i] = src[i];
}

}  // namespace

void page_plan(const PageBatch& b, uint64_t max_rows, hipStream_t s) {
    if (!b.ncol) return;
    const uint64_t words = (max_rows + 31) / 32;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(256, std::max<uint64_t>(1, words / 2048));
    hipLaunchKernelGGL(k_page_valid, dim3(blocks, b.ncol), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_page_plan, dim3(b.ncol), dim3(64), 0, s, b);
    if (b.jump_rows) {  // VARCHAR columns planned by jumps; jump_rows = the longest's rows
        const uint64_t jn = b.jump_rows;
        const dim3 g1((uint32_t)((jn + kJBlk - 1) / kJBlk), b.ncol);
        const dim3 gr((uint32_t)((jn + 1 + 255) / 256), b.ncol);
        const uint64_t coarse = (jn + 1) / 1024 + 2;  // chain nodes at every 1024th page
        hipLaunchKernelGGL(k_pj_cost, g1, dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_pj_scan, dim3(b.ncol), dim3(1024), 0, s, b);
        hipLaunchKernelGGL(k_pj_next, gr, dim3(256), 0, s, b);
        for (int step = 0; step < 10; ++step)
            hipLaunchKernelGGL(k_pj_sq, gr, dim3(256), 0, s, b, step);
        hipLaunchKernelGGL(k_pj_walk, dim3(b.ncol), dim3(64), 0, s, b);
        hipLaunchKernelGGL(k_pj_mid, dim3((uint32_t)((coarse + 63) / 64), b.ncol), dim3(64), 0, s, b);
        hipLaunchKernelGGL(k_pj_fine, dim3((uint32_t)((coarse * 32 + 255) / 256), b.ncol), dim3(256),
                           0, s, b);
        hipLaunchKernelGGL(k_pj_refs, dim3(b.ncol), dim3(1024), 0, s, b);
    }
}

void page_write(const PageBatch& b, uint64_t p0, uint64_t count, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_page_write, dim3((uint32_t)count), dim3(kWriteBlock), 0, s, b, p0);
}

}  // namespace hj
