// pipeline.hip - builds and the fused probe of the join pipeline (gfx950, wave64).
//
// Build, all build sides in the same three launches:
//   filter  rows passing the scan's filters with a non-NULL key -> staged (key, row) at the
//           start of the block's row range                          reads key + filter columns
//   clear   the job's tuple count, then its 2^k key slots (2^k >= 2n) and Bloom words
//   insert  a 32-bit compare-and-swap of the key into a slot of its home bucket (then the other
//           slots, then the next buckets), the row next to it; a key found twice flags the
//           build (the caller then falls back to the general join path)
// Probe: one fused pass over the source rows (a wave takes U consecutive 64-row groups, one row
// per lane). The keys of every streamed join are loaded first (coalesced, row-space mirror),
// then each join in turn: validity, the LDS Bloom filter of the first (smallest) build, and
// for the lanes still alive a 16-byte load of the two entries at the key's home slot, which
// settles nearly every probe. Survivors are appended through a wave-aggregated cursor.
// Joins keyed by a column of a build table run afterwards over the survivor list, marking
// failures, and an ordered compaction writes the output row ids.
#include "pipeline.hpp"
#include "gindex.hpp"
#include "wavecol.hpp"

#include <hip/hip_ext.h>

#include <algorithm>

namespace hj {

namespace {

constexpr int kB = 256;
constexpr uint32_t kNW = kB / HJ_WAVE;
static_assert(kGroupIndexBlock == (uint32_t)kB, "index blocks of the filter launch: one group per thread");
static_assert(kProbeIdxCols <= (int)kNW, "the fused probe's page tables: one wave per column");
static_assert(kPageTable == HJ_WAVE, "a page table: one entry per lane");

__device__ inline uint32_t log2ceil(uint32_t n) { return n <= 1 ? 0u : 32u - __clz(n - 1u); }

// Blocked Bloom filter of a build side, held in LDS by the probe: one 32-bit word per key with
// 4 bits set, about 16 bits per key; only used when it gets >= 8 bits per key within
// kBloomWords words. Word and bits come from the high bits of two more multiplicative hashes of
// the key's Fibonacci hash (high bits of a product depend on every bit of the operand).
__device__ inline bool bf_enabled(uint32_t n) { return n && (uint64_t)n * 8u <= kBloomWords * 32u; }
__device__ inline uint32_t bf_log2w(uint32_t n) {
    uint32_t b = log2ceil((n + 1u) / 2u);  // n/2 words = 16 bits per key
    return b < 5u ? 5u : (b > 12u ? 12u : b);
}
__device__ inline uint32_t bf_word(uint32_t h, uint32_t lw) {
    return ((h ^ (h >> 15)) * 0x2C1B3C6Du) >> (32u - lw);
}
__device__ inline uint32_t bf_mask(uint32_t h) {
    const uint32_t g = (h ^ (h >> 13)) * 0x297A2D39u;
    return (1u << (g >> 27)) | (1u << ((g >> 22) & 31u)) | (1u << ((g >> 17) & 31u)) |
           (1u << ((g >> 12) & 31u));
}

template <int NW>
__device__ inline uint32_t bsum_u32(uint32_t v, uint32_t* lds) {
    uint32_t s = wave_sum(v);
    if (lane_id() == 0) lds[threadIdx.x / HJ_WAVE] = s;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) t += lds[i];
    __syncthreads();
    return t;
}

__device__ inline uint32_t ballot_prefix(bool pred, uint32_t* lds, uint32_t* total) {
    const uint32_t w = threadIdx.x / HJ_WAVE;
    const uint64_t b = __ballot(pred);
    if (lane_id() == 0) lds[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < (int)kNW; ++i) {
        off += i < (int)w ? lds[i] : 0u;
        tot += lds[i];
    }
    __syncthreads();
    *total = tot;
    return off + (uint32_t)__popcll(b & lanemask_lt());
}


// ---------------------------------------------------------------- build
// The filter and insert passes run one 1-D grid over all jobs: job j owns blocks
// [blk0, blk0 + nblk), block x of a job takes its rows [x, x + 1) * kBuildRows and stages the
// tuples that pass at the start of that range.
constexpr int kFU = kBuildRows / (kB * 2);  // 64-row groups per wave and iteration (2 rounds)
static_assert(kFU * kB * 2 == (int)kBuildRows, "build rows per block");
constexpr int kIU = 4;                       // tuples per thread and round of the insert pass

// job of block b: the number of jobs after the first whose blocks start at or before b (the
// first blocks of all jobs are one contiguous array, unused entries ~0)
__device__ inline int build_job_of(const BuildBatch* bb, uint32_t b) {
    int j = 0;
#pragma unroll
    for (int i = 1; i < kMaxStages; ++i) j += b >= bb->blk0[i] ? 1 : 0;
    return j;
}

// filter pass: rows with a non-NULL key that pass the scan's range filters -> (key, row)
__global__ __launch_bounds__(kB) void k_sb_filter(const BuildBatch* __restrict__ bb) {
    if (blockIdx.x >= bb->grid) {  // the probe side's group index
        group_index_thread(bb->gib, blockIdx.x - bb->grid, threadIdx.x);
        return;
    }
    __shared__ uint32_t s_cnt;
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    const BuildJob& jb = bb->job[build_job_of(bb, blockIdx.x)];
    const uint32_t x = blockIdx.x - jb.blk0;
    const uint32_t rows = jb.rows;
    const uint32_t r0 = x * kBuildRows;
    const uint32_t r1 = r0 + kBuildRows < rows ? r0 + kBuildRows : rows;
    if (t == 0) s_cnt = 0;
    __syncthreads();
    const IdxCol wk = idx_col(jb.key, rows);
    const int nf = jb.nf;
    // lane u < kFU reads the record of group u of the wave's rows
    const uint32_t lu = (lane < (uint32_t)kFU ? lane : (uint32_t)kFU - 1u) * 4u;
    for (uint32_t base = r0; base < r1; base += kB * kFU) {
        // wave w takes groups base/64 + 4u + w, u < kFU
        const uint32_t grp0 = base / 64u + w;
        const u32x4 kr = idx_rec_any(wk, grp0 + lu);
        int32_t key[kFU];
        uint64_t live[kFU];
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            const uint64_t vw = rec_valid(kr, u);
            key[u] = idx_load(wk, kr, u, vw, lane);
            const uint32_t g0 = base + u * 256u + w * 64u;
            live[u] = (g0 < r1 ? range_mask(r1 - g0) : 0ull) & vw;
        }
        for (int f = 0; f < nf; ++f) {
            const DCol& fc = jb.fcol[f];
            const int32_t lo = jb.flo[f], hi = jb.fhi[f];
            int32_t v[kFU];
            uint64_t fvw[kFU];
            if (fc.flat) {
                const FlatCol ff = flat_col(fc, rows);
                const u32x2 vv = __builtin_amdgcn_raw_buffer_load_b64(ff.vb, (grp0 + lu) * 8u, 0, 0);
#pragma unroll
                for (int u = 0; u < kFU; ++u) {
                    v[u] = __builtin_amdgcn_raw_buffer_load_b32(
                        ff.val, ((grp0 + 4u * u) * 64u + lane) * 4u, 0, 0);
                    fvw[u] = lane_word(vv.x | ff.all_valid, vv.y | ff.all_valid, u);
                }
            } else {
                const IdxCol wf = idx_col(fc, rows);
                const u32x4 fr = idx_rec_any(wf, grp0 + lu);
#pragma unroll
                for (int u = 0; u < kFU; ++u) {
                    fvw[u] = rec_valid(fr, u);
                    v[u] = idx_load(wf, fr, u, fvw[u], lane);
                }
            }
#pragma unroll
            for (int u = 0; u < kFU; ++u) live[u] &= __ballot(v[u] >= lo && v[u] <= hi) & fvw[u];
        }
        uint32_t pos[kFU];
        block_append_masks<kFU>(&s_cnt, live, pos);
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            if (in_mask(live[u])) {
                gp(jb.st_key)[r0 + pos[u]] = (uint32_t)key[u];
                gp(jb.st_row)[r0 + pos[u]] = base + u * 256u + w * 64u + lane;
            }
        }
    }
    __syncthreads();
    if (t == 0) gp(jb.blk)[x] = s_cnt;
}

// ---------------------------------------------------------------- partitioned build
// Filter pass of the partitioned build: as k_sb_filter, and the block's tuples are staged in
// order of their hash partition (top pbits bits of the Fibonacci hash), counted with LDS
// atomics; poff records where each partition starts in the block's range. The first block of a
// job also clears the job's small Bloom filter (set by the segment pass).
constexpr int kFR = (int)(kBuildRows / (kB * kFU));  // rounds per block
static_assert(kFR * kB * kFU == (int)kBuildRows, "build rows per block");

__global__ __launch_bounds__(kB) void k_sb_filter_part(const BuildBatch* __restrict__ bb) {
    if (blockIdx.x >= bb->grid) {  // the probe side's group index
        group_index_thread(bb->gib, blockIdx.x - bb->grid, threadIdx.x);
        return;
    }
    __shared__ uint32_t hist[(1u << kPartBits) + 1];
    __shared__ uint32_t lds[kNW];
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    const int job = build_job_of(bb, blockIdx.x);
    const BuildJob& jb = bb->job[job];
    const uint32_t x = blockIdx.x - jb.blk0;
    const uint32_t rows = jb.rows;
    const uint32_t r0 = x * kBuildRows;
    const uint32_t r1 = r0 + kBuildRows < rows ? r0 + kBuildRows : rows;
    const uint32_t fb = jb.pbits, P = 1u << fb;
    if (x == 0)
        for (uint32_t i = t; i < kBloomWords; i += kB) gp(jb.bf)[i] = 0;
    for (uint32_t i = t; i <= P; i += kB) hist[i] = 0;
    __syncthreads();
    const IdxCol wk = idx_col(jb.key, rows);
    const int nf = jb.nf;
    // the first filter streams with the keys when it has a group index
    const int f0 = nf && !jb.fcol[0].flat ? 1 : 0;
    const uint32_t lu = (lane < (uint32_t)kFU ? lane : (uint32_t)kFU - 1u) * 4u;
    int32_t key[kFR][kFU];
    uint32_t rank[kFR][kFU];
    uint32_t alive = 0;  // bit r * kFU + u: this lane's row of round r, group u passes
    // every round's group records at once (key column and first filter column), then every
    // round's keys and filter values, then the other filters' columns
    uint64_t live[kFR][kFU];
    {
        const IdxCol wf = idx_col(f0 ? jb.fcol[0] : jb.key, rows);
        const int32_t lo = f0 ? jb.flo[0] : INT32_MIN, hi = f0 ? jb.fhi[0] : INT32_MAX;
        u32x4 kr[kFR], fr[kFR];
#pragma unroll
        for (int rd = 0; rd < kFR; ++rd) {
            const uint32_t grp0 = (r0 + rd * (kB * kFU)) / 64u + w;
            kr[rd] = idx_rec_any(wk, grp0 + lu);
            fr[rd] = idx_rec_any(wf, grp0 + lu);
        }
        int32_t v[kFR][kFU];
        uint64_t kvw[kFR][kFU], fvw[kFR][kFU];
#pragma unroll
        for (int rd = 0; rd < kFR; ++rd)
#pragma unroll
            for (int u = 0; u < kFU; ++u) {
                kvw[rd][u] = rec_valid(kr[rd], u);
                fvw[rd][u] = rec_valid(fr[rd], u);
                key[rd][u] = idx_load(wk, kr[rd], u, kvw[rd][u], lane);
                v[rd][u] = idx_load(wf, fr[rd], u, fvw[rd][u], lane);
            }
#pragma unroll
        for (int rd = 0; rd < kFR; ++rd)
#pragma unroll
            for (int u = 0; u < kFU; ++u) {
                const uint32_t g0 = r0 + rd * (kB * kFU) + u * 256u + w * 64u;
                live[rd][u] = (g0 < r1 ? range_mask(r1 - g0) : 0ull) & kvw[rd][u] & fvw[rd][u] &
                              __ballot(v[rd][u] >= lo && v[rd][u] <= hi);
            }
    }
    for (int f = f0; f < nf; ++f) {
        const DCol& fc = jb.fcol[f];
        const int32_t lo = jb.flo[f], hi = jb.fhi[f];
        int32_t v[kFR][kFU];
        uint64_t fvw[kFR][kFU];
        if (fc.flat) {
            const FlatCol ff = flat_col(fc, rows);
#pragma unroll
            for (int rd = 0; rd < kFR; ++rd) {
                const uint32_t grp0 = (r0 + rd * (kB * kFU)) / 64u + w;
                const u32x2 vv = __builtin_amdgcn_raw_buffer_load_b64(ff.vb, (grp0 + lu) * 8u, 0, 0);
#pragma unroll
                for (int u = 0; u < kFU; ++u) {
                    v[rd][u] = __builtin_amdgcn_raw_buffer_load_b32(
                        ff.val, ((grp0 + 4u * u) * 64u + lane) * 4u, 0, 0);
                    fvw[rd][u] = lane_word(vv.x | ff.all_valid, vv.y | ff.all_valid, u);
                }
            }
        } else {
            const IdxCol wf = idx_col(fc, rows);
            u32x4 fr[kFR];
#pragma unroll
            for (int rd = 0; rd < kFR; ++rd) fr[rd] = idx_rec_any(wf, (r0 + rd * (kB * kFU)) / 64u + w + lu);
#pragma unroll
            for (int rd = 0; rd < kFR; ++rd)
#pragma unroll
                for (int u = 0; u < kFU; ++u) {
                    fvw[rd][u] = rec_valid(fr[rd], u);
                    v[rd][u] = idx_load(wf, fr[rd], u, fvw[rd][u], lane);
                }
        }
#pragma unroll
        for (int rd = 0; rd < kFR; ++rd)
#pragma unroll
            for (int u = 0; u < kFU; ++u)
                live[rd][u] &= __ballot(v[rd][u] >= lo && v[rd][u] <= hi) & fvw[rd][u];
    }
#pragma unroll
    for (int rd = 0; rd < kFR; ++rd)
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            rank[rd][u] = 0;
            if (in_mask(live[rd][u])) {
                const uint32_t fp = fb ? fib32((uint32_t)key[rd][u]) >> (32u - fb) : 0u;
                rank[rd][u] = atomicAdd(&hist[fp], 1u);
                alive |= 1u << (rd * kFU + u);
            }
        }
    __syncthreads();
    // partition starts: exclusive scan of the P counts, kPB per thread; hist[P] = the total
    {
        constexpr int kPB = (1 << kPartBits) / kB;
        const uint32_t per = (P + kB - 1) / kB;
        uint32_t loc[kPB], sum = 0;
#pragma unroll
        for (int c = 0; c < kPB; ++c) {
            const uint32_t bin = t * per + c;
            const uint32_t v = (uint32_t)c < per && bin < P ? hist[bin] : 0u;
            loc[c] = sum;
            sum += v;
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kB>(sum, lds, &tot);
#pragma unroll
        for (int c = 0; c < kPB; ++c) {
            const uint32_t bin = t * per + c;
            if ((uint32_t)c < per && bin < P) hist[bin] = ex + loc[c];
        }
        if (t == 0) hist[P] = tot;
    }
    __syncthreads();
    HJ_GLOBAL uint16_t* po = gp(jb.poff) + (uint64_t)x * (P + 1u);
    for (uint32_t i = t; i <= P; i += kB) po[i] = (uint16_t)hist[i];
#pragma unroll
    for (int rd = 0; rd < kFR; ++rd)
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            if (!((alive >> (rd * kFU + u)) & 1u)) continue;
            const uint32_t fp = fb ? fib32((uint32_t)key[rd][u]) >> (32u - fb) : 0u;
            const uint32_t pos = hist[fp] + rank[rd][u];
            gp(jb.st_key)[r0 + pos] = (uint32_t)key[rd][u];
            gp(jb.st_row)[r0 + pos] = r0 + rd * (kB * kFU) + u * 256u + w * 64u + lane;
        }
    if (t == 0) {
        gp(jb.blk)[x] = hist[P];
        // the job's tuple count, in one of its kCountWays partial counters
        if (hist[P]) atomicAdd((uint32_t*)&gp(bb->pcnt)[job * kCountWays + x % kCountWays], hist[P]);
    }
}

// Payload pass (jobs with pay_on, after either filter pass): the kPaySplit blocks of filter
// block x replace the rows staged in x's range (rows of x's own kBuildRows-row range, so the
// value reads stay in a few pages) by their values of the job's payload column, which the table
// then holds in place of rows; each thread takes kPayRows rows, their loads in flight together.
// A pass of its own keeps the filter kernels' registers as they are.
constexpr int kPayRows = 8;
constexpr uint32_t kPaySplit = kBuildRows / (kB * kPayRows);
__global__ __launch_bounds__(kB) void k_sb_pay(const BuildBatch* __restrict__ bb) {
    const int job = build_job_of(bb, blockIdx.x / kPaySplit);
    const BuildJob& jb = bb->job[job];
    if (!jb.pay_on) return;
    const uint32_t x = blockIdx.x / kPaySplit - jb.blk0;
    const uint32_t n = gp(jb.blk)[x];
    const uint32_t j0 = (blockIdx.x % kPaySplit) * (kB * kPayRows);
    HJ_GLOBAL uint32_t* sr = gp(jb.st_row) + (uint64_t)x * kBuildRows;
    uint32_t row[kPayRows];
    int32_t v[kPayRows];
#pragma unroll
    for (int k = 0; k < kPayRows; ++k) {
        const uint32_t j = j0 + k * kB + threadIdx.x;
        row[k] = j < n ? sr[j] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPayRows; ++k) {
        v[k] = 0;
        if (j0 + k * kB + threadIdx.x < n) (void)get_int32(jb.pay, row[k], &v[k]);
    }
#pragma unroll
    for (int k = 0; k < kPayRows; ++k) {
        const uint32_t j = j0 + k * kB + threadIdx.x;
        if (j < n) sr[j] = (uint32_t)v[k];
    }
}

// Segment pass: the table segments of every job (2^fine_bits(k) per job, k from the tuple
// count), a grid-stride loop over them. A block gathers a segment's tuples from every filter
// block's staged range (its partitions are contiguous there; a partition that holds several
// segments is read by each of them, each keeping its own tuples), inserts them into the segment
// held in LDS, and writes the segment out whole (the table needs no clearing). A segment that
// runs out of slots flags the build, as a repeated key does. Loop index i maps to segment
// (i % 8) * C + i / 8 (C = ceil(segments / 8)): with a grid that is a multiple of 8, the blocks
// on one XCD (blocks are dealt to the 8 XCDs in turn) take neighbouring segments, whose staged
// tuples and partition starts share cache lines.
__global__ __launch_bounds__(kB) void k_sb_segbuild(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t T[8u << kSegBits];  // the segment: [4 keys | 4 rows] per bucket
    __shared__ uint32_t s_lo[kMaxPartBlocks], s_pre[kMaxPartBlocks + 1];
    __shared__ uint32_t lds[kNW];
    const uint32_t t = threadIdx.x;
    const int nj = bb->njobs;
    uint32_t first[kMaxStages + 1];  // first flat segment of each job
    uint32_t nj_t[kMaxStages];       // tuples per job (all partial counters: independent loads)
#pragma unroll
    for (int j = 0; j < kMaxStages; ++j) {
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < (int)kCountWays; ++i) c += gp(bb->pcnt)[j * kCountWays + i];
        nj_t[j] = c;
    }
    // the count for the kernels after this one
    if (blockIdx.x == 0 && t == 0)
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j)
            if (j < nj) gp(bb->job[j].meta)[0] = nj_t[j];
    first[0] = 0;
#pragma unroll
    for (int j = 0; j < kMaxStages; ++j)
        first[j + 1] = first[j] + (j < nj ? 1u << fine_bits(table_bits(nj_t[j])) : 0u);
    const uint32_t F = first[kMaxStages], C = (F + 7u) / 8u;
    for (uint32_t i = blockIdx.x; i < 8u * C; i += gridDim.x) {
        const uint32_t f = (i & 7u) * C + (i >> 3);
        if (f >= F) continue;
        int j = 0;
#pragma unroll
        for (int i = 1; i < kMaxStages; ++i) j += f >= first[i] ? 1 : 0;
        const BuildJob& jb = bb->job[j];
        HJ_GLOBAL uint32_t* meta = gp(jb.meta);
        uint32_t n = 0;  // the job's tuple count, loaded above
#pragma unroll
        for (int jj = 0; jj < kMaxStages; ++jj) n = jj == j ? nj_t[jj] : n;
        const uint32_t fb = jb.pbits, P = 1u << fb;
        const uint32_t k = table_bits(n), sb = fine_bits(k), S = 1u << (k - 2 - sb);
        const uint32_t seg = f - first[j];
        // partitions [y, y + per) hold the segment; with fewer partitions than segments, keep
        // only the tuples whose top sb hash bits are seg
        const bool pick = sb > fb;
        const uint32_t per = pick ? 1u : 1u << (fb - sb);
        const uint32_t y = pick ? seg >> (sb - fb) : seg * per;
        const uint32_t nb = jb.nblk;
        __syncthreads();  // the previous segment's LDS is written out
        // this segment's range in each filter block (thread t: blocks 2t, 2t + 1), their
        // exclusive scan, so that the segment's tuples are spread over all threads
        uint32_t cnt[2] = {0, 0};
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const uint32_t b = 2u * t + i;
            if (b < nb) {
                const HJ_GLOBAL uint16_t* po = gp(jb.poff) + (uint64_t)b * (P + 1u);
                const uint32_t lo = po[y], hi = po[y + per];
                s_lo[b] = lo;
                cnt[i] = hi - lo;
            }
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kB>(cnt[0] + cnt[1], lds, &tot);
        if (2u * t < nb) s_pre[2u * t] = ex;
        if (2u * t + 1u < nb) s_pre[2u * t + 1u] = ex + cnt[0];
        if (t == 0) s_pre[nb] = tot;
        for (uint32_t i = t; i < 8u * S; i += kB) T[i] = (i & 4u) ? 0u : kEmptyKey;
        __syncthreads();
        const uint32_t smask = S - 1u;
        const bool use_bf = bf_enabled(n);
        const uint32_t lw = bf_log2w(n);
        HJ_GLOBAL uint32_t* K = gp(jb.table);
        constexpr int kSU = 4;  // items per thread and round: their gathers in flight at once
        for (uint32_t i0 = 0; i0 < tot; i0 += kB * kSU) {
            uint32_t key[kSU], row[kSU];
            bool have[kSU];
#pragma unroll
            for (int u = 0; u < kSU; ++u) {
                const uint32_t i = i0 + u * kB + t;
                have[u] = i < tot;
                key[u] = row[u] = 0;
                if (!have[u]) continue;
                // filter block holding item i: the last b with s_pre[b] <= i
                uint32_t lo = 0, hi = nb;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_pre[mid] <= i)
                        lo = mid;
                    else
                        hi = mid;
                }
                const uint64_t src = (uint64_t)lo * kBuildRows + s_lo[lo] + (i - s_pre[lo]);
                key[u] = gp(jb.st_key)[src];
                row[u] = gp(jb.st_row)[src];
            }
#pragma unroll
            for (int u = 0; u < kSU; ++u) {
                if (!have[u]) continue;
                const uint32_t h = fib32(key[u]);
                if (pick && (h >> (32u - sb)) != seg) continue;
                if (use_bf) atomicOr((uint32_t*)&gp(jb.bf)[bf_word(h, lw)], bf_mask(h));
                if (key[u] == kEmptyKey) {  // kept aside in slot 2^k
                    if (atomicCAS((uint32_t*)&meta[2], 0u, 1u) != 0u)
                        meta[1] = 1;
                    else
                        K[row_word(1u << k)] = row[u];
                    continue;
                }
                const uint32_t s0 = h >> (32u - k);  // home slot
                uint32_t bk = (s0 >> 2) & smask;
                const uint32_t j0 = s0 & 3u;
                bool done = false;
                for (uint32_t m = 0; m <= smask && !done; ++m, bk = (bk + 1u) & smask) {
                    for (uint32_t jj = 0; jj < kBucketKeys && !done; ++jj) {
                        const uint32_t sl = m == 0 ? (j0 + jj) & 3u : jj;
                        const uint32_t old = atomicCAS(&T[8u * bk + sl], kEmptyKey, key[u]);
                        if (old == kEmptyKey) {
                            T[8u * bk + 4u + sl] = row[u];
                            done = true;
                        } else if (old == key[u]) {
                            meta[1] = 1;
                            done = true;
                        }
                    }
                }
                if (!done) meta[1] = 1;  // the segment is full
            }
        }
        __syncthreads();
        HJ_GLOBAL u32x4* dst = (HJ_GLOBAL u32x4*)(K + 8u * (seg * S));
        const u32x4* srcv = (const u32x4*)T;
        for (uint32_t i = t; i < 2u * S; i += kB) dst[i] = srcv[i];
    }
}

// tuple count of a job (sum of its filter blocks' counts; every block of the clear pass
// computes it, block 0 records it), then the table's 2^k key slots and the Bloom words
__global__ __launch_bounds__(kB) void k_sb_clear(const BuildBatch* __restrict__ bb) {
    const BuildJob& jb = bb->job[blockIdx.y];
    __shared__ uint32_t s_part[kNW];
    {
        uint32_t c = 0;
        for (uint32_t i = threadIdx.x; i < jb.nblk; i += kB) c += gp(jb.blk)[i];
        c = wave_sum(c);
        if (lane_id() == 0) s_part[threadIdx.x / HJ_WAVE] = c;
    }
    __syncthreads();
    uint32_t n = 0;
#pragma unroll
    for (int i = 0; i < (int)kNW; ++i) n += s_part[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) gp(jb.meta)[0] = n;
    const uint32_t nbk = (1u << table_bits(n)) / kBucketKeys;
    const uint32_t g = blockIdx.x * kB + threadIdx.x, stride = gridDim.x * kB;
    HJ_GLOBAL u32x4* T = (HJ_GLOBAL u32x4*)gp(jb.table);
    const u32x4 empty = {kEmptyKey, kEmptyKey, kEmptyKey, kEmptyKey};
    for (uint32_t i = g; i < 2u * nbk; i += stride) T[i] = empty;  // whole buckets: full lines
    if (bf_enabled(n)) {
        const uint32_t W = 1u << bf_log2w(n);
        for (uint32_t i = g; i < W; i += stride) gp(jb.bf)[i] = 0;
    }
}

// home slot is the top k bits of its Fibonacci hash (bucket = the top k - 2, then the slot in
// the bucket); kIU independent compare-and-swaps per thread are in flight at once.
__global__ __launch_bounds__(kB) void k_sb_insert(const BuildBatch* __restrict__ bb) {
    const BuildJob& jb = bb->job[build_job_of(bb, blockIdx.x)];
    const uint32_t x = blockIdx.x - jb.blk0;
    HJ_GLOBAL uint32_t* meta = gp(jb.meta);
    const uint32_t n = meta[0];
    const uint32_t k = table_bits(n), smask = seg_mask(k);
    const bool use_bf = bf_enabled(n);
    const uint32_t lw = bf_log2w(n);
    const uint32_t r0 = x * kBuildRows;
    const uint32_t cnt = gp(jb.blk)[x];
    HJ_GLOBAL uint32_t* K = gp(jb.table);
    for (uint32_t i0 = 0; i0 < cnt; i0 += kB * kIU) {
        uint32_t key[kIU], row[kIU], s[kIU], old[kIU];
        bool ok[kIU];
#pragma unroll
        for (int u = 0; u < kIU; ++u) {
            const uint32_t i = i0 + u * kB + threadIdx.x;
            ok[u] = i < cnt;
            key[u] = ok[u] ? gp(jb.st_key)[r0 + i] : 0u;
            row[u] = ok[u] ? gp(jb.st_row)[r0 + i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kIU; ++u) {
            s[u] = fib32(key[u]) >> (32u - k);
            old[u] = ~0u;
            if (ok[u] && key[u] != kEmptyKey)
                old[u] = atomicCAS((uint32_t*)&K[8u * (s[u] >> 2) + (s[u] & 3u)], kEmptyKey, key[u]);
        }
#pragma unroll
        for (int u = 0; u < kIU; ++u) {
            if (!ok[u]) continue;
            if (key[u] == kEmptyKey) {  // kept aside in slot 2^k
                if (atomicCAS((uint32_t*)&meta[2], 0u, 1u) != 0u)
                    meta[1] = 1;
                else
                    K[row_word(1u << k)] = row[u];
            } else if (old[u] == kEmptyKey) {
                K[row_word(s[u])] = row[u];
            } else if (old[u] == key[u]) {
                meta[1] = 1;
            } else {
                insert_rest(K, smask, s[u], key[u], row[u], meta);
            }
            if (use_bf) {
                const uint32_t h = fib32(key[u]);
                atomicOr((uint32_t*)&gp(jb.bf)[bf_word(h, lw)], bf_mask(h));
            }
        }
    }
}

// ---------------------------------------------------------------- fused probe
// added to the block's LDS totals
template <int NS>
__device__ inline void count_bytes(const PipeDesc& pd, uint64_t* s_bytes, uint32_t row,
                                   const uint32_t (&brow)[NS]) {
    for (int c = 0; c < pd.nostr; ++c) {
        const int sl = pd.ostr_slot[c];
        uint32_t rid = row;
#pragma unroll
        for (int q = 0; q < NS; ++q) rid = sl == 1 + q ? brow[q] : rid;
        const uint64_t ref = string_ref(pd.ostr[c], rid);
        if (ref == kNullRef) continue;
        const uint64_t len =
            (ref & kLongRef) ? long_length(pd.ostr[c], (uint32_t)ref) : (ref & 0xffffu);
        if (len) atomicAdd((unsigned long long*)&s_bytes[c], (unsigned long long)len);
    }
}

// NS streamed joins, U 64-row groups per wave and iteration
constexpr uint32_t kQueue = 128;  // per-wave queue of rows that passed the LDS filter

// look up the joins for the first n (<= 64) entries of a wave's queue, one entry per lane, and
// stage the rows that match all of them (row, then the build row of each join in probe order).
// The joins run as a cascade, each on the lanes the previous ones kept: first those without an
// LDS filter (probe order), last the LDS-filtered first join, whose filter already dropped most
// non-matching rows, so its table is read only for the final candidates.
template <int NS>
__device__ inline void queue_flush(uint32_t (&qb)[NS + 1][kQueue], uint32_t n, uint32_t lane,
                                   const HJ_GLOBAL uint32_t* const (&tkey)[NS],
                                   const uint32_t (&tk)[NS], const uint32_t (&tsp)[NS],
                                   uint32_t* s_cnt, uint64_t* s_bytes, const PipeDesc& pd,
                                   uint32_t st0) {
    __builtin_amdgcn_wave_barrier();
    uint64_t alive = n >= HJ_WAVE ? ~0ull : (1ull << n) - 1ull;
    const uint32_t row = qb[0][lane];
    uint32_t slot[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const int q = i + 1 < NS ? i + 1 : 0;  // joins 1 .. NS-1, then join 0
        if (!alive) return;
        uint32_t sl = kDead;
        if (in_mask(alive)) sl = table_lookup(tkey[q], tk[q], tsp[q], qb[1 + q][lane]);
        slot[q] = sl;
        alive &= __ballot(sl != kDead);
    }
    if (!alive) return;
    uint64_t am[1] = {alive};
    uint32_t pos[1];
    block_append_masks<1>(s_cnt, am, pos);
    if (in_mask(alive)) {
        uint32_t brow[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) brow[q] = tkey[q][row_word(slot[q])];
        gp(pd.stage[0])[st0 + pos[0]] = row;
#pragma unroll
        for (int q = 0; q < NS; ++q) gp(pd.stage[1 + q])[st0 + pos[0]] = brow[q];
        count_bytes<NS>(pd, s_bytes, row, brow);
    }
}

// move the queue's entries [64, 64 + rem) to the front
template <int NS>
__device__ inline void queue_shift(uint32_t (&qb)[NS + 1][kQueue], uint32_t rem, uint32_t lane) {
    __builtin_amdgcn_wave_barrier();
    uint32_t e[NS + 1];
#pragma unroll
    for (int q = 0; q <= NS; ++q) e[q] = qb[q][HJ_WAVE + lane];
    __builtin_amdgcn_wave_barrier();
    if (lane < rem)
#pragma unroll
        for (int q = 0; q <= NS; ++q) qb[q][lane] = e[q];
    __builtin_amdgcn_wave_barrier();
}

template <int NS, int U, bool FILT>
__global__ __launch_bounds__(kB) void k_sp_probe(const PipeDesc* __restrict__ pdp) {
    __shared__ uint32_t bfl[kBloomWords];
    __shared__ uint32_t qbuf[kNW][NS + 1][kQueue];  // [wave][row, key of each join][entry]
    __shared__ uint32_t s_cnt;
    __shared__ uint64_t s_bytes[kMaxOutStr];  // string bytes of this block's survivors
    const PipeDesc& pd = *pdp;
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    // a build with a duplicate key voids the pass (the counters of all kMaxStages jobs follow
    // the pipeline's in one zero-initialized block: independent loads)
    uint32_t bad = 0;
#pragma unroll
    for (int j = 0; j < kMaxStages; ++j) bad |= gp(pd.meta)[4 + 4 * j + 1];
    const uint32_t n0 = gp(pd.sj[0].meta)[0];
    const bool use_bf = bf_enabled(n0);
    const uint32_t lw = bf_log2w(n0);
    if (pd.nsidx) {
        // the source columns' group index for this block's groups (also when the pass is
        // voided below: the query's other paths read it), written before the waves read it;
        // the page tables live in the queue's LDS, unused until the stream
        static_assert(sizeof(qbuf) >= kProbeIdxCols * kPageTableWords * 4, "page tables");
        uint32_t* pts = &qbuf[0][0][0];
        const uint32_t gpb = pd.probe_gpb, gb0 = blockIdx.x * gpb;
        const uint32_t r_lo = gb0 * 64u;
        if ((int)w < pd.nsidx) {
            const DCol& c = pd.sidx_col[w];
            const uint64_t end = (uint64_t)(gb0 + gpb) * 64u;
            const uint32_t r_hi = (uint32_t)(end < c.num_rows ? end : c.num_rows);
            if (r_lo < r_hi)
                page_table_load(c, r_lo, r_hi, pts + w * kPageTableWords, lane, pd.sidx_table);
        }
        __syncthreads();
        for (int c = 0; c < pd.nsidx; ++c)
            page_table_groups(pd.sidx_col[c], pd.sidx_out[c], gb0, gpb, t, kB,
                              pts + c * kPageTableWords);
        __syncthreads();
    }
    if (bad) {
        if (t == 0) {
            if (blockIdx.x == 0) gp(pd.meta)[0] = 1;
            gp(pd.blk_cnt)[blockIdx.x] = 0;
        }
        return;
    }
    {
        // the whole Bloom word array (kBloomWords, used or not): independent 16-byte loads
        constexpr int kC = kBloomWords / 4 / kB;
        const HJ_GLOBAL u32x4* src = (const HJ_GLOBAL u32x4*)gp(pd.sj[0].bf);
        u32x4 c[kC];
#pragma unroll
        for (int i = 0; i < kC; ++i) c[i] = src[t + i * kB];
#pragma unroll
        for (int i = 0; i < kC; ++i) ((u32x4*)bfl)[t + i * kB] = c[i];
    }
    if (t == 0) s_cnt = 0;
    if (t < (uint32_t)kMaxOutStr) s_bytes[t] = 0;
    __syncthreads();
    const uint32_t nsrc = pd.src_rows;
    const uint32_t G = (nsrc + 63u) >> 6;
    IdxCol wc[NS];  // the key columns, read from their pages
    const HJ_GLOBAL uint32_t* tkey[NS];
    uint32_t tk[NS], tsp[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const StreamJoin& sj = pd.sj[q];
        wc[q] = idx_col(sj.key, nsrc);
        tkey[q] = gp(sj.table);
        tk[q] = table_bits(gp(sj.meta)[0]);
        tsp[q] = gp(sj.meta)[2];
    }
    const int nf = FILT ? pd.nf : 0;  // filters of the source scan (their loads are not
                                       // prefetched: a separate instantiation)
    // this block's groups [g_lo, g_hi) (g_lo a multiple of U); survivors are staged from row
    // g_lo * 64 of the stage arrays on
    const uint32_t g_lo = blockIdx.x * pd.probe_gpb;
    const uint32_t g_hi = g_lo + pd.probe_gpb < G ? g_lo + pd.probe_gpb : G;
    const uint32_t st0 = g_lo * 64u;
    uint32_t qc = 0;  // entries in this wave's queue (uniform)
    // Every streamed join's group records are fetched two iterations ahead and its keys one
    // iteration ahead (the keys' page offsets come from the records), each into one of two
    // register sets (alternating), so that both are in flight while the current groups are
    // filtered and looked up. Lane u < U holds the record and the validity word of group u.
    const uint32_t lu = lane < (uint32_t)U ? lane : (uint32_t)U - 1u;
    int32_t ka[NS][U], kb[NS][U];
    u32x2 va[NS], vb[NS];
    u32x4 ra[NS], rb[NS];
    auto fetch_rec = [&](u32x4 (&rx)[NS], uint32_t gx) {
#pragma unroll
        for (int q = 0; q < NS; ++q) rx[q] = idx_rec(wc[q], gx * U + lu);
    };
    auto fetch = [&](int32_t (&kx)[NS][U], u32x2 (&vx)[NS], const u32x4 (&rx)[NS]) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            vx[q] = u32x2{rx[q].x, rx[q].y};
#pragma unroll
            for (int u = 0; u < U; ++u) kx[q][u] = idx_load(wc[q], rx[q], u, rec_valid(rx[q], u), lane);
        }
    };
    // one iteration: the U groups of wave-iteration gx, whose keys and validity words are in
    // key / vw
    auto step = [&](const int32_t (&key)[NS][U], const u32x2 (&vw)[NS], uint32_t gx) {
        uint32_t vl = ~0u, vh = ~0u;
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            vl &= vw[q].x;
            vh &= vw[q].y;
        }
        // live rows of each group: in this block's range, non-NULL in every streamed join
        uint64_t live[U];
        uint32_t r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t g = gx * U + u;
            const uint32_t rem = g < g_hi ? nsrc - g * 64u : 0u;
            live[u] = (rem >= 64u ? ~0ull : (1ull << rem) - 1ull) & lane_word(vl, vh, u);
            r[u] = g * 64u + lane;
        }
        for (int f = 0; f < nf; ++f) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                bool ok = false;
                if (in_mask(live[u])) {
                    int32_t v = 0;
                    ok = get_int32(pd.fcol[f], r[u], &v) && v >= pd.flo[f] && v <= pd.fhi[f];
                }
                live[u] &= __ballot(ok);
            }
        }
        if (use_bf) {
            // a strong LDS filter on the first join: rows that pass it go to this wave's queue;
            // full 64-lane batches of the queue then look up every join
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t h = fib32((uint32_t)key[0][u]);
                const uint32_t m = bf_mask(h);
                live[u] &= __ballot((bfl[bf_word(h, lw)] & m) == m);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!live[u]) continue;
                if (in_mask(live[u])) {
                    const uint32_t slot = qc + mask_rank(live[u]);
                    qbuf[w][0][slot] = r[u];
#pragma unroll
                    for (int q = 0; q < NS; ++q) qbuf[w][1 + q][slot] = (uint32_t)key[q][u];
                }
                qc += (uint32_t)__popcll(live[u]);
                if (qc >= HJ_WAVE) {
                    queue_flush<NS>(qbuf[w], HJ_WAVE, lane, tkey, tk, tsp, &s_cnt, s_bytes, pd, st0);
                    queue_shift<NS>(qbuf[w], qc - HJ_WAVE, lane);
                    qc -= HJ_WAVE;
                }
            }
            return;
        }
        // joins in order, each on the lanes the previous ones kept
        uint32_t brow[NS][U];  // key slot of the match, then its build row
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            uint64_t any = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                any |= live[u];
                brow[q][u] = kDead;
            }
            if (any) {  // uniform
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    uint32_t sl = kDead;
                    if (in_mask(live[u]))
                        sl = table_lookup(tkey[q], tk[q], tsp[q], (uint32_t)key[q][u]);
                    brow[q][u] = sl;
                    live[u] &= __ballot(sl != kDead);
                }
                // the build rows of the matches, read while their buckets are still cached
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (in_mask(live[u])) brow[q][u] = tkey[q][row_word(brow[q][u])];
            }
        }
        uint64_t any = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) any |= live[u];
        if (!any) return;
        uint32_t pos[U];
        block_append_masks<U>(&s_cnt, live, pos);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (in_mask(live[u])) {
                uint32_t br[NS];
#pragma unroll
                for (int q = 0; q < NS; ++q) br[q] = brow[q][u];
                gp(pd.stage[0])[st0 + pos[u]] = r[u];
#pragma unroll
                for (int q = 0; q < NS; ++q) gp(pd.stage[1 + q])[st0 + pos[u]] = br[q];
                count_bytes<NS>(pd, s_bytes, r[u], br);
            }
        }
    };
    // Issue order per iteration: the records two iterations ahead, then the keys of the next
    // iteration (waiting only for that iteration's records, which were issued before the
    // current keys: vector loads complete in order), then the current groups.
    uint32_t gb = g_lo / U + w;
    fetch_rec(ra, gb);
    fetch_rec(rb, gb + kNW);
    fetch(ka, va, ra);
    while (gb * U < g_hi) {
        fetch_rec(ra, gb + 2u * kNW);
        fetch(kb, vb, rb);
        step(ka, va, gb);
        gb += kNW;
        if (gb * U >= g_hi) break;
        fetch_rec(rb, gb + 2u * kNW);
        fetch(ka, va, ra);
        step(kb, vb, gb);
        gb += kNW;
    }
    if (qc) queue_flush<NS>(qbuf[w], qc, lane, tkey, tk, tsp, &s_cnt, s_bytes, pd, st0);
    __syncthreads();
    if (t == 0) gp(pd.blk_cnt)[blockIdx.x] = s_cnt;
    if (pd.blk_bytes && t < (uint32_t)pd.nostr)
        gp(pd.blk_bytes)[(uint64_t)blockIdx.x * kMaxOutStr + t] = s_bytes[t];
}

// ---------------------------------------------------------------- sliced probe
// One streamed join over a table larger than an XCD's L2 (the two-way shape): the table's 8
// slices (the top 3 bits of the key hash, i.e. of the home bucket) are probed by the XCDs one
// each. Workgroups are dispatched to the 8 XCDs in turn (block b runs on XCD b % 8), so a probe
// block b with b % 8 == x reads only slice x, which stays in that XCD's L2.
//   k_sp_part: block p takes the source rows [p * kPartRows, (p + 1) * kPartRows), keeps the
//     live ones (valid key, in range, source filters) and writes their (key, row) pairs to the
//     same range ordered by slice, with the 8 slice counts in pcnt[p * 8 ..].
//   k_sp_probe_sl: block b = c * 8 + x takes slice x of the partition blocks
//     [c * part_k, (c + 1) * part_k) (one run each), looks the keys up and stages the survivors
//     from blk_origin[b] on: the position of slice x in that chunk of rows if its pairs were
//     ordered by slice, which has room for every pair of the run set.
constexpr uint32_t kPartGroupsPerWave = kPartRows / 64 / kNW;
constexpr int kPartU = 4;  // groups per wave and round of the partition pass
static_assert(kPartGroupsPerWave % kPartU == 0, "partition rounds");

template <bool FILT>
__global__ __launch_bounds__(kB) void k_sp_part(const PipeDesc* __restrict__ pdp) {
    __shared__ uint32_t s_key[kPartRows];
    __shared__ uint8_t s_sl[kPartRows];    // slice of a live row, 0xff for any other row
    __shared__ uint32_t o_key[kPartRows];  // the live pairs ordered by slice ...
    __shared__ uint16_t o_row[kPartRows];  // ... (row - first row of the block)
    __shared__ uint32_t s_wc[kNW][8], s_wb[kNW][8], s_tot;  // per wave: slice counts, starts
    const PipeDesc& pd = *pdp;
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    const uint32_t nsrc = pd.src_rows;
    const IdxCol wc = idx_col(pd.sj[0].key, nsrc);
    const uint32_t r0 = blockIdx.x * kPartRows;
    const int nf = FILT ? pd.nf : 0;
    // the wave's groups in rounds of kPartU: one record load (lane u: group u), then the
    // round's key loads at once
    uint32_t c8[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) c8[x] = 0;
    const uint32_t lu = lane < (uint32_t)kPartU ? lane : (uint32_t)kPartU - 1u;
    for (uint32_t it = 0; it < kPartGroupsPerWave / kPartU; ++it) {
        const uint32_t lg0 = w * kPartGroupsPerWave + it * kPartU;  // block-local group
        const uint32_t g0 = r0 / 64u + lg0;
        const u32x4 rec = idx_rec(wc, g0 + lu);
        int32_t key[kPartU];
        uint64_t vw[kPartU];
#pragma unroll
        for (int u = 0; u < kPartU; ++u) {
            vw[u] = rec_valid(rec, u);
            key[u] = idx_load(wc, rec, u, vw[u], lane);
        }
#pragma unroll
        for (int u = 0; u < kPartU; ++u) {
            const uint32_t g = g0 + u;
            const uint32_t rem = g * 64u < nsrc ? nsrc - g * 64u : 0u;
            uint64_t live = (rem >= 64u ? ~0ull : (1ull << rem) - 1ull) & vw[u];
            const uint32_t r = g * 64u + lane;
            for (int f = 0; f < nf; ++f) {
                bool ok = false;
                if (in_mask(live)) {
                    int32_t v = 0;
                    ok = get_int32(pd.fcol[f], r, &v) && v >= pd.flo[f] && v <= pd.fhi[f];
                }
                live &= __ballot(ok);
            }
            const uint32_t sl = fib32((uint32_t)key[u]) >> 29;
            const bool on = in_mask(live);
            const uint32_t i = (lg0 + u) * 64u + lane;
            s_key[i] = (uint32_t)key[u];
            s_sl[i] = on ? (uint8_t)sl : (uint8_t)0xff;
#pragma unroll
            for (int x = 0; x < 8; ++x)
                c8[x] += (uint32_t)__popcll(__ballot(on && sl == (uint32_t)x));
        }
    }
    // each wave's slice counts, then every wave's start in each slice (slices in order, waves
    // in order within a slice): the wave places its pairs with ballots alone
    if (lane == 0)
#pragma unroll
        for (int x = 0; x < 8; ++x) s_wc[w][x] = c8[x];
    __syncthreads();
    if (t == 0) {
        uint32_t o = 0;
        for (int x = 0; x < 8; ++x) {
            uint32_t n = 0;
            for (uint32_t v = 0; v < kNW; ++v) {
                s_wb[v][x] = o + n;
                n += s_wc[v][x];
            }
            gp(pd.pcnt)[blockIdx.x * 8u + x] = n;
            o += n;
        }
        s_tot = o;
    }
    __syncthreads();
    uint32_t base[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) base[x] = s_wb[w][x];
    // order the live pairs by slice in LDS, then store them with coalesced writes (stores
    // scattered over 8 runs would cost a store instruction per run and 64 rows)
    for (uint32_t i0 = w * (kPartRows / kNW); i0 < (w + 1) * (kPartRows / kNW); i0 += HJ_WAVE) {
        const uint32_t i = i0 + lane;
        const uint32_t sl = s_sl[i];
        const uint32_t key = s_key[i];
#pragma unroll
        for (uint32_t x = 0; x < 8; ++x) {
            const uint64_t m = __ballot(sl == x);
            if (sl == x) {
                const uint32_t pos = base[x] + mask_rank(m);
                o_key[pos] = key;
                o_row[pos] = (uint16_t)i;
            }
            base[x] += (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
    HJ_GLOBAL uint32_t* tk = gp(pd.ptk) + r0;
    HJ_GLOBAL uint32_t* tr = gp(pd.ptr) + r0;
    const uint32_t tot = s_tot;
    if (!pd.ppay_on) {
        for (uint32_t j = t; j < tot; j += kB) {
            tk[j] = o_key[j];
            tr[j] = r0 + o_row[j];
        }
        return;
    }
    // the source's only output, an INT32 column without NULL, travels with the key in place of
    // the row: every value of the block's range loaded first (one per thread and round), then
    // the stores
    constexpr int kPR = (int)(kPartRows / kB);
    uint32_t pv[kPR];
#pragma unroll
    for (int k = 0; k < kPR; ++k) {
        const uint32_t j = t + k * kB;
        int32_t v = 0;
        if (j < tot) (void)get_int32(pd.ppay, r0 + o_row[j], &v);
        pv[k] = (uint32_t)v;
    }
#pragma unroll
    for (int k = 0; k < kPR; ++k) {
        const uint32_t j = t + k * kB;
        if (j < tot) {
            tk[j] = o_key[j];
            tr[j] = pv[k];
        }
    }
}

constexpr int kSlU = 4;  // pairs per thread and round of the sliced probe

__global__ __launch_bounds__(kB) void k_sp_probe_sl(const PipeDesc* __restrict__ pdp) {
    __shared__ uint32_t s_pc[kSliceMaxK * 8];  // the chunk's slice counts
    __shared__ uint32_t s_pre[kSliceMaxK + 1];  // slice x's pairs before each run
    __shared__ uint32_t s_run[kSliceMaxK];      // first pair of each run
    __shared__ uint32_t s_cnt, s_origin;
    __shared__ uint64_t s_bytes[kMaxOutStr];
    const PipeDesc& pd = *pdp;
    const uint32_t t = threadIdx.x;
    uint32_t bad = 0;
#pragma unroll
    for (int j = 0; j < kMaxStages; ++j) bad |= gp(pd.meta)[4 + 4 * j + 1];
    if (bad) {
        if (t == 0) {
            if (blockIdx.x == 0) gp(pd.meta)[0] = 1;
            gp(pd.blk_cnt)[blockIdx.x] = 0;
            gp(pd.blk_origin)[blockIdx.x] = 0;
        }
        return;
    }
    const uint32_t x = blockIdx.x & 7u, c = blockIdx.x >> 3;
    const uint32_t b0 = c * pd.part_k;
    const uint32_t b1 = b0 + pd.part_k < pd.part_blocks ? b0 + pd.part_k : pd.part_blocks;
    const uint32_t nr = b1 > b0 ? b1 - b0 : 0u;  // runs (partition blocks) of this chunk
    for (uint32_t i = t; i < nr * 8u; i += kB) s_pc[i] = gp(pd.pcnt)[b0 * 8u + i];
    if (t < (uint32_t)kMaxOutStr) s_bytes[t] = 0;
    __syncthreads();
    if (t == 0) {
        uint32_t pre = 0, before = 0;
        for (uint32_t j = 0; j < nr; ++j) {
            uint32_t off = 0;
            for (uint32_t y = 0; y < x; ++y) off += s_pc[j * 8u + y];
            before += off;
            s_run[j] = (b0 + j) * kPartRows + off;
            s_pre[j] = pre;
            pre += s_pc[j * 8u + x];
        }
        s_pre[nr] = pre;
        s_origin = b0 * kPartRows + before;
        s_cnt = 0;
    }
    __syncthreads();
    const uint32_t total = s_pre[nr], origin = s_origin;
    const HJ_GLOBAL uint32_t* tkey = gp(pd.sj[0].table);
    const uint32_t tk = table_bits(gp(pd.sj[0].meta)[0]);
    const uint32_t tsp = gp(pd.sj[0].meta)[2];
    const HJ_GLOBAL uint32_t* pk = gp(pd.ptk);
    const HJ_GLOBAL uint32_t* pr = gp(pd.ptr);
    for (uint32_t f0 = 0; f0 < total; f0 += kB * kSlU) {
        uint32_t key[kSlU], row[kSlU], brow[kSlU];
        uint64_t live[kSlU];
#pragma unroll
        for (int u = 0; u < kSlU; ++u) {
            const uint32_t f = f0 + u * kB + t;
            const bool have = f < total;
            uint32_t lo = 0, hi = nr;  // the run holding pair f: the last j with s_pre[j] <= f
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pre[mid] <= f)
                    lo = mid;
                else
                    hi = mid;
            }
            const uint32_t idx = have ? s_run[lo] + (f - s_pre[lo]) : 0u;
            key[u] = have ? pk[idx] : 0u;
            row[u] = have ? pr[idx] : 0u;
            live[u] = __ballot(have);
        }
        // each home bucket's keys and row words together (one 32-byte line): the build row of
        // a match in its home bucket needs no further load
        u32x4 kv[kSlU], rv[kSlU];
#pragma unroll
        for (int u = 0; u < kSlU; ++u) {
            const uint32_t hb = home_bucket(key[u], tk);
            kv[u] = u32x4{0u, 0u, 0u, 0u};
            rv[u] = kv[u];
            if (in_mask(live[u])) {
                kv[u] = load_bucket(tkey, hb);
                rv[u] = *(const HJ_GLOBAL u32x4*)(tkey + 8u * hb + 4u);
            }
        }
#pragma unroll
        for (int u = 0; u < kSlU; ++u) {
            const uint32_t hb = home_bucket(key[u], tk);
            uint32_t sl = kDead;
            if (in_mask(live[u])) sl = resolve_home(kv[u], key[u], hb, tk, tsp);
            const uint32_t j = sl & 3u;
            uint32_t br = j == 0 ? rv[u].x : (j == 1 ? rv[u].y : (j == 2 ? rv[u].z : rv[u].w));
            const bool far = sl == kDead - 1u || (sl != kDead && (sl >> 2) != hb);
            if (in_mask(live[u]) && far) {  // beyond the home bucket (rare), or INT32_MIN
                if (sl == kDead - 1u) {
                    const uint32_t smask = seg_mask(tk);
                    const uint32_t nb = next_bucket(hb, smask);
                    sl = bucket_resolve(tkey, smask, nb, key[u], load_bucket(tkey, nb));
                }
                if (sl != kDead) br = tkey[row_word(sl)];
            }
            brow[u] = br;
            live[u] &= __ballot(sl != kDead);
        }
        uint32_t pos[kSlU];
        block_append_masks<kSlU>(&s_cnt, live, pos);
#pragma unroll
        for (int u = 0; u < kSlU; ++u) {
            if (!in_mask(live[u])) continue;
            gp(pd.stage[0])[origin + pos[u]] = row[u];
            gp(pd.stage[1])[origin + pos[u]] = brow[u];
            const uint32_t br1[1] = {brow[u]};
            count_bytes<1>(pd, s_bytes, row[u], br1);
        }
    }
    __syncthreads();
    if (t == 0) {
        gp(pd.blk_cnt)[blockIdx.x] = s_cnt;
        gp(pd.blk_origin)[blockIdx.x] = origin;
    }
    if (pd.blk_bytes && t < (uint32_t)pd.nostr)
        gp(pd.blk_bytes)[(uint64_t)blockIdx.x * kMaxOutStr + t] = s_bytes[t];
}

// the fused pass's staged survivors -> the outputs (no joins keyed by a build table) or the
// dense survivor list, at the block's offset from the scan of the block counts
__global__ __launch_bounds__(kB) void k_sp_compact(const PipeDesc* __restrict__ pdp) {
    const PipeDesc& pd = *pdp;
    if (gp(pd.meta)[0]) return;
    const uint32_t b = blockIdx.x;
    const uint32_t cnt = gp(pd.blk_cnt)[b], off = gp(pd.blk_off)[b];
    const uint32_t st0 = b * pd.probe_gpb * 64u;
    for (uint32_t i = threadIdx.x; i < cnt; i += kB) {
        if (!pd.ndep) {
            for (int o = 0; o < pd.nout; ++o) {
                const int sl = pd.out_slot[o];
                gp(pd.final_out[o])[off + i] = gp(pd.stage[sl ? 1 + pd.qpos[sl - 1] : 0])[st0 + i];
            }
        } else {
            gp(pd.l_row)[off + i] = gp(pd.stage[0])[st0 + i];
            for (int q = 0; q < pd.nstream; ++q)
                gp(pd.l_b[pd.order[q]])[off + i] = gp(pd.stage[1 + q])[st0 + i];
        }
    }
}

// a join keyed by a column of an earlier build table, over the survivor list; previous = the
// stage whose kDead marks propagate (-1: none)
__global__ __launch_bounds__(kB) void k_pp_stage(const PipeDesc* __restrict__ pdp, int s,
                                                 int previous) {
    const PipeDesc& pd = *pdp;
    if (gp(pd.meta)[0]) return;
    const uint32_t n = gp(pd.meta)[2];
    const PipeStage& ps = pd.st[s];
    const HJ_GLOBAL BuildJob* jp = gp(pd.jobs) + ps.job;
    const HJ_GLOBAL uint32_t* tkey = gp(jp->table);
    const uint32_t k = table_bits(gp(jp->meta)[0]), smask = seg_mask(k);
    const uint32_t special = gp(jp->meta)[2] ? (1u << k) : kDead;
    const HJ_GLOBAL uint32_t* prev = previous < 0 ? nullptr : gp(pd.l_b[previous]);
    const HJ_GLOBAL uint32_t* src =
        ps.key_slot == 0 ? gp(pd.l_row) : gp(pd.l_b[ps.key_slot - 1]);
    HJ_GLOBAL uint32_t* outb = gp(pd.l_b[s]);
    constexpr int U = 4;
    const uint32_t stride = gridDim.x * kB * U;
    for (uint32_t base = blockIdx.x * kB * U; base < n; base += stride) {
        uint32_t jx[U], r[U], hb[U];
        int32_t key[U];
        bool ok[U];
        u32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            jx[u] = base + u * kB + threadIdx.x;
            ok[u] = jx[u] < n && (!prev || prev[jx[u]] != kDead);
            r[u] = ok[u] ? src[jx[u]] : kDead;
            ok[u] = ok[u] && r[u] != kDead;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            key[u] = 0;
            ok[u] = ok[u] && get_int32(ps.keycol, r[u], &key[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            hb[u] = home_bucket((uint32_t)key[u], k);
            v[u] = u32x4{kEmptyKey, kEmptyKey, kEmptyKey, kEmptyKey};
            if (ok[u]) v[u] = load_bucket(tkey, hb[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (jx[u] >= n) continue;
            uint32_t slot = kDead;
            if (ok[u]) {
                const uint32_t kk = (uint32_t)key[u];
                slot = kk == kEmptyKey ? special : bucket_resolve(tkey, smask, hb[u], kk, v[u]);
            }
            outb[jx[u]] = slot == kDead ? kDead : tkey[row_word(slot)];
        }
    }
}

// exclusive scan of n u32 by one block (n = *n_dev when n_dev is given, rounded up to tiles)
__global__ __launch_bounds__(1024) void k_scan1_u32(const uint32_t* in, uint32_t* out, uint32_t n,
                                                    const uint32_t* n_dev, uint32_t* total) {
    __shared__ uint32_t lds[1024 / HJ_WAVE];
    constexpr int IT = 8;
    if (n_dev) n = (*n_dev + 4095u) / 4096u;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += 1024 * IT) {
        uint32_t v[IT], acc = 0;
        const uint32_t s0 = b0 + threadIdx.x * IT;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            v[i] = s0 + i < n ? in[s0 + i] : 0u;
            acc += v[i];
        }
        uint32_t tot;
        uint32_t ex = carry + block_excl_scan<1024>(acc, lds, &tot);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            if (s0 + i < n) out[s0 + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// final compaction of the survivors that passed every join: per 4096-entry tile counts ...
__global__ __launch_bounds__(kB) void k_pp_fcount(const PipeDesc* __restrict__ pdp, int lastj) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const PipeDesc& pd = *pdp;
    const uint32_t n = gp(pd.meta)[2];
    const uint32_t nt = (n + 4095u) / 4096u;
    const HJ_GLOBAL uint32_t* last = lastj < 0 ? nullptr : gp(pd.l_b[lastj]);
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        uint32_t c = 0;
        for (uint32_t i = t * 4096u + threadIdx.x; i < (t + 1) * 4096u && i < n; i += kB)
            c += !last || last[i] != kDead;
        const uint32_t tot = bsum_u32<kB / HJ_WAVE>(c, lds);
        if (threadIdx.x == 0) gp(pd.tile_cnt)[t] = tot;
    }
}

// ... then an ordered write of each output slot
__global__ __launch_bounds__(kB) void k_pp_fwrite(const PipeDesc* __restrict__ pdp, int lastj) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const PipeDesc& pd = *pdp;
    const uint32_t n = gp(pd.meta)[2];
    const uint32_t nt = (n + 4095u) / 4096u;
    const HJ_GLOBAL uint32_t* last = lastj < 0 ? nullptr : gp(pd.l_b[lastj]);
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        uint32_t run = gp(pd.tile_off)[t];
        for (uint32_t c0 = 0; c0 < 4096u; c0 += kB) {
            const uint32_t i = t * 4096u + c0 + threadIdx.x;
            const bool alive = i < n && (!last || last[i] != kDead);
            uint32_t tot;
            const uint32_t o = run + ballot_prefix(alive, lds, &tot);
            if (alive) {
                for (int q = 0; q < pd.nout; ++q) {
                    const int sl = pd.out_slot[q];
                    gp(pd.final_out[q])[o] = sl == 0 ? gp(pd.l_row)[i] : gp(pd.l_b[sl - 1])[i];
                }
            }
            run += tot;
        }
    }
}

template <class K>
void launch_timed(K kernel, const PipeDesc& hp, const PipeDesc* dp, hipStream_t s, hipEvent_t a,
                  hipEvent_t b) {
    if (a && b)
        hipExtLaunchKernelGGL(kernel, dim3(hp.probe_blocks), dim3(kB), 0, s, a, b, 0, dp);
    else
        hipLaunchKernelGGL(kernel, dim3(hp.probe_blocks), dim3(kB), 0, s, dp);
}

// the sliced probe's two launches, the first event before the partition pass and the second
// after the probe
void launch_sliced(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s, hipEvent_t a,
                   hipEvent_t b) {
    const dim3 pg(hp.part_blocks), sg(hp.probe_blocks), bs(kB);
    if (hp.nf)
        hipExtLaunchKernelGGL(k_sp_part<true>, pg, bs, 0, s, a, nullptr, 0, dp);
    else
        hipExtLaunchKernelGGL(k_sp_part<false>, pg, bs, 0, s, a, nullptr, 0, dp);
    hipExtLaunchKernelGGL(k_sp_probe_sl, sg, bs, 0, s, nullptr, b, 0, dp);
}

template <int NS>
void launch_probe(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s, hipEvent_t a,
                  hipEvent_t b) {
    constexpr int U = probe_unroll(NS);
    if constexpr (NS == 1) {
        if (hp.xslices) {
            launch_sliced(hp, dp, s, a, b);
            return;
        }
    }
    if (hp.nf)
        launch_timed(k_sp_probe<NS, U, true>, hp, dp, s, a, b);
    else
        launch_timed(k_sp_probe<NS, U, false>, hp, dp, s, a, b);
}

template <int NS>
uint32_t resident_blocks(uint32_t num_cus) {
    // the smaller occupancy of the two instantiations (with and without source filters),
    // queried once per process (a function-local static's initializer runs exactly once, also
    // when several contexts' threads arrive together)
    static const uint32_t per_cu = [] {
        int n0 = 0, n1 = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n0, k_sp_probe<NS, probe_unroll(NS), false>,
                                                         kB, 0) != hipSuccess)
            n0 = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n1, k_sp_probe<NS, probe_unroll(NS), true>,
                                                         kB, 0) != hipSuccess)
            n1 = 1;
        return (uint32_t)std::max(1, std::min(n0, n1));
    }();
    return per_cu * num_cus;
}

}  // namespace

uint32_t probe_resident_blocks(int ns, uint32_t num_cus) {
    switch (ns) {
        case 1: return resident_blocks<1>(num_cus);
        case 2: return resident_blocks<2>(num_cus);
        case 3: return resident_blocks<3>(num_cus);
        case 4: return resident_blocks<4>(num_cus);
        case 5: return resident_blocks<5>(num_cus);
        case 6: return resident_blocks<6>(num_cus);
        case 7: return resident_blocks<7>(num_cus);
        default: return resident_blocks<8>(num_cus);
    }
}

void pipe_build(const BuildBatch& hb, const BuildBatch* db, hipStream_t s) {
    if (hb.njobs == 0) return;
    bool pay = false;
    for (int j = 0; j < hb.njobs; ++j) pay = pay || hb.job[j].pay_on;
    if (hb.part) {
        hipLaunchKernelGGL(k_sb_filter_part, dim3(hb.grid + hb.igrid), dim3(kB), 0, s, db);
        if (pay) hipLaunchKernelGGL(k_sb_pay, dim3(hb.grid * kPaySplit), dim3(kB), 0, s, db);
        hipLaunchKernelGGL(k_sb_segbuild, dim3(hb.sgrid), dim3(kB), 0, s, db);
        return;
    }
    hipLaunchKernelGGL(k_sb_filter, dim3(hb.grid + hb.igrid), dim3(kB), 0, s, db);
    if (pay) hipLaunchKernelGGL(k_sb_pay, dim3(hb.grid * kPaySplit), dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_sb_clear, dim3(hb.cgrid, hb.njobs), dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_sb_insert, dim3(hb.grid), dim3(kB), 0, s, db);
}

void pipe_probe_fused(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s, hipEvent_t ev_a,
                      hipEvent_t ev_b) {
    switch (hp.nstream) {
        case 1: launch_probe<1>(hp, dp, s, ev_a, ev_b); break;
        case 2: launch_probe<2>(hp, dp, s, ev_a, ev_b); break;
        case 3: launch_probe<3>(hp, dp, s, ev_a, ev_b); break;
        case 4: launch_probe<4>(hp, dp, s, ev_a, ev_b); break;
        case 5: launch_probe<5>(hp, dp, s, ev_a, ev_b); break;
        case 6: launch_probe<6>(hp, dp, s, ev_a, ev_b); break;
        case 7: launch_probe<7>(hp, dp, s, ev_a, ev_b); break;
        case 8: launch_probe<8>(hp, dp, s, ev_a, ev_b); break;
        default: return;
    }
}

void pipe_probe_tail(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s) {
    if (hp.nstream < 1 || hp.nstream > kMaxStages) return;
    hipLaunchKernelGGL(k_scan1_u32, dim3(1), dim3(1024), 0, s, hp.blk_cnt, hp.blk_off,
                       hp.probe_blocks, (const uint32_t*)nullptr, hp.meta + (hp.ndep ? 2 : 3));
    hipLaunchKernelGGL(k_sp_compact, dim3(hp.probe_blocks), dim3(kB), 0, s, dp);
    if (hp.ndep == 0) return;
    int previous = -1;
    for (int st = 0; st < hp.nst; ++st) {
        if (hp.st[st].key_slot == 0) continue;
        hipLaunchKernelGGL(k_pp_stage, dim3(hp.grid), dim3(kB), 0, s, dp, st, previous);
        previous = st;
    }
    hipLaunchKernelGGL(k_pp_fcount, dim3(hp.grid), dim3(kB), 0, s, dp, previous);
    hipLaunchKernelGGL(k_scan1_u32, dim3(1), dim3(1024), 0, s, hp.tile_cnt, hp.tile_off, 0u,
                       (const uint32_t*)hp.meta + 2, hp.meta + 3);
    hipLaunchKernelGGL(k_pp_fwrite, dim3(hp.grid), dim3(kB), 0, s, dp, previous);
}

}  // namespace hj
