// pipeline.hip - batched sort-free builds and the fused probe pipeline (gfx950, wave64).
//
// Build (all build sides of the chain in the same launches, blocks mapped to jobs):
//   count   rows passing the scan's filters with a non-NULL key            reads key+filter cols
//   clear   slot counts / Bloom words of the 2^k slots, k = ceil(log2 n)    writes 8 B / slot
//   insert  per tuple: atomicAdd on its slot count (= its position inside the slot), atomicOr of
//           its Bloom bits, staging of (hash, row, position)               reads key+filter cols
//   scan    exclusive scan of the slot counts -> dir[s] = start << 16 | Bloom, dir[D] = n << 16
//   scatter tuple -> adjacency array at dir start + position              12 B read, 8 B write
// Probe: stage 0 streams the probe source's first key column, one block per page and 8 rows per
// thread, with a Bloom-filtered directory lookup per row; survivors (row, build row) are staged at
// the page's first row, then compacted into a dense list by a scan of the per-page counts. Each
// later join runs over that list (grid-stride, 4 survivors per thread, independent gathers of
// the join key), marking failures; a final ordered compaction writes the output row ids.
#include "pipeline.hpp"

namespace hj {

namespace {

constexpr int kB = 256;
constexpr int kRows = 8;                 // rows per thread per 2048-row chunk
constexpr uint32_t kChunk = kB * kRows;  // 2048
constexpr uint32_t kSlotTile = 4096;
constexpr uint32_t kTupleTile = 1024;

__device__ inline uint32_t log2ceil(uint32_t n) { return n <= 1 ? 0u : 32u - __clz(n - 1u); }

// Blocked Bloom filter of a build side, held in LDS by the probe: one 32-bit word per key with
// 4 bits set, about 16 bits per key; only used when it gets >= 8 bits per key within
// kBloomWords words. The word and bit hashes apply the "lowbias32" integer mixer (C. Wellons,
// hash-prospector) to the key's Fibonacci hash.
__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ inline bool bf_enabled(uint32_t n) { return n && (uint64_t)n * 8u <= kBloomWords * 32u; }
__device__ inline uint32_t bf_log2w(uint32_t n) {
    uint32_t b = log2ceil((n + 1u) / 2u);  // n/2 words = 16 bits per key
    return b < 5u ? 5u : (b > 12u ? 12u : b);
}
__device__ inline uint32_t bf_word(uint32_t h, uint32_t lw) { return mix32(h) >> (32u - lw); }
__device__ inline uint32_t bf_mask(uint32_t h) {
    const uint32_t g = mix32(h ^ 0x9E3779B9u);
    return (1u << (g & 31u)) | (1u << ((g >> 5) & 31u)) | (1u << ((g >> 10) & 31u)) |
           (1u << ((g >> 15) & 31u));
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
    constexpr int NW = kB / HJ_WAVE;
    const uint32_t w = threadIdx.x / HJ_WAVE;
    const uint64_t b = __ballot(pred);
    if (lane_id() == 0) lds[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        off += i < (int)w ? lds[i] : 0u;
        tot += lds[i];
    }
    __syncthreads();
    *total = tot;
    return off + (uint32_t)__popcll(b & lanemask_lt());
}

__device__ inline int job_by(const BuildBatch& bb, uint32_t blk, int) {
    int j = 0;
    for (;;) {
        if (j + 1 >= bb.njobs) break;
        if (blk < bb.job[j + 1].page_block0) break;
        ++j;
    }
    return j;
}

// key of local row l of build page p after the job's filters; false = row dropped
__device__ inline bool build_row_key(const BuildJob& jb, uint32_t p, const uint8_t* pg,
                                     uint32_t hdr, uint32_t l, uint32_t r, uint32_t* key) {
    int32_t k;
    if (!get_int32_in_page(jb.key, p, pg, hdr, l, &k)) return false;
    for (int f = 0; f < jb.nf; ++f) {
        int32_t v;
        if (!get_int32(jb.fcol[f], r, &v) || v < jb.flo[f] || v > jb.fhi[f]) return false;
    }
    *key = (uint32_t)k;
    return true;
}

// ---------------------------------------------------------------- build kernels
__global__ void k_pb_zero_meta(const BuildBatch* __restrict__ bb) {
    const int j = threadIdx.x / 4, i = threadIdx.x % 4;
    if (j < bb->njobs) bb->job[j].meta[i] = 0;
}

// extract: one block per key page; passing rows -> (hash, row) staged at the page's first row
__global__ __launch_bounds__(kB) void k_pb_extract(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const int j = job_by(*bb, blockIdx.x, 0);
    const BuildJob& jb = bb->job[j];
    const uint32_t p = blockIdx.x - jb.page_block0;
    const uint8_t* pg = page_ptr(jb.key, p);
    const uint32_t hdr = *(const uint32_t*)pg;
    const uint32_t nr = hdr & 0xffffu;
    const uint32_t r0 = jb.key.row_start[p];
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < nr; c0 += kChunk) {
        uint32_t keys[kRows], m = 0;
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            const uint32_t l = c0 + i * kB + threadIdx.x;
            uint32_t kk = 0;
            if (l < nr && build_row_key(jb, p, pg, hdr, l, r0 + l, &kk)) m |= 1u << i;
            keys[i] = kk;
        }
        uint32_t tot;
        uint32_t o = r0 + run + block_excl_scan<kB>((uint32_t)__popc(m), lds, &tot);
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            if ((m >> i) & 1u) {
                jb.st_h[o] = fib32(keys[i]);
                jb.st_row[o] = r0 + c0 + i * kB + threadIdx.x;
                ++o;
            }
        }
        run += tot;
    }
    if (threadIdx.x == 0) jb.blk[p] = run;
}

// one block per job: exclusive scan of the per-page counts, total -> meta[0]
__global__ __launch_bounds__(1024) void k_pb_blkscan(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t lds[1024 / HJ_WAVE];
    const BuildJob& jb = bb->job[blockIdx.x];
    const uint32_t n = jb.key.num_pages;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < n ? jb.blk[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<1024>(v, lds, &tot);
        if (i < n) jb.blk_off[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) jb.meta[0] = carry;
}

__global__ __launch_bounds__(kB) void k_pb_clear(const BuildBatch* __restrict__ bb) {
    const uint32_t g = blockIdx.x * kB + threadIdx.x, stride = gridDim.x * kB;
    for (int j = 0; j < bb->njobs; ++j) {
        const BuildJob& jb = bb->job[j];
        const uint32_t D = 1u << log2ceil(jb.meta[0]);
        for (uint32_t i = g; i < D; i += stride) {
            jb.cnt[i] = 0;
            jb.bloom[i] = 0;
        }
        if (bf_enabled(jb.meta[0])) {
            const uint32_t W = 1u << bf_log2w(jb.meta[0]);
            for (uint32_t i = g; i < W; i += stride) jb.bf[i] = 0;
        }
    }
}

// slot counts (the returned count is the tuple's position in its slot) and Bloom bits
__global__ __launch_bounds__(kB) void k_pb_hist(const BuildBatch* __restrict__ bb) {
    const int j = job_by(*bb, blockIdx.x, 0);
    const BuildJob& jb = bb->job[j];
    const uint32_t p = blockIdx.x - jb.page_block0;
    const uint32_t n = jb.blk[p];
    const uint32_t r0 = jb.key.row_start[p];
    const uint32_t k = log2ceil(jb.meta[0]);
    const bool use_bf = bf_enabled(jb.meta[0]);
    const uint32_t lw = bf_log2w(jb.meta[0]);
    for (uint32_t i = threadIdx.x; i < n; i += kB) {
        const uint32_t h = jb.st_h[r0 + i];
        const uint32_t s = slot_of(h, k);
        jb.st_pos[r0 + i] = atomicAdd(&jb.cnt[s], 1u);
        atomicOr(&jb.bloom[s], bloom_mask(bloom_tag(h)));
        if (use_bf) atomicOr(&jb.bf[bf_word(h, lw)], bf_mask(h));
    }
}

// slot scan phase A: per 4096-slot tile sum -> part, longest slot -> meta[2]
__global__ __launch_bounds__(kB) void k_pb_scan_a(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    for (int j = 0; j < bb->njobs; ++j) {
        const BuildJob& jb = bb->job[j];
        const uint32_t D = 1u << log2ceil(jb.meta[0]);
        const uint32_t nt = (D + kSlotTile - 1) / kSlotTile;
        for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
            const uint32_t base = t * kSlotTile;
            uint32_t s = 0, mx = 0;
            for (uint32_t i = base + threadIdx.x; i < base + kSlotTile && i < D; i += kB) {
                const uint32_t c = jb.cnt[i];
                s += c;
                mx = c > mx ? c : mx;
            }
#pragma unroll
            for (int d = HJ_WAVE / 2; d > 0; d >>= 1) {
                const uint32_t o = __shfl_xor(mx, d, HJ_WAVE);
                mx = o > mx ? o : mx;
            }
            if (lane_id() == 0 && mx) atomicMax(&jb.meta[2], mx);
            const uint32_t tot = bsum_u32<kB / HJ_WAVE>(s, lds);
            if (threadIdx.x == 0) jb.part[t] = tot;
        }
    }
}

// phase B: one block per job scans its tile partials in place (exclusive)
__global__ __launch_bounds__(1024) void k_pb_scan_b(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t lds[1024 / HJ_WAVE];
    const BuildJob& jb = bb->job[blockIdx.x];
    const uint32_t D = 1u << log2ceil(jb.meta[0]);
    const uint32_t nt = (D + kSlotTile - 1) / kSlotTile;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nt; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nt ? jb.part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<1024>(v, lds, &tot);
        if (i < nt) jb.part[i] = carry + ex;
        carry += tot;
    }
}

// phase C: dir words
__global__ __launch_bounds__(kB) void k_pb_scan_c(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    constexpr int IT = kSlotTile / kB;
    for (int j = 0; j < bb->njobs; ++j) {
        const BuildJob& jb = bb->job[j];
        const uint32_t n = jb.meta[0];
        const uint32_t D = 1u << log2ceil(n);
        const uint32_t nt = (D + kSlotTile - 1) / kSlotTile;
        for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
            const uint32_t base = t * kSlotTile;
            const uint32_t s0 = base + threadIdx.x * IT;
            uint32_t v[IT];
            uint32_t acc = 0;
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                v[i] = s0 + i < D ? jb.cnt[s0 + i] : 0u;
                acc += v[i];
            }
            uint32_t tot;
            uint32_t ex = jb.part[t] + block_excl_scan<kB>(acc, lds, &tot);
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                const uint32_t sl = s0 + i;
                if (sl < D)
                    jb.dir[sl] = ((uint64_t)ex << 32) | ((uint64_t)(v[i] < 0xffffu ? v[i] : 0xffffu) << 16) |
                                 (uint64_t)(jb.bloom[sl] & 0xffffu);
                ex += v[i];
            }
            if (t == nt - 1 && threadIdx.x == 0) jb.dir[D] = (uint64_t)n << 32;
        }
    }
}

__global__ __launch_bounds__(kB) void k_pb_scatter(const BuildBatch* __restrict__ bb) {
    const int j = job_by(*bb, blockIdx.x, 0);
    const BuildJob& jb = bb->job[j];
    const uint32_t p = blockIdx.x - jb.page_block0;
    const uint32_t n = jb.blk[p];
    const uint32_t r0 = jb.key.row_start[p];
    const uint32_t k = log2ceil(jb.meta[0]);
    for (uint32_t i = threadIdx.x; i < n; i += kB) {
        const uint32_t h = jb.st_h[r0 + i];
        const uint32_t pos = (uint32_t)(jb.dir[slot_of(h, k)] >> 32) + jb.st_pos[r0 + i];
        jb.adj[pos] = ((uint64_t)jb.st_row[r0 + i] << 32) | h;
    }
}

// ---------------------------------------------------------------- probe
// Directory word w of a slot: Bloom reject, then the slot's entries (row << 32 | hash). The
// first kFirst entries of all rows a thread holds are loaded together; longer slots (rare with
// a directory of >= n slots) finish in a loop. Returns the match count; *row = first match.
constexpr uint32_t kFirst = 2;

__device__ inline bool bloom_pass(uint64_t w, uint32_t h) {
    const uint32_t m = bloom_mask(bloom_tag(h));
    return ((uint32_t)w & m) == m && ((w >> 16) & 0xffffu) != 0;
}

__device__ inline uint32_t slot_rest(const uint64_t* adj, uint64_t w, uint32_t h, uint32_t hits,
                                     uint32_t* row) {
    const uint32_t b = (uint32_t)(w >> 32), c = (uint32_t)((w >> 16) & 0xffffu);
    for (uint32_t q = kFirst; q < c; ++q) {
        const uint64_t e = adj[b + q];
        if ((uint32_t)e == h) {
            if (!hits) *row = (uint32_t)(e >> 32);
            ++hits;
        }
    }
    return hits;
}

// A page staged in LDS: header, packed values and the validity bitmap, plus the rank of every
// 64-row group (non-NULL rows before it), so that row -> value needs LDS reads only.
struct LdsPage {
    uint64_t w[HJ_PAGE_BYTES / 8];
    uint16_t grank[1024];  // a page holds at most 65535 rows = 1024 groups
};

// stage the page with 16-byte loads and rank its 64-row groups; returns the header word
__device__ inline uint32_t lds_stage_page(LdsPage& lp, const uint8_t* pg, uint32_t* lds_scan) {
    const uint4* src = (const uint4*)pg;
    uint4* dst = (uint4*)lp.w;
#pragma unroll
    for (int i = 0; i < (int)(HJ_PAGE_BYTES / 16 / kB); ++i) dst[i * kB + threadIdx.x] = src[i * kB + threadIdx.x];
    __syncthreads();
    const uint32_t hdr = (uint32_t)lp.w[0];
    const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
    if (nr != nn) {
        const uint32_t nb = (nr + 7) >> 3, ng = (nr + 63) >> 6;
        const uint32_t base_byte = HJ_PAGE_BYTES - nb, sh = (base_byte & 7u) * 8u;
        // each thread ranks 4 consecutive groups (1024 groups max)
        uint32_t pc[4], acc = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t g = threadIdx.x * 4 + q;
            uint32_t c = 0;
            if (g < ng) {
                const uint32_t wi = (base_byte + 8u * g) >> 3;
                const uint64_t w0 = lp.w[wi];
                const uint64_t w1 = (sh && wi + 1 < HJ_PAGE_BYTES / 8) ? lp.w[wi + 1] : 0ull;
                uint64_t bits = sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0;
                if (g == ng - 1 && (nr & 63u)) bits &= (1ull << (nr & 63u)) - 1ull;
                c = (uint32_t)__popcll(bits);
            }
            pc[q] = c;
            acc += c;
        }
        uint32_t tot;
        uint32_t ex = block_excl_scan<kB>(acc, lds_scan, &tot);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t g = threadIdx.x * 4 + q;
            if (g < ng) lp.grank[g] = (uint16_t)ex;
            ex += pc[q];
        }
        __syncthreads();
    }
    return hdr;
}

// INT32 key of local row l (< nr) of a page staged in LDS; branch-free
__device__ inline bool lds_key(const LdsPage& lp, uint32_t nr, uint32_t nn, uint32_t l,
                               int32_t* key) {
    uint32_t vi = l;
    bool valid = true;
    if (nr != nn) {
        const uint32_t nb = (nr + 7) >> 3;
        const uint32_t base_byte = HJ_PAGE_BYTES - nb, sh = (base_byte & 7u) * 8u;
        const uint32_t g = l >> 6, b = l & 63u;
        const uint32_t wi = (base_byte + 8u * g) >> 3;
        const uint64_t w0 = lp.w[wi];
        const uint64_t w1 = lp.w[wi + 1 < HJ_PAGE_BYTES / 8 ? wi + 1 : wi];
        const uint64_t bits = sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0;
        valid = (bits >> b) & 1ull;
        vi = (uint32_t)lp.grank[g] + (uint32_t)__popcll(bits & ((1ull << b) - 1ull));
    }
    vi = vi < nn ? vi : 0u;
    *key = ((const int32_t*)lp.w)[1 + vi];
    return valid && nn != 0;
}

// stage 0: one block per page of the driver column; the page is staged in LDS, each thread
// takes 4 rows, and every row's directory word and first slot entries are loaded without
// branches (indices are clamped, results masked), so all loads of a thread are in flight
// together. Survivors (row, build row) are staged at the page's first row.
constexpr int kS0Rows = 4;
__global__ __launch_bounds__(kB) void k_pp_stage0(const PipeDesc* __restrict__ pdp) {
    __shared__ LdsPage lp;
    __shared__ uint32_t lds[kB / HJ_WAVE];
    __shared__ uint32_t s_bad;
    const PipeDesc& pd = *pdp;
    const uint32_t p = blockIdx.x, t = threadIdx.x;
    if (t == 0) {
        uint32_t bad = 0;
        for (int s = 0; s < pd.nst; ++s) bad |= pd.jobs[pd.st[s].job].meta[2] > kMaxFusedSlot;
        s_bad = bad;
    }
    const DCol& dc = pd.driver;
    const uint32_t hdr = lds_stage_page(lp, page_ptr(dc, p), lds);  // contains a barrier
    if (s_bad) {
        if (t == 0) {
            pd.meta[1] = 1;
            pd.page_cnt[p] = 0;
        }
        return;
    }
    const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
    const uint32_t r0 = dc.row_start[p];
    const BuildJob& j0 = pd.jobs[pd.st[0].job];
    const uint64_t* dir0 = j0.dir;
    const uint64_t* adj0 = j0.adj;
    const uint32_t n0 = j0.meta[0];
    const uint32_t k0 = log2ceil(n0);
    const uint32_t amax = n0 ? n0 - 1 : 0;
    const int nf = pd.nf;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < nr; c0 += kB * kS0Rows) {
        uint32_t hs[kS0Rows], okm = 0;
        uint64_t ws[kS0Rows];
#pragma unroll
        for (int i = 0; i < kS0Rows; ++i) {
            const uint32_t l = c0 + i * kB + t;
            int32_t key;
            bool ok = lds_key(lp, nr, nn, l < nr ? l : nr - 1, &key) && l < nr;
            if (nf) {
                for (int f = 0; ok && f < nf; ++f) {
                    int32_t v;
                    ok = get_int32(pd.fcol[f], r0 + l, &v) && v >= pd.flo[f] && v <= pd.fhi[f];
                }
            }
            hs[i] = fib32((uint32_t)key);
            okm |= ok ? (1u << i) : 0u;
        }
#pragma unroll
        for (int i = 0; i < kS0Rows; ++i) ws[i] = dir0[slot_of(hs[i], k0)];
        uint64_t e[kS0Rows][kFirst];
#pragma unroll
        for (int i = 0; i < kS0Rows; ++i) {
            const uint32_t b = (uint32_t)(ws[i] >> 32);
#pragma unroll
            for (uint32_t q = 0; q < kFirst; ++q) e[i][q] = adj0[b + q < amax ? b + q : amax];
        }
        uint32_t alive = 0, bv[kS0Rows];
        uint32_t longm = 0;
#pragma unroll
        for (int i = 0; i < kS0Rows; ++i) {
            const bool pass = ((okm >> i) & 1u) && bloom_pass(ws[i], hs[i]);
            const uint32_t c = (uint32_t)((ws[i] >> 16) & 0xffffu);
            uint32_t hits = 0, row = 0;
#pragma unroll
            for (uint32_t q = 0; q < kFirst; ++q) {
                const bool m = pass && q < c && (uint32_t)e[i][q] == hs[i];
                row = (m && !hits) ? (uint32_t)(e[i][q] >> 32) : row;
                hits += m ? 1u : 0u;
            }
            bv[i] = row;
            alive |= hits ? (1u << i) : 0u;
            longm |= (pass && c > kFirst) ? (1u << i) : 0u;
            if (hits > 1) pd.meta[0] = 1;
        }
        if (longm) {  // slots longer than kFirst: finish them (uncommon)
#pragma unroll
            for (int i = 0; i < kS0Rows; ++i) {
                if ((longm >> i) & 1u) {
                    uint32_t row = bv[i];
                    const uint32_t hits =
                        slot_rest(adj0, ws[i], hs[i], (alive >> i) & 1u, &row);
                    if (hits > 1) pd.meta[0] = 1;
                    if (hits) {
                        alive |= 1u << i;
                        bv[i] = row;
                    }
                }
            }
        }
        uint32_t tot;
        uint32_t o = r0 + run + block_excl_scan<kB>((uint32_t)__popc(alive), lds, &tot);
#pragma unroll
        for (int i = 0; i < kS0Rows; ++i) {
            if ((alive >> i) & 1u) {
                pd.s0_row[o] = r0 + c0 + i * kB + t;
                pd.s0_b[o] = bv[i];
                ++o;
            }
        }
        run += tot;
    }
    if (t == 0) pd.page_cnt[p] = run;
}

// multi-join stage keyed by the source: stream the key column's pages (persistent blocks,
// page staged in LDS, the build side's Bloom filter in LDS), keep rows set in the input bitmap
// whose key is found, and set their bits in the output bitmap. Source filters apply in the
// first streamed stage.
__global__ __launch_bounds__(kB) void k_pp_stream(const PipeDesc* __restrict__ pdp, int s,
                                                  int first, const uint32_t* __restrict__ in_bm,
                                                  uint32_t* __restrict__ out_bm) {
    __shared__ LdsPage lp;
    __shared__ uint32_t bfl[kBloomWords];
    __shared__ uint32_t lds[kB / HJ_WAVE];
    __shared__ uint32_t s_bad;
    const PipeDesc& pd = *pdp;
    const uint32_t t = threadIdx.x;
    const PipeStage& ps = pd.st[s];
    const BuildJob& jb = pd.jobs[ps.job];
    const uint32_t n = jb.meta[0];
    const uint32_t k = log2ceil(n);
    const bool use_bf = bf_enabled(n);
    const uint32_t lw = bf_log2w(n);
    const uint64_t* dir = jb.dir;
    const uint64_t* adj = jb.adj;
    const uint32_t amax = n ? n - 1 : 0;
    if (t == 0) {
        uint32_t bad = 0;
        for (int q = 0; q < pd.nst; ++q) bad |= pd.jobs[pd.st[q].job].meta[2] > kMaxFusedSlot;
        s_bad = bad;
    }
    if (use_bf)
        for (uint32_t i = t; i < (1u << lw); i += kB) bfl[i] = jb.bf[i];
    __syncthreads();
    if (s_bad) {
        if (t == 0 && blockIdx.x == 0) pd.meta[1] = 1;
        return;
    }
    const DCol& dc = ps.keycol;
    const int nf = first ? pd.nf : 0;
    const uint32_t w = t / HJ_WAVE;
    for (uint32_t p = blockIdx.x; p < dc.num_pages; p += gridDim.x) {
        const uint32_t hdr = lds_stage_page(lp, page_ptr(dc, p), lds);
        const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
        const uint32_t r0 = dc.row_start[p];
        for (uint32_t c0 = 0; c0 < nr; c0 += kB * kS0Rows) {
            uint32_t hs[kS0Rows], pm = 0;
            uint64_t ws[kS0Rows];
#pragma unroll
            for (int i = 0; i < kS0Rows; ++i) {
                const uint32_t l = c0 + i * kB + t;
                const uint32_t r = r0 + l;
                int32_t key;
                bool ok = lds_key(lp, nr, nn, l < nr ? l : nr - 1, &key) && l < nr;
                if (in_bm) ok = ok && ((in_bm[r >> 5] >> (r & 31u)) & 1u);
                for (int f = 0; ok && f < nf; ++f) {
                    int32_t v;
                    ok = get_int32(pd.fcol[f], r, &v) && v >= pd.flo[f] && v <= pd.fhi[f];
                }
                hs[i] = fib32((uint32_t)key);
                if (ok && use_bf) {
                    const uint32_t m = bf_mask(hs[i]);
                    ok = (bfl[bf_word(hs[i], lw)] & m) == m;
                }
                pm |= ok ? (1u << i) : 0u;
            }
#pragma unroll
            for (int i = 0; i < kS0Rows; ++i)
                ws[i] = ((pm >> i) & 1u) ? dir[slot_of(hs[i], k)] : 0ull;
            uint32_t qm = 0;
#pragma unroll
            for (int i = 0; i < kS0Rows; ++i) qm |= (((pm >> i) & 1u) && bloom_pass(ws[i], hs[i])) ? 1u << i : 0u;
            uint64_t e[kS0Rows][kFirst];
#pragma unroll
            for (int i = 0; i < kS0Rows; ++i) {
                const uint32_t b = (uint32_t)(ws[i] >> 32);
                const uint32_t c = (uint32_t)((ws[i] >> 16) & 0xffffu);
#pragma unroll
                for (uint32_t q = 0; q < kFirst; ++q)
                    e[i][q] = (((qm >> i) & 1u) && q < c) ? adj[b + q < amax ? b + q : amax] : ~0ull;
            }
#pragma unroll
            for (int i = 0; i < kS0Rows; ++i) {
                uint32_t hits = 0, row = 0;
                if ((qm >> i) & 1u) {
#pragma unroll
                    for (uint32_t q = 0; q < kFirst; ++q)
                        hits += ((uint32_t)e[i][q] == hs[i] && e[i][q] != ~0ull) ? 1u : 0u;
                    if (((ws[i] >> 16) & 0xffffu) > kFirst) hits = slot_rest(adj, ws[i], hs[i], hits, &row);
                    if (hits > 1) pd.meta[0] = 1;
                }
                // rows [rw, rw + 64) of this wave; their hit bits go to the output bitmap
                const uint64_t bits = __ballot(hits != 0);
                if (lane_id() == 0 && bits) {
                    const uint32_t rw = r0 + c0 + i * kB + w * HJ_WAVE;
                    const uint32_t w0 = rw >> 5, sh = rw & 31u;
                    const uint64_t a = bits << sh;
                    const uint32_t hiw = sh ? (uint32_t)(bits >> (64u - sh)) : 0u;
                    if ((uint32_t)a) atomicOr(&out_bm[w0], (uint32_t)a);
                    if ((uint32_t)(a >> 32)) atomicOr(&out_bm[w0 + 1], (uint32_t)(a >> 32));
                    if (hiw) atomicOr(&out_bm[w0 + 2], hiw);
                }
            }
        }
        __syncthreads();  // the page buffer is refilled for the next page
    }
}

// survivor bitmap -> dense row list: per tile of kB words (kB * 32 rows) counts, then ordered
// writes
__global__ __launch_bounds__(kB) void k_bm_count(const PipeDesc* __restrict__ pdp,
                                                 const uint32_t* __restrict__ bm, uint32_t nwords) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const PipeDesc& pd = *pdp;
    const uint32_t nt = (nwords + kB - 1) / kB;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const uint32_t wi = t * kB + threadIdx.x;
        const uint32_t c = wi < nwords ? (uint32_t)__popc(bm[wi]) : 0u;
        const uint32_t tot = bsum_u32<kB / HJ_WAVE>(c, lds);
        if (threadIdx.x == 0) pd.tile_cnt[t] = tot;
    }
}

__global__ __launch_bounds__(kB) void k_bm_write(const PipeDesc* __restrict__ pdp,
                                                 const uint32_t* __restrict__ bm, uint32_t nwords) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const PipeDesc& pd = *pdp;
    const uint32_t nt = (nwords + kB - 1) / kB;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const uint32_t wi = t * kB + threadIdx.x;
        uint32_t word = wi < nwords ? bm[wi] : 0u;
        uint32_t tot;
        uint32_t o = pd.tile_off[t] + block_excl_scan<kB>((uint32_t)__popc(word), lds, &tot);
        while (word) {
            const uint32_t b = (uint32_t)__ffs(word) - 1u;
            pd.l_row[o++] = wi * 32u + b;
            word &= word - 1u;
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

// stage-0 survivors -> dense list (or straight into the outputs for a single join)
__global__ __launch_bounds__(kB) void k_pp_compact0(const PipeDesc* __restrict__ pdp) {
    const PipeDesc& pd = *pdp;
    const uint32_t p = blockIdx.x;
    if (p == 0 && threadIdx.x == 0 && pd.nst == 1) pd.meta[3] = pd.meta[2];
    const uint32_t n = pd.page_cnt[p];
    if (!n) return;
    const uint32_t r0 = pd.driver.row_start[p];
    const uint32_t off = pd.page_off[p];
    for (uint32_t i = threadIdx.x; i < n; i += kB) {
        const uint32_t row = pd.s0_row[r0 + i], b = pd.s0_b[r0 + i];
        if (pd.nst == 1) {
            for (int q = 0; q < pd.nout; ++q) pd.final_out[q][off + i] = pd.out_slot[q] ? b : row;
        } else {
            pd.l_row[off + i] = row;
            pd.l_b[0][off + i] = b;
        }
    }
}

// stage s >= 1 over the survivor list: 4 survivors per thread, grid-stride
// stage s over the survivor list: previous = stage whose kDead marks propagate (-1: none)
__global__ __launch_bounds__(kB) void k_pp_stage(const PipeDesc* __restrict__ pdp, int s,
                                                 int previous) {
    const PipeDesc& pd = *pdp;
    if (pd.meta[1]) return;
    const uint32_t n = pd.meta[2];
    const PipeStage& ps = pd.st[s];
    const BuildJob& jb = pd.jobs[ps.job];
    const uint32_t k = log2ceil(jb.meta[0]);
    const uint32_t* prev = previous < 0 ? nullptr : pd.l_b[previous];
    const uint32_t* src = ps.key_slot == 0 ? pd.l_row : pd.l_b[ps.key_slot - 1];
    uint32_t* outb = pd.l_b[s];
    constexpr int U = 4;
    const uint32_t stride = gridDim.x * kB * U;
    for (uint32_t base = blockIdx.x * kB * U; base < n; base += stride) {
        uint32_t jx[U], r[U], h[U], sl[U];
        uint64_t w[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            jx[u] = base + u * kB + threadIdx.x;
            ok[u] = jx[u] < n && (!prev || prev[jx[u]] != kDead);
            r[u] = ok[u] ? src[jx[u]] : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int32_t key = 0;
            ok[u] = ok[u] && get_int32(ps.keycol, r[u], &key);
            h[u] = fib32((uint32_t)key);
            sl[u] = slot_of(h[u], k);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = ok[u] ? jb.dir[sl[u]] : 0ull;
        uint64_t e[U][kFirst];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = ok[u] && bloom_pass(w[u], h[u]);
            const uint32_t b = (uint32_t)(w[u] >> 32), c = (uint32_t)((w[u] >> 16) & 0xffffu);
#pragma unroll
            for (uint32_t q = 0; q < kFirst; ++q) e[u][q] = ok[u] && q < c ? jb.adj[b + q] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (jx[u] >= n) continue;
            uint32_t res = kDead;
            if (ok[u]) {
                uint32_t hits = 0, row = 0;
#pragma unroll
                for (uint32_t q = 0; q < kFirst; ++q) {
                    if ((uint32_t)e[u][q] == h[u] && e[u][q] != ~0ull) {
                        if (!hits) row = (uint32_t)(e[u][q] >> 32);
                        ++hits;
                    }
                }
                if (((w[u] >> 16) & 0xffffu) > kFirst) hits = slot_rest(jb.adj, w[u], h[u], hits, &row);
                if (hits > 1) pd.meta[0] = 1;
                if (hits) res = row;
            }
            outb[jx[u]] = res;
        }
    }
}

// final compaction of the survivors that passed every join: per 4096-entry tile counts ...
__global__ __launch_bounds__(kB) void k_pp_fcount(const PipeDesc* __restrict__ pdp, int lastj) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const PipeDesc& pd = *pdp;
    const uint32_t n = pd.meta[2];
    const uint32_t nt = (n + 4095u) / 4096u;
    const uint32_t* last = lastj < 0 ? nullptr : pd.l_b[lastj];
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        uint32_t c = 0;
        for (uint32_t i = t * 4096u + threadIdx.x; i < (t + 1) * 4096u && i < n; i += kB)
            c += !last || last[i] != kDead;
        const uint32_t tot = bsum_u32<kB / HJ_WAVE>(c, lds);
        if (threadIdx.x == 0) pd.tile_cnt[t] = tot;
    }
}

// ... then an ordered write of each output slot
__global__ __launch_bounds__(kB) void k_pp_fwrite(const PipeDesc* __restrict__ pdp, int lastj) {
    __shared__ uint32_t lds[kB / HJ_WAVE];
    const PipeDesc& pd = *pdp;
    const uint32_t n = pd.meta[2];
    const uint32_t nt = (n + 4095u) / 4096u;
    const uint32_t* last = lastj < 0 ? nullptr : pd.l_b[lastj];
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        uint32_t run = pd.tile_off[t];
        for (uint32_t c0 = 0; c0 < 4096u; c0 += kB) {
            const uint32_t i = t * 4096u + c0 + threadIdx.x;
            const bool alive = i < n && (!last || last[i] != kDead);
            uint32_t tot;
            const uint32_t o = run + ballot_prefix(alive, lds, &tot);
            if (alive) {
                for (int q = 0; q < pd.nout; ++q) {
                    const int sl = pd.out_slot[q];
                    pd.final_out[q][o] = sl == 0 ? pd.l_row[i] : pd.l_b[sl - 1][i];
                }
            }
            run += tot;
        }
    }
}

}  // namespace

void pipe_build(const BuildBatch& hb, const BuildBatch* db, hipStream_t s) {
    if (hb.njobs == 0) return;
    hipLaunchKernelGGL(k_pb_zero_meta, dim3(1), dim3(4 * kMaxStages), 0, s, db);
    if (hb.page_blocks)
        hipLaunchKernelGGL(k_pb_extract, dim3(hb.page_blocks), dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_pb_blkscan, dim3(hb.njobs), dim3(1024), 0, s, db);
    hipLaunchKernelGGL(k_pb_clear, dim3(hb.grid), dim3(kB), 0, s, db);
    if (hb.page_blocks) hipLaunchKernelGGL(k_pb_hist, dim3(hb.page_blocks), dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_pb_scan_a, dim3(hb.grid), dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_pb_scan_b, dim3(hb.njobs), dim3(1024), 0, s, db);
    hipLaunchKernelGGL(k_pb_scan_c, dim3(hb.grid), dim3(kB), 0, s, db);
    if (hb.page_blocks)
        hipLaunchKernelGGL(k_pb_scatter, dim3(hb.page_blocks), dim3(kB), 0, s, db);
}

void pipe_probe(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s) {
    if (hp.nst == 1) {
        const uint32_t np = hp.driver.num_pages;
        if (np) hipLaunchKernelGGL(k_pp_stage0, dim3(np), dim3(kB), 0, s, dp);
        hipLaunchKernelGGL(k_scan1_u32, dim3(1), dim3(1024), 0, s, hp.page_cnt, hp.page_off, np,
                           (const uint32_t*)nullptr, hp.meta + 2);
        if (np) hipLaunchKernelGGL(k_pp_compact0, dim3(np), dim3(kB), 0, s, dp);
        return;
    }
    // streamed joins, bitmaps alternating
    const uint32_t* in = nullptr;
    for (int q = 0; q < hp.nstream; ++q) {
        uint32_t* out = hp.bm[q & 1];
        zero_u32(out, hp.bm_words, s);
        hipLaunchKernelGGL(k_pp_stream, dim3(hp.grid), dim3(kB), 0, s, dp, hp.order[q],
                           q == 0 ? 1 : 0, in, out);
        in = out;
    }
    // survivors -> list
    hipLaunchKernelGGL(k_bm_count, dim3(hp.grid), dim3(kB), 0, s, dp, in, hp.bm_words);
    hipLaunchKernelGGL(k_scan1_u32, dim3(1), dim3(1024), 0, s, hp.tile_cnt, hp.tile_off,
                       (hp.bm_words + kB - 1) / kB, (const uint32_t*)nullptr, hp.meta + 2);
    hipLaunchKernelGGL(k_bm_write, dim3(hp.grid), dim3(kB), 0, s, dp, in, hp.bm_words);
    // build rows of the streamed joins (every survivor matches), then the joins keyed by a
    // build table, in plan order, propagating failures
    for (int q = 0; q < hp.nstream; ++q)
        hipLaunchKernelGGL(k_pp_stage, dim3(hp.grid), dim3(kB), 0, s, dp, hp.order[q], -1);
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
