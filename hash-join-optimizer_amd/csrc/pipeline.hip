// pipeline.hip - builds and the fused probe of the join pipeline (gfx950, wave64).
//
// Build, all build sides in the same launches (blockIdx.y = job):
//   filter  rows passing the scan's filters with a non-NULL key -> staged (key, row), appended
//           through a wave-aggregated cursor                        reads key + filter columns
//   clear   the 2^k table entries (2^k >= 2n for n staged tuples) and the Bloom words
//   insert  linear probing with a 64-bit compare-and-swap of row << 32 | key; a key found twice
//           flags the build (the caller then falls back to the general join path)
// Probe: one fused pass over the source rows (a wave takes U consecutive 64-row groups, one row
// per lane). The keys of every streamed join are loaded first (coalesced, row-space mirror),
// then each join in turn: validity, the LDS Bloom filter of the first (smallest) build, and
// for the lanes still alive a 16-byte load of the two entries at the key's home slot, which
// settles nearly every probe. Survivors are appended through a wave-aggregated cursor.
// Joins keyed by a column of a build table run afterwards over the survivor list, marking
// failures, and an ordered compaction writes the output row ids.
#include "pipeline.hpp"

#include <algorithm>

namespace hj {

namespace {

constexpr int kB = 256;
constexpr uint32_t kNW = kB / HJ_WAVE;
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));  // one 16-byte load or store

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

// ---------------------------------------------------------------- tables
// home slot of a key in a table of 2^k entries (k >= kMinTableBits)
__device__ inline uint32_t home_slot(uint32_t key, uint32_t k) { return fib32(key) >> (32u - k); }

// build row of `key`, kDead when absent, given the two entries e0, e1 at the even slot at or
// below its home slot s; a run past them (rare at load factor <= 1/2) continues entry by entry
__device__ inline uint32_t tab_resolve(const HJ_GLOBAL uint64_t* T, uint32_t mask, uint32_t s, uint32_t key,
                                       uint64_t e0, uint64_t e1) {
    if (!(s & 1u)) {
        if (e0 == kEmpty) return kDead;
        if ((uint32_t)e0 == key) return (uint32_t)(e0 >> 32);
    }
    if (e1 == kEmpty) return kDead;
    if ((uint32_t)e1 == key) return (uint32_t)(e1 >> 32);
    for (uint32_t q = ((s | 1u) + 1u) & mask;; q = (q + 1u) & mask) {
        const uint64_t e = T[q];
        if (e == kEmpty) return kDead;
        if ((uint32_t)e == key) return (uint32_t)(e >> 32);
    }
}

__device__ inline void tab_pair(const HJ_GLOBAL uint64_t* T, uint32_t s, uint64_t* e0, uint64_t* e1) {
    const u64x2 v = *(const HJ_GLOBAL u64x2*)(T + (s & ~1u));
    *e0 = v.x;
    *e1 = v.y;
}

// ---------------------------------------------------------------- build
constexpr int kBU = kBuildTile / kB;  // rows per thread per iteration of the filter pass

// block-local append of U predicates per lane through an LDS counter (one LDS atomic per wave)
template <int U>
__device__ inline void block_append(uint32_t* lds_cnt, const bool (&pred)[U], uint32_t (&pos)[U]) {
    uint64_t b[U];
    uint32_t tot = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        b[u] = __ballot(pred[u]);
        tot += (uint32_t)__popcll(b[u]);
    }
    if (!tot) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(lds_cnt, tot);
    base = __shfl(base, 0, HJ_WAVE);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        pos[u] = base + (uint32_t)__popcll(b[u] & lanemask_lt());
        base += (uint32_t)__popcll(b[u]);
    }
}

// filter pass: block x of job y takes rows [x * range, (x + 1) * range) and stages the tuples
// that pass at the start of that range
__global__ __launch_bounds__(kB) void k_sb_filter(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t s_cnt;
    const BuildJob& jb = bb->job[blockIdx.y];
    const uint32_t rows = jb.rows;
    const uint32_t range = build_range(rows, gridDim.x);
    const uint32_t r0 = blockIdx.x * range;
    const uint32_t r1 = r0 + range < rows ? r0 + range : rows;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (uint32_t base = r0; base < r1; base += kB * kBU) {
        bool ok[kBU];
        int32_t key[kBU];
        uint32_t r[kBU], pos[kBU];
#pragma unroll
        for (int u = 0; u < kBU; ++u) {
            r[u] = base + u * kB + threadIdx.x;
            key[u] = 0;
            ok[u] = r[u] < r1 && get_int32(jb.key, r[u], &key[u]);
        }
        for (int f = 0; f < jb.nf; ++f) {
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                int32_t v = 0;
                ok[u] = ok[u] && get_int32(jb.fcol[f], r[u], &v) && v >= jb.flo[f] && v <= jb.fhi[f];
            }
        }
        block_append<kBU>(&s_cnt, ok, pos);
#pragma unroll
        for (int u = 0; u < kBU; ++u) {
            if (ok[u]) {
                gp(jb.st_key)[r0 + pos[u]] = (uint32_t)key[u];
                gp(jb.st_row)[r0 + pos[u]] = r[u];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) gp(jb.blk)[blockIdx.x] = s_cnt;
}

// tuple count of a job = sum of the filter pass's block counts
__device__ inline uint32_t job_tuples(const BuildJob& jb, uint32_t blocks, uint32_t* lds) {
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < blocks; i += kB) c += gp(jb.blk)[i];
    return bsum_u32<kNW>(c, lds);
}

__global__ __launch_bounds__(kB) void k_sb_clear(const BuildBatch* __restrict__ bb) {
    __shared__ uint32_t lds[kNW];
    const BuildJob& jb = bb->job[blockIdx.y];
    const uint32_t n = job_tuples(jb, gridDim.x, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) gp(jb.meta)[0] = n;
    const uint32_t half = (1u << table_bits(n)) / 2u;
    const uint32_t g = blockIdx.x * kB + threadIdx.x, stride = gridDim.x * kB;
    HJ_GLOBAL u64x2* T = (HJ_GLOBAL u64x2*)gp(jb.table);
    const u64x2 empty = {kEmpty, kEmpty};
    for (uint32_t i = g; i < half; i += stride) T[i] = empty;
    if (bf_enabled(n)) {
        const uint32_t W = 1u << bf_log2w(n);
        for (uint32_t i = g; i < W; i += stride) gp(jb.bf)[i] = 0;
    }
}

// insert: block x of job y inserts the tuples staged by block x of the filter pass
__global__ __launch_bounds__(kB) void k_sb_insert(const BuildBatch* __restrict__ bb) {
    const BuildJob& jb = bb->job[blockIdx.y];
    const uint32_t n = gp(jb.meta)[0];
    const uint32_t k = table_bits(n), mask = (1u << k) - 1u;
    const bool use_bf = bf_enabled(n);
    const uint32_t lw = bf_log2w(n);
    const uint32_t r0 = blockIdx.x * build_range(jb.rows, gridDim.x);
    const uint32_t cnt = gp(jb.blk)[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < cnt; i += kB) {
        const uint32_t key = gp(jb.st_key)[r0 + i], row = gp(jb.st_row)[r0 + i];
        const unsigned long long e = ((unsigned long long)row << 32) | key;
        for (uint32_t s = home_slot(key, k);; s = (s + 1u) & mask) {
            const unsigned long long old =
                atomicCAS((unsigned long long*)&gp(jb.table)[s], (unsigned long long)kEmpty, e);
            if (old == kEmpty) break;
            if ((uint32_t)old == key) {
                gp(jb.meta)[1] = 1;
                break;
            }
        }
        if (use_bf) {
            const uint32_t h = fib32(key);
            atomicOr((uint32_t*)&gp(jb.bf)[bf_word(h, lw)], bf_mask(h));
        }
    }
}

// ---------------------------------------------------------------- fused probe
// NS streamed joins, U 64-row groups per wave and iteration
template <int NS, int U>
__global__ __launch_bounds__(kB) void k_sp_probe(const PipeDesc* __restrict__ pdp) {
    __shared__ uint32_t bfl[kBloomWords];
    __shared__ uint32_t s_bad, s_cnt;
    const PipeDesc& pd = *pdp;
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    if (t == 0) {
        s_cnt = 0;
        uint32_t bad = 0;
        for (int j = 0; j < pd.nst; ++j) bad |= gp((gp(pd.jobs) + pd.st[j].job)->meta)[1];
        s_bad = bad;
    }
    const HJ_GLOBAL BuildJob* j0p = gp(pd.jobs) + pd.st[pd.order[0]].job;
    const uint32_t n0 = gp(j0p->meta)[0];
    const bool use_bf = bf_enabled(n0);
    const uint32_t lw = bf_log2w(n0);
    if (use_bf)
        for (uint32_t i = t; i < (1u << lw); i += kB) bfl[i] = gp(j0p->bf)[i];
    __syncthreads();
    if (s_bad) {
        if (t == 0) {
            if (blockIdx.x == 0) gp(pd.meta)[0] = 1;
            gp(pd.blk_cnt)[blockIdx.x] = 0;
        }
        return;
    }
    const HJ_GLOBAL int32_t* kf[NS];
    const HJ_GLOBAL uint64_t* kv[NS];
    const HJ_GLOBAL uint64_t* tab[NS];
    uint32_t tk[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const PipeStage& ps = pd.st[pd.order[q]];
        kf[q] = gp(ps.keycol.flat);
        kv[q] = gp(ps.keycol.vbits);
        const HJ_GLOBAL BuildJob* jp = gp(pd.jobs) + ps.job;
        tab[q] = gp(jp->table);
        tk[q] = table_bits(gp(jp->meta)[0]);
    }
    const uint32_t nsrc = pd.src_rows;
    const uint32_t G = (nsrc + 63u) >> 6;
    const int nf = pd.nf;
    // this block's groups [g_lo, g_hi); survivors are staged from row g_lo * 64 of the stage
    // arrays on
    const uint32_t g_lo = blockIdx.x * pd.probe_gpb;
    const uint32_t g_hi = g_lo + pd.probe_gpb < G ? g_lo + pd.probe_gpb : G;
    const uint32_t st0 = g_lo * 64u;
    for (uint32_t gb = g_lo / U + w; gb * U < g_hi; gb += kNW) {
        uint32_t r[U];
        bool alive[U];
        int32_t key[NS][U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            r[u] = (gb * U + u) * 64u + lane;
            alive[u] = r[u] < nsrc && gb * U + u < g_hi;
        }
        // keys of every streamed join first: NS * U coalesced loads in flight
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int u = 0; u < U; ++u) key[q][u] = alive[u] ? kf[q][r[u]] : 0;
        for (int f = 0; f < nf; ++f) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int32_t v = 0;
                alive[u] = alive[u] && get_int32(pd.fcol[f], r[u], &v) && v >= pd.flo[f] &&
                           v <= pd.fhi[f];
            }
        }
        uint32_t brow[NS][U];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                brow[q][u] = kDead;
                const uint32_t g = gb * U + u;
                if (kv[q] && g < g_hi) alive[u] = alive[u] && ((kv[q][g] >> lane) & 1ull);
            }
            if (q == 0 && use_bf) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t h = fib32((uint32_t)key[0][u]);
                    const uint32_t m = bf_mask(h);
                    alive[u] = alive[u] && (bfl[bf_word(h, lw)] & m) == m;
                }
            }
            bool any = false;
#pragma unroll
            for (int u = 0; u < U; ++u) any = any || alive[u];
            if (!__ballot(any)) break;
            uint64_t e0[U], e1[U];
            uint32_t hs[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                hs[u] = home_slot((uint32_t)key[q][u], tk[q]);
                e0[u] = e1[u] = kEmpty;
                if (alive[u]) tab_pair(tab[q], hs[u], &e0[u], &e1[u]);
            }
            const uint32_t mask = (1u << tk[q]) - 1u;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (alive[u]) {
                    brow[q][u] = tab_resolve(tab[q], mask, hs[u], (uint32_t)key[q][u], e0[u], e1[u]);
                    alive[u] = brow[q][u] != kDead;
                }
            }
        }
        uint32_t pos[U];
        block_append<U>(&s_cnt, alive, pos);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!alive[u]) continue;
            gp(pd.stage[0])[st0 + pos[u]] = r[u];
#pragma unroll
            for (int q = 0; q < NS; ++q) gp(pd.stage[1 + q])[st0 + pos[u]] = brow[q][u];
        }
    }
    __syncthreads();
    if (t == 0) gp(pd.blk_cnt)[blockIdx.x] = s_cnt;
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
    const HJ_GLOBAL uint64_t* table = gp(jp->table);
    const uint32_t k = table_bits(gp(jp->meta)[0]), mask = (1u << k) - 1u;
    const HJ_GLOBAL uint32_t* prev = previous < 0 ? nullptr : gp(pd.l_b[previous]);
    const HJ_GLOBAL uint32_t* src =
        ps.key_slot == 0 ? gp(pd.l_row) : gp(pd.l_b[ps.key_slot - 1]);
    HJ_GLOBAL uint32_t* outb = gp(pd.l_b[s]);
    constexpr int U = 4;
    const uint32_t stride = gridDim.x * kB * U;
    for (uint32_t base = blockIdx.x * kB * U; base < n; base += stride) {
        uint32_t jx[U], r[U], hs[U];
        int32_t key[U];
        bool ok[U];
        uint64_t e0[U], e1[U];
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
            hs[u] = home_slot((uint32_t)key[u], k);
            e0[u] = e1[u] = kEmpty;
            if (ok[u]) tab_pair(table, hs[u], &e0[u], &e1[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (jx[u] >= n) continue;
            outb[jx[u]] = ok[u] ? tab_resolve(table, mask, hs[u], (uint32_t)key[u], e0[u], e1[u])
                                : kDead;
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

template <int NS>
void launch_probe(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s) {
    constexpr int U = probe_unroll(NS);
    hipLaunchKernelGGL((k_sp_probe<NS, U>), dim3(hp.probe_blocks), dim3(kB), 0, s, dp);
}

}  // namespace

void pipe_build(const BuildBatch& hb, const BuildBatch* db, hipStream_t s) {
    if (hb.njobs == 0) return;
    const dim3 grid(hb.grid, hb.njobs);
    hipLaunchKernelGGL(k_sb_filter, grid, dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_sb_clear, grid, dim3(kB), 0, s, db);
    hipLaunchKernelGGL(k_sb_insert, grid, dim3(kB), 0, s, db);
}

void pipe_probe(const PipeDesc& hp, const PipeDesc* dp, hipStream_t s) {
    switch (hp.nstream) {
        case 1: launch_probe<1>(hp, dp, s); break;
        case 2: launch_probe<2>(hp, dp, s); break;
        case 3: launch_probe<3>(hp, dp, s); break;
        case 4: launch_probe<4>(hp, dp, s); break;
        case 5: launch_probe<5>(hp, dp, s); break;
        case 6: launch_probe<6>(hp, dp, s); break;
        case 7: launch_probe<7>(hp, dp, s); break;
        case 8: launch_probe<8>(hp, dp, s); break;
        default: return;
    }
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
