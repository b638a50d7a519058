// batch.hip - the batch kernels (gfx950, wave64): the index, insert and probe items of every
// query of a batch, one launch per item kind; each launch is a grid of resident workgroups that
// take the kind's items in a fixed grid-stride order. Launch boundaries order the kinds (indexes
// before the inserts that read them, tables before the probes), so no workgroup waits for
// another. Every value that steers a branch around a barrier is made wave-uniform explicitly.
#include "batch.hpp"
#include "gindex.hpp"
#include "wavecol.hpp"

namespace hj {

namespace {

constexpr uint32_t kBNW = kBatchBlock / HJ_WAVE;
constexpr uint32_t kInsChunk = 1024;  // build rows per round of an insert item (4 groups/wave)
constexpr uint32_t kInsU = kInsChunk / (kBatchBlock);  // 64-row groups per wave and round
static_assert(kInsU * kBatchBlock == kInsChunk, "insert round");
constexpr uint32_t kMaxRadixBits = 12;
static_assert(2 * kInsChunk + (1u << kMaxRadixBits) <= kStageWords, "insert sort fits the stage");

__device__ inline uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// The rare continuations of a lookup (the home bucket is full) and of an insert (the home slot
// is taken) are calls, not inlined copies: the persistent kernel holds every item kind, and one
// inlined copy per unrolled row group would set its register count for all of them.
__device__ __noinline__ uint32_t lookup_rest(const HJ_GLOBAL uint32_t* K, uint32_t k, uint32_t hb,
                                             uint32_t kk) {
    const uint32_t smask = seg_mask(k);
    const uint32_t b = next_bucket(hb, smask);
    return bucket_resolve(K, smask, b, kk, load_bucket(K, b));
}
__device__ __noinline__ void insert_rest_call(HJ_GLOBAL uint32_t* K, uint32_t smask, uint32_t s,
                                              uint32_t key, uint32_t row, HJ_GLOBAL uint32_t* qctr) {
    insert_rest(K, smask, s, key, row, qctr);
}

// slot of kk in table K (2^k key slots), kDead when absent: table_lookup with its continuation
// called
__device__ inline uint32_t batch_lookup(const HJ_GLOBAL uint32_t* K, uint32_t k, uint32_t sp,
                                        uint32_t kk) {
    const uint32_t hb = home_bucket(kk, k);
    uint32_t sl = resolve_home(load_bucket(K, hb), kk, hb, k, sp);
    if (sl == kDead - 1u) sl = lookup_rest(K, k, hb, kk);
    return sl;
}

// ---------------------------------------------------------------- insert
// insert one tuple (key != kEmptyKey) into table K of 2^k slots at its home slot, then the rest
// of its bucket and segment; qctr[1] flags a repeated key or a full segment
__device__ inline void insert_tuple(HJ_GLOBAL uint32_t* K, uint32_t k, uint32_t key, uint32_t row,
                                    HJ_GLOBAL uint32_t* qctr) {
    const uint32_t s = fib32(key) >> (32u - k);
    const uint32_t old = atomicCAS((uint32_t*)&K[8u * (s >> 2) + (s & 3u)], kEmptyKey, key);
    if (old == kEmptyKey)
        K[row_word(s)] = row;
    else if (old == key)
        qctr[1] = 1;
    else
        insert_rest_call(K, seg_mask(k), s, key, row, qctr);
}

// rows [lo, hi) of join j's build table: filters, then inserts (radix_bits > 0: each round's
// tuples ordered by the top radix_bits bits of their hash first, so that consecutive threads
// insert into neighbouring table lines)
__device__ inline void do_insert(const BatchDesc& d, const BatchQuery& Q, uint32_t j, uint32_t lo,
                                 uint32_t hi, uint32_t* lds, uint32_t* s_n) {
    const BatchJoin& J = Q.j[j];
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    const uint32_t rows = uni(J.brows), k = uni(J.k);
    HJ_GLOBAL uint32_t* K = gp(J.table);
    HJ_GLOBAL uint32_t* qctr = gp(Q.ctr);
    const IdxCol wk = idx_col(J.bkey, rows);
    const uint32_t lu = (lane < kInsU ? lane : kInsU - 1u);
    const uint32_t rb = uni(d.radix_bits);
    const int nbf = uni(J.nbf);
    uint32_t* sk = lds;
    uint32_t* sr = lds + kInsChunk;
    uint32_t* hist = lds + 2 * kInsChunk;
    for (uint32_t base = lo; base < hi; base += kInsChunk) {
        // wave w takes groups base/64 + w * kInsU + u
        const uint32_t grp0 = base / 64u + w * kInsU;
        const u32x4 kr = idx_rec_any(wk, grp0 + lu);
        int32_t key[kInsU];
        uint64_t live[kInsU];
#pragma unroll
        for (int u = 0; u < (int)kInsU; ++u) {
            const uint64_t vw = rec_valid(kr, u);
            key[u] = idx_load(wk, kr, u, vw, lane);
            const uint32_t g0 = (grp0 + u) * 64u;
            live[u] = (g0 < hi ? range_mask(hi - g0) : 0ull) & vw;
        }
        for (int f = 0; f < nbf; ++f) {
            const IdxCol wf = idx_col(J.bfcol[f], rows);
            const u32x4 fr = idx_rec_any(wf, grp0 + lu);
            const int32_t flo = J.bflo[f], fhi = J.bfhi[f];
#pragma unroll
            for (int u = 0; u < (int)kInsU; ++u) {
                const uint64_t fvw = rec_valid(fr, u);
                const int32_t v = idx_load(wf, fr, u, fvw, lane);
                live[u] &= __ballot(v >= flo && v <= fhi) & fvw;
            }
        }
        if (rb == 0) {
#pragma unroll
            for (int u = 0; u < (int)kInsU; ++u) {
                if (!in_mask(live[u])) continue;
                const uint32_t kk = (uint32_t)key[u], row = (grp0 + u) * 64u + lane;
                if (kk == kEmptyKey) {  // kept aside in slot 2^k
                    if (atomicCAS((uint32_t*)&qctr[4 + j], 0u, 1u) != 0u)
                        qctr[1] = 1;
                    else
                        K[row_word(1u << k)] = row;
                } else {
                    insert_tuple(K, k, kk, row, qctr);
                }
            }
            continue;
        }
        // ordered inserts: LDS histogram of the round's tuples by their top rb hash bits, scan,
        // scatter, then insert in that order
        const uint32_t nb = 1u << rb;
        for (uint32_t i = t; i < nb; i += kBatchBlock) hist[i] = 0;
        __syncthreads();
        uint32_t bin[kInsU];
#pragma unroll
        for (int u = 0; u < (int)kInsU; ++u) {
            bin[u] = fib32((uint32_t)key[u]) >> (32u - rb);
            if (in_mask(live[u])) atomicAdd(&hist[bin[u]], 1u);
        }
        __syncthreads();
        {
            // thread t scans bins [t * per, (t + 1) * per)
            const uint32_t per = (nb + kBatchBlock - 1) / kBatchBlock;
            uint32_t sum = 0;
            for (uint32_t i = 0; i < per; ++i) {
                const uint32_t b = t * per + i;
                if (b < nb) sum += hist[b];
            }
            uint32_t total;
            uint32_t run = block_excl_scan<kBatchBlock>(sum, s_n, &total);
            for (uint32_t i = 0; i < per; ++i) {
                const uint32_t b = t * per + i;
                if (b < nb) {
                    const uint32_t c = hist[b];
                    hist[b] = run;
                    run += c;
                }
            }
            __syncthreads();
            s_n[kBNW] = total;
        }
        __syncthreads();
        const uint32_t n = uni(s_n[kBNW]);
#pragma unroll
        for (int u = 0; u < (int)kInsU; ++u) {
            if (in_mask(live[u])) {
                const uint32_t p = atomicAdd(&hist[bin[u]], 1u);
                sk[p] = (uint32_t)key[u];
                sr[p] = (grp0 + u) * 64u + lane;
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < n; i += kBatchBlock) {
            const uint32_t kk = sk[i], row = sr[i];
            if (kk == kEmptyKey) {
                if (atomicCAS((uint32_t*)&qctr[4 + j], 0u, 1u) != 0u)
                    qctr[1] = 1;
                else
                    K[row_word(1u << k)] = row;
            } else {
                insert_tuple(K, k, kk, row, qctr);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- probe
// Cooperative lookup: four lanes per row, each reading one key word of the row's bucket (16 rows
// per pass, 4 passes per wave); the slot (or kDead) comes back to the row's own lane.
__device__ inline uint32_t coop_lookup(const HJ_GLOBAL uint32_t* K, uint32_t k, uint32_t sp,
                                       uint32_t key, uint64_t live) {
    const uint32_t lane = lane_id(), sub = lane & 3u;
    const uint32_t smask = seg_mask(k);
    uint32_t result = kDead;
#pragma unroll 1
    for (uint32_t p = 0; p < 4; ++p) {
        const uint32_t src = p * 16u + (lane >> 2);
        const uint32_t kk = (uint32_t)__shfl((int)key, (int)src, HJ_WAVE);
        bool done = !((live >> src) & 1ull) || kk == kEmptyKey;
        uint32_t found = kDead;
        uint32_t b = home_bucket(kk, k);
        for (uint32_t m = 0; m <= smask; ++m) {
            if (!__ballot(!done)) break;
            const uint32_t wv = done ? 0u : K[8u * b + sub];
            const uint64_t mb = __ballot(!done && wv == kk);
            const uint64_t eb = __ballot(!done && wv == kEmptyKey);
            const uint32_t sh = lane & ~3u;
            const uint32_t gm = (uint32_t)(mb >> sh) & 0xfu, ge = (uint32_t)(eb >> sh) & 0xfu;
            if (!done && gm) {
                found = 4u * b + (uint32_t)__builtin_ctz(gm);
                done = true;
            } else if (!done && ge) {
                done = true;
            }
            b = next_bucket(b, smask);
        }
        if (((live >> src) & 1ull) && kk == kEmptyKey) found = sp ? (1u << k) : kDead;
        const uint32_t v = (uint32_t)__shfl((int)found, (int)(((lane - p * 16u) & 15u) * 4u), HJ_WAVE);
        if (lane >= p * 16u && lane < p * 16u + 16u) result = v;
    }
    return ((live >> lane) & 1ull) ? result : kDead;
}

// the staged survivors' output columns, at a range of the query's output reserved with one
// atomic; the stage is emptied
__device__ inline void flush_stage(const BatchQuery& Q, const uint32_t* stage, uint32_t S,
                                   uint32_t* s_n, uint32_t* s_base) {
    __syncthreads();
    const uint32_t n = uni(*s_n);
    if (n == 0) return;
    if (threadIdx.x == 0) *s_base = atomicAdd(&Q.ctr[2], n);
    __syncthreads();
    const uint32_t base = uni(*s_base);
    const int nout = uni(Q.nout);
    for (uint32_t i = threadIdx.x; i < n; i += kBatchBlock) {
        for (int c = 0; c < nout; ++c) {
            const uint32_t row = stage[(uint32_t)Q.oslot[c] * S + i];
            int32_t v = 0;
            const bool ok = get_int32(Q.ocol[c], row, &v);
            gp(Q.out[c])[base + i] = ok ? v : 0;
            if (((Q.onull >> c) & 1u) && ok)
                atomicOr(&Q.ovalid[c][(base + i) >> 5], 1u << ((base + i) & 31u));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *s_n = 0;
    __syncthreads();
}

// 64-row groups [lo, hi) of query Q's source, NJ joins: U groups per wave and round
template <int U, int NJ, bool COOP>
__device__ inline void do_probe_n(const BatchQuery& Q, uint32_t lo, uint32_t hi, uint32_t* stage,
                                  uint32_t* s_n, uint32_t* s_base) {
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / HJ_WAVE);
    constexpr int nj = NJ;
    const uint32_t nsrc = uni(Q.src_rows);
    const int nsf = uni(Q.nsf);
    const uint32_t S = kStageWords / (1u + (uint32_t)nj);
    const uint32_t lu = lane < (uint32_t)U ? lane : (uint32_t)U - 1u;
    constexpr uint32_t kRound = kBNW * U * 64u;
    for (uint32_t g0 = lo; g0 < hi; g0 += kBNW * U) {
        if (uni(*s_n) + kRound > S) flush_stage(Q, stage, S, s_n, s_base);
        const uint32_t gw = g0 + w * U;
        uint64_t live[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t g = gw + u;
            live[u] = (g < hi && g * 64u < nsrc) ? range_mask(nsrc - g * 64u) : 0ull;
        }
        int32_t key[NJ][U];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const IdxCol wc = idx_col(Q.j[j].pkey, nsrc);
            const u32x4 rc = idx_rec_any(wc, gw + lu);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t vw = rec_valid(rc, u);
                key[j][u] = idx_load(wc, rc, u, vw, lane);
                live[u] &= vw;
            }
        }
        for (int f = 0; f < nsf; ++f) {
            const IdxCol wf = idx_col(Q.sfcol[f], nsrc);
            const u32x4 fr = idx_rec_any(wf, gw + lu);
            const int32_t flo = Q.sflo[f], fhi = Q.sfhi[f];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t fvw = rec_valid(fr, u);
                const int32_t v = idx_load(wf, fr, u, fvw, lane);
                live[u] &= __ballot(v >= flo && v <= fhi) & fvw;
            }
        }
        uint32_t brow[NJ][U];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const HJ_GLOBAL uint32_t* K = gp(Q.j[j].table);
            const uint32_t k = uni(Q.j[j].k);
            const uint32_t sp = uni(Q.ctr[4 + j]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                uint32_t sl = kDead;
                if (COOP) {
                    if (live[u]) sl = coop_lookup(K, k, sp, (uint32_t)key[j][u], live[u]);
                } else if (in_mask(live[u])) {
                    sl = batch_lookup(K, k, sp, (uint32_t)key[j][u]);
                }
                live[u] &= __ballot(sl != kDead);
                brow[j][u] = sl;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (in_mask(live[u])) brow[j][u] = K[row_word(brow[j][u])];
        }
        uint32_t pos[U];
        block_append_masks<U>(s_n, live, pos);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (in_mask(live[u])) {
                stage[pos[u]] = (gw + u) * 64u + lane;
#pragma unroll
                for (int j = 0; j < NJ; ++j) stage[(1u + (uint32_t)j) * S + pos[u]] = brow[j][u];
            }
        }
        __syncthreads();
    }
    flush_stage(Q, stage, S, s_n, s_base);
}

// at most 8 row groups x joins unrolled per wave and round: U groups for up to two joins, at
// most 2 for three or four
template <int U, bool COOP>
__device__ inline void do_probe(const BatchQuery& Q, uint32_t lo, uint32_t hi, uint32_t* stage,
                                uint32_t* s_n, uint32_t* s_base) {
    constexpr int U2 = U > 2 ? 2 : U;
    switch (uni(Q.nj)) {
        case 1: do_probe_n<U, 1, COOP>(Q, lo, hi, stage, s_n, s_base); break;
        case 2: do_probe_n<U, 2, COOP>(Q, lo, hi, stage, s_n, s_base); break;
        case 3: do_probe_n<U2, 3, COOP>(Q, lo, hi, stage, s_n, s_base); break;
        default: do_probe_n<U2, 4, COOP>(Q, lo, hi, stage, s_n, s_base); break;
    }
}

// ---------------------------------------------------------------- the launches
// item it of the batch, its fields wave-uniform
struct Item {
    uint32_t kind, q, sub, lo, hi;
};
__device__ inline Item get_item(const BatchDesc& d, uint32_t it) {
    const BatchItem b = d.items[it];
    return Item{uni(b.kind_q >> 24), uni(b.kind_q & 0xffffffu), uni(b.sub), uni(b.lo), uni(b.hi)};
}

// index items [0, nidx): 64-row groups of a column, one thread per group
__global__ __launch_bounds__(kBatchBlock) void k_batch_index(const BatchDesc* __restrict__ dp) {
    const BatchDesc& d = *dp;
    const uint32_t n = uni(d.nidx);
    for (uint32_t it = blockIdx.x; it < n; it += gridDim.x) {
        const Item I = get_item(d, it);
        for (uint32_t g = I.lo + threadIdx.x; g < I.hi; g += kBatchBlock)
            group_index_group(d.icol[I.sub], d.iout[I.sub], g);
    }
}

// insert items [nidx, probe0)
__global__ __launch_bounds__(kBatchBlock) void k_batch_insert(const BatchDesc* __restrict__ dp) {
    __shared__ uint32_t lds[kStageWords];
    __shared__ uint32_t s_n[kBNW + 1];
    const BatchDesc& d = *dp;
    const uint32_t i0 = uni(d.nidx), n = uni(d.probe0);
    for (uint32_t it = i0 + blockIdx.x; it < n; it += gridDim.x) {
        const Item I = get_item(d, it);
        do_insert(d, d.q[I.q], I.sub, I.lo, I.hi, lds, s_n);
        __syncthreads();  // the next item reuses the LDS
    }
}

// probe items [probe0, nitems)
template <int U, bool COOP>
__global__ __launch_bounds__(kBatchBlock, 5) void k_batch_probe(const BatchDesc* __restrict__ dp) {
    __shared__ uint32_t lds[kStageWords];
    __shared__ uint32_t s_base;
    __shared__ uint32_t s_n[kBNW + 1];
    const BatchDesc& d = *dp;
    const uint32_t i0 = uni(d.probe0), n = uni(d.nitems);
    for (uint32_t it = i0 + blockIdx.x; it < n; it += gridDim.x) {
        const Item I = get_item(d, it);
        if (threadIdx.x == 0) s_n[0] = 0;
        __syncthreads();
        do_probe<U, COOP>(d.q[I.q], I.lo, I.hi, lds, &s_n[0], &s_base);
    }
}

// the clear pass: every fill segment (tables, counters, bitmaps), 16-byte stores
__global__ __launch_bounds__(256) void k_batch_clear(const BatchDesc* __restrict__ dp) {
    const FillSeg s = dp->fill[blockIdx.y];
    const uint64_t n4 = s.words / 4;
    const u32x4 v = {s.v, s.v, s.v, s.v};
    HJ_GLOBAL u32x4* p = (HJ_GLOBAL u32x4*)gp(s.p);
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n4;
         i += (uint64_t)gridDim.x * 256u)
        p[i] = v;
}

template <class K>
uint32_t per_cu_of(K kernel) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kBatchBlock, 0) != hipSuccess ||
        n < 1)
        n = 1;
    return (uint32_t)n;
}

}  // namespace

BatchGrids batch_grids(uint32_t unroll, uint32_t coop, uint32_t num_cus) {
    // thread-safe one-time occupancy queries
    static const uint32_t idx = per_cu_of(k_batch_index);
    static const uint32_t ins = per_cu_of(k_batch_insert);
    static const uint32_t p[4] = {per_cu_of(k_batch_probe<2, false>), per_cu_of(k_batch_probe<4, false>),
                                  per_cu_of(k_batch_probe<2, true>), per_cu_of(k_batch_probe<4, true>)};
    return BatchGrids{idx * num_cus, ins * num_cus, p[(coop ? 2 : 0) + (unroll >= 4 ? 1 : 0)] * num_cus};
}

void batch_launch(const BatchDesc* d_desc, const BatchLaunch& L, hipStream_t s, int first,
                  int last) {
    if (first <= 0 && last >= 0 && L.nfill) {
        const uint64_t x = std::min<uint64_t>(1024, (L.max_fill_words / 4 + 1023) / 1024 + 1);
        hipLaunchKernelGGL(k_batch_clear, dim3((uint32_t)x, L.nfill), dim3(256), 0, s, d_desc);
    }
    if (first <= 1 && last >= 1 && L.nidx)
        hipLaunchKernelGGL(k_batch_index, dim3(std::min(L.nidx, L.grids.index)), dim3(kBatchBlock), 0,
                           s, d_desc);
    if (first <= 2 && last >= 2 && L.nins)
        hipLaunchKernelGGL(k_batch_insert, dim3(std::min(L.nins, L.grids.insert)), dim3(kBatchBlock),
                           0, s, d_desc);
    if (first > 3 || last < 3 || !L.nprobe) return;
    const dim3 g(std::min(L.nprobe, L.grids.probe)), b(kBatchBlock);
    if (L.coop) {
        if (L.unroll >= 4)
            hipLaunchKernelGGL((k_batch_probe<4, true>), g, b, 0, s, d_desc);
        else
            hipLaunchKernelGGL((k_batch_probe<2, true>), g, b, 0, s, d_desc);
    } else {
        if (L.unroll >= 4)
            hipLaunchKernelGGL((k_batch_probe<4, false>), g, b, 0, s, d_desc);
        else
            hipLaunchKernelGGL((k_batch_probe<2, false>), g, b, 0, s, d_desc);
    }
}

}  // namespace hj
