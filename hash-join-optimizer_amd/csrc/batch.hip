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
constexpr uint32_t kWaveStage = 2048;  // LDS survivor stage of one wave of a probe item (8 KiB)
static_assert(2 * kInsChunk + (1u << kMaxRadixBits) <= kStageWords, "insert sort fits the stage");

__device__ inline uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
template <class T>
__device__ inline T* uni_ptr(T* p) {
    return (T*)uni64((uint64_t)p);
}

// idx_col with every field made wave-uniform: the descriptors are read inside loops that also
// store to memory, so the compiler cannot prove them unchanged and would keep the buffer
// resources in vector registers (a loop of lane reads around every buffer load)
__device__ inline IdxCol uidx_col(const DCol& c, uint32_t rows) {
    const uint64_t* gidx = uni_ptr(c.gidx);
    return IdxCol{make_rsrc(uni_ptr(c.pages), uni(c.num_pages) * HJ_PAGE_BYTES),
                  make_rsrc(gidx, gidx ? ((rows + 63u) >> 6) * 16u : 0u),
                  gidx ? 0ull : uni64(c.umagic), uni(c.uniform_rows)};
}

// The rare continuations of a lookup (the home bucket is full) and of an insert (the home slot
// is taken) are calls, not inlined copies: one inlined copy per unrolled row group would set the
// kernel's register count for its common path too.
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
    HJ_GLOBAL uint32_t* K = gp(uni_ptr(J.table));
    HJ_GLOBAL uint32_t* qctr = gp(uni_ptr(Q.ctr));
    const IdxCol wk = uidx_col(J.bkey, rows);
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
            const IdxCol wf = uidx_col(J.bfcol[f], rows);
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

// An output column read by row id, its access path chosen once per flush (wave-uniform):
// a dense column of uniform pages by arithmetic, an indexed one through its group records,
// any other by page search
struct OutCol {
    const uint8_t* pages;
    const uint64_t* gidx;
    uint64_t umagic;
    uint32_t upr;
};
__device__ inline OutCol out_col(const DCol& c) {
    return OutCol{uni_ptr(c.pages), uni_ptr(c.gidx), uni64(c.umagic), uni(c.uniform_rows)};
}

// This wave's staged survivors (cnt of them, slot-major at stride Sw) written to the query's
// output at a range reserved with one atomic: each lane writes every 64th survivor's output
// columns (64 consecutive rows per store instruction). A column that may hold NULL gets its
// validity bits as the wave's ballot, OR-ed into the (at most three) bitmap words it spans.
__device__ inline void flush_wave(const BatchQuery& Q, const uint32_t* stage, uint32_t Sw,
                                  uint32_t cnt) {
    if (cnt == 0) return;
    __builtin_amdgcn_wave_barrier();
    const uint32_t lane = lane_id();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&Q.ctr[2], cnt);
    base = uni(base);
    const int nout = uni(Q.nout);
    const uint32_t onull = uni(Q.onull);
    for (int c = 0; c < nout; ++c) {
        const DCol& dc = Q.ocol[c];
        const OutCol oc = out_col(dc);
        const uint32_t slot = uni((uint32_t)Q.oslot[c]);
        HJ_GLOBAL int32_t* out = gp(uni_ptr(Q.out[c]));
        const bool nullable = (onull >> c) & 1u;
        for (uint32_t i0 = 0; i0 < cnt; i0 += HJ_WAVE) {
            const uint32_t i = i0 + lane;
            const bool in = i < cnt;
            const uint32_t row = in ? stage[slot * Sw + i] : 0u;
            int32_t v = 0;
            bool ok = false;
            if (in) {
                if (oc.umagic) {
                    const uint32_t p = dense_page(oc.umagic, row);
                    v = *(const HJ_GLOBAL int32_t*)(gp(oc.pages) + (size_t)p * HJ_PAGE_BYTES + 4u +
                                                    4u * (row - p * oc.upr));
                    ok = true;
                } else if (oc.gidx) {
                    ok = gidx_int32(oc.pages, oc.gidx, row, &v);
                } else {
                    ok = get_int32(dc, row, &v);
                }
                out[base + i] = ok ? v : 0;
            }
            if (nullable) {
                // bits base + i0 .. base + i0 + 63: the valid lanes, shifted into place
                const uint64_t m = __ballot(ok);
                const uint32_t b0 = base + i0, w0 = b0 >> 5, sh = b0 & 31u;
                HJ_GLOBAL uint32_t* vb = gp(uni_ptr(Q.ovalid[c]));
                if (lane < 3) {
                    const uint32_t part = lane == 0 ? (uint32_t)(m << sh)
                                         : lane == 1 ? (uint32_t)(m >> (32u - sh))
                                                     : (sh ? (uint32_t)(m >> (64u - sh)) : 0u);
                    if (part) atomicOr((uint32_t*)&vb[w0 + lane], part);
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// 64-row groups [lo, hi) of query Q's source, NJ joins. Every wave works on its own: iteration i
// takes groups lo + i * kBNW * U + w * U + u (u < U), one row per lane; survivors go to the
// wave's own LDS stage (source row, then the build row of each join) and are written out when
// the next iteration might not fit. The group records of every join's key column are loaded two
// iterations ahead and the keys one iteration ahead (two register sets, alternating), so that
// they are in flight while the current groups are looked up.
template <int U, int NJ, bool COOP>
__device__ inline void do_probe_n(const BatchQuery& Q, uint32_t lo, uint32_t hi, uint32_t* lds) {
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint32_t w = uni(t / HJ_WAVE);
    uint32_t* stage = lds + w * kWaveStage;
    const uint32_t Sw = kWaveStage / (1u + (uint32_t)NJ);
    const uint32_t nsrc = uni(Q.src_rows);
    const int nsf = uni(Q.nsf);
    const uint32_t lu = lane < (uint32_t)U ? lane : (uint32_t)U - 1u;
    constexpr uint32_t kStep = kBNW * U;  // groups per iteration of the block
    IdxCol wc[NJ];
    const HJ_GLOBAL uint32_t* K[NJ];
    uint32_t tk[NJ], sp[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        wc[j] = uidx_col(Q.j[j].pkey, nsrc);
        K[j] = gp(uni_ptr(Q.j[j].table));
        tk[j] = uni(Q.j[j].k);
        sp[j] = uni(Q.ctr[4 + j]);
    }
    uint32_t cnt = 0;  // survivors in this wave's stage (wave-uniform)
    int32_t ka[NJ][U], kb[NJ][U];
    u32x2 va[NJ], vb[NJ];
    u32x4 ra[NJ], rb[NJ];
    auto first = [&](uint32_t it) { return lo + it * kStep + w * U; };
    auto fetch_rec = [&](u32x4 (&rx)[NJ], uint32_t it) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) rx[j] = idx_rec_any(wc[j], first(it) + lu);
    };
    auto fetch = [&](int32_t (&kx)[NJ][U], u32x2 (&vx)[NJ], const u32x4 (&rx)[NJ]) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            vx[j] = u32x2{rx[j].x, rx[j].y};
#pragma unroll
            for (int u = 0; u < U; ++u) kx[j][u] = idx_load(wc[j], rx[j], u, rec_valid(rx[j], u), lane);
        }
    };
    auto step = [&](const int32_t (&key)[NJ][U], const u32x2 (&vw)[NJ], uint32_t it) {
        if (cnt + U * 64u > Sw) {
            flush_wave(Q, stage, Sw, cnt);
            cnt = 0;
        }
        uint32_t vl = ~0u, vh = ~0u;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            vl &= vw[j].x;
            vh &= vw[j].y;
        }
        const uint32_t g0 = first(it);
        uint64_t live[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t g = g0 + u;
            live[u] = ((g < hi && g * 64u < nsrc) ? range_mask(nsrc - g * 64u) : 0ull) &
                      lane_word(vl, vh, u);
        }
        for (int f = 0; f < nsf; ++f) {
            const IdxCol wf = uidx_col(Q.sfcol[f], nsrc);
            const u32x4 fr = idx_rec_any(wf, g0 + lu);
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
#pragma unroll
            for (int u = 0; u < U; ++u) {
                uint32_t sl = kDead;
                if (COOP) {
                    if (live[u]) sl = coop_lookup(K[j], tk[j], sp[j], (uint32_t)key[j][u], live[u]);
                } else if (in_mask(live[u])) {
                    sl = batch_lookup(K[j], tk[j], sp[j], (uint32_t)key[j][u]);
                }
                live[u] &= __ballot(sl != kDead);
                brow[j][u] = sl;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (in_mask(live[u])) brow[j][u] = K[j][row_word(brow[j][u])];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (in_mask(live[u])) {
                const uint32_t pos = cnt + mask_rank(live[u]);
                stage[pos] = (g0 + u) * 64u + lane;
#pragma unroll
                for (int j = 0; j < NJ; ++j) stage[(1u + (uint32_t)j) * Sw + pos] = brow[j][u];
            }
            cnt += (uint32_t)__popcll(live[u]);
        }
    };
    // issue order per iteration: the records two iterations ahead, the keys of the next
    // iteration (whose records were issued before the current keys: vector loads complete in
    // order), then the current groups
    uint32_t it = 0;
    fetch_rec(ra, 0);
    fetch_rec(rb, 1);
    fetch(ka, va, ra);
    while (first(it) < hi) {
        fetch_rec(ra, it + 2);
        fetch(kb, vb, rb);
        step(ka, va, it);
        ++it;
        if (first(it) >= hi) break;
        fetch_rec(rb, it + 2);
        fetch(ka, va, ra);
        step(kb, vb, it);
        ++it;
    }
    flush_wave(Q, stage, Sw, cnt);
}

// at most 8 row groups x joins unrolled per wave and round: U groups for up to two joins, at
// most 2 for three or four
// queries of one or two joins (WIDE false) and of three or four (WIDE true, at most 2 groups
// per wave and iteration) run in separate launches, so that the register count of the wide
// variant does not limit the common one
template <int U, bool COOP, bool WIDE>
__device__ inline void do_probe(const BatchQuery& Q, uint32_t lo, uint32_t hi, uint32_t* lds) {
    constexpr int U2 = U > 2 ? 2 : U;
    const int nj = uni(Q.nj);
    if (!WIDE) {
        if (nj == 1)
            do_probe_n<U, 1, COOP>(Q, lo, hi, lds);
        else
            do_probe_n<U, 2, COOP>(Q, lo, hi, lds);
    } else {
        if (nj == 3)
            do_probe_n<U2, 3, COOP>(Q, lo, hi, lds);
        else
            do_probe_n<U2, 4, COOP>(Q, lo, hi, lds);
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

// probe items [probe0, wide0) of queries with one or two joins, [wide0, nitems) of three or four
template <int U, bool COOP, bool WIDE>
__global__ __launch_bounds__(kBatchBlock, 5) void k_batch_probe(const BatchDesc* __restrict__ dp) {
    __shared__ uint32_t lds[kBNW * kWaveStage];
    const BatchDesc& d = *dp;
    const uint32_t i0 = uni(WIDE ? d.wide0 : d.probe0), n = uni(WIDE ? d.nitems : d.wide0);
    for (uint32_t it = i0 + blockIdx.x; it < n; it += gridDim.x) {
        const Item I = get_item(d, it);
        do_probe<U, COOP, WIDE>(d.q[I.q], I.lo, I.hi, lds);
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
    static const uint32_t p[4] = {per_cu_of(k_batch_probe<2, false, false>),
                                  per_cu_of(k_batch_probe<4, false, false>),
                                  per_cu_of(k_batch_probe<2, true, false>),
                                  per_cu_of(k_batch_probe<4, true, false>)};
    static const uint32_t wide[2] = {per_cu_of(k_batch_probe<2, false, true>),
                                     per_cu_of(k_batch_probe<2, true, true>)};
    return BatchGrids{idx * num_cus, ins * num_cus,
                      p[(coop ? 2 : 0) + (unroll >= 4 ? 1 : 0)] * num_cus,
                      wide[coop ? 1 : 0] * num_cus};
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
    if (first > 3 || last < 3) return;
    const dim3 b(kBatchBlock);
    if (L.nprobe) {
        const dim3 g(std::min(L.nprobe, L.grids.probe));
        if (L.coop) {
            if (L.unroll >= 4)
                hipLaunchKernelGGL((k_batch_probe<4, true, false>), g, b, 0, s, d_desc);
            else
                hipLaunchKernelGGL((k_batch_probe<2, true, false>), g, b, 0, s, d_desc);
        } else {
            if (L.unroll >= 4)
                hipLaunchKernelGGL((k_batch_probe<4, false, false>), g, b, 0, s, d_desc);
            else
                hipLaunchKernelGGL((k_batch_probe<2, false, false>), g, b, 0, s, d_desc);
        }
    }
    if (L.nwide) {
        const dim3 g(std::min(L.nwide, L.grids.wide));
        if (L.coop)
            hipLaunchKernelGGL((k_batch_probe<2, true, true>), g, b, 0, s, d_desc);
        else
            hipLaunchKernelGGL((k_batch_probe<2, false, true>), g, b, 0, s, d_desc);
    }
}

}  // namespace hj
