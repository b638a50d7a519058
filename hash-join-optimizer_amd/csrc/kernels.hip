// kernels.hip - CDNA4 (gfx950, wave64) kernels of the general join path.
//
// Integer / byte work only: no MFMA, no floating point. Every kernel is HBM-bound; the comments
// give the bytes a launch must move (the roofline the kernel is held to).
#include "kernels.hpp"
#include "gindex.hpp"

#include <algorithm>

namespace hj {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kRelTile = 1024;  // rows per block for intermediate relations

struct OpSum {
    template <class T>
    __device__ static T op(T a, T b) {
        return a + b;
    }
};
struct OpMax {
    template <class T>
    __device__ static T op(T a, T b) {
        return a > b ? a : b;
    }
};

template <class Op, class T>
__device__ inline T wave_incl_op(T x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < HJ_WAVE; d <<= 1) {
        T y = __shfl_up(x, d, HJ_WAVE);
        if (lane >= (uint32_t)d) x = Op::op(x, y);
    }
    return x;
}

// exclusive block scan with Op (identity 0 for unsigned sum/max)
template <class Op, class T>
__device__ inline T block_excl_op(T v, T* lds, T* total) {
    constexpr int NW = kBlock / HJ_WAVE;
    const uint32_t w = threadIdx.x / HJ_WAVE;
    T incl = wave_incl_op<Op>(v);
    T excl_w = __shfl_up(incl, 1, HJ_WAVE);
    if (lane_id() == 0) excl_w = 0;
    if (lane_id() == HJ_WAVE - 1) lds[w] = incl;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        T t = lds[i];
        if (i < (int)w) off = Op::op(off, t);
        tot = Op::op(tot, t);
    }
    __syncthreads();
    *total = tot;
    return Op::op(off, excl_w);
}

// wave-ballot stream compaction: position of this lane among the block's `pred` lanes
__device__ inline uint32_t block_ballot_prefix(bool pred, uint32_t* lds, uint32_t* total) {
    constexpr int NW = kBlock / HJ_WAVE;
    const uint32_t w = threadIdx.x / HJ_WAVE;
    uint64_t b = __ballot(pred);
    if (lane_id() == 0) lds[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint32_t t = lds[i];
        off += i < (int)w ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + (uint32_t)__popcll(b & lanemask_lt());
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (!n) return;
    uint64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(kBlock), 0, s, p, n, v);
}

}  // namespace

void zero_u32(uint32_t* p, uint64_t n, hipStream_t s) { fill_u32(p, n, 0u, s); }

// =====================================================================================
// Page index of a resident column: rank[group_start[p] + g] = non-NULL rows of page p before
// 64-row group g. One wave per page; reads the bitmaps (1 bit per row).
// =====================================================================================
__global__ __launch_bounds__(64) void k_rank_index(DCol c, uint16_t* rank_out) {
    const uint32_t p = blockIdx.x;
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
    if (c.type == 3 /*VARCHAR*/ && (nr == 0xffffu || nr == 0xfffeu)) return;
    if (nr == nn) return;
    const uint32_t nb = (nr + 7) >> 3;
    const uint32_t ngroups = (nr + 63) >> 6;
    const uint32_t gs = c.group_start[p];
    uint32_t carry = 0;
    for (uint32_t g0 = 0; g0 < ngroups; g0 += HJ_WAVE) {
        uint32_t g = g0 + lane_id();
        uint32_t pc = g < ngroups ? (uint32_t)__popcll(group_bits(pg, nb, g)) : 0u;
        uint32_t incl = wave_incl_scan(pc);
        if (g < ngroups) rank_out[gs + g] = (uint16_t)(carry + incl - pc);
        carry += __shfl(incl, HJ_WAVE - 1, HJ_WAVE);
    }
}

void launch_rank_index(const DCol& c, uint16_t* rank_out, hipStream_t s) {
    if (c.num_pages == 0 || c.all_dense) return;
    hipLaunchKernelGGL(k_rank_index, dim3(c.num_pages), dim3(64), 0, s, c, rank_out);
}

// INT32 pages -> row-space mirror: one block per page, one row per thread; each wave's 64
// validity bits are or-ed into the (zeroed) bitmap at the page's first row, which need not be
// 64-aligned, so a wave touches up to two words
__global__ __launch_bounds__(kBlock) void k_decode_int32(DCol c, int32_t* flat, uint64_t* vbits) {
    const uint32_t p = blockIdx.x;
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    const uint32_t nr = hdr & 0xffffu;
    const uint32_t r0 = c.row_start[p];
    for (uint32_t l0 = 0; l0 < nr; l0 += kBlock) {
        const uint32_t l = l0 + threadIdx.x;
        int32_t v = 0;
        const bool ok = l < nr && get_int32_in_page(c, p, pg, hdr, l, &v);
        if (l < nr) flat[r0 + l] = ok ? v : 0;
        if (vbits) {
            const uint64_t bits = __ballot(ok);
            if (lane_id() == 0 && bits) {
                const uint32_t rb = r0 + l0 + (threadIdx.x & ~(uint32_t)(HJ_WAVE - 1));
                const uint32_t w0 = rb >> 6, sh = rb & 63u;
                atomicOr((unsigned long long*)&vbits[w0], (unsigned long long)(bits << sh));
                if (sh && (bits >> (64u - sh)))
                    atomicOr((unsigned long long*)&vbits[w0 + 1],
                             (unsigned long long)(bits >> (64u - sh)));
            }
        }
    }
}

// VARCHAR pages -> one reference per row: one block per page, one row per thread
__global__ __launch_bounds__(kBlock) void k_decode_varchar(DCol c, uint64_t* sref) {
    const uint32_t p = blockIdx.x;
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
    const uint32_t r0 = c.row_start[p];
    if (nr == 0xfffeu) return;  // continuation page: no row of its own
    if (nr == 0xffffu) {
        if (threadIdx.x == 0) sref[r0] = kLongRef | p;
        return;
    }
    const HJ_GLOBAL uint16_t* offs = (const HJ_GLOBAL uint16_t*)(pg + 4);
    const uint64_t base = (uint64_t)p * HJ_PAGE_BYTES + 4u + 2u * nn;
    for (uint32_t l = threadIdx.x; l < nr; l += kBlock) {
        uint32_t vi;
        uint64_t ref = kNullRef;
        if (value_index(c, p, pg, hdr, l, &vi)) {
            const uint32_t start = vi ? offs[vi - 1] : 0u;
            ref = ((base + start) << 16) | (uint64_t)(offs[vi] - start);
        }
        sref[r0 + l] = ref;
    }
}

static_assert(kGroupIndexBlock == kBlock, "one group per thread");
__global__ __launch_bounds__(kBlock) void k_group_index(GroupIndexBatch b) {
    group_index_thread(b, blockIdx.x, threadIdx.x);
}

void launch_group_index(const GroupIndexBatch& b, hipStream_t s) {
    if (b.n <= 0 || b.blk0[b.n] == 0) return;
    hipLaunchKernelGGL(k_group_index, dim3(b.blk0[b.n]), dim3(kBlock), 0, s, b);
}

void launch_decode_varchar(const DCol& c, uint64_t* sref, hipStream_t s) {
    if (c.num_pages == 0) return;
    hipLaunchKernelGGL(k_decode_varchar, dim3(c.num_pages), dim3(kBlock), 0, s, c, sref);
}

void launch_decode_int32(const DCol& c, int32_t* flat, uint64_t* vbits, hipStream_t s) {
    if (c.num_pages == 0) return;
    hipLaunchKernelGGL(k_decode_int32, dim3(c.num_pages), dim3(kBlock), 0, s, c, flat, vbits);
}

// =====================================================================================
// Exclusive scans: reduce tiles -> scan tile partials in one block -> rescan tiles.
// A tile is 4096 elements (256 threads x 16), staged through LDS for coalesced access.
// =====================================================================================
namespace {
constexpr int kScanItems = 16;
constexpr uint64_t kScanTile = (uint64_t)kBlock * kScanItems;

template <class Op, class T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T* in, uint64_t n, T* partial) {
    __shared__ T lds[kBlock / HJ_WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
        if (idx < n) acc = Op::op(acc, in[idx]);
    }
    T tot;
    block_excl_op<Op>(acc, lds, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// in-place exclusive scan of partial[0..np) by one block
template <class Op, class T>
__global__ __launch_bounds__(kBlock) void k_scan_partials(T* partial, uint64_t np) {
    __shared__ T lds[kBlock / HJ_WAVE];
    T carry = 0;
    for (uint64_t base = 0; base < np; base += kScanTile) {
        T v[kScanItems];
        T acc = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + i;
            v[i] = idx < np ? partial[idx] : (T)0;
            acc = Op::op(acc, v[i]);
        }
        T tot;
        T ex = block_excl_op<Op>(acc, lds, &tot);
        ex = Op::op(carry, ex);
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + i;
            if (idx < np) partial[idx] = ex;
            ex = Op::op(ex, v[i]);
        }
        carry = Op::op(carry, tot);
    }
}

template <class T>
struct OutArray {
    T* out;
    T* total;
    __device__ void put(uint64_t i, T v) const { out[i] = v; }
    __device__ void last(uint64_t n, T incl) const {
        if (total) *total = incl;
    }
};

struct OutDir {
    uint64_t* dir;
    const uint32_t* bloom;
    uint64_t nbuild;
    __device__ void put(uint64_t i, uint32_t start) const {
        dir[i] = ((uint64_t)start << 16) | (uint64_t)(bloom[i] & 0xffffu);
    }
    __device__ void last(uint64_t n, uint32_t) const { dir[n] = nbuild << 16; }
};

template <class Op, class T, class Out>
__global__ __launch_bounds__(kBlock) void k_scan_down(const T* in, uint64_t n, const T* partial,
                                                      Out out) {
    __shared__ T buf[kScanTile];
    __shared__ T lds[kBlock / HJ_WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
        buf[i * kBlock + threadIdx.x] = idx < n ? in[idx] : (T)0;
    }
    __syncthreads();
    T v[kScanItems];
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = buf[threadIdx.x * kScanItems + i];
        acc = Op::op(acc, v[i]);
    }
    T tot;
    T ex = block_excl_op<Op>(acc, lds, &tot);
    ex = Op::op(partial ? partial[blockIdx.x] : (T)0, ex);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        buf[threadIdx.x * kScanItems + i] = ex;
        if (base + (uint64_t)threadIdx.x * kScanItems + i == n - 1) out.last(n, Op::op(ex, v[i]));
        ex = Op::op(ex, v[i]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
        if (idx < n) out.put(idx, buf[i * kBlock + threadIdx.x]);
    }
}

template <class Op, class T, class Out>
void scan_generic(const T* in, uint64_t n, T* partial, Out out, hipStream_t s) {
    if (n == 0) return;
    uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles > 1) {
        hipLaunchKernelGGL((k_scan_reduce<Op, T>), dim3((uint32_t)tiles), dim3(kBlock), 0, s, in,
                           n, partial);
        hipLaunchKernelGGL((k_scan_partials<Op, T>), dim3(1), dim3(kBlock), 0, s, partial,
                           tiles);
        hipLaunchKernelGGL((k_scan_down<Op, T, Out>), dim3((uint32_t)tiles), dim3(kBlock), 0, s,
                           in, n, (const T*)partial, out);
    } else {
        hipLaunchKernelGGL((k_scan_down<Op, T, Out>), dim3(1), dim3(kBlock), 0, s, in, n,
                           (const T*)nullptr, out);
    }
}
}  // namespace

size_t scan_partials_needed(uint64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

void scan_excl_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partial,
                       uint32_t* total, hipStream_t s) {
    scan_generic<OpSum>(in, n, partial, OutArray<uint32_t>{out, total}, s);
}
void scan_excl_sum_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* partial,
                       uint64_t* total, hipStream_t s) {
    scan_generic<OpSum>(in, n, partial, OutArray<uint64_t>{out, total}, s);
}
void scan_dir_words(const uint32_t* e, const uint32_t* bloom, uint64_t* dir, uint64_t D,
                    uint64_t n, uint32_t* partial, hipStream_t s) {
    scan_generic<OpMax>(e, D, partial, OutDir{dir, bloom, n}, s);
}

// =====================================================================================
// LSD radix sort of (key, value) pairs, 8-bit digits. Per pass: tile histograms, one scan over
// the digit-major histogram, then a stable scatter: each wave owns a contiguous 512-key
// segment of the 2048-key tile, ranks keys by digit with 8 ballots (peer masks), and the tile
// is reordered by digit in LDS before the coalesced write-out. Moves 4 B read (hist) +
// 16 B read/write (scatter) per pair per pass.
// =====================================================================================
namespace {
constexpr uint32_t kSortTile = 2048;

__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* keys, uint64_t n,
                                                       uint32_t shift, uint32_t* hist,
                                                       uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int i = 0; i < (int)(kSortTile / kBlock); ++i) {
        uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* kin, const uint32_t* vin,
                                                          uint32_t* kout, uint32_t* vout,
                                                          uint64_t n, uint32_t shift,
                                                          const uint32_t* gofs, uint32_t nblocks) {
    constexpr int NW = kBlock / HJ_WAVE;                 // 4 waves
    constexpr int ITEMS = kSortTile / kBlock;            // 8 keys per lane
    constexpr uint32_t SEG = kSortTile / NW;             // 512 keys per wave
    __shared__ uint32_t cnt[NW][256];
    __shared__ uint32_t woff[NW][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t kst[kSortTile];
    __shared__ uint32_t vst[kSortTile];
    __shared__ uint32_t scan_lds[NW];
    const uint32_t t = threadIdx.x, w = t / HJ_WAVE, lane = lane_id();
    for (uint32_t i = t; i < NW * 256; i += kBlock) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    uint32_t k[ITEMS], v[ITEMS], rk[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)w * SEG + (uint64_t)j * HJ_WAVE + lane;
        const bool valid = idx < n;
        k[j] = valid ? kin[idx] : 0u;
        v[j] = valid ? vin[idx] : 0u;
        const uint32_t d = (k[j] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(m & lanemask_lt());
        uint32_t pre = 0;
        if (valid) pre = cnt[w][d];
        if (valid && below == 0) cnt[w][d] = pre + (uint32_t)__popcll(m);
        rk[j] = pre + below;
    }
    __syncthreads();
    {
        const uint32_t d = t;
        uint32_t run = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            woff[i][d] = run;
            run += cnt[i][d];
        }
        uint32_t blocktot;
        dstart[d] = block_excl_op<OpSum>(run, scan_lds, &blocktot);
        gbase[d] = gofs[(uint64_t)d * nblocks + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)w * SEG + (uint64_t)j * HJ_WAVE + lane;
        if (idx < n) {
            const uint32_t d = (k[j] >> shift) & 255u;
            const uint32_t pos = dstart[d] + woff[w][d] + rk[j];
            kst[pos] = k[j];
            vst[pos] = v[j];
        }
    }
    __syncthreads();
    const uint32_t tile_n = (uint32_t)(n - base < kSortTile ? n - base : kSortTile);
    for (uint32_t i = t; i < tile_n; i += kBlock) {
        const uint32_t kk = kst[i];
        const uint32_t d = (kk >> shift) & 255u;
        const uint64_t g = (uint64_t)gbase[d] + (i - dstart[d]);
        kout[g] = kk;
        vout[g] = vst[i];
    }
}
}  // namespace

size_t radix_sort_scratch_bytes(uint64_t n) {
    uint64_t nblocks = (n + kSortTile - 1) / kSortTile;
    uint64_t hist = 256 * nblocks;
    return (size_t)(hist * 4 + scan_partials_needed(hist) * 4 + 512);
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                      uint64_t n, void* scratch, hipStream_t s) {
    if (n <= 1) return;
    const uint32_t nblocks = (uint32_t)((n + kSortTile - 1) / kSortTile);
    const uint64_t nh = 256ull * nblocks;
    uint32_t* hist = (uint32_t*)scratch;
    uint32_t* partial = hist + ((nh + 63) & ~63ull);
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kBlock), 0, s, ki, n, shift, hist,
                           nblocks);
        scan_excl_sum_u32(hist, hist, nh, partial, nullptr, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kBlock), 0, s, ki, vi, ko, vo, n,
                           shift, hist, nblocks);
        uint32_t* tk = ki;
        ki = ko;
        ko = tk;
        uint32_t* tv = vi;
        vi = vo;
        vo = tv;
    }
    // 4 passes: the sorted data is back in keys/vals
}

// =====================================================================================
// Unchained directory over hashes sorted ascending: e[s] = end of slot s (written by the slot's
// last tuple), bloom[s] |= mask of each distinct key of the slot (written by the first tuple of
// each equal-hash run, so a heavy key costs one atomic); then dir = exclusive prefix-max of e
// (start of every slot, empty slots included) packed with the Bloom bits.
// =====================================================================================
namespace {
__global__ void k_dir_marks(const uint32_t* sh, uint64_t n, uint32_t k, uint32_t* e,
                            uint32_t* bloom) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = sh[i];
    const uint32_t s = slot_of(h, k);
    if (i + 1 == n || slot_of(sh[i + 1], k) != s) e[s] = (uint32_t)(i + 1);
    if (i == 0 || sh[i - 1] != h) atomicOr(&bloom[s], bloom_mask(bloom_tag(h)));
}
}  // namespace

void build_directory(const uint32_t* sh, uint64_t n, uint32_t k, uint64_t* dir, uint32_t* e,
                     uint32_t* bloom, uint32_t* partial, hipStream_t s) {
    const uint64_t D = 1ull << k;
    fill_u32(e, D, 0, s);
    fill_u32(bloom, D, 0, s);
    if (n) {
        hipLaunchKernelGGL(k_dir_marks, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, s, sh, n, k, e, bloom);
    }
    scan_dir_words(e, bloom, dir, D, n, partial, s);
}

// =====================================================================================
// Key lookup in the directory: Bloom reject, then the run of equal hashes inside the slot
// (linear for short slots, binary search for long ones, so a Zipf heavy slot costs O(log n)).
// Returns the match count and the first matching tuple.
// =====================================================================================
__device__ inline uint32_t dir_lookup(const DirView& d, uint32_t key, uint32_t* lo) {
    const uint32_t h = fib32(key);
    const uint32_t s = slot_of(h, d.k);
    const uint64_t w = d.dir[s];
    const uint32_t m = bloom_mask(bloom_tag(h));
    if (((uint32_t)w & m) != m) return 0;
    const uint64_t b = w >> 16, e = d.dir[s + 1] >> 16;
    if (e - b <= 8) {
        for (uint64_t j = b; j < e; ++j) {
            const uint32_t x = d.sh[j];
            if (x == h) {
                uint64_t c = 1;
                while (j + c < e && d.sh[j + c] == h) ++c;
                *lo = (uint32_t)j;
                return (uint32_t)c;
            }
            if (x > h) break;
        }
        return 0;
    }
    uint64_t L = b, R = e;
    while (L < R) {
        uint64_t mid = (L + R) >> 1;
        if (d.sh[mid] < h)
            L = mid + 1;
        else
            R = mid;
    }
    if (L == e || d.sh[L] != h) return 0;
    uint64_t L2 = L, R2 = e;
    while (L2 < R2) {
        uint64_t mid = (L2 + R2) >> 1;
        if (d.sh[mid] <= h)
            L2 = mid + 1;
        else
            R2 = mid;
    }
    *lo = (uint32_t)L;
    return (uint32_t)(L2 - L);
}

// key of local row `local` of key page p, after the scan's filters; false = row dropped
__device__ inline bool scan_row_key(const ScanSrc& s, uint32_t p, const HJ_GLOBAL uint8_t* pg, uint32_t hdr,
                                    uint32_t local, uint32_t r, uint32_t* key) {
    int32_t k;
    // through the query's group index where the column has one: one record and one value load
    if (!(s.key.gidx ? gidx_int32(s.key.pages, s.key.gidx, r, &k)
                     : get_int32_in_page(s.key, p, pg, hdr, local, &k)))
        return false;
    for (int f = 0; f < s.nf; ++f) {
        int32_t v;
        if (!get_int32(s.fcol[f], r, &v)) return false;
        if (v < s.flo[f] || v > s.fhi[f]) return false;
    }
    *key = (uint32_t)k;
    return true;
}

__device__ inline bool rel_row_key(const RelSrc& s, uint64_t i, uint32_t* key) {
    int32_t k;
    if (!get_int32(s.key, s.rid[i], &k)) return false;
    *key = (uint32_t)k;
    return true;
}

// output row pos of probe row prow and the build row at directory position dpos
__device__ inline void write_out(const OutSlots& o, const DirView& d, uint64_t pos, uint32_t prow,
                                 uint32_t dpos) {
    uint32_t brow = 0;
    bool have_brow = false;
    for (int s = 0; s < o.n; ++s) {
        if (o.values && o.side[s] && o.bval[s]) {  // in directory order: read in place
            o.dst[s][pos] = (uint32_t)o.bval[s][dpos];
            if (o.valid8[s]) o.valid8[s][pos] = o.bvalid[s][dpos];
            continue;
        }
        if (o.side[s] && !have_brow) {
            brow = d.srow[dpos];
            have_brow = true;
        }
        const uint32_t src_row = o.side[s] ? brow : prow;
        const uint32_t* m = o.src[s];
        const uint32_t row = m ? m[src_row] : src_row;
        if (!o.values) {
            o.dst[s][pos] = row;
            continue;
        }
        int32_t v = 0;
        bool ok;
        ok = get_int32(o.col[s], row, &v);
        o.dst[s][pos] = ok ? (uint32_t)v : 0u;
        if (o.valid8[s]) o.valid8[s][pos] = ok ? 1 : 0;  // none for a column without NULL
    }
}

__device__ inline void emit(const OutSlots& o, const DirView& d, uint32_t prow, uint32_t lo,
                            uint32_t cnt, uint64_t off, HeavyEntry* heavy, uint32_t* heavy_count) {
    if (cnt <= kLightMax) {
        for (uint32_t j = 0; j < cnt; ++j) write_out(o, d, off + j, prow, lo + j);
    } else {
        const uint32_t nch = (cnt + kHeavyChunk - 1) / kHeavyChunk;
        const uint32_t idx = atomicAdd(heavy_count, nch);
        for (uint32_t c = 0; c < nch; ++c) {
            HeavyEntry h;
            h.prow = prow;
            h.lo = lo + c * kHeavyChunk;
            h.cnt = cnt - c * kHeavyChunk < kHeavyChunk ? cnt - c * kHeavyChunk : kHeavyChunk;
            h.pad = 0;
            h.out = off + (uint64_t)c * kHeavyChunk;
            heavy[idx + c] = h;
        }
    }
}

// The light matches of a wave's 64 probe rows, written by the whole wave: match e of the wave
// (in row order) goes to lane e mod 64, which finds its row by a binary search over the rows'
// inclusive match counts in LDS. Output positions of one row are consecutive and rows follow
// each other, so the writes are coalesced whatever the rows' match counts; heavy rows go to the
// heavy list as before. Called by every thread of the block (it holds block barriers).
// probe-side INT32 output columns whose values a wave reads once per probe row (value mode)
constexpr int kWaveVals = 4;
struct WaveRows {
    uint32_t incl[HJ_WAVE];
    uint32_t lo[HJ_WAVE];
    uint32_t prow[HJ_WAVE];
    uint64_t off[HJ_WAVE];
    int32_t pv[kWaveVals][HJ_WAVE];  // the row's value of each of the first kWaveVals probe-side
    uint8_t pok[kWaveVals][HJ_WAVE];  // output columns, and its validity
};

// write_out for match dpos of the wave's row i, whose probe-side values are in wr
__device__ inline void write_out_wave(const OutSlots& o, const DirView& d, uint64_t pos,
                                      uint32_t prow, uint32_t dpos, const WaveRows* wr,
                                      uint32_t i) {
    uint32_t brow = 0;
    bool have_brow = false;
    int pc = 0;  // probe-side value columns so far
    for (int s = 0; s < o.n; ++s) {
        if (o.values && !o.side[s] && pc < kWaveVals) {
            o.dst[s][pos] = (uint32_t)wr->pv[pc][i];
            if (o.valid8[s]) o.valid8[s][pos] = wr->pok[pc][i];
            ++pc;
            continue;
        }
        if (o.values && !o.side[s]) ++pc;
        if (o.values && o.side[s] && o.bval[s]) {
            o.dst[s][pos] = (uint32_t)o.bval[s][dpos];
            if (o.valid8[s]) o.valid8[s][pos] = o.bvalid[s][dpos];
            continue;
        }
        if (o.side[s] && !have_brow) {
            brow = d.srow[dpos];
            have_brow = true;
        }
        const uint32_t src_row = o.side[s] ? brow : prow;
        const uint32_t* m = o.src[s];
        const uint32_t row = m ? m[src_row] : src_row;
        if (!o.values) {
            o.dst[s][pos] = row;
            continue;
        }
        int32_t v = 0;
        const bool ok = get_int32(o.col[s], row, &v);
        o.dst[s][pos] = ok ? (uint32_t)v : 0u;
        if (o.valid8[s]) o.valid8[s][pos] = ok ? 1 : 0;
    }
}

__device__ inline void emit_wave(const OutSlots& o, const DirView& d, uint32_t prow, uint32_t lo,
                                 uint32_t cnt, uint64_t off, HeavyEntry* heavy,
                                 uint32_t* heavy_count, WaveRows* wr) {
    const uint32_t lane = lane_id();
    if (cnt > kLightMax) emit(o, d, prow, lo, cnt, off, heavy, heavy_count);
    const uint32_t cl = cnt <= kLightMax ? cnt : 0u;
    const uint32_t incl = wave_incl_scan(cl);
    const uint32_t total = __shfl(incl, HJ_WAVE - 1, HJ_WAVE);
    wr->incl[lane] = incl;
    wr->lo[lane] = lo;
    wr->prow[lane] = prow;
    wr->off[lane] = off;
    if (o.values) {  // each light row's probe-side values, read once by its own lane
        int pc = 0;
        for (int s = 0; s < o.n && pc < kWaveVals; ++s) {
            if (o.side[s]) continue;
            int32_t v = 0;
            bool ok = false;
            if (cl) {
                const uint32_t* m = o.src[s];
                ok = get_int32(o.col[s], m ? m[prow] : prow, &v);
            }
            wr->pv[pc][lane] = ok ? v : 0;
            wr->pok[pc][lane] = ok ? 1 : 0;
            ++pc;
        }
    }
    __syncthreads();
    for (uint32_t e = lane; e < total; e += HJ_WAVE) {
        uint32_t i = 0;  // rows whose matches all come before match e
#pragma unroll
        for (uint32_t step = HJ_WAVE / 2; step; step >>= 1)
            if (wr->incl[i + step - 1] <= e) i += step;
        const uint32_t before = i ? wr->incl[i - 1] : 0u;
        const uint32_t j = e - before;
        write_out_wave(o, d, wr->off[i] + j, wr->prow[i], wr->lo[i] + j, wr, i);
    }
    __syncthreads();
}

// =====================================================================================
// Build-side key extraction: one block per key page (base scan) or per 1024 rows (relation).
// Count pass writes per-block survivors; write pass compacts (h, row) with wave ballots.
// =====================================================================================
namespace {
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_extract_scan(ScanSrc src, uint32_t* counts,
                                                         const uint32_t* offsets, uint32_t* out_h,
                                                         uint32_t* out_row) {
    __shared__ uint32_t lds[kBlock / HJ_WAVE];
    const uint32_t p = blockIdx.x;
    const HJ_GLOBAL uint8_t* pg = page_ptr(src.key, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    const uint32_t nr = hdr & 0xffffu;
    const uint32_t r0 = src.key.row_start[p];
    uint32_t run = WRITE ? offsets[p] : 0u;
    for (uint32_t l0 = 0; l0 < nr; l0 += kBlock) {
        const uint32_t l = l0 + threadIdx.x;
        uint32_t key = 0;
        const bool keep = l < nr && scan_row_key(src, p, pg, hdr, l, r0 + l, &key);
        uint32_t tot;
        const uint32_t pos = block_ballot_prefix(keep, lds, &tot);
        if (WRITE && keep) {
            out_h[run + pos] = fib32(key);
            out_row[run + pos] = r0 + l;
        }
        run += tot;
    }
    if (!WRITE && threadIdx.x == 0) counts[p] = run;
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_extract_rel(RelSrc src, uint32_t* counts,
                                                        const uint32_t* offsets, uint32_t* out_h,
                                                        uint32_t* out_row) {
    __shared__ uint32_t lds[kBlock / HJ_WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * kRelTile;
    uint32_t run = WRITE ? offsets[blockIdx.x] : 0u;
    for (uint32_t l0 = 0; l0 < kRelTile; l0 += kBlock) {
        const uint64_t i = base + l0 + threadIdx.x;
        uint32_t key = 0;
        const bool keep = i < src.n && rel_row_key(src, i, &key);
        uint32_t tot;
        const uint32_t pos = block_ballot_prefix(keep, lds, &tot);
        if (WRITE && keep) {
            out_h[run + pos] = fib32(key);
            out_row[run + pos] = (uint32_t)i;
        }
        run += tot;
    }
    if (!WRITE && threadIdx.x == 0) counts[blockIdx.x] = run;
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_select_scan(ScanSrc src, uint32_t* counts,
                                                        const uint32_t* offsets,
                                                        uint32_t* out_row) {
    __shared__ uint32_t lds[kBlock / HJ_WAVE];
    const uint32_t p = blockIdx.x;
    const HJ_GLOBAL uint8_t* pg = page_ptr(src.key, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    uint32_t nr = hdr & 0xffffu;
    if (src.key.type == 3 && nr == 0xffffu) nr = 1;
    if (src.key.type == 3 && nr == 0xfffeu) nr = 0;
    const uint32_t r0 = src.key.row_start[p];
    uint32_t run = WRITE ? offsets[p] : 0u;
    for (uint32_t l0 = 0; l0 < nr; l0 += kBlock) {
        const uint32_t l = l0 + threadIdx.x;
        bool keep = l < nr;
        for (int f = 0; keep && f < src.nf; ++f) {
            int32_t v;
            keep = get_int32(src.fcol[f], r0 + l, &v) && v >= src.flo[f] && v <= src.fhi[f];
        }
        uint32_t tot;
        const uint32_t pos = block_ballot_prefix(keep, lds, &tot);
        if (WRITE && keep) out_row[run + pos] = r0 + l;
        run += tot;
    }
    if (!WRITE && threadIdx.x == 0) counts[p] = run;
}
}  // namespace

uint32_t extract_scan_blocks(const ScanSrc& src) { return src.key.num_pages; }
uint32_t rel_blocks(uint64_t n) { return (uint32_t)((n + kRelTile - 1) / kRelTile); }

void extract_scan_count(const ScanSrc& src, uint32_t* counts, hipStream_t s) {
    if (!src.key.num_pages) return;
    hipLaunchKernelGGL(k_extract_scan<false>, dim3(src.key.num_pages), dim3(kBlock), 0, s, src,
                       counts, nullptr, nullptr, nullptr);
}
void extract_scan_write(const ScanSrc& src, const uint32_t* offsets, uint32_t* out_h,
                        uint32_t* out_row, hipStream_t s) {
    if (!src.key.num_pages) return;
    hipLaunchKernelGGL(k_extract_scan<true>, dim3(src.key.num_pages), dim3(kBlock), 0, s, src,
                       nullptr, offsets, out_h, out_row);
}
void extract_rel_count(const RelSrc& src, uint32_t* counts, hipStream_t s) {
    if (!src.n) return;
    hipLaunchKernelGGL(k_extract_rel<false>, dim3(rel_blocks(src.n)), dim3(kBlock), 0, s, src,
                       counts, nullptr, nullptr, nullptr);
}
void extract_rel_write(const RelSrc& src, const uint32_t* offsets, uint32_t* out_h,
                       uint32_t* out_row, hipStream_t s) {
    if (!src.n) return;
    hipLaunchKernelGGL(k_extract_rel<true>, dim3(rel_blocks(src.n)), dim3(kBlock), 0, s, src,
                       nullptr, offsets, out_h, out_row);
}
void select_scan_count(const ScanSrc& src, uint32_t* counts, hipStream_t s) {
    if (!src.key.num_pages) return;
    hipLaunchKernelGGL(k_select_scan<false>, dim3(src.key.num_pages), dim3(kBlock), 0, s, src,
                       counts, nullptr, nullptr);
}
void select_scan_write(const ScanSrc& src, const uint32_t* offsets, uint32_t* out_row,
                       hipStream_t s) {
    if (!src.key.num_pages) return;
    hipLaunchKernelGGL(k_select_scan<true>, dim3(src.key.num_pages), dim3(kBlock), 0, s, src,
                       nullptr, offsets, out_row);
}

// =====================================================================================
// Probe: one lane per probe tuple. Count pass: per-block output totals (u64). Write pass: block
// prefix sums of the per-lane match counts give each lane its output offset; lanes with <= 64
// matches write their rows directly, heavier lanes hand 64K-output chunks to expand_heavy.
// =====================================================================================
namespace {
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_probe_scan(ScanSrc src, DirView d, uint64_t* counts,
                                                       const uint64_t* offsets, OutSlots o,
                                                       HeavyEntry* heavy, uint32_t* heavy_count,
                                                       uint64_t* match) {
    __shared__ uint64_t lds[kBlock / HJ_WAVE];
    const uint32_t p = blockIdx.x;
    const HJ_GLOBAL uint8_t* pg = page_ptr(src.key, p);
    const uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    const uint32_t nr = hdr & 0xffffu;
    const uint32_t r0 = src.key.row_start[p];
    if (!WRITE) {
        // kCU rows per thread and round: their keys first, then their lookups, then the
        // stores, so that the rows' load chains are in flight together
        constexpr int kCU = 4;
        uint64_t acc = 0;
        for (uint32_t l0 = threadIdx.x; l0 < nr; l0 += kBlock * kCU) {
            uint32_t key[kCU], lo[kCU], cnt[kCU];
            bool ok[kCU];
#pragma unroll
            for (int u = 0; u < kCU; ++u) {
                const uint32_t l = l0 + u * kBlock;
                key[u] = 0;
                ok[u] = l < nr && scan_row_key(src, p, pg, hdr, l, r0 + l, &key[u]);
            }
#pragma unroll
            for (int u = 0; u < kCU; ++u) {
                lo[u] = 0;
                cnt[u] = ok[u] ? dir_lookup(d, key[u], &lo[u]) : 0u;
            }
#pragma unroll
            for (int u = 0; u < kCU; ++u) {
                const uint32_t l = l0 + u * kBlock;
                acc += cnt[u];
                if (match && l < nr) match[r0 + l] = cnt[u] ? ((uint64_t)lo[u] << 32 | cnt[u]) : 0ull;
            }
        }
        const uint64_t tot = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) counts[p] = tot;
        return;
    }
    __shared__ WaveRows wr[kBlock / HJ_WAVE];
    uint64_t run = offsets[p];
    for (uint32_t l0 = 0; l0 < nr; l0 += kBlock) {
        const uint32_t l = l0 + threadIdx.x;
        uint32_t key, lo = 0, cnt = 0;
        if (l < nr) {
            if (match) {
                const uint64_t m = match[r0 + l];
                lo = (uint32_t)(m >> 32);
                cnt = (uint32_t)m;
            } else if (scan_row_key(src, p, pg, hdr, l, r0 + l, &key)) {
                cnt = dir_lookup(d, key, &lo);
            }
        }
        uint64_t tot;
        const uint64_t off = run + block_excl_op<OpSum>((uint64_t)cnt, lds, &tot);
        emit_wave(o, d, r0 + l, lo, cnt, off, heavy, heavy_count, &wr[threadIdx.x / HJ_WAVE]);
        run += tot;
    }
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_probe_rel(RelSrc src, DirView d, uint64_t* counts,
                                                      const uint64_t* offsets, OutSlots o,
                                                      HeavyEntry* heavy, uint32_t* heavy_count,
                                                      uint64_t* match) {
    __shared__ uint64_t lds[kBlock / HJ_WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * kRelTile;
    if (!WRITE) {
        uint64_t acc = 0;
        for (uint32_t l = threadIdx.x; l < kRelTile; l += kBlock) {
            const uint64_t i = base + l;
            if (i >= src.n) break;
            uint32_t key, lo = 0, cnt = 0;
            if (rel_row_key(src, i, &key)) cnt = dir_lookup(d, key, &lo);
            acc += cnt;
            if (match) match[i] = cnt ? ((uint64_t)lo << 32 | cnt) : 0ull;
        }
        const uint64_t tot = block_sum<kBlock>(acc, lds);
        if (threadIdx.x == 0) counts[blockIdx.x] = tot;
        return;
    }
    __shared__ WaveRows wr[kBlock / HJ_WAVE];
    uint64_t run = offsets[blockIdx.x];
    for (uint32_t l0 = 0; l0 < kRelTile; l0 += kBlock) {
        const uint64_t i = base + l0 + threadIdx.x;
        uint32_t key, lo = 0, cnt = 0;
        if (i < src.n) {
            if (match) {
                const uint64_t m = match[i];
                lo = (uint32_t)(m >> 32);
                cnt = (uint32_t)m;
            } else if (rel_row_key(src, i, &key)) {
                cnt = dir_lookup(d, key, &lo);
            }
        }
        uint64_t tot;
        const uint64_t off = run + block_excl_op<OpSum>((uint64_t)cnt, lds, &tot);
        emit_wave(o, d, (uint32_t)i, lo, cnt, off, heavy, heavy_count, &wr[threadIdx.x / HJ_WAVE]);
        run += tot;
    }
}

// heavy chunks: one block per chunk at a time, grid-stride over the chunk list. In value mode
// the chunk's probe row's values are read once per chunk (every thread reads the same words)
// and each output takes them from registers.
__global__ __launch_bounds__(kBlock) void k_expand_heavy(DirView d, OutSlots o,
                                                         const HeavyEntry* heavy,
                                                         const uint32_t* heavy_count) {
    __shared__ WaveRows wr;  // entry 0: the chunk's probe row
    const uint32_t nh = *heavy_count;
    for (uint32_t e = blockIdx.x; e < nh; e += gridDim.x) {
        const HeavyEntry h = heavy[e];
        if (!o.values) {
            for (uint32_t j = threadIdx.x; j < h.cnt; j += kBlock)
                write_out(o, d, h.out + j, h.prow, h.lo + j);
            continue;
        }
        if (threadIdx.x == 0) {
            int pc = 0;
            for (int s = 0; s < o.n && pc < kWaveVals; ++s) {
                if (o.side[s]) continue;
                int32_t v = 0;
                const uint32_t* m = o.src[s];
                const bool ok = get_int32(o.col[s], m ? m[h.prow] : h.prow, &v);
                wr.pv[pc][0] = ok ? v : 0;
                wr.pok[pc][0] = ok ? 1 : 0;
                ++pc;
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < h.cnt; j += kBlock)
            write_out_wave(o, d, h.out + j, h.prow, h.lo + j, &wr, 0);
        __syncthreads();
    }
}
}  // namespace

namespace {
// a value-mode output column of the build side in the directory's row order
__global__ __launch_bounds__(kBlock) void k_dir_gather(DirView d, OutSlots o, int c, int32_t* bval,
                                                       uint8_t* bvalid) {
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < d.n;
         p += (uint64_t)gridDim.x * kBlock) {
        const uint32_t brow = d.srow[p];
        const uint32_t row = o.src[c] ? o.src[c][brow] : brow;
        int32_t v = 0;
        const bool ok = get_int32(o.col[c], row, &v);
        bval[p] = ok ? v : 0;
        if (bvalid) bvalid[p] = ok ? 1 : 0;
    }
}
}  // namespace

void dir_gather_values(const DirView& d, const OutSlots& o, int c, int32_t* bval, uint8_t* bvalid,
                       hipStream_t s) {
    if (!d.n) return;
    const uint64_t blocks = std::min<uint64_t>((d.n + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_dir_gather, dim3((uint32_t)blocks), dim3(kBlock), 0, s, d, o, c, bval,
                       bvalid);
}

void probe_scan_count(const ScanSrc& src, const DirView& d, uint64_t* counts, uint64_t* match,
                      hipStream_t s) {
    if (!src.key.num_pages) return;
    hipLaunchKernelGGL(k_probe_scan<false>, dim3(src.key.num_pages), dim3(kBlock), 0, s, src, d,
                       counts, nullptr, OutSlots{}, nullptr, nullptr, match);
}
void probe_scan_write(const ScanSrc& src, const DirView& d, const uint64_t* offsets,
                      const OutSlots& o, HeavyEntry* heavy, uint32_t* heavy_count,
                      uint64_t* match, hipStream_t s) {
    if (!src.key.num_pages) return;
    hipLaunchKernelGGL(k_probe_scan<true>, dim3(src.key.num_pages), dim3(kBlock), 0, s, src, d,
                       nullptr, offsets, o, heavy, heavy_count, match);
}
void probe_rel_count(const RelSrc& src, const DirView& d, uint64_t* counts, uint64_t* match,
                     hipStream_t s) {
    if (!src.n) return;
    hipLaunchKernelGGL(k_probe_rel<false>, dim3(rel_blocks(src.n)), dim3(kBlock), 0, s, src, d,
                       counts, nullptr, OutSlots{}, nullptr, nullptr, match);
}
void probe_rel_write(const RelSrc& src, const DirView& d, const uint64_t* offsets,
                     const OutSlots& o, HeavyEntry* heavy, uint32_t* heavy_count, uint64_t* match,
                     hipStream_t s) {
    if (!src.n) return;
    hipLaunchKernelGGL(k_probe_rel<true>, dim3(rel_blocks(src.n)), dim3(kBlock), 0, s, src, d,
                       nullptr, offsets, o, heavy, heavy_count, match);
}
namespace {
// thread t: rows 8t .. 8t + 7, read as one 8-byte word, written as bitmap byte t; without
// validity bytes (a column without NULL) every row below n is set
__global__ __launch_bounds__(kBlock) void k_pack_valid8(const uint8_t* v8, uint64_t n,
                                                        uint8_t* bits) {
    const uint64_t nbytes = (n + 63) / 64 * 8;
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < nbytes;
         t += (uint64_t)gridDim.x * kBlock) {
        const uint64_t r = 8 * t;
        uint32_t byte = 0;
        if (r < n) {
            if (!v8) {
                byte = 0xffu;
            } else if (r + 8 <= n) {
                const uint64_t w = *(const uint64_t*)(v8 + r);  // 8-byte aligned: r = 8t
#pragma unroll
                for (int k = 0; k < 8; ++k) byte |= (uint32_t)((w >> (8 * k)) & 0xffu ? 1u : 0u) << k;
            } else {
                for (uint64_t k = 0; r + k < n; ++k) byte |= (v8[r + k] ? 1u : 0u) << k;
            }
            if (n - r < 8) byte &= (1u << (n - r)) - 1u;
        }
        bits[t] = (uint8_t)byte;
    }
}
}  // namespace

void pack_valid8(const uint8_t* valid8, uint64_t n, uint8_t* bits, hipStream_t s) {
    if (!n) return;
    const uint64_t nbytes = (n + 63) / 64 * 8;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, (nbytes + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_pack_valid8, dim3(grid), dim3(kBlock), 0, s, valid8, n, bits);
}

void expand_heavy(const DirView& d, const OutSlots& o, const HeavyEntry* heavy,
                  const uint32_t* heavy_count, hipStream_t s) {
    hipLaunchKernelGGL(k_expand_heavy, dim3(2048), dim3(kBlock), 0, s, d, o, heavy, heavy_count);
}

// =====================================================================================
// Late materialization (execute.cpp:134-156 + 431-573 without the page writer): one lane per
// output row; validity is written as a bitmap, one 64-bit ballot per wave.
// =====================================================================================
namespace {
}  // namespace

// =====================================================================================
// Directory test surface (raw u32 keys)
// =====================================================================================
namespace {
__global__ void k_hash_keys(const uint32_t* keys, uint64_t n, uint32_t* h, uint32_t* row) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    h[i] = fib32(keys[i]);
    row[i] = (uint32_t)i;
}
__global__ void k_dir_probe_keys(DirView d, const uint32_t* keys, uint64_t m, uint64_t* counts,
                                 uint64_t* first) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t lo = 0, c = d.n ? dir_lookup(d, keys[i], &lo) : 0u;
    counts[i] = c;
    if (first) first[i] = c ? (uint64_t)d.srow[lo] : ~0ull;
}
__global__ void k_dir_probe_pairs(DirView d, const uint32_t* keys, uint64_t m,
                                  const uint64_t* offsets, uint64_t* pairs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t lo = 0, c = d.n ? dir_lookup(d, keys[i], &lo) : 0u;
    for (uint32_t j = 0; j < c; ++j) {
        pairs[2 * (offsets[i] + j)] = i;
        pairs[2 * (offsets[i] + j) + 1] = d.srow[lo + j];
    }
}
}  // namespace

void hash_keys(const uint32_t* keys, uint64_t n, uint32_t* h, uint32_t* row, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_hash_keys, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, keys, n, h, row);
}
void dir_probe_keys(const DirView& d, const uint32_t* keys, uint64_t m, uint64_t* counts,
                    uint64_t* first, hipStream_t s) {
    if (!m) return;
    hipLaunchKernelGGL(k_dir_probe_keys, dim3((uint32_t)((m + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, d, keys, m, counts, first);
}
void dir_probe_pairs(const DirView& d, const uint32_t* keys, uint64_t m, const uint64_t* offsets,
                     uint64_t* pairs, hipStream_t s) {
    if (!m) return;
    hipLaunchKernelGGL(k_dir_probe_pairs, dim3((uint32_t)((m + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, d, keys, m, offsets, pairs);
}

}  // namespace hj
