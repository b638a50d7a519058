// materialize.hip - batched gathers of a join result's output columns (gfx950, wave64).
#include "materialize.hpp"

namespace hj {
namespace {

constexpr int kB = 256;
constexpr int kRowsPerThread = kMatTile / kB;

// copy n bytes in 64-byte rounds: 17 independent aligned dword loads (the source pages carry
// slack past their end), realigned in registers, then byte stores
__device__ inline void copy_span(HJ_GLOBAL char* dst, const HJ_GLOBAL uint8_t* src, uint32_t n) {
    for (uint32_t j0 = 0; j0 < n; j0 += 64) {
        const uintptr_t a = (uintptr_t)(src + j0);
        const HJ_GLOBAL uint32_t* w = (const HJ_GLOBAL uint32_t*)(a & ~(uintptr_t)3);
        const uint32_t sh = (uint32_t)(a & 3u);
        uint32_t d[17];
#pragma unroll
        for (int k = 0; k < 17; ++k) d[k] = w[k];
        uint32_t o[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) o[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], sh);
        const uint32_t m = n - j0 < 64u ? n - j0 : 64u;
#pragma unroll
        for (int k = 0; k < 64; ++k)
            if ((uint32_t)k < m) dst[j0 + k] = (char)(o[k >> 2] >> (8 * (k & 3)));
    }
}

// one tile of column blockIdx.y; rows i = tile * kMatTile + k * 256 + threadIdx.x, so each
// wave's 64 rows of a k-slice are consecutive and its validity ballot is one aligned word
__global__ __launch_bounds__(kB) void k_mat_a(MatBatch mb) {
    __shared__ uint64_t lds[kB / HJ_WAVE];
    const MatCol& mc = mb.c[blockIdx.y];
    const uint64_t n = mb.n;
    const uint64_t base = (uint64_t)blockIdx.x * kMatTile;
    uint64_t bytes = 0;
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const uint64_t i = base + (uint64_t)k * kB + threadIdx.x;
        bool ok = false;
        if (i < n) {
            const uint32_t r = mc.rid ? gp(mc.rid)[i] : (uint32_t)i;
            if (mc.type == 0) {
                int32_t v = 0;
                ok = get_int32(mc.col, r, &v);
                gp(mc.values)[i] = ok ? v : 0;
            } else {
                const uint64_t ref = string_ref(mc.col, r);
                ok = ref != kNullRef;
                uint64_t l = 0;
                if (ok) l = (ref & kLongRef) ? long_length(mc.col, (uint32_t)ref) : (ref & 0xffffu);
                gp(mc.offsets)[i + 1] = l;
                gp(mc.src)[i] = ref;
                bytes += l;
            }
        }
        const uint64_t bits = __ballot(ok);
        if (lane_id() == 0 && i < n) *(HJ_GLOBAL uint64_t*)(gp(mc.validity) + (i >> 3)) = bits;
    }
    if (mc.type != 0) {
        const uint64_t tot = block_sum<kB>(bytes, lds);
        if (threadIdx.x == 0) gp(mc.tile_sum)[blockIdx.x] = tot;
    }
}

// exclusive scan of a VARCHAR column's tile byte counts; the total goes to tile_sum[tiles]
__global__ __launch_bounds__(1024) void k_mat_b(MatBatch mb) {
    __shared__ uint64_t lds[1024 / HJ_WAVE];
    const MatCol& mc = mb.c[blockIdx.x];
    if (mc.type == 0) return;
    HJ_GLOBAL uint64_t* ts = gp(mc.tile_sum);
    const uint32_t nt = mb.tiles;
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < nt; b0 += 1024) {
        const uint32_t j = b0 + threadIdx.x;
        const uint64_t v = j < nt ? ts[j] : 0ull;
        uint64_t tot;
        const uint64_t ex = block_excl_scan<1024>(v, lds, &tot);
        if (j < nt) ts[j] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) ts[nt] = carry;
}

// offsets and bytes of one tile of VARCHAR column blockIdx.y
__global__ __launch_bounds__(kB) void k_mat_c(MatBatch mb) {
    __shared__ uint64_t lds[kB / HJ_WAVE];
    const MatCol& mc = mb.c[blockIdx.y];
    if (mc.type == 0) return;
    const uint64_t n = mb.n;
    const uint64_t base = (uint64_t)blockIdx.x * kMatTile;
    HJ_GLOBAL uint64_t* off = gp(mc.offsets);
    uint64_t run = gp(mc.tile_sum)[blockIdx.x];
    if (blockIdx.x == 0 && threadIdx.x == 0) off[0] = 0;
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const uint64_t i = base + (uint64_t)k * kB + threadIdx.x;
        const uint64_t len = i < n ? off[i + 1] : 0ull;
        uint64_t tot;
        const uint64_t start = run + block_excl_scan<kB>(len, lds, &tot);
        run += tot;
        if (i >= n) continue;
        off[i + 1] = start + len;
        if (!len) continue;
        const uint64_t ref = gp(mc.src)[i];
        HJ_GLOBAL char* dst = gp(mc.chars) + start;
        if (!(ref & kLongRef)) {
            copy_span(dst, page_ptr(mc.col, 0) + (ref >> 16), (uint32_t)len);
            continue;
        }
        // long string: the 0xffff page and its 0xfffe continuation pages, in order
        const uint32_t lp = (uint32_t)ref;
        uint32_t q = lp;
        uint64_t done = 0;
        while (done < len && q < mc.col.num_pages) {
            const HJ_GLOBAL uint8_t* pg = page_ptr(mc.col, q);
            const uint32_t h = *(const HJ_GLOBAL uint32_t*)pg;
            if (q != lp && (h & 0xffffu) != 0xfffeu) break;
            const uint32_t pl = (h >> 16) < len - done ? (h >> 16) : (uint32_t)(len - done);
            copy_span(dst + done, pg + 4, pl);
            done += pl;
            ++q;
        }
    }
}

// ---------------------------------------------------------------- direct output
// Block g covers the fused pass's blocks [g, g + 1) * group: its output rows start after the
// survivors of the blocks before them and its string bytes after theirs (the pass counted both
// per block), summed by the whole block with independent loads; within the block, rows are
// taken in output order (pass block, then staged row) 256 at a time.
constexpr uint32_t kMaxGroup = 16;
constexpr int kPU = 8;  // blocks of the fused pass <= kPU * 256
static_assert(kPU * kB == kEmitMaxBlocks, "k_emit's loads cover every block of the fused pass");

// block-wide sums of K values per thread in one pass: the K wave reductions are interleaved
// (one wait per shuffle step for all of them), then two barriers; the result is in out[k]
template <int K>
__device__ inline void block_sums(uint64_t (&v)[K], uint64_t (*part)[K], uint64_t* out) {
    const uint32_t w = threadIdx.x / HJ_WAVE;
#pragma unroll
    for (int dd = HJ_WAVE / 2; dd > 0; dd >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], dd, HJ_WAVE);
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) part[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < (uint32_t)K) {
        uint64_t x = 0;
#pragma unroll
        for (int i = 0; i < (int)(kB / HJ_WAVE); ++i) x += part[i][threadIdx.x];
        out[threadIdx.x] = x;
    }
    __syncthreads();
}

// block (g, c): column c of the rows of the fused pass's blocks [g, g + 1) * group
__global__ __launch_bounds__(kB) void k_emit(EmitDesc d) {
    // [0] rows before this group, [1] own rows, [2] string bytes before, [3] own bytes
    constexpr int KP = 4;
    __shared__ uint64_t lds[kB / HJ_WAVE];
    __shared__ uint64_t s_part[kB / HJ_WAVE][KP];
    __shared__ uint64_t s_sum[KP];
    __shared__ uint32_t s_rb[kMaxGroup + 1];
    __shared__ uint32_t s_org[kMaxGroup];  // first stage row of each block of the group
    const uint32_t t = threadIdx.x, G = d.group;
    const uint32_t j0 = blockIdx.x * G;
    const uint32_t j1 = j0 + G < d.blocks ? j0 + G : d.blocks;
    const EmitCol& ec = d.c[blockIdx.y];
    const bool str = ec.type != 0;
    const bool voided = gp(d.meta)[0] != 0;  // a duplicate build key voids the pass
    {
        // every thread's share of the blocks [0, j1) at once: kPU independent loads each (the
        // pass has at most kPU * kB blocks)
        uint32_t r[kPU];
        uint64_t v[kPU];
#pragma unroll
        for (int u = 0; u < kPU; ++u) {
            const uint32_t j = t + u * kB;
            const bool in = j < j1;
            r[u] = in ? gp(d.blk_cnt)[j] : 0u;
            v[u] = in && str ? gp(d.blk_bytes)[(uint64_t)j * d.bytes_stride + ec.sidx] : 0ull;
        }
        uint64_t acc[KP] = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < kPU; ++u) {
            const int h = t + u * kB >= j0 ? 1 : 0;  // own or before
            acc[h] += r[u];
            acc[2 + h] += v[u];
        }
        if (t < G) {
            s_rb[1 + t] = j0 + t < j1 ? gp(d.blk_cnt)[j0 + t] : 0u;
            s_org[t] = d.blk_origin ? (j0 + t < j1 ? gp(d.blk_origin)[j0 + t] : 0u)
                                    : (j0 + t) * d.rows_per_block;
        }
        block_sums<KP>(acc, s_part, s_sum);
        if (t == 0) {
            s_rb[0] = 0;
            for (uint32_t jj = 1; jj <= G; ++jj) s_rb[jj] += s_rb[jj - 1];
        }
        __syncthreads();
    }
    if (voided) return;
    const uint64_t off = s_sum[0];
    const uint32_t R = s_rb[G];
    if (ec.nonull && blockIdx.x + 1 == gridDim.x) {
        // the last group knows the row count: the whole bitmap of a column without NULL
        const uint64_t n = off + R < d.cap_rows ? off + R : d.cap_rows;
        HJ_GLOBAL uint64_t* bits = (HJ_GLOBAL uint64_t*)gp(ec.validity);
        for (uint64_t w = t; w < (n + 63) / 64; w += kB)
            bits[w] = n - 64 * w >= 64 ? ~0ull : (1ull << (n - 64 * w)) - 1ull;
    }
    if (blockIdx.x + 1 == gridDim.x && t == 0) {
        if (blockIdx.y == 0) gp(d.out)[0] = off + R;
        if (str) {
            const uint64_t tot = s_sum[2] + s_sum[3];
            gp(d.out)[1 + blockIdx.y] = tot;
            if (tot > ec.chars_cap) gp(d.out)[1 + kEmitCols] = 1;
        }
    }
    const HJ_GLOBAL uint32_t* rids = gp(d.stage[ec.slot]);
    HJ_GLOBAL uint64_t* offs = gp(ec.offsets);
    if (str && blockIdx.x == 0 && t == 0) offs[0] = 0;
    if (!str) {
        // INT32: kIntRows rows per thread and step, their loads in flight together (a group of
        // a large output holds tens of thousands of rows)
        constexpr int kIntRows = 8;
        for (uint32_t f0 = 0; f0 < R; f0 += kB * kIntRows) {
            uint32_t rid[kIntRows];
#pragma unroll
            for (int u = 0; u < kIntRows; ++u) {
                const uint32_t f = f0 + u * kB + t;
                uint32_t jj = 0;
                for (uint32_t x = 1; x < G; ++x) jj = f >= s_rb[x] ? x : jj;
                rid[u] = f < R ? rids[(uint64_t)s_org[jj] + (f - s_rb[jj])] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kIntRows; ++u) {
                const uint32_t f = f0 + u * kB + t;
                if (f >= R) continue;
                int32_t v = (int32_t)rid[u];
                const bool ok = ec.staged || get_int32(ec.col, rid[u], &v);
                gp(ec.values)[off + f] = ok ? v : 0;
                if (!ec.nonull) gp(ec.valid8)[off + f] = ok ? 1 : 0;
            }
        }
        return;
    }
    uint64_t run = s_sum[2];  // string byte offset of the next row (same in all threads)
    for (uint32_t f0 = 0; f0 < R; f0 += kB) {
        const uint32_t f = f0 + t;
        const bool have = f < R;
        // pass block jj (local) and staged row of output row f
        uint32_t jj = 0;
        for (uint32_t x = 1; x < G; ++x) jj = f >= s_rb[x] ? x : jj;
        const uint64_t o = off + f;
        const uint32_t rid = have ? rids[(uint64_t)s_org[jj] + (f - s_rb[jj])] : 0u;
        if (!str) {
            if (!have) continue;
            int32_t v = 0;
            const bool ok = get_int32(ec.col, rid, &v);
            gp(ec.values)[o] = ok ? v : 0;
            gp(ec.valid8)[o] = ok ? 1 : 0;
            continue;
        }
        uint64_t ref = kNullRef, len = 0;
        if (have) {
            ref = string_ref(ec.col, rid);
            if (ref != kNullRef)
                len = (ref & kLongRef) ? long_length(ec.col, (uint32_t)ref) : (ref & 0xffffu);
        }
        uint64_t tot;
        const uint64_t start = run + block_excl_scan<kB>(len, lds, &tot);
        run += tot;
        if (!have) continue;
        offs[o + 1] = start + len;
        if (!ec.nonull) gp(ec.valid8)[o] = ref != kNullRef ? 1 : 0;
        if (!len || start + len > ec.chars_cap) continue;
        HJ_GLOBAL char* dst = gp(ec.chars) + start;
        if (!(ref & kLongRef)) {
            copy_span(dst, page_ptr(ec.col, 0) + (ref >> 16), (uint32_t)len);
            continue;
        }
        const uint32_t lp = (uint32_t)ref;
        uint32_t q = lp;
        uint64_t done = 0;
        while (done < len && q < ec.col.num_pages) {
            const HJ_GLOBAL uint8_t* pg = page_ptr(ec.col, q);
            const uint32_t h = *(const HJ_GLOBAL uint32_t*)pg;
            if (q != lp && (h & 0xffffu) != 0xfffeu) break;
            const uint32_t pl = (h >> 16) < len - done ? (h >> 16) : (uint32_t)(len - done);
            copy_span(dst + done, pg + 4, pl);
            done += pl;
            ++q;
        }
    }
}

// validity bitmaps from the validity bytes: word w of every column, one wave per word
__global__ __launch_bounds__(kB) void k_emit_bits(EmitDesc d) {
    if (gp(d.meta)[0] || gp(d.out)[1 + kEmitCols]) return;
    const uint64_t n = gp(d.out)[0] < d.cap_rows ? gp(d.out)[0] : d.cap_rows;
    const uint64_t words = (n + 63) / 64;
    const uint32_t lane = lane_id();
    const uint64_t wstep = (uint64_t)gridDim.x * (kB / HJ_WAVE);
    for (uint64_t wd = (uint64_t)blockIdx.x * (kB / HJ_WAVE) + threadIdx.x / HJ_WAVE; wd < words;
         wd += wstep) {
        const uint64_t r = wd * 64 + lane;
#pragma unroll
        for (int c = 0; c < kEmitCols; ++c) {
            if (c >= d.ncol) break;
            if (d.c[c].nonull) continue;  // set whole by k_emit
            const uint64_t bits = __ballot(r < n && gp(d.c[c].valid8)[r < n ? r : 0] != 0);
            if (lane == 0) ((HJ_GLOBAL uint64_t*)gp(d.c[c].validity))[wd] = bits;
        }
    }
}

}  // namespace

bool emit_rows(const EmitDesc& d, hipStream_t s) {
    if (d.blocks > kEmitMaxBlocks || d.group > kMaxGroup) return false;
    if (!d.blocks || !d.ncol || !d.group) return true;
    hipLaunchKernelGGL(k_emit, dim3((d.blocks + d.group - 1) / d.group, d.ncol), dim3(kB), 0, s, d);
    bool bytes = false;
    for (int c = 0; c < d.ncol; ++c) bytes = bytes || !d.c[c].nonull;
    if (bytes) hipLaunchKernelGGL(k_emit_bits, dim3(512), dim3(kB), 0, s, d);
    return true;
}

void mat_gather(const MatBatch& b, hipStream_t s) {
    if (!b.n || !b.ncol) return;
    hipLaunchKernelGGL(k_mat_a, dim3(b.tiles, b.ncol), dim3(kB), 0, s, b);
    hipLaunchKernelGGL(k_mat_b, dim3(b.ncol), dim3(1024), 0, s, b);
}

void mat_strings(const MatBatch& b, hipStream_t s) {
    if (!b.n || !b.ncol) return;
    hipLaunchKernelGGL(k_mat_c, dim3(b.tiles, b.ncol), dim3(kB), 0, s, b);
}

}  // namespace hj
