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

// total bytes of the long string starting at its 0xffff page p
__device__ inline uint64_t long_length(const DCol& c, uint32_t p) {
    uint64_t total = *(const HJ_GLOBAL uint32_t*)page_ptr(c, p) >> 16;
    for (uint32_t q = p + 1; q < c.num_pages; ++q) {
        const uint32_t h = *(const HJ_GLOBAL uint32_t*)page_ptr(c, q);
        if ((h & 0xffffu) != 0xfffeu) break;
        total += h >> 16;
    }
    return total;
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
                const uint64_t ref = gp(mc.col.sref)[r];
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

}  // namespace

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
