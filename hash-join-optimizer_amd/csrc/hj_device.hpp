// hj_device.hpp - device-side data layout and helpers shared by the join kernels (gfx950, wave64).
//
// HBM layout of a resident input column (one per ColumnarTable column, include/plan.h:106-166):
//   pages       num_pages * 8192 B, the reference's page bytes unchanged
//   row_start   u32[num_pages + 1]   first global row of each page (0xffff page = 1 row,
//                                    0xfffe continuation page = 0 rows), last = row count
//   rank        u16[groups]          per 64-row group of a page with NULLs: the number of
//                                    non-NULL rows of the page before the group
//   group_start u32[num_pages]       index of the page's first group in rank
//   chunk_page  u32[rows/2048 + 2]   page holding the first row of each 2048-row chunk
// Every query then reads the pages themselves. Before its first kernel, a query indexes the
// columns it reads in 64-row groups of row space (k_group_index, in the query's arena):
//   gidx        u64[2 * groups]      [2g] the validity bits of rows 64g .. 64g + 63 (0 past the
//                                    last row), [2g + 1] where the group's values are, from the
//                                    page of row 64g, the index of the group's first row among
//                                    that page's non-NULL values and how many of the group's
//                                    rows are in that page (`split`; the rest open the next page):
//                                    INT32: the byte offset A of the group's first value in the
//                                    column's pages, and the step D from a value index in the
//                                    first page to the same index in the next (gidx_offset);
//                                    VARCHAR: page | first index << 32 | split << 48
// so that 64 consecutive rows are one 16-byte record load and one load of packed values per
// page they touch. A column with a page of fewer than 64 rows before its last page has no such
// index: an INT32 column is decoded into a row-space array for the query (flat, vbits), and a
// VARCHAR row is found by page search.
// A row r is found by page search (or r / uniform_rows when every page but the last holds the
// same number of rows, the dense 1984-row case of include/column_t.h:14), then its value index
// among the page's packed non-NULL values is rank + popcount of the bitmap bits below it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HJ_PAGE_BYTES 8192u
#define HJ_WAVE 64

namespace hj {

// Pointers read from descriptors in device memory are generic ("flat") to the compiler, and
// flat loads count against both the vector-memory and the LDS/scalar wait counters, so a wait
// for an LDS read also waits for every HBM load in flight. gp() marks such a pointer as global
// so that its loads and stores are global_* instructions.
#define HJ_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline HJ_GLOBAL T* gp(T* p) {
    return (HJ_GLOBAL T*)p;
}

struct DCol {
    const uint8_t* pages;
    const uint32_t* row_start;
    const uint16_t* rank;
    const uint32_t* group_start;
    const uint32_t* chunk_page;  // page holding row c * 2048, for c = 0 .. ceil(rows/2048)
    const uint64_t* gidx;        // the query's group index (see above), or nullptr
    const int32_t* flat;         // INT32 row-space array of the query (a decoded column without
                                 // a group index, or a predicate's 0 / 1 results), or nullptr
    const uint64_t* vbits;       // its validity per row; nullptr when it has no NULL
    const uint64_t* sref;        // VARCHAR: per row, kNullRef, kLongRef | first page, or
                                 // (byte offset into pages) << 16 | length; nullptr: none
    uint32_t num_pages;
    int32_t type;
    uint32_t uniform_rows;
    uint32_t all_dense;
    uint64_t num_rows;
    // INT32 with every page dense and every page but the last of uniform_rows (>= 64) rows:
    // ceil(2^64 / uniform_rows), so that the page of row r is umul64hi(r, umagic) and no index
    // is needed; 0 otherwise
    uint64_t umagic;
};

// ---------------------------------------------------------------- hashing
// Directory prefix: 32-bit Fibonacci (multiplicative) hash, h = key * floor(2^32/phi). Odd
// multiplier => a bijection on u32, so sorting tuples by h orders them by directory slot (the
// top bits) and keeps equal keys adjacent; equal h <=> equal key, so the adjacency array stores
// h alone. The Bloom filter follows include/unchained.h:107-128 (16-bit mask, one bit per nibble
// of a 16-bit tag); the tag is bits 40..55 of h * floor(2^64/phi), because the low bits of a
// multiplicative hash depend only on the low key bits (unchained.h:75-77 takes the low 16).
__host__ __device__ inline uint32_t fib32(uint32_t key) { return key * 0x9E3779B9u; }
__host__ __device__ inline uint16_t bloom_tag(uint32_t h) {
    return (uint16_t)(((uint64_t)h * 0x9E3779B97F4A7C15ull) >> 40);
}
__host__ __device__ inline uint32_t bloom_mask(uint16_t s) {
    return (1u << (s & 0xF)) | (1u << ((s >> 4) & 0xF)) | (1u << ((s >> 8) & 0xF)) |
           (1u << ((s >> 12) & 0xF));
}
__host__ __device__ inline uint32_t slot_of(uint32_t h, uint32_t k) {
    return k ? (h >> (32u - k)) : 0u;
}

constexpr uint64_t kNullRef = ~0ull;        // sref of a NULL string
constexpr uint64_t kLongRef = 1ull << 63;    // sref of a long string: | its 0xffff page

// ---------------------------------------------------------------- row-space group index
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

__host__ __device__ inline uint64_t gloc_pack(uint32_t page, uint32_t rank0, uint32_t split) {
    return (uint64_t)page | ((uint64_t)rank0 << 32) | ((uint64_t)split << 48);
}
__device__ inline uint32_t gloc_page(uint64_t d) { return (uint32_t)d; }
__device__ inline uint32_t gloc_rank0(uint64_t d) { return (uint32_t)(d >> 32) & 0xffffu; }
__device__ inline uint32_t gloc_split(uint64_t d) { return (uint32_t)(d >> 48) & 0xffu; }

// value index (among its page's non-NULL values) and page of row l (< 64) of a group with
// validity bits vw and location d
__device__ inline uint32_t gloc_value(uint64_t vw, uint64_t d, uint32_t l, uint32_t* page) {
    const uint32_t split = gloc_split(d);
    const uint64_t below = vw & ((1ull << l) - 1ull);
    if (l < split) {
        *page = gloc_page(d);
        return gloc_rank0(d) + (uint32_t)__popcll(below);
    }
    *page = gloc_page(d) + 1u;
    return (uint32_t)__popcll(below >> split);  // split <= l < 64
}

// ---------------------------------------------------------------- page access
__device__ inline const HJ_GLOBAL uint8_t* page_ptr(const DCol& c, uint32_t p) {
    return gp(c.pages) + (size_t)p * HJ_PAGE_BYTES;
}

#define HJ_CHUNK_SHIFT 11
// last page p with row_start[p] <= r (never a 0-row page, see the layout note above): the
// chunk index narrows the search to the pages overlapping r's 2048-row chunk
__device__ inline uint32_t find_page(const DCol& c, uint32_t r) {
    if (c.uniform_rows) {
        uint32_t p = r / c.uniform_rows;
        return p < c.num_pages ? p : c.num_pages - 1;
    }
    const uint32_t ch = r >> HJ_CHUNK_SHIFT;
    uint32_t lo = gp(c.chunk_page)[ch], hi = gp(c.chunk_page)[ch + 1] + 1;
    if (hi > c.num_pages) hi = c.num_pages;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (gp(c.row_start)[mid] <= r)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// the 64 validity bits of 64-row group g of a page whose bitmap (nb bytes) ends at the page end:
// two aligned 8-byte loads and a funnel shift (bits past the page's last row are garbage and are
// never used: callers test bit (local & 63) and count bits below it). The second word is only
// loaded while it lies inside the page.
__device__ inline uint64_t group_bits(const HJ_GLOBAL uint8_t* pg, uint32_t nb, uint32_t g) {
    // 64-bit address arithmetic on a generic pointer (flat loads). Forms that loaded the word
    // through a global pointer at a 32-bit page offset returned wrong words for some lanes now
    // and then on gfx950 when all lanes of a wave read the same word (reproduced outside the
    // engine, with and without forced memory waits); this form was exact in every run.
    const uintptr_t a = (uintptr_t)((const uint8_t*)pg + HJ_PAGE_BYTES - nb + 8u * g);
    const uint64_t* w = (const uint64_t*)(a & ~(uintptr_t)7);
    const uint32_t sh = (uint32_t)(a & 7) * 8u;
    const uint64_t w0 = w[0];
    if (sh == 0) return w0;
    const uint64_t w1 =
        ((uintptr_t)(w + 1) < (uintptr_t)((const uint8_t*)pg + HJ_PAGE_BYTES)) ? w[1] : 0ull;
    return (w0 >> sh) | (w1 << (64u - sh));
}

// value index of local row `local` in page p (header word hdr = nrows | nnonnull << 16);
// returns false for a NULL row. In a sequential scan the 64 lanes of a wave share one group,
// so the group word and its rank are one broadcast load each.
__device__ inline bool value_index(const DCol& c, uint32_t p, const HJ_GLOBAL uint8_t* pg, uint32_t hdr,
                                   uint32_t local, uint32_t* vi) {
    const uint32_t nr = hdr & 0xffffu, nn = hdr >> 16;
    if (nn == nr) {
        *vi = local;
        return true;
    }
    const uint32_t g = local >> 6, b = local & 63u;
    const uint64_t bits = group_bits(pg, (nr + 7) >> 3, g);
    if (!((bits >> b) & 1ull)) return false;
    *vi = (uint32_t)gp(c.rank)[gp(c.group_start)[p] + g] +
          (uint32_t)__popcll(bits & ((1ull << b) - 1ull));
    return true;
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

// validity of row r of an INT32 column's row-space array
__device__ inline bool flat_valid(const DCol& c, uint32_t r) {
    return !c.vbits || ((gp(c.vbits)[r >> 6] >> (r & 63u)) & 1ull);
}

// INT32 group record: A = byte offset of the value of the group's first non-NULL row in its
// page (page * 8192 + 4 + 4 * first index), E = D | split << 16 with D = 8192 - 4 * (the page's
// non-NULL values up to the group's first row + those among the group's first split rows):
// the value of non-NULL row l of the group sits at A + 4 * (valid rows of the group before l),
// plus D when l >= split (the next page; 32-bit arithmetic wraps to the right offset)
__host__ __device__ inline uint64_t gint_pack(uint32_t A, uint32_t D, uint32_t split) {
    return (uint64_t)A | ((uint64_t)(D | (split << 16)) << 32);
}
__device__ inline uint32_t gidx_offset(uint64_t vw, uint64_t d, uint32_t l) {
    const uint32_t E = (uint32_t)(d >> 32);
    uint32_t off = (uint32_t)d + 4u * (uint32_t)__popcll(vw & ((1ull << l) - 1ull));
    if (l >= (E >> 16)) off += E & 0xffffu;
    return off;
}

// INT32 value of global row r through the group index (one 16-byte record, then the value)
__device__ inline bool gidx_int32(const uint8_t* pages, const uint64_t* gidx, uint32_t r,
                                  int32_t* v) {
    const u64x2 e = *(const HJ_GLOBAL u64x2*)(gp(gidx) + 2u * (r >> 6));
    const uint32_t l = r & 63u;
    if (!((e.x >> l) & 1ull)) return false;
    *v = *(const HJ_GLOBAL int32_t*)(gp(pages) + gidx_offset(e.x, e.y, l));
    return true;
}

// page of row r of a dense column of uniform pages (exact for r < 2^32: the error of the
// rounded-up reciprocal stays below 2^-32 < 1 / uniform_rows)
__device__ inline uint32_t dense_page(uint64_t umagic, uint32_t r) {
    return (uint32_t)__umul64hi((uint64_t)r, umagic);
}

// INT32 value of global row r; false for NULL
__device__ inline bool get_int32(const DCol& c, uint32_t r, int32_t* v) {
    if (c.flat) {
        *v = gp(c.flat)[r];
        return flat_valid(c, r);
    }
    if (c.umagic) {
        const uint32_t p = dense_page(c.umagic, r);
        *v = *(const HJ_GLOBAL int32_t*)(page_ptr(c, p) + 4u + 4u * (r - p * c.uniform_rows));
        return true;
    }
    if (c.gidx) return gidx_int32(c.pages, c.gidx, r, v);
    uint32_t p = find_page(c, r);
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    uint32_t vi;
    if (!value_index(c, p, pg, hdr, r - gp(c.row_start)[p], &vi)) return false;
    *v = *(const HJ_GLOBAL int32_t*)(pg + 4 + 4 * vi);
    return true;
}

// INT32 value of local row `local` of a known page
__device__ inline bool get_int32_in_page(const DCol& c, uint32_t p, const HJ_GLOBAL uint8_t* pg,
                                         uint32_t hdr, uint32_t local, int32_t* v) {
    uint32_t vi;
    if (!value_index(c, p, pg, hdr, local, &vi)) return false;
    *v = *(const HJ_GLOBAL int32_t*)(pg + 4 + 4 * vi);
    return true;
}

// VARCHAR string of global row r: returns false for NULL; for a normal page *src points at the
// bytes and *len is their count; for a long string (0xffff page) *first_page is set and *len is
// the total over the page and its 0xfffe continuation pages.
__device__ inline bool get_string(const DCol& c, uint32_t r, const uint8_t** src, uint32_t* len,
                                  uint32_t* long_page) {
    uint32_t p = find_page(c, r);
    const HJ_GLOBAL uint8_t* pg = page_ptr(c, p);
    uint32_t hdr = *(const HJ_GLOBAL uint32_t*)pg;
    uint32_t nr = hdr & 0xffffu;
    if (nr == 0xffffu) {
        uint32_t total = hdr >> 16;
        for (uint32_t q = p + 1; q < c.num_pages; ++q) {
            uint32_t h2 = *(const HJ_GLOBAL uint32_t*)page_ptr(c, q);
            if ((h2 & 0xffffu) != 0xfffeu) break;
            total += h2 >> 16;
        }
        *src = (const uint8_t*)(pg + 4);
        *len = total;
        *long_page = p;
        return true;
    }
    uint32_t vi;
    if (!value_index(c, p, pg, hdr, r - gp(c.row_start)[p], &vi)) return false;
    uint32_t nn = hdr >> 16;
    const HJ_GLOBAL uint16_t* offs = (const HJ_GLOBAL uint16_t*)(pg + 4);
    uint32_t start = vi ? offs[vi - 1] : 0u;
    *src = (const uint8_t*)(pg + 4 + 2 * nn + start);
    *len = (uint32_t)offs[vi] - start;
    *long_page = 0xffffffffu;
    return true;
}

// the reference of VARCHAR row r in the encoding of DCol::sref: kNullRef, kLongRef | its 0xffff
// page, or (byte offset of its bytes in pages) << 16 | length
__device__ inline uint64_t string_ref(const DCol& c, uint32_t r) {
    if (c.sref) return gp(c.sref)[r];
    uint32_t page, vi;
    const HJ_GLOBAL uint8_t* pg;
    uint32_t hdr;
    if (c.gidx) {
        const u64x2 e = *(const HJ_GLOBAL u64x2*)(gp(c.gidx) + 2u * (r >> 6));
        const uint32_t l = r & 63u;
        if (!((e.x >> l) & 1ull)) return kNullRef;
        vi = gloc_value(e.x, e.y, l, &page);
        pg = page_ptr(c, page);
        hdr = *(const HJ_GLOBAL uint32_t*)pg;
    } else {
        page = find_page(c, r);
        pg = page_ptr(c, page);
        hdr = *(const HJ_GLOBAL uint32_t*)pg;
        if ((hdr & 0xffffu) == 0xffffu) return kLongRef | page;
        if (!value_index(c, page, pg, hdr, r - gp(c.row_start)[page], &vi)) return kNullRef;
    }
    const HJ_GLOBAL uint16_t* offs = (const HJ_GLOBAL uint16_t*)(pg + 4);
    const uint32_t start = vi ? offs[vi - 1] : 0u;
    const uint64_t at = (uint64_t)page * HJ_PAGE_BYTES + 4u + 2u * (hdr >> 16) + start;
    return (at << 16) | (uint64_t)(offs[vi] - start);
}

// ---------------------------------------------------------------- wave / block primitives
__device__ inline uint32_t lane_id() { return threadIdx.x & (HJ_WAVE - 1); }
__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

template <typename T>
__device__ inline T wave_incl_scan(T x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < HJ_WAVE; d <<= 1) {
        T y = __shfl_up(x, d, HJ_WAVE);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

template <typename T>
__device__ inline T wave_sum(T x) {
#pragma unroll
    for (int d = HJ_WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, HJ_WAVE);
    return x;
}

// exclusive scan over a block of BLOCK threads; lds needs BLOCK/64 entries; returns the
// exclusive prefix and the block total. Contains two __syncthreads.
template <int BLOCK, typename T>
__device__ inline T block_excl_scan(T v, T* lds, T* total) {
    constexpr int NW = BLOCK / HJ_WAVE;
    const uint32_t w = threadIdx.x / HJ_WAVE;
    T incl = wave_incl_scan(v);
    if (lane_id() == HJ_WAVE - 1) lds[w] = incl;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        T t = lds[i];
        off += (i < (int)w) ? t : (T)0;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// block sum, result valid in every thread
template <int BLOCK, typename T>
__device__ inline T block_sum(T v, T* lds) {
    constexpr int NW = BLOCK / HJ_WAVE;
    T s = wave_sum(v);
    if (lane_id() == 0) lds[threadIdx.x / HJ_WAVE] = s;
    __syncthreads();
    T tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) tot += lds[i];
    __syncthreads();
    return tot;
}

}  // namespace hj
