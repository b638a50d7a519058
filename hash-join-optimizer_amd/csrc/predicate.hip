// predicate.hip - load-time scan predicates on the device (see predicate.hpp).
//
// Semantics restated from src/statement.cpp:46-230 and include/inner_column.h:250-560: a
// comparison of a NULL row is 0 (NOT_LIKE and NEQ included), IS_NULL / IS_NOT_NULL read the
// validity, NOT flips its operand whatever it is. VARCHAR comparisons are byte-lexicographic
// (std::string_view), LIKE is Comparison::like_match (include/statement.h:118-161): the pattern
// as a regular expression with '%' -> ".*" and '_' -> ".", fully matched, '.' being one UTF-8
// character other than '\n'. The match runs as a bit-parallel automaton over the pattern's
// tokens: bit j = "the first j tokens match the text read so far".
#include <hip/hip_runtime.h>

#include "../../include/hjoin.h"
#include "predicate.hpp"

namespace hj {
namespace {

// sequential reader of one VARCHAR value: a span of a page, or a long string's 0xffff page and
// its 0xfffe continuation pages. Bytes come 8 at a time from aligned 8-byte loads (the page
// buffer is 8-byte aligned with slack after its end, so such a load stays inside it).
struct StrCursor {
    const HJ_GLOBAL uint8_t* p;  // next byte not yet in buf
    uint32_t seg;   // bytes of the current span not yet in buf
    uint32_t page;  // current page (long strings)
    uint64_t rem;   // bytes left in the value
    uint64_t buf;   // buffered bytes, next in the low byte
    uint32_t nbuf;
    __device__ inline uint32_t next(const DCol& c) {
        if (!nbuf) {
            if (!seg) {
                ++page;
                const HJ_GLOBAL uint8_t* pg = page_ptr(c, page);
                const uint32_t h = *(const HJ_GLOBAL uint32_t*)pg;
                p = pg + 4;
                seg = h >> 16;
            }
            const uintptr_t a = (uintptr_t)p;
            const uint32_t sh = (uint32_t)(a & 7u);
            buf = *(const HJ_GLOBAL uint64_t*)(a - sh) >> (8u * sh);
            nbuf = 8u - sh < seg ? 8u - sh : seg;
            p += nbuf;
            seg -= nbuf;
        }
        const uint32_t b = (uint32_t)(buf & 0xffu);
        buf >>= 8;
        --nbuf;
        --rem;
        return b;
    }
};

__device__ inline bool open_string(const DCol& c, uint32_t row, StrCursor* cur) {
    const uint64_t ref = string_ref(c, row);
    if (ref == kNullRef) return false;
    cur->nbuf = 0;
    if (ref & kLongRef) {
        const uint32_t q = (uint32_t)ref;
        const HJ_GLOBAL uint8_t* pg = page_ptr(c, q);
        cur->p = pg + 4;
        cur->seg = *(const HJ_GLOBAL uint32_t*)pg >> 16;
        cur->page = q;
        cur->rem = long_length(c, q);
    } else {
        cur->p = page_ptr(c, 0) + (ref >> 16);
        cur->seg = (uint32_t)(ref & 0xffffu);
        cur->page = 0;
        cur->rem = ref & 0xffffu;
    }
    return true;
}

// one UTF-8 character from the cursor; a byte that does not start a complete sequence (lead
// 0xC2-0xF4 and its continuation bytes) is a character of its own, 0x110000 + the byte. An
// incomplete sequence gives back its lead byte alone: the bytes after it are read again.
__device__ inline uint32_t next_char(const DCol& c, StrCursor* cur) {
    const uint32_t b = cur->next(c);
    uint32_t need = 0, cp = b;
    if (b >= 0xC2 && b <= 0xDF) { need = 1; cp = b & 0x1F; }
    else if (b >= 0xE0 && b <= 0xEF) { need = 2; cp = b & 0x0F; }
    else if (b >= 0xF0 && b <= 0xF4) { need = 3; cp = b & 0x07; }
    else if (b >= 0x80) return 0x110000u + b;
    if (!need) return cp;
    StrCursor save = *cur;
    for (uint32_t k = 0; k < need; ++k) {
        if (!cur->rem) {
            *cur = save;
            return 0x110000u + b;
        }
        const uint32_t x = cur->next(c);
        if ((x & 0xC0) != 0x80) {
            *cur = save;
            return 0x110000u + b;
        }
        cp = (cp << 6) | (x & 0x3F);
    }
    return cp;
}

// '%' tokens match the empty run: pattern prefixes ending in them extend for free; runs = the
// longest run of consecutive '%' tokens
__device__ inline uint64_t like_closure(uint64_t d, uint64_t pct, uint32_t runs) {
    for (uint32_t i = 0; i < runs; ++i) d |= (d << 1) & pct;
    return d;
}

// ascii: for characters below 128, the pattern positions they match (from LDS)
__device__ bool like_match(const DCol& c, StrCursor cur, const HJ_GLOBAL uint32_t* tok, uint32_t m,
                           uint64_t pct, uint64_t any, uint32_t runs, const uint64_t* ascii) {
    uint64_t d = like_closure(1ull, pct, runs);
    while (cur.rem && d) {
        const uint32_t ch = next_char(c, &cur);
        uint64_t lit = 0;
        if (ascii && ch < 128)
            lit = ascii[ch];
        else
            for (uint32_t j = 0; j < m; ++j) lit |= (uint64_t)(tok[j] == ch) << (j + 1);
        const uint64_t nl = ch == '\n' ? 0ull : ~0ull;
        d = like_closure(((d << 1) & (lit | (any & nl))) | (d & pct & nl), pct, runs);
    }
    return (d >> m) & 1ull;
}

// The same automaton over bytes, for patterns of '%' and ASCII characters only (at most
// kMaxByteTokens tokens, so the state fits 32 bits). No literal matches a byte >= 0x80 and '%'
// takes every character but '\n', so the bytes of a multi-byte character (or an invalid byte)
// leave the state one step over the whole character would: closure(d & pct), which further
// such bytes do not change. tab[b]: low word the pattern positions byte b matches, high word
// pct, or 0 for '\n'. Eight bytes per aligned 8-byte load; their table entries are read before
// the state steps through them.
__device__ inline uint32_t like_closure32(uint32_t d, uint32_t pct, uint32_t runs) {
    for (uint32_t i = 0; i < runs; ++i) d |= (d << 1) & pct;
    return d;
}

__device__ bool like_match_bytes(const DCol& c, StrCursor cur, uint32_t m, uint32_t pct,
                                 uint32_t runs, const uint64_t* tab) {
    uint32_t d = like_closure32(1u, pct, runs);
    while (cur.rem && d) {
        if (!cur.seg) {
            ++cur.page;
            const HJ_GLOBAL uint8_t* pg = page_ptr(c, cur.page);
            cur.p = pg + 4;
            cur.seg = *(const HJ_GLOBAL uint32_t*)pg >> 16;
        }
        const uintptr_t a = (uintptr_t)cur.p;
        const uint32_t sh = (uint32_t)(a & 7u);
        const uint64_t w = *(const HJ_GLOBAL uint64_t*)(a - sh) >> (8u * sh);
        uint32_t nb = 8u - sh < cur.seg ? 8u - sh : cur.seg;
        if ((uint64_t)nb > cur.rem) nb = (uint32_t)cur.rem;
        cur.p += nb;
        cur.seg -= nb;
        cur.rem -= nb;
        uint64_t e[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) e[k] = tab[(uint32_t)(w >> (8 * k)) & 0xffu];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if ((uint32_t)k < nb)
                d = like_closure32(((d << 1) & (uint32_t)e[k]) | (d & (uint32_t)(e[k] >> 32)),
                                   pct, runs);
    }
    return (d >> m) & 1u;
}

// -1 / 0 / 1: the value against lit[0, ll), bytes as unsigned, a proper prefix first
__device__ int compare_string(const DCol& c, StrCursor cur, const HJ_GLOBAL uint8_t* lit, uint32_t ll) {
    uint32_t i = 0;
    while (cur.rem && i < ll) {
        const uint32_t a = cur.next(c), b = lit[i++];
        if (a != b) return a < b ? -1 : 1;
    }
    if (cur.rem) return 1;
    return i < ll ? -1 : 0;
}

constexpr int kLikeTables = 4;  // LIKE entries with a table in LDS

__global__ __launch_bounds__(256) void k_predicate(PredProg p) {
    // per LIKE entry (the first kLikeTables): byte patterns, the byte automaton's table over
    // all 256 bytes; other patterns, the positions each ASCII character matches (first 128)
    __shared__ uint64_t ascii[kLikeTables][256];
    {
        int nt = 0;  // LIKE entries seen so far: the i-th one uses table i
        for (int k = 0; k < p.nops && nt < kLikeTables; ++k) {
            const PredOp& o = p.op[k];
            if (o.op != HJ_P_LIKE && o.op != HJ_P_NOT_LIKE) continue;
            const uint32_t ch = threadIdx.x;
            if (o.ival || ch < 128) {
                const HJ_GLOBAL uint32_t* tok = (const HJ_GLOBAL uint32_t*)gp(p.lit + o.loff);
                uint64_t lit = 0;
                for (uint32_t j = 0; j < o.llen; ++j) lit |= (uint64_t)(tok[j] == ch) << (j + 1);
                if (o.ival) lit |= (uint64_t)(ch == '\n' ? 0u : (uint32_t)o.pct) << 32;
                ascii[nt][ch] = lit;
            }
            ++nt;
        }
    }
    __syncthreads();
    const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= p.n) return;
    int nl = 0;  // LIKE entries evaluated so far
    const uint32_t r = (uint32_t)row;
    uint32_t st = 0;  // bit stack, top = bit 0
    for (int k = 0; k < p.nops; ++k) {
        const PredOp& o = p.op[k];
        if (o.op == HJ_P_AND) {
            const uint32_t a = st & 1u;
            st >>= 1;
            st = (st & ~1u) | (st & a);
            continue;
        }
        if (o.op == HJ_P_OR) {
            const uint32_t a = st & 1u;
            st >>= 1;
            st |= a;
            continue;
        }
        if (o.op == HJ_P_NOT) {
            st ^= 1u;
            continue;
        }
        const DCol& c = p.col[o.col];
        const bool like = o.op == HJ_P_LIKE || o.op == HJ_P_NOT_LIKE;
        const int li = nl;  // this entry's ASCII table, when it has one
        nl += like ? 1 : 0;
        bool b = false;
        if (c.type == HJ_INT32) {
            int32_t v = 0;
            const bool ok = get_int32(c, r, &v);
            switch (o.op) {
                case HJ_P_IS_NULL: b = !ok; break;
                case HJ_P_IS_NOT_NULL: b = ok; break;
                case HJ_P_EQ: b = ok && v == o.ival; break;
                case HJ_P_NEQ: b = ok && v != o.ival; break;
                case HJ_P_LT: b = ok && v < o.ival; break;
                case HJ_P_GT: b = ok && v > o.ival; break;
                case HJ_P_LEQ: b = ok && v <= o.ival; break;
                default: b = ok && v >= o.ival; break;
            }
        } else {
            StrCursor cur;
            const bool ok = open_string(c, r, &cur);
            if (o.op == HJ_P_IS_NULL) b = !ok;
            else if (o.op == HJ_P_IS_NOT_NULL) b = ok;
            else if (!ok) b = false;
            else if (like) {
                const bool mt =
                    o.ival && li < kLikeTables
                        ? like_match_bytes(c, cur, o.llen, (uint32_t)o.pct, o.pad, ascii[li])
                        : like_match(c, cur, (const HJ_GLOBAL uint32_t*)gp(p.lit + o.loff), o.llen,
                                     o.pct, o.any, o.pad, li < kLikeTables ? ascii[li] : nullptr);
                b = o.op == HJ_P_LIKE ? mt : !mt;
            } else {
                const int cmp = compare_string(c, cur, gp(p.lit + o.loff), o.llen);
                switch (o.op) {
                    case HJ_P_EQ: b = cmp == 0; break;
                    case HJ_P_NEQ: b = cmp != 0; break;
                    case HJ_P_LT: b = cmp < 0; break;
                    case HJ_P_GT: b = cmp > 0; break;
                    case HJ_P_LEQ: b = cmp <= 0; break;
                    default: b = cmp >= 0; break;
                }
            }
        }
        st = (st << 1) | (b ? 1u : 0u);
    }
    gp(p.out)[row] = (int32_t)(st & 1u);
}

}  // namespace

void eval_predicate(const PredProg& p, hipStream_t s) {
    if (!p.n) return;
    hipLaunchKernelGGL(k_predicate, dim3((uint32_t)((p.n + 255) / 256)), dim3(256), 0, s, p);
}

}  // namespace hj
