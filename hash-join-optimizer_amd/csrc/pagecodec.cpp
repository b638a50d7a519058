// pagecodec.cpp - host page writer/reader of the reference's 8192-byte column pages.
//
// Writer rules (one page fills until the next row would not fit):
//   INT32   u16 rows @0, u16 values @2, packed int32 values @4, validity bitmap at the page end;
//           a non-NULL row fits while 4 + 4*(values+1) + rows/8 + 1 <= 8192, a NULL row while
//           4 + 4*values + rows/8 + 1 <= 8192 (include/plan.h:265-282,
//           src/build_table.cpp:668-720, src/execute.cpp:447-491).
//   VARCHAR u16 rows @0, u16 strings @2, u16 end offsets @4, chars, bitmap at the page end;
//           a string longer than 8185 bytes goes to its own 0xffff page followed by 0xfffe
//           pages of <= 8188 bytes each (include/plan.h:291-396, src/build_table.cpp:752-833,
//           src/execute.cpp:492-569).
// Reader rules: src/build_table.cpp:513-637.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/hjoin.h"
#include "pagecodec.hpp"

namespace hj {

namespace {
constexpr uint32_t P = HJ_PAGE_SIZE;

struct PageSink {
    std::vector<uint8_t> bytes;  // num_pages * P
    uint8_t* fresh() {
        size_t off = bytes.size();
        bytes.resize(off + P, 0);
        return bytes.data() + off;
    }
    uint32_t pages() const { return (uint32_t)(bytes.size() / P); }
};

void put16(uint8_t* p, uint32_t v) {
    uint16_t x = (uint16_t)v;
    std::memcpy(p, &x, 2);
}
uint32_t get16(const uint8_t* p) {
    uint16_t x;
    std::memcpy(&x, p, 2);
    return x;
}

void finish(PageSink& s, int32_t type, hj_out_column* out) {
    out->type = type;
    out->num_pages = s.pages();
    out->pages = nullptr;
    if (out->num_pages) {
        out->pages = (uint8_t*)std::malloc(s.bytes.size());
        if (!out->pages) throw std::bad_alloc();
        std::memcpy(out->pages, s.bytes.data(), s.bytes.size());
    }
}
}  // namespace

void encode_int32(const int32_t* values, const ValidFn& valid, uint64_t n, hj_out_column* out) {
    PageSink sink;
    uint32_t rows = 0, nvals = 0;
    int32_t vals[2048];
    uint8_t bits[P];  // a page of NULLs holds up to 8 * (P - 5) rows
    std::memset(bits, 0, sizeof bits);
    auto flush = [&]() {
        uint8_t* pg = sink.fresh();
        put16(pg, rows);
        put16(pg + 2, nvals);
        std::memcpy(pg + 4, vals, (size_t)nvals * 4);
        uint32_t nb = (rows + 7) / 8;
        std::memcpy(pg + P - nb, bits, nb);
        rows = nvals = 0;
        std::memset(bits, 0, sizeof bits);
    };
    for (uint64_t i = 0; i < n; ++i) {
        if (valid(i)) {
            if (4 + (nvals + 1) * 4 + (rows / 8 + 1) > P) flush();
            bits[rows >> 3] |= (uint8_t)(1u << (rows & 7));
            vals[nvals++] = values[i];
        } else {
            if (4 + nvals * 4 + (rows / 8 + 1) > P) flush();
        }
        ++rows;
    }
    if (rows) flush();
    finish(sink, HJ_INT32, out);
}

void encode_varchar(const uint64_t* offsets, const char* chars, const ValidFn& valid, uint64_t n,
                    hj_out_column* out) {
    PageSink sink;
    uint32_t rows = 0, nstr = 0, nchars = 0;
    uint16_t ends[P / 2];
    uint8_t data[P];
    uint8_t bits[P];  // a page of NULLs holds up to 8 * (P - 5) rows
    std::memset(bits, 0, sizeof bits);
    auto flush = [&]() {
        uint8_t* pg = sink.fresh();
        put16(pg, rows);
        put16(pg + 2, nstr);
        std::memcpy(pg + 4, ends, (size_t)nstr * 2);
        std::memcpy(pg + 4 + (size_t)nstr * 2, data, nchars);
        uint32_t nb = (rows + 7) / 8;
        std::memcpy(pg + P - nb, bits, nb);
        rows = nstr = nchars = 0;
        std::memset(bits, 0, sizeof bits);
    };
    for (uint64_t i = 0; i < n; ++i) {
        if (!valid(i)) {
            if (4 + nstr * 2 + nchars + (rows / 8 + 1) > P) flush();
            ++rows;
            continue;
        }
        const char* s = chars + offsets[i];
        const uint64_t len = offsets[i + 1] - offsets[i];
        if (len > P - 7) {
            if (rows) flush();
            for (uint64_t off = 0; off < len;) {
                uint8_t* pg = sink.fresh();
                put16(pg, off == 0 ? 0xffffu : 0xfffeu);
                uint64_t l = len - off < P - 4 ? len - off : P - 4;
                put16(pg + 2, (uint32_t)l);
                std::memcpy(pg + 4, s + off, l);
                off += l;
            }
            continue;
        }
        if (4 + (nstr + 1) * 2 + (nchars + len) + (rows / 8 + 1) > P) flush();
        bits[rows >> 3] |= (uint8_t)(1u << (rows & 7));
        std::memcpy(data + nchars, s, len);
        nchars += (uint32_t)len;
        ends[nstr++] = (uint16_t)nchars;
        ++rows;
    }
    if (rows) flush();
    finish(sink, HJ_VARCHAR, out);
}

uint64_t column_rows(const hj_column* col) {
    uint64_t rows = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        uint32_t nr = get16(col->pages[p]);
        if (col->type == HJ_VARCHAR && nr == 0xffffu)
            rows += 1;
        else if (col->type == HJ_VARCHAR && nr == 0xfffeu)
            continue;
        else
            rows += nr;
    }
    return rows;
}

uint64_t column_chars(const hj_column* col) {
    if (col->type != HJ_VARCHAR) return 0;
    uint64_t c = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        uint32_t nr = get16(pg);
        if (nr == 0xffffu || nr == 0xfffeu) {
            c += get16(pg + 2);
        } else {
            uint32_t nn = get16(pg + 2);
            if (nn) c += get16(pg + 4 + 2 * (nn - 1));
        }
    }
    return c;
}

int decode_int32(const hj_column* col, int32_t* values, uint8_t* validity) {
    if (col->type != HJ_INT32) return HJ_E_TYPE;
    uint64_t row = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        const uint32_t nr = get16(pg);
        const uint8_t* bm = pg + P - (nr + 7) / 8;
        const uint8_t* data = pg + 4;
        uint32_t vi = 0;
        for (uint32_t i = 0; i < nr; ++i, ++row) {
            const bool ok = (bm[i >> 3] >> (i & 7)) & 1;
            if (ok) std::memcpy(&values[row], data + 4 * (vi++), 4);
            else values[row] = 0;
            validity[row] = ok;
        }
    }
    return 0;
}

int decode_varchar(const hj_column* col, uint64_t* offsets, char* chars, uint8_t* validity) {
    if (col->type != HJ_VARCHAR) return HJ_E_TYPE;
    uint64_t row = 0, pos = 0;
    offsets[0] = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        const uint32_t nr = get16(pg);
        if (nr == 0xffffu || nr == 0xfffeu) {
            if (nr == 0xfffeu && row == 0) return HJ_E_INVALID;
            const uint32_t l = get16(pg + 2);
            std::memcpy(chars + pos, pg + 4, l);
            pos += l;
            if (nr == 0xffffu) {
                validity[row] = 1;
                ++row;
            }
            offsets[row] = pos;
            continue;
        }
        const uint32_t nn = get16(pg + 2);
        const uint8_t* ends = pg + 4;
        const uint8_t* data = pg + 4 + 2 * (size_t)nn;
        const uint8_t* bm = pg + P - (nr + 7) / 8;
        uint32_t vi = 0, start = 0;
        for (uint32_t i = 0; i < nr; ++i) {
            const bool ok = (bm[i >> 3] >> (i & 7)) & 1;
            if (ok) {
                const uint32_t end = get16(ends + 2 * (vi++));
                std::memcpy(chars + pos, data + start, end - start);
                pos += end - start;
                start = end;
            }
            validity[row] = ok;
            offsets[++row] = pos;
        }
    }
    return 0;
}

}  // namespace hj
