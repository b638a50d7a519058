// ref_shim.cpp - extern "C" access to the reference's own self-contained headers, compiled from
// where they lie (REF_INCLUDE = /root/reference/include) into oracle/_ref/libhjref.so by
// oracle/Makefile. TEST INFRASTRUCTURE ONLY: it exists so tests can check the oracle's
// restatement against the reference code itself.
//
// Headers used (each compiles with the standard library alone):
//   partitioned_directory_hash.h + slab_allocator.h  (the hash table execute() uses)
//   unchained.h                                       (the directory design of the GPU path)
//   value_t.h                                         (value_t / StrRef bit layout)
// execute.cpp, plan.h, column_t.h and table.h need re2 / fmt / range-v3, which this image
// does not have, so they are not built here.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "partitioned_directory_hash.h"
#include "unchained.h"
#include "value_t.h"

#include <cstdio>

extern "C" {

uint64_t ref_splitmix64(uint64_t x) { return Contest::splitmix64(x); }

// StrRef(is_long, table, col, page, off).ref; returns -1 in *err when the constructor throws
uint64_t ref_strref(int is_long, uint64_t table, uint64_t col, uint64_t page, uint64_t off,
                    int* err) {
    try {
        StrRef s(is_long != 0, table, col, page, off);
        *err = 0;
        return s.ref;
    } catch (const std::out_of_range&) {
        *err = -1;
        return 0;
    }
}

// the message StrRef's constructor throws for these fields ("" when it does not throw)
void ref_strref_message(int is_long, uint64_t table, uint64_t col, uint64_t page, uint64_t off,
                        char* buf, uint64_t n) {
    if (!n) return;
    buf[0] = 0;
    try {
        StrRef s(is_long != 0, table, col, page, off);
        (void)s;
    } catch (const std::out_of_range& e) {
        std::snprintf(buf, n, "%s", e.what());
    }
}

uint64_t ref_value_int32(int32_t v) {
    value_t x;
    x.parse_int32(v);
    return x.value;
}

uint64_t ref_value_strref(uint64_t ref) {
    StrRef s;
    s.ref = ref;
    value_t x;
    x.parse_strref(s);
    return x.value;
}

int ref_value_tag(uint64_t v) {
    value_t x;
    x.value = v;
    return x.is_int32() ? 1 : x.is_strref() ? 2 : x.is_null() ? 0 : 3;
}

// PartitionedDirectoryHash build over build_keys (build_valid[i]==0 -> NULL, skipped) with
// `threads` threads, then probe every valid probe key; writes (probe index, build row) pairs.
// Returns the number of pairs, or UINT64_MAX if cap was too small.
uint64_t ref_pdh_join(const uint32_t* build_keys, const uint8_t* build_valid, uint64_t nb,
                      const uint32_t* probe_keys, const uint8_t* probe_valid, uint64_t np,
                      int threads, uint64_t* pairs, uint64_t cap) {
    struct OptU32 {
        bool has;
        uint32_t v;
        bool has_value() const { return has; }
        uint32_t value() const { return v; }
    };
    Contest::PartitionedDirectoryHash index;
    index.init((size_t)threads, 16);
    index.build((size_t)nb, (size_t)threads, [&](size_t row) -> OptU32 {
        if (build_valid && !build_valid[row]) return {false, 0};
        return {true, build_keys[row]};
    });
    uint64_t cnt = 0;
    bool overflow = false;
    for (uint64_t r = 0; r < np; ++r) {
        if (probe_valid && !probe_valid[r]) continue;
        index.probe(probe_keys[r], [&](size_t build_idx) {
            if (cnt < cap) {
                pairs[2 * cnt] = r;
                pairs[2 * cnt + 1] = build_idx;
            } else {
                overflow = true;
            }
            ++cnt;
        });
    }
    return overflow ? UINT64_MAX : cnt;
}

// UnchainedHash<uint64_t, uint64_t>: insert n (key, payload), finalize, probe m keys.
// counts[i] = number of callbacks for probe i; payloads of all callbacks are appended to out
// (up to cap). Returns the total number of callbacks.
uint64_t ref_unchained(const uint64_t* keys, const uint64_t* payloads, uint64_t n, int finalize,
                       const uint64_t* probes, uint64_t m, uint64_t* counts, uint64_t* out,
                       uint64_t cap) {
    Contest::UnchainedHash<uint64_t, uint64_t> h;
    for (uint64_t i = 0; i < n; ++i) h.build_insert(keys[i], payloads[i]);
    if (finalize) h.finalize_build();
    uint64_t total = 0;
    for (uint64_t j = 0; j < m; ++j) {
        uint64_t c = 0;
        h.probe(probes[j], [&](uint64_t p) {
            if (total < cap) out[total] = p;
            ++total;
            ++c;
        });
        counts[j] = c;
    }
    return total;
}

}  // extern "C"
