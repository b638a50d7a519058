/*
 * oracle.h - CPU restatement of the reference's execute() path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code, and
 * only as the checker / the timed CPU baseline. The engine under hash-join-optimizer_amd/ never
 * calls it.
 *
 * What each function restates (reference file:line) is given at its definition in oracle.c.
 * Pinning: every function is checked by tests/test_oracle_*.py against the reference's own
 * known-answer tests (tests/unit_tests.cpp, valuet_tests.cpp, columnt_tests.cpp,
 * unchained_tests.cpp), against the reference headers compiled from their own sources
 * (oracle/_ref, built by oracle/Makefile), against a brute-force nested-loop join, and against
 * closed forms (histogram output counts, page capacities).
 */
#ifndef HJ_ORACLE_H
#define HJ_ORACLE_H

#include "../include/hjoin.h"

#ifdef __cplusplus
extern "C" {
#endif

/* value_t / StrRef bit layout (include/value_t.h) */
uint64_t oracle_value_int32(int32_t v);
uint64_t oracle_value_strref(uint64_t ref);
/* returns 0 and the packed ref, or HJ_E_RANGE on field overflow */
int oracle_strref_encode(int is_long, uint64_t table, uint64_t col, uint64_t page, uint64_t off,
                         uint64_t* out);

/* column_t (include/column_t.h:16-82, include/column_merge.h:7-31), for its unit tests:
 * pages of 1024 value_t, direct mode over a dense INT32 input column */
void* oracle_vcol_new(int32_t type);
void oracle_vcol_append(void* c, uint64_t value);
uint64_t oracle_vcol_get(const void* c, uint64_t row);
uint64_t oracle_vcol_rows(const void* c);
uint64_t oracle_vcol_pages(const void* c);
int32_t oracle_vcol_type(const void* c);
void oracle_vcol_use_input(void* c, const hj_column* in, uint64_t rows);
void oracle_vcol_merge(void* dst, void* src);
void oracle_vcol_free(void* c);
uint64_t oracle_splitmix64(uint64_t x);

/* Contest::execute (src/execute.cpp:598-608). Filters on scan nodes are applied first, outside
 * the timed region, the way the reference filters at load time (src/build_table.cpp:319-371).
 * threads = 0 -> SIGMOD_THREADS or the online CPU count (partitioned_directory_hash.h:29-36).
 * exec_ns (optional) receives the steady-clock time of execute() alone. */
int oracle_execute(const hj_plan* plan, int threads, hj_result* out, uint64_t* exec_ns,
                   char* err, size_t errlen);
/* nested-loop join over decoded values: an algorithm independent of the hash path, for tiny
 * inputs only */
int oracle_nested_loop_execute(const hj_plan* plan, hj_result* out, char* err, size_t errlen);
void oracle_result_free(hj_result* r);

/* Table::to_columnar writer rules (src/build_table.cpp:657-882) */
int oracle_encode_int32(const int32_t* values, const uint8_t* validity, uint64_t n,
                        hj_out_column* out);
int oracle_encode_varchar(const uint64_t* offsets, const char* chars, const uint8_t* validity,
                          uint64_t n, hj_out_column* out);
void oracle_out_column_free(hj_out_column* c);
/* Table::from_columnar reader rules (src/build_table.cpp:513-637) */
uint64_t oracle_column_rows(const hj_column* col);
uint64_t oracle_column_chars(const hj_column* col);
int oracle_decode_int32(const hj_column* col, int32_t* values, uint8_t* validity);
int oracle_decode_varchar(const hj_column* col, uint64_t* offsets, char* chars, uint8_t* validity);

/* Comparison::like_match (include/statement.h:118-161): 1 iff str matches the LIKE pattern */
int oracle_like(const char* str, uint64_t slen, const char* pat, uint64_t plen);

/* .tbl cache files: DumpTable::dump (include/table.h:128-211) and Table::from_cache
 * (src/build_table.cpp:147-203). The header page is TableMeta at byte 0 - num_rows (u64) at 0,
 * num_cols (u64) at 8, types (16 x int32) at 16, num_pages (16 x u64) at 80, 208 bytes - and
 * zeros to 8192; the pages of column 0, 1, ... follow. The reader returns -1 for a file that is
 * shorter than its header, has more than 16 columns or a size other than 8192 x (1 + pages). */
int oracle_cache_write(const hj_table* t, const char* path);
int oracle_cache_read(const char* path, hj_result* out);

/* PartitionedDirectoryHash build + probe (include/partitioned_directory_hash.h:68-215) over raw
 * keys; writes (probe index, build index) pairs in probe-chunk order. Returns the pair count, or
 * (uint64_t)-1 when cap is too small (pairs then holds the first cap pairs). */
uint64_t oracle_pdh_join(const uint32_t* build_keys, const uint8_t* build_valid, uint64_t nb,
                         const uint32_t* probe_keys, const uint8_t* probe_valid, uint64_t np,
                         int threads, uint64_t* pairs, uint64_t cap);

/* UnchainedHash<uint64_t, uint64_t> (include/unchained.h:14-247) */
typedef struct oracle_unchained oracle_unchained;
oracle_unchained* oracle_unchained_new(void);
void oracle_unchained_insert(oracle_unchained* h, uint64_t key, uint64_t payload);
void oracle_unchained_finalize(oracle_unchained* h);
void oracle_unchained_clear(oracle_unchained* h);
/* writes up to cap payloads of matching tuples; returns the match count */
uint64_t oracle_unchained_probe(const oracle_unchained* h, uint64_t key, uint64_t* payloads,
                                uint64_t cap);
uint32_t oracle_unchained_dir_size(const oracle_unchained* h);
void oracle_unchained_free(oracle_unchained* h);

#ifdef __cplusplus
}
#endif
#endif
