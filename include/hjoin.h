/*
 * hjoin.h - C ABI of the MI355X multi-way INT32 equi-join engine.
 *
 * The engine is a drop-in for the reference's `Contest::execute(const Plan&, void*)`
 * (include/plan.h:398-405, src/execute.cpp:598-614). Every type below is a plain-C
 * restatement of the reference's data model (include/plan.h:73-210, include/attribute.h:8-13)
 * so that C, C++, Python (ctypes) or any FFI can drive it. No torch types, no C++ types.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative HJ_E* code. The message of the
 *     last failure on a context is returned by hj_last_error(ctx) (or hj_global_error()
 *     when no context exists yet). Messages match the reference's where one exists
 *     ("Join key isn't of type int32_t", src/execute.cpp:81).
 *   - Pages are the reference's 8192-byte pages (include/plan.h:100-104), given as an
 *     array of page pointers exactly like Column::pages (include/plan.h:106-109).
 *   - All sizes are 64-bit. Row ids inside the engine are 32-bit (a base table holds < 2^32 rows).
 *   - One context = one HIP stream + one HBM arena on the current device. Calls on one
 *     context must not overlap; different contexts may be driven from different threads.
 */
#ifndef HJOIN_H
#define HJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HJ_PAGE_SIZE 8192u

/* DataType, same order as include/attribute.h:8-13 */
enum { HJ_INT32 = 0, HJ_INT64 = 1, HJ_FP64 = 2, HJ_VARCHAR = 3 };

/* PlanNode kind (std::variant<ScanNode, JoinNode> index order of include/plan.h:91) */
enum { HJ_NODE_SCAN = 0, HJ_NODE_JOIN = 1 };

/* error codes */
enum {
    HJ_OK = 0,
    HJ_E_INVALID = -1,     /* malformed plan / argument */
    HJ_E_TYPE = -2,        /* non-INT32 join key, unsupported column type */
    HJ_E_OOM = -3,         /* arena / device allocation exhausted */
    HJ_E_HIP = -4,         /* a HIP runtime call failed */
    HJ_E_RANGE = -5,       /* StrRef-style field overflow (include/value_t.h:36-47) */
    HJ_E_BUSY = -6,        /* serving admission refused (memory reservation) */
    HJ_E_NODEV = -7        /* no usable GPU */
};

/* Column (include/plan.h:106-161): type + pages. */
typedef struct hj_column {
    int32_t type;
    uint32_t num_pages;
    const uint8_t* const* pages; /* num_pages pointers to HJ_PAGE_SIZE-byte pages */
} hj_column;

/* ColumnarTable (include/plan.h:163-166) */
typedef struct hj_table {
    uint64_t num_rows;
    uint32_t num_columns;
    uint32_t reserved;
    const hj_column* columns;
} hj_table;

/* one entry of PlanNode::output_attrs (include/plan.h:92): (index, DataType) */
typedef struct hj_attr {
    uint64_t index;
    int32_t type;
    int32_t reserved;
} hj_attr;

/* Engine extension with no counterpart in Plan: a load-time range filter pushed into a scan.
 * Keeps a row iff column `column` (index into the base table) is non-NULL and lo <= v <= hi,
 * the semantics of `attr BETWEEN lo AND hi` evaluated by the reference before execute()
 * (tests/read_sql.cpp:585-591, include/inner_column.h:258-259, src/build_table.cpp:319-371). */
typedef struct hj_filter {
    uint32_t column;
    int32_t lo;
    int32_t hi;
    int32_t reserved;
} hj_filter;

/* Engine extension with no counterpart in Plan: a load-time predicate pushed into a scan, the
 * reference's Comparison / LogicalOperation trees (include/statement.h:47-270) written in postfix
 * order and evaluated per row as the reference's column-level eval does (src/statement.cpp:46-200,
 * include/inner_column.h:250-560): a comparison of a NULL row is false, NOT negates whatever it
 * gets (NOT (x = 1) keeps NULL rows), AND / OR take the two results below them. INT32 columns
 * compare with (int32_t)ival; VARCHAR columns compare bytes lexicographically with sval[0,slen)
 * and LIKE matches the whole string, '%' any run and '_' one character (UTF-8, neither crossing a
 * newline), as the reference's RE2 translation does (include/statement.h:118-161). A scan keeps a
 * row iff its filters pass and the program leaves 1. At most 32 entries, stack depth 32, LIKE
 * patterns of at most 63 characters. */
enum {
    HJ_P_EQ = 0, HJ_P_NEQ = 1, HJ_P_LT = 2, HJ_P_GT = 3, HJ_P_LEQ = 4, HJ_P_GEQ = 5,
    HJ_P_LIKE = 6, HJ_P_NOT_LIKE = 7, HJ_P_IS_NULL = 8, HJ_P_IS_NOT_NULL = 9,
    HJ_P_AND = 16, HJ_P_OR = 17, HJ_P_NOT = 18
};
typedef struct hj_pred {
    int32_t op;
    uint32_t column;   /* comparisons: column of the scanned table */
    int64_t ival;      /* INT32 column: the literal */
    const char* sval;  /* VARCHAR column: the literal or LIKE pattern, slen bytes */
    uint64_t slen;
} hj_pred;

/* PlanNode (include/plan.h:78-98). kind selects ScanNode or JoinNode fields. */
typedef struct hj_node {
    int32_t kind;
    int32_t build_left;          /* JoinNode::build_left */
    uint64_t base_table_id;      /* ScanNode::base_table_id */
    uint64_t left, right;        /* JoinNode::left/right (node indices) */
    uint64_t left_attr, right_attr;
    uint64_t num_output_attrs;
    const hj_attr* output_attrs;
    uint64_t num_filters;        /* scan only; 0 for a plain reference plan */
    const hj_filter* filters;
    uint64_t num_preds;          /* scan only: a predicate program (hj_pred), 0 for none */
    const hj_pred* preds;
} hj_node;

/* Plan (include/plan.h:173-210) */
typedef struct hj_plan {
    uint64_t num_nodes;
    const hj_node* nodes;
    uint64_t num_inputs;
    const hj_table* inputs;
    uint64_t root;
} hj_plan;

/* Result column/table: pages are one contiguous block of num_pages*HJ_PAGE_SIZE bytes of host
 * memory, owned by the library: release it with hj_result_free / hj_out_column_free only (result
 * pages are pinned blocks that the library recycles for later results). */
typedef struct hj_out_column {
    int32_t type;
    uint32_t num_pages;
    uint8_t* pages;
} hj_out_column;

typedef struct hj_result {
    uint64_t num_rows;
    uint32_t num_columns;
    uint32_t reserved;
    hj_out_column* columns;
} hj_result;

typedef struct hj_context hj_context;     /* HIP stream + HBM arena + resident tables */
typedef struct hj_dtable hj_dtable;       /* a ColumnarTable resident in HBM */
typedef struct hj_dresult hj_dresult;     /* a join result resident in HBM (flat columns) */

/* ---- hash-table selection: the reference's CMake option HASH_ALGO (CMakeLists.txt:173-187,
 * include/hash_config.h:3-22) picks its table at configure time. This library is built with
 * HASH_ALGO=hip (every other value is refused by its build) and returns "hip". ---- */
const char* hj_hash_algo(void);

/* ---- context: replaces Contest::build_context / destroy_context (src/execute.cpp:610-614) ---- */
/* arena_bytes = 0 picks the default (env HJ_ARENA_BYTES, else 16 GiB). */
int hj_context_create(uint64_t arena_bytes, hj_context** out);
int hj_context_destroy(hj_context* ctx);
const char* hj_last_error(const hj_context* ctx);
const char* hj_global_error(void);

/* ---- one-call drop-in: replaces Contest::execute (include/plan.h:403, src/execute.cpp:598-608).
 * Uploads the plan's inputs, runs every scan/join on the GPU, materializes the root's
 * output_attrs into pages. result must be released with hj_result_free. ---- */
int hj_execute(hj_context* ctx, const hj_plan* plan, hj_result* result);
void hj_result_free(hj_result* result);

/* ---- resident inputs: the ColumnarTable pages of plan.inputs kept in HBM across calls.
 * Replaces the reference's borrowed, read-only plan.inputs (include/plan.h:175). ---- */
int hj_table_upload(hj_context* ctx, const hj_table* table, hj_dtable** out);
int hj_table_free(hj_context* ctx, hj_dtable* table);
uint64_t hj_table_num_rows(const hj_dtable* table);

/* Run the plan over resident tables (tables[i] stands for plan->inputs[i]; plan->inputs may be
 * NULL). The result stays in HBM as flat columns; it is valid until hj_dresult_free. The
 * device part of execute(): scan/filter, build, probe, late materialization (src/execute.cpp:35-185,
 * 265-369, 431-573 minus the page encoding). */
int hj_execute_device(hj_context* ctx, const hj_plan* plan, const hj_dtable* const* tables,
                      hj_dresult** out);
uint64_t hj_dresult_num_rows(const hj_dresult* r);
/* copy the device result to host and encode it into pages with the reference's writer rules
 * (src/execute.cpp:431-573) */
int hj_dresult_to_host(hj_context* ctx, const hj_dresult* r, hj_result* result);
/* device pointers of output column c: INT32 -> values (int32[n]) and validity bitmap (bit i of
 * byte i/8); VARCHAR -> values = uint64 offsets[n+1] into chars, chars, validity. */
int hj_dresult_column(const hj_dresult* r, uint32_t c, int32_t* type, const void** values,
                      const void** chars, const uint8_t** validity);
/* The device columns are complete for work ordered after the query on the context's stream
 * (hj_context_stream), or after hj_synchronize. Freeing does not wait: the memory is reused by
 * the context's next query, in stream order. */
int hj_dresult_free(hj_context* ctx, hj_dresult* r);
/* stream the context launches on (hipStream_t), for event timing by the caller */
void* hj_context_stream(hj_context* ctx);
int hj_synchronize(hj_context* ctx);

/* ---- page codec: ColumnInserter / Table::to_columnar writer rules (include/plan.h:212-396,
 * src/build_table.cpp:657-882) and Table::from_columnar reader rules (src/build_table.cpp:513-637).
 * validity: NULL means all rows valid; otherwise byte i != 0 means row i is valid. ---- */
int hj_encode_int32(const int32_t* values, const uint8_t* validity, uint64_t n, hj_out_column* out);
/* offsets: n+1 entries into chars (row i = chars[offsets[i], offsets[i+1])) */
int hj_encode_varchar(const uint64_t* offsets, const char* chars, const uint8_t* validity,
                      uint64_t n, hj_out_column* out);
void hj_out_column_free(hj_out_column* col);
/* decode: rows counted over the column's pages; values/validity sized by the caller from
 * hj_column_num_rows. VARCHAR: offsets (n+1) and chars sized by hj_column_num_chars. */
uint64_t hj_column_num_rows(const hj_column* col);
uint64_t hj_column_num_chars(const hj_column* col);
int hj_decode_int32(const hj_column* col, int32_t* values, uint8_t* validity);
int hj_decode_varchar(const hj_column* col, uint64_t* offsets, char* chars, uint8_t* validity);

/* ---- .tbl table cache files (host only, no GPU needed): Table::from_cache
 * (src/build_table.cpp:147-203) maps a file written by DumpTable::dump (include/table.h:128-212).
 * Layout: an 8192-byte header page holding TableMeta {uint64 num_rows; uint64 num_cols;
 * int32 types[16]; uint64 num_pages[16]} (208 bytes, then zeros), then the pages of column 0,
 * column 1, ... Errors carry the reference's messages ("Failed to open file: <path>", "File is
 * empty: <path>"); a file with more than 16 columns, an unknown type or a size other than
 * 8192 x (1 + its page count) is refused. ---- */
typedef struct hj_cache hj_cache;
int hj_cache_open(const char* path, hj_cache** out);
/* the file's table: page pointers into the read-only mapping, valid until hj_cache_close */
const hj_table* hj_cache_table(const hj_cache* cache);
void hj_cache_close(hj_cache* cache);
/* DumpTable::dump (include/table.h:200-211): the header page, then every column's pages */
int hj_cache_write(const hj_table* table, const char* path);
/* Table::from_cache straight into HBM: the mapped pages go to the device through a pinned
 * double buffer, then as hj_table_upload. */
int hj_table_upload_cache(hj_context* ctx, const char* path, hj_dtable** out);

/* ---- serving (no counterpart in the reference, whose execute() runs one query at a time): a
 * pool of `slots` contexts on the device, each with its own stream, arena (arena_bytes, 0 = 4 GiB)
 * and host thread, sharing one directory cache. Submitted queries wait in one FIFO queue; the
 * first is admitted when a slot is free and its reserved memory (its scans' staging and build
 * tables, and its output at the widest scan's rows) fits what the running queries leave of the
 * server's budget, so queries arriving while others run are admitted as slots and memory free up
 * and overlap on the GPU. A query whose reservation exceeds a slot's arena or the whole budget is
 * refused with HJ_E_BUSY. Tables are uploaded once with any context of the device
 * (hj_table_upload) and shared read-only by every query. ---- */
typedef struct hj_server hj_server;
int hj_server_create(uint32_t slots, uint64_t arena_bytes, hj_server** out);
int hj_server_destroy(hj_server* server);
/* queue a query over resident tables (tables[i] stands for plan->inputs[i]); plan and tables
 * must stay valid until hj_server_wait has returned for the ticket */
int hj_server_submit(hj_server* server, const hj_plan* plan, const hj_dtable* const* tables,
                     uint64_t* ticket);
/* wait for a query: its result (free with hj_result_free) and its time in the queue and running
 * (nanoseconds); a failed query returns its error code (message: hj_server_last_error). One
 * waiter per ticket: the first call claims it, and any other call for the same ticket returns
 * HJ_E_INVALID at once. */
int hj_server_wait(hj_server* server, uint64_t ticket, hj_result* result, uint64_t* queued_ns,
                   uint64_t* run_ns);
uint32_t hj_server_slots(const hj_server* server);
const char* hj_server_last_error(hj_server* server);
/* server options, set before the first submit:
 *   "budget_bytes"    bytes the running queries may reserve together (default: slots x arena);
 *                     a query waits in the queue until its reservation fits what the running
 *                     ones leave, and is refused (HJ_E_BUSY) only if it exceeds the whole budget
 *                     or a slot's arena
 *   "dir_cache_bytes" HBM for the slots' shared directory cache (default 1 GiB, 0: none) */
int hj_server_set_option(hj_server* server, const char* name, int64_t value);
/* the shared directory cache: lookups that found a table, lookups that did not, tables held and
 * their bytes */
int hj_server_cache_stats(hj_server* server, uint64_t* hits, uint64_t* misses, uint64_t* entries,
                          uint64_t* bytes);

/* ---- batched mode (no counterpart in the reference, whose execute() runs one plan at a time and
 * spreads one join's probe over threads by fetch_add on a 4096-row chunk counter,
 * src/execute.cpp:87-117): many independent queries over resident tables run in four launches
 * shared by all of them: counters and tables cleared, group indexes of the columns they read,
 * build inserts, probes. Each launch is one list of work items (of every query, interleaved)
 * that its workgroups walk grid-stride, with no wait between workgroups: the launch order is
 * what makes a query's inserts complete before its probes. The four launches are captured once
 * as a HIP graph and replayed by every hj_batch_run. A query is batched when it is a left-deep chain whose build
 * sides are base scans keyed by INT32 columns of the chain's source scan (a star join), with
 * range filters at most and INT32 outputs; every other query (and a batched one whose build keys
 * repeat) runs on its own through hj_execute_device after the launch. Results are the same either
 * way. Tables must stay resident and unchanged while the batch exists. ---- */
typedef struct hj_batch hj_batch;
typedef struct hj_batch_params {
    uint32_t probe_item_groups; /* 64-row groups of a query's source per probe item (0: 256) */
    uint32_t insert_item_rows;  /* build rows per insert item, a multiple of 1024 (0: 8192) */
    uint32_t radix_bits;        /* 0-12: each 1024-tuple round of an insert item is ordered by this
                                   many top hash bits before its inserts (0: unordered) */
    uint32_t load_shift;        /* a build of n rows gets 2^k >= n << load_shift key slots
                                   (0 or 1: load <= 1/2; 2: <= 1/4; 3: <= 1/8) */
    uint32_t large_build_rows;  /* builds of at least this many rows use large_load_shift (0: none) */
    uint32_t large_load_shift;
    uint32_t unroll;            /* 2 or 4 (0: 4): 64-row groups per wave and round of a probe item */
    uint32_t coop;              /* probe lookups: 0 one lane per probe row, 1 four lanes per row */
    uint32_t use_graph;         /* 1: the launch is a HIP graph instantiated on the first run */
    uint32_t plain_probe_rows;  /* star joins whose source has at least this many rows run one at
                                   a time through the fused pipeline after the launches, their
                                   results written into the batch's outputs (0: every star join
                                   in the launches) */
    uint32_t plain_build_rows;  /* likewise star joins whose build sides hold at least this many
                                   rows in all (0: no such route): the pipeline builds their
                                   tables in LDS segments, the launches by one compare-and-swap
                                   per tuple */
    uint32_t reserved[1];
} hj_batch_params;
/* plans[i] runs over tables[i][0 .. plans[i]->num_inputs); plans and tables stay owned by the
 * caller and must outlive the batch. params NULL: the defaults above with use_graph = 1. */
int hj_batch_create(hj_context* ctx, uint32_t n, const hj_plan* const* plans,
                    const hj_dtable* const* const* tables, const hj_batch_params* params,
                    hj_batch** out);
/* run every query of the batch: the four launches, the star joins routed past them on up to four
 * helper contexts beside them (created by the first run that has such queries: own stream, host
 * thread and an arena of at most 8 GiB each, freed with the batch), then the queries that run on
 * their own on ctx;
 * returns when every result is complete. Results of the previous run are replaced. */
int hj_batch_run(hj_context* ctx, hj_batch* b);
uint32_t hj_batch_size(const hj_batch* b);
/* after a run: whether query i ran in the batch's launches, and its result rows */
int hj_batch_info(const hj_batch* b, uint32_t i, int32_t* batched, uint64_t* rows);
/* the device result of query i (owned by the batch, valid until its next run or
 * hj_batch_free; hj_dresult_free refuses it): every star join has one, whether it ran in the
 * launches or on its own; HJ_E_INVALID for any other query (its result is in host pages) */
int hj_batch_result(const hj_batch* b, uint32_t i, const hj_dresult** out);
/* the result of query i as pages in host memory (free with hj_result_free); a query that ran on
 * its own hands over its result once */
int hj_batch_result_to_host(hj_context* ctx, hj_batch* b, uint32_t i, hj_result* out);
uint64_t hj_batch_device_bytes(const hj_batch* b);
int hj_batch_free(hj_context* ctx, hj_batch* b);

/* ---- engine options (no counterpart in the reference):
 *   "pipeline" 1 (default) / 0: run left-deep chains whose build sides are base scans as one
 *              fused probe pipeline; 0 forces the join-at-a-time path for every plan.
 *   "partitioned_build" 1 (default) / 0: the pipeline builds its hash tables segment by segment
 *              in LDS (build sides of up to 4M rows); 0 uses device-wide compare-and-swap inserts
 *              into cleared tables.
 *   "gpu_pages" 1 (default) / 0: hj_execute and hj_dresult_to_host write the result pages on the
 *              device and copy them out; 0 copies the flat columns and pages them on the host.
 *              Both give the same bytes.
 *   "xcd_slices" 1 (default) / 0: one-join pipelines over large tables partition the probe rows
 *              by hash slice and probe each slice on its own XCD; 0 probes every row in place.
 *   "index_tables" 1 (default) / 0: the fused probe's blocks index their own source rows in
 *              64-row groups through a table of the pages those rows span, held in LDS; 0 finds
 *              every group's page by its own search (what a block does when its rows span more
 *              pages than a table holds). Both give the same rows.
 *   "match_stash" 1 (default) / 0: the join-at-a-time path keeps each probe row's lookup from
 *              the count pass (8 bytes a row of scratch) for the write pass; 0 looks up again.
 *              Both give the same rows.
 *   "dir_cache_bytes" N > 0: keep the tables the fused pipeline builds for joins keyed by the
 *              source scan (a dimension's key column with its scan's filters and predicate
 *              program) in a directory cache of N bytes of HBM, so that a later query over the
 *              same resident table, key and filters probes the cached table instead of building
 *              it (least recently used tables make room); 0 (default): no cache. A table freed
 *              and uploaded again is a new table to the cache. ---- */
int hj_set_option(hj_context* ctx, const char* name, int64_t value);
/* the context's directory cache: lookups that found a table, lookups that did not, tables held
 * and their bytes (all 0 without a cache) */
int hj_dir_cache_stats(hj_context* ctx, uint64_t* hits, uint64_t* misses, uint64_t* entries,
                       uint64_t* bytes);

/* ---- kernel timing (no counterpart in the reference): when enabled, every engine kernel launch
 * whose name starts with `prefix` ("" = all, NULL = off) is bracketed by HIP events on the
 * context's stream. hj_profile_report synchronizes and writes one line per kernel name,
 * "name launches total_ms\n", then clears the records. ---- */
int hj_profile_enable(hj_context* ctx, const char* prefix);
int hj_profile_report(hj_context* ctx, char* buf, uint64_t len);

/* ---- directory (hash index) test surface: the GPU unchained directory of one key column,
 * the design of include/unchained.h:14-247. Used by the ported unchained_tests. ---- */
typedef struct hj_directory hj_directory;
int hj_directory_build(hj_context* ctx, const uint32_t* keys, uint64_t n, hj_directory** out);
/* for each probe key: count of build rows with an equal key and the smallest build row id among
 * them (UINT64_MAX when count == 0); row ids are positions in the build key array */
int hj_directory_probe(hj_context* ctx, const hj_directory* dir, const uint32_t* probe_keys,
                       uint64_t m, uint64_t* counts, uint64_t* first_row);
/* No counterpart in the reference: hj_directory_probe for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops or harnesses that forbid hipMalloc). It
 * allocates nothing: the probe keys, counts and first rows are staged in one fixed device buffer
 * of HJ_NOALLOC_SCRATCH_BYTES that each device holds from its first context until the process
 * ends. At most HJ_NOALLOC_MAX_PROBE_KEYS keys per call (more: HJ_E_INVALID, nothing run). Every
 * call on a device shares that device's buffer: calls that overlap on one device race and must
 * be serialized by the caller. Results equal hj_directory_probe's. */
#define HJ_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define HJ_NOALLOC_MAX_PROBE_KEYS 3300000ull
int hj_directory_probe_noalloc(hj_context* ctx, const hj_directory* dir,
                               const uint32_t* probe_keys, uint64_t m, uint64_t* counts,
                               uint64_t* first_row);
/* all matches of probe key i as (probe index, build row id) pairs, sorted by probe index */
int hj_directory_probe_pairs(hj_context* ctx, const hj_directory* dir, const uint32_t* probe_keys,
                             uint64_t m, uint64_t** pairs, uint64_t* num_pairs);
int hj_directory_info(const hj_directory* dir, uint64_t* num_slots, uint64_t* num_tuples,
                      uint64_t* max_slot_len);
int hj_directory_free(hj_context* ctx, hj_directory* dir);
void hj_free(void* p);

#ifdef __cplusplus
}
#endif
#endif /* HJOIN_H */
