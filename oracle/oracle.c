/*
 * oracle.c - CPU restatement of the reference's execute() path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() (as the checker) and bench.py's cpu_baseline leg
 * (as the timed CPU baseline), never by the engine.
 *
 * Restated (reference files):
 *   value_t / StrRef .................. include/value_t.h:7-127
 *   column_t (paged + direct mode) .... include/column_t.h:14-82
 *   copy_scan / col_is_dense .......... src/execute.cpp:251-369
 *   JoinAlgorithm::run ................ src/execute.cpp:35-185
 *   PartitionedDirectoryHash .......... include/partitioned_directory_hash.h:16-215
 *   merge_column_pages ................ include/column_merge.h:7-31
 *   deref_str_ref / table_to_columnar . src/execute.cpp:371-573
 *   execute / execute_impl ............ src/execute.cpp:584-608
 *   Table::to_columnar / from_columnar  src/build_table.cpp:513-882
 *   load-time INT32 range filter ...... src/build_table.cpp:319-371, include/inner_column.h:250-300
 *   scan predicates (Comparison, LIKE,  src/statement.cpp:46-230, include/statement.h:47-270,
 *     AND/OR/NOT) ..................... include/inner_column.h:330-560
 *   UnchainedHash ..................... include/unchained.h:14-247
 *   DumpTable::dump / Table::from_cache include/table.h:30-37,128-211, src/build_table.cpp:147-203
 * Not restated: the slab allocator's chunk order (include/slab_allocator.h), which only fixes the
 * output row order; parity is on the sorted multiset of rows (tests/read_sql.cpp:1293-1300).
 * Divergence: with one partition the reference shifts a 64-bit hash by 64 (undefined,
 * partitioned_directory_hash.h:100,111); here one partition means partition 0.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define PAGE 8192u
#define DENSE_ROWS_PER_PAGE 1984u /* include/column_t.h:14 */

/* ------------------------------------------------------------------ errors */
typedef struct {
    atomic_int failed;
    int code;
    char msg[256];
    pthread_mutex_t mu;
} errstate;

static void err_set(errstate* e, int code, const char* fmt, ...) {
    pthread_mutex_lock(&e->mu);
    if (!atomic_load(&e->failed)) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(e->msg, sizeof e->msg, fmt, ap);
        va_end(ap);
        e->code = code;
        atomic_store(&e->failed, 1);
    }
    pthread_mutex_unlock(&e->mu);
}

static void* xmalloc(size_t n) {
    void* p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static void* xcalloc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s ? s : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}
static void* xrealloc(void* p, size_t n) {
    p = realloc(p, n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* ------------------------------------------------------------------ value_t / StrRef */
/* include/value_t.h:95-127: tag 01 = INT32 (value in the low 32 bits), 10 = StrRef, 00 = NULL */
uint64_t oracle_value_int32(int32_t v) { return (uint64_t)(uint32_t)v | (1ull << 62); }
uint64_t oracle_value_strref(uint64_t ref) { return (ref & ~(3ull << 62)) | (1ull << 63); }
static inline int vt_is_int32(uint64_t v) { return (v >> 62) == 1; }
static inline int vt_is_strref(uint64_t v) { return (v >> 62) == 2; }
static inline int vt_is_null(uint64_t v) { return (v >> 62) == 0; }

/* include/value_t.h:7-47: long bit 61, table 10 bits @51, column 10 @41, page 18 @23,
 * offset 16 @7, 7 padding bits; a field overflow throws std::out_of_range */
#define SR_LONG (1ull << 61)
#define SR_TABLE_SHIFT 51
#define SR_COLUMN_SHIFT 41
#define SR_PAGE_SHIFT 23
#define SR_OFFSET_SHIFT 7
#define SR_TABLE_MASK ((1ull << 10) - 1)
#define SR_COLUMN_MASK ((1ull << 10) - 1)
#define SR_PAGE_MASK ((1ull << 18) - 1)
#define SR_OFFSET_MASK ((1ull << 16) - 1)

int oracle_strref_encode(int is_long, uint64_t table, uint64_t col, uint64_t page, uint64_t off,
                         uint64_t* out) {
    if (table > SR_TABLE_MASK || col > SR_COLUMN_MASK || page > SR_PAGE_MASK ||
        off > SR_OFFSET_MASK)
        return HJ_E_RANGE;
    uint64_t r = 0;
    if (is_long) r |= SR_LONG;
    r |= (table & SR_TABLE_MASK) << SR_TABLE_SHIFT;
    r |= (col & SR_COLUMN_MASK) << SR_COLUMN_SHIFT;
    r |= (page & SR_PAGE_MASK) << SR_PAGE_SHIFT;
    r |= (off & SR_OFFSET_MASK) << SR_OFFSET_SHIFT;
    *out = r;
    return 0;
}
/* the message the reference's constructor throws for the first field that overflows
 * (include/value_t.h:36-47: table, column, page, offset in that order) */
static void strref_overflow(errstate* e, uint64_t table, uint64_t col, uint64_t page, uint64_t off) {
    if (table > SR_TABLE_MASK)
        err_set(e, HJ_E_RANGE, "StrRef table overflow: %llu > %llu", (unsigned long long)table,
                (unsigned long long)SR_TABLE_MASK);
    else if (col > SR_COLUMN_MASK)
        err_set(e, HJ_E_RANGE, "StrRef column overflow: %llu > %llu", (unsigned long long)col,
                (unsigned long long)SR_COLUMN_MASK);
    else if (page > SR_PAGE_MASK)
        err_set(e, HJ_E_RANGE, "StrRef page overflow: %llu > %llu", (unsigned long long)page,
                (unsigned long long)SR_PAGE_MASK);
    else
        err_set(e, HJ_E_RANGE, "StrRef offset overflow: %llu > %llu", (unsigned long long)off,
                (unsigned long long)SR_OFFSET_MASK);
}
static inline uint64_t sr_table(uint64_t r) { return (r >> SR_TABLE_SHIFT) & SR_TABLE_MASK; }
static inline uint64_t sr_column(uint64_t r) { return (r >> SR_COLUMN_SHIFT) & SR_COLUMN_MASK; }
static inline uint64_t sr_page(uint64_t r) { return (r >> SR_PAGE_SHIFT) & SR_PAGE_MASK; }
static inline uint64_t sr_offset(uint64_t r) { return (r >> SR_OFFSET_SHIFT) & SR_OFFSET_MASK; }

/* ------------------------------------------------------------------ hashing */
/* partitioned_directory_hash.h:16-21: SplitMix64 finalizer (Steele, Lea, Flood 2014) */
uint64_t oracle_splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

static size_t next_pow2_sz(size_t x) {
    size_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

/* partitioned_directory_hash.h:29-36 */
static int default_threads(void) {
    const char* s = getenv("SIGMOD_THREADS");
    if (s) {
        long v = strtol(s, NULL, 10);
        if (v > 0) return (int)v;
    }
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

/* ------------------------------------------------------------------ page access */
static inline uint16_t rd16(const uint8_t* p) {
    uint16_t v;
    memcpy(&v, p, 2);
    return v;
}
static inline int32_t rd32(const uint8_t* p) {
    int32_t v;
    memcpy(&v, p, 4);
    return v;
}
static inline int bit_get(const uint8_t* bm, uint32_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }

/* ------------------------------------------------------------------ column_t */
/* include/column_t.h:16-82: values in pages of 1024 value_t (a page is allocated by the append
 * that finds the last one full); direct mode reads a dense INT32 input column in place, at page
 * row / 1984, slot row % 1984. */
#define COLUMN_PAGE_LEN 1024u /* include/column_t.h:11-12 */
typedef struct {
    uint64_t rows;
    int32_t type;
    uint64_t** pages;
    uint64_t npages, cap_pages;
    int direct;
    const hj_column* in;
} vcol;

static void vcol_init(vcol* c, int32_t type) {
    memset(c, 0, sizeof *c);
    c->type = type;
}
static void vcol_free(vcol* c) {
    for (uint64_t i = 0; i < c->npages; ++i) free(c->pages[i]);
    free(c->pages);
    c->pages = NULL;
    c->npages = c->cap_pages = c->rows = 0;
}
/* column_t::append_row, include/column_t.h:30-36 */
static inline void vcol_append(vcol* c, uint64_t v) {
    const uint64_t slot = c->rows & (COLUMN_PAGE_LEN - 1);
    if (slot == 0) {
        if (c->npages == c->cap_pages) {
            c->cap_pages = c->cap_pages ? c->cap_pages * 2 : 16;
            c->pages = (uint64_t**)xrealloc(c->pages, c->cap_pages * sizeof(uint64_t*));
        }
        c->pages[c->npages++] = (uint64_t*)xcalloc(COLUMN_PAGE_LEN, sizeof(uint64_t));
    }
    c->pages[c->npages - 1][slot] = v;
    c->rows++;
}
/* column_t::get_row_value, include/column_t.h:37-66 */
static inline uint64_t vcol_get(const vcol* c, uint64_t row) {
    if (c->direct) {
        uint64_t p = row / DENSE_ROWS_PER_PAGE, o = row % DENSE_ROWS_PER_PAGE;
        const uint8_t* pg = c->in->pages[p];
        return oracle_value_int32(rd32(pg + 4 + 4 * o));
    }
    return c->pages[row >> 10][row & (COLUMN_PAGE_LEN - 1)];
}
/* column_t::use_input_col, include/column_t.h:76-80 */
static void vcol_use_input(vcol* c, const hj_column* in, uint64_t rows) {
    c->direct = 1;
    c->in = in;
    c->rows = rows;
}

/* merge_column_pages, include/column_merge.h:7-31: an empty destination takes the source's
 * pages, otherwise the source's rows are appended one by one */
static void vcol_merge(vcol* dst, vcol* src) {
    if (!src->direct && dst->rows == 0 && dst->npages == 0 && !dst->direct) {
        vcol_free(dst);
        dst->pages = src->pages;
        dst->npages = src->npages;
        dst->cap_pages = src->cap_pages;
        dst->rows = src->rows;
        src->pages = NULL;
        src->npages = src->cap_pages = src->rows = 0;
        return;
    }
    for (uint64_t i = 0; i < src->rows; ++i) vcol_append(dst, vcol_get(src, i));
}

/* column_t exported for its unit tests (tests/columnt_tests.cpp) */
void* oracle_vcol_new(int32_t type) {
    vcol* c = (vcol*)xcalloc(1, sizeof(vcol));
    vcol_init(c, type);
    return c;
}
void oracle_vcol_append(void* c, uint64_t v) { vcol_append((vcol*)c, v); }
uint64_t oracle_vcol_get(const void* c, uint64_t row) { return vcol_get((const vcol*)c, row); }
uint64_t oracle_vcol_rows(const void* c) { return ((const vcol*)c)->rows; }
uint64_t oracle_vcol_pages(const void* c) { return ((const vcol*)c)->npages; }
int32_t oracle_vcol_type(const void* c) { return ((const vcol*)c)->type; }
void oracle_vcol_use_input(void* c, const hj_column* in, uint64_t rows) {
    vcol_use_input((vcol*)c, in, rows);
}
void oracle_vcol_merge(void* dst, void* src) { vcol_merge((vcol*)dst, (vcol*)src); }
void oracle_vcol_free(void* c) {
    if (!c) return;
    vcol_free((vcol*)c);
    free(c);
}

/* ------------------------------------------------------------------ copy_scan */
/* src/execute.cpp:251-262 */
static int col_is_dense(const hj_column* col) {
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        uint16_t nr = rd16(pg);
        int16_t nn;
        memcpy(&nn, pg + 2, 2);
        if ((int)nr != (int)nn) return 0;
    }
    return 1;
}

typedef struct {
    const hj_table* table;
    uint64_t table_id;
    const hj_attr* attrs;
    uint64_t nattrs;
    vcol* out;
    uint64_t begin, end;
    errstate* err;
} scan_task;

/* src/execute.cpp:277-365 (one task = a range of output columns) */
static void* scan_worker(void* arg) {
    scan_task* t = (scan_task*)arg;
    for (uint64_t ci = t->begin; ci < t->end; ++ci) {
        uint64_t in_idx = t->attrs[ci].index;
        if (in_idx >= t->table->num_columns) {
            err_set(t->err, HJ_E_INVALID, "scan column %llu out of range",
                    (unsigned long long)in_idx);
            return NULL;
        }
        const hj_column* col = &t->table->columns[in_idx];
        vcol* res = &t->out[ci];
        if (col->type == HJ_INT32 && col_is_dense(col)) {
            vcol_use_input(res, col, t->table->num_rows);
            continue;
        }
        for (uint32_t p = 0; p < col->num_pages; ++p) {
            const uint8_t* pg = col->pages[p];
            if (col->type == HJ_INT32) {
                uint16_t nr = rd16(pg);
                const uint8_t* bm = pg + PAGE - (nr + 7) / 8;
                uint32_t di = 0;
                for (uint32_t i = 0; i < nr; ++i) {
                    if (bit_get(bm, i))
                        vcol_append(res, oracle_value_int32(rd32(pg + 4 + 4 * (di++))));
                    else
                        vcol_append(res, 0);
                }
            } else if (col->type == HJ_VARCHAR) {
                uint16_t nr = rd16(pg);
                if (nr == 0xffff) {
                    uint64_t ref;
                    if (oracle_strref_encode(1, t->table_id, in_idx, p, 0, &ref)) {
                        strref_overflow(t->err, t->table_id, in_idx, p, 0);
                        return NULL;
                    }
                    vcol_append(res, oracle_value_strref(ref));
                } else if (nr == 0xfffe) {
                    continue;
                } else {
                    const uint8_t* bm = pg + PAGE - (nr + 7) / 8;
                    uint32_t di = 0;
                    for (uint32_t i = 0; i < nr; ++i) {
                        if (bit_get(bm, i)) {
                            uint64_t ref;
                            if (oracle_strref_encode(0, t->table_id, in_idx, p, di, &ref)) {
                                strref_overflow(t->err, t->table_id, in_idx, p, di);
                                return NULL;
                            }
                            ++di;
                            vcol_append(res, oracle_value_strref(ref));
                        } else {
                            vcol_append(res, 0);
                        }
                    }
                }
            }
            /* INT64 / FP64: the reference's switch has no case (src/execute.cpp:294-362) */
        }
    }
    return NULL;
}

#define SCAN_POOL 12 /* FilterThreadPool filter_tp(12), include/inner_column.h:107 */

static vcol* copy_scan(const hj_table* table, uint64_t table_id, const hj_attr* attrs,
                       uint64_t nattrs, errstate* err) {
    vcol* out = (vcol*)xcalloc(nattrs, sizeof(vcol));
    for (uint64_t i = 0; i < nattrs; ++i) vcol_init(&out[i], attrs[i].type);
    /* FilterThreadPool::run splits ntasks evenly over its threads (inner_column.h:85-104) */
    uint64_t nth = nattrs < SCAN_POOL ? nattrs : SCAN_POOL;
    if (nth == 0) return out;
    pthread_t th[SCAN_POOL];
    scan_task tasks[SCAN_POOL];
    uint64_t per = (nattrs + nth - 1) / nth;
    for (uint64_t t = 0; t < nth; ++t) {
        tasks[t] = (scan_task){table, table_id, attrs, nattrs, out, t * per,
                               (t + 1) * per < nattrs ? (t + 1) * per : nattrs, err};
        if (tasks[t].begin > tasks[t].end) tasks[t].begin = tasks[t].end;
        pthread_create(&th[t], NULL, scan_worker, &tasks[t]);
    }
    for (uint64_t t = 0; t < nth; ++t) pthread_join(th[t], NULL);
    return out;
}

/* ------------------------------------------------------------------ PartitionedDirectoryHash */
typedef struct {
    uint64_t hash;
    uint32_t key;
    uint64_t row;
} dirtuple; /* DirTuple, partitioned_directory_hash.h:48-52 (24 bytes) */

typedef struct {
    dirtuple* v;
    uint64_t n, cap;
} tvec;

static inline void tvec_push(tvec* a, dirtuple t) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 512;
        a->v = (dirtuple*)xrealloc(a->v, a->cap * sizeof(dirtuple));
    }
    a->v[a->n++] = t;
}

typedef struct {
    uint64_t* dir_store; /* dir_size + 1; directory = dir_store + 1 so directory[-1] exists */
    uint64_t* dir;
    dirtuple* tuples;
    uint64_t ntuples;
    uint64_t num_parts, dir_bits, dir_size, shift;
} pdh;

/* key source for the build: either a column_t or a raw key array */
typedef struct {
    const vcol* col;
    const uint32_t* keys;
    const uint8_t* valid;
    uint64_t n;
} keysrc;

/* returns 1 = valid key, 0 = NULL, -1 = not an int32 */
static inline int keysrc_get(const keysrc* s, uint64_t row, uint32_t* key) {
    if (s->col) {
        uint64_t v = vcol_get(s->col, row);
        if (vt_is_null(v)) return 0;
        if (!vt_is_int32(v)) return -1;
        *key = (uint32_t)(int32_t)(uint32_t)v;
        return 1;
    }
    if (s->valid && !s->valid[row]) return 0;
    *key = s->keys[row];
    return 1;
}

typedef struct {
    const keysrc* src;
    uint64_t begin, end;
    tvec* parts; /* [num_parts] for this thread */
    uint64_t num_parts, part_shift;
    errstate* err;
} p1_task;

/* phase 1, partitioned_directory_hash.h:96-118 */
static void* pdh_phase1(void* arg) {
    p1_task* t = (p1_task*)arg;
    for (uint64_t row = t->begin; row < t->end; ++row) {
        uint32_t key;
        int ok = keysrc_get(t->src, row, &key);
        if (ok == 0) continue;
        if (ok < 0) {
            err_set(t->err, HJ_E_TYPE, "Join key isn't of type int32_t");
            return NULL;
        }
        uint64_t h = oracle_splitmix64((uint64_t)key);
        uint64_t part = t->num_parts == 1 ? 0 : (h >> t->part_shift);
        tvec_push(&t->parts[part], (dirtuple){h, key, row});
    }
    return NULL;
}

typedef struct {
    pdh* ix;
    tvec** tls; /* [thread][part] */
    uint64_t nthreads;
    const uint64_t* part_counts;
    const uint64_t* part_offsets;
    uint64_t pbegin, pend;
} p3_task;

/* phase 3, partitioned_directory_hash.h:137-197 */
static void* pdh_phase3(void* arg) {
    p3_task* t = (p3_task*)arg;
    pdh* ix = t->ix;
    for (uint64_t part = t->pbegin; part < t->pend; ++part) {
        if (t->part_counts[part] == 0) continue;
        const uint64_t k = ix->dir_bits;
        const uint64_t start = (part << k) / ix->num_parts;
        const uint64_t end = ((part + 1) << k) / ix->num_parts;
        for (uint64_t th = 0; th < t->nthreads; ++th) {
            const tvec* tv = &t->tls[th][part];
            for (uint64_t i = 0; i < tv->n; ++i) {
                uint64_t slot = tv->v[i].hash >> ix->shift;
                ix->dir[slot] += (1ull << 16);
                ix->dir[slot] |= (uint16_t)tv->v[i].hash; /* computeTag */
            }
        }
        uint64_t cur = t->part_offsets[part];
        for (uint64_t i = start; i < end; ++i) {
            uint64_t cnt = ix->dir[i] >> 16;
            ix->dir[i] = (cur << 16) | (uint16_t)ix->dir[i];
            cur += cnt;
        }
        for (uint64_t th = 0; th < t->nthreads; ++th) {
            const tvec* tv = &t->tls[th][part];
            for (uint64_t i = 0; i < tv->n; ++i) {
                uint64_t slot = tv->v[i].hash >> ix->shift;
                uint64_t target = ix->dir[slot] >> 16;
                ix->tuples[target] = tv->v[i];
                ix->dir[slot] += (1ull << 16);
            }
        }
    }
    return NULL;
}

/* split_range, partitioned_directory_hash.h:40-45 */
static void split_range(uint64_t n, uint64_t tid, uint64_t nt, uint64_t* b, uint64_t* e) {
    uint64_t chunk = (n + nt - 1) / nt;
    uint64_t bb = tid * chunk;
    uint64_t ee = bb + chunk < n ? bb + chunk : n;
    if (bb > n) bb = n;
    *b = bb;
    *e = ee < bb ? bb : ee;
}

/* init(num_parts, 16) + build(), partitioned_directory_hash.h:68-198 */
static int pdh_build(pdh* ix, const keysrc* src, uint64_t nthreads, errstate* err) {
    ix->num_parts = next_pow2_sz(nthreads);
    ix->dir_bits = 16; /* execute.cpp:68 */
    ix->dir_size = 1ull << ix->dir_bits;
    ix->shift = 64 - ix->dir_bits;
    ix->dir_store = (uint64_t*)xcalloc(ix->dir_size + 1, sizeof(uint64_t));
    ix->dir = ix->dir_store + 1;

    uint64_t log2p = (uint64_t)__builtin_ctzll(ix->num_parts);
    uint64_t part_shift = 64 - log2p;
    tvec** tls = (tvec**)xcalloc(nthreads, sizeof(tvec*));
    for (uint64_t t = 0; t < nthreads; ++t) tls[t] = (tvec*)xcalloc(ix->num_parts, sizeof(tvec));
    pthread_t* th = (pthread_t*)xcalloc(nthreads, sizeof(pthread_t));
    p1_task* t1 = (p1_task*)xcalloc(nthreads, sizeof(p1_task));
    for (uint64_t t = 0; t < nthreads; ++t) {
        t1[t].src = src;
        split_range(src->n, t, nthreads, &t1[t].begin, &t1[t].end);
        t1[t].parts = tls[t];
        t1[t].num_parts = ix->num_parts;
        t1[t].part_shift = part_shift;
        t1[t].err = err;
        pthread_create(&th[t], NULL, pdh_phase1, &t1[t]);
    }
    for (uint64_t t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);

    /* phase 2, partitioned_directory_hash.h:120-135 */
    uint64_t* pc = (uint64_t*)xcalloc(ix->num_parts, sizeof(uint64_t));
    uint64_t* po = (uint64_t*)xcalloc(ix->num_parts + 1, sizeof(uint64_t));
    for (uint64_t p = 0; p < ix->num_parts; ++p) {
        uint64_t s = 0;
        for (uint64_t t = 0; t < nthreads; ++t) s += tls[t][p].n;
        pc[p] = s;
    }
    for (uint64_t p = 0; p < ix->num_parts; ++p) po[p + 1] = po[p] + pc[p];
    ix->ntuples = po[ix->num_parts];
    ix->tuples = (dirtuple*)xmalloc(ix->ntuples * sizeof(dirtuple));
    ix->dir[-1] = 0;

    if (!atomic_load(&err->failed)) {
        p3_task* t3 = (p3_task*)xcalloc(nthreads, sizeof(p3_task));
        for (uint64_t t = 0; t < nthreads; ++t) {
            t3[t] = (p3_task){ix, tls, nthreads, pc, po, 0, 0};
            split_range(ix->num_parts, t, nthreads, &t3[t].pbegin, &t3[t].pend);
            pthread_create(&th[t], NULL, pdh_phase3, &t3[t]);
        }
        for (uint64_t t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
        free(t3);
    }
    for (uint64_t t = 0; t < nthreads; ++t) {
        for (uint64_t p = 0; p < ix->num_parts; ++p) free(tls[t][p].v);
        free(tls[t]);
    }
    free(tls);
    free(th);
    free(t1);
    free(pc);
    free(po);
    return atomic_load(&err->failed) ? -1 : 0;
}

static void pdh_free(pdh* ix) {
    free(ix->dir_store);
    free(ix->tuples);
    memset(ix, 0, sizeof *ix);
}

/* probe, partitioned_directory_hash.h:200-215: returns [begin,end) of the slot's range */
static inline void pdh_range(const pdh* ix, uint32_t key, uint64_t* b, uint64_t* e) {
    uint64_t h = oracle_splitmix64((uint64_t)key);
    uint64_t slot = h >> ix->shift;
    *e = ix->dir[slot] >> 16;
    *b = (slot == 0) ? (ix->dir[-1] >> 16) : (ix->dir[slot - 1] >> 16);
}

/* ------------------------------------------------------------------ JoinAlgorithm::run */
#define PROBE_CHUNK 4096 /* execute.cpp:87 */

typedef struct {
    const pdh* ix;
    const vcol* build_side;
    uint64_t nbuild;
    const vcol* probe_side;
    uint64_t nprobe_cols;
    const vcol* probe_key;
    const hj_attr* out_attrs;
    uint64_t nout;
    int build_left;
    uint64_t left_size;
    vcol** chunk_out; /* [num_chunks][nout] */
    uint64_t num_chunks;
    atomic_ullong* next_chunk;
    errstate* err;
} probe_task;

/* probe loop with atomic work stealing, src/execute.cpp:111-162 */
static void* probe_worker(void* arg) {
    probe_task* t = (probe_task*)arg;
    const uint64_t nprobe = t->probe_key->rows;
    for (;;) {
        uint64_t cid = atomic_fetch_add(t->next_chunk, 1);
        if (cid >= t->num_chunks) break;
        uint64_t begin = cid * PROBE_CHUNK;
        uint64_t end = begin + PROBE_CHUNK < nprobe ? begin + PROBE_CHUNK : nprobe;
        vcol* out = (vcol*)xcalloc(t->nout, sizeof(vcol));
        for (uint64_t c = 0; c < t->nout; ++c) vcol_init(&out[c], t->out_attrs[c].type);
        t->chunk_out[cid] = out;
        for (uint64_t row = begin; row < end; ++row) {
            uint64_t v = vcol_get(t->probe_key, row);
            if (vt_is_null(v)) continue;
            if (!vt_is_int32(v)) {
                err_set(t->err, HJ_E_TYPE, "Join key isn't of type int32_t");
                return NULL;
            }
            uint32_t key = (uint32_t)(int32_t)(uint32_t)v;
            uint64_t b, e;
            pdh_range(t->ix, key, &b, &e);
            for (uint64_t i = b; i < e; ++i) {
                const dirtuple* tup = &t->ix->tuples[i];
                if (tup->key != key) continue;
                uint64_t bidx = tup->row;
                for (uint64_t c = 0; c < t->nout; ++c) {
                    uint64_t ci = t->out_attrs[c].index, val;
                    /* execute.cpp:139-152: output index into concat(left, right) */
                    if (t->build_left) {
                        val = ci < t->left_size ? vcol_get(&t->build_side[ci], bidx)
                                                : vcol_get(&t->probe_side[ci - t->left_size], row);
                    } else {
                        val = ci < t->left_size ? vcol_get(&t->probe_side[ci], row)
                                                : vcol_get(&t->build_side[ci - t->left_size], bidx);
                    }
                    vcol_append(&out[c], val);
                }
            }
        }
    }
    return NULL;
}

static vcol* join_run(int build_left, vcol* left, uint64_t nleft, vcol* right, uint64_t nright,
                      uint64_t left_attr, uint64_t right_attr, const hj_attr* out_attrs,
                      uint64_t nout, int nthreads, errstate* err) {
    vcol* build_side = build_left ? left : right;
    vcol* probe_side = build_left ? right : left;
    uint64_t nbuild = build_left ? nleft : nright;
    uint64_t nprobe_cols = build_left ? nright : nleft;
    uint64_t build_col = build_left ? left_attr : right_attr;
    uint64_t probe_col = build_left ? right_attr : left_attr;
    if (build_col >= nbuild || probe_col >= nprobe_cols) {
        err_set(err, HJ_E_INVALID, "join attribute out of range");
        return NULL;
    }
    for (uint64_t c = 0; c < nout; ++c) {
        if (out_attrs[c].index >= nleft + nright) {
            err_set(err, HJ_E_INVALID, "output attribute out of range");
            return NULL;
        }
    }
    vcol* results = (vcol*)xcalloc(nout ? nout : 1, sizeof(vcol));
    for (uint64_t c = 0; c < nout; ++c) vcol_init(&results[c], out_attrs[c].type);

    pdh ix;
    memset(&ix, 0, sizeof ix);
    keysrc src = {&build_side[build_col], NULL, NULL, build_side[build_col].rows};
    if (pdh_build(&ix, &src, (uint64_t)nthreads, err) != 0) {
        pdh_free(&ix);
        return results;
    }

    const vcol* pk = &probe_side[probe_col];
    uint64_t num_chunks = (pk->rows + PROBE_CHUNK - 1) / PROBE_CHUNK;
    vcol** chunk_out = (vcol**)xcalloc(num_chunks ? num_chunks : 1, sizeof(vcol*));
    atomic_ullong next = 0;
    pthread_t* th = (pthread_t*)xcalloc((size_t)nthreads, sizeof(pthread_t));
    probe_task task = {&ix,       build_side, nbuild,     probe_side, nprobe_cols, pk,
                       out_attrs, nout,       build_left, nleft,      chunk_out,   num_chunks,
                       &next,     err};
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, probe_worker, &task);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);

    /* merge in chunk order, execute.cpp:167-184 + column_merge.h:7-31 */
    for (uint64_t cid = 0; cid < num_chunks; ++cid) {
        vcol* co = chunk_out[cid];
        if (!co) continue;
        for (uint64_t c = 0; c < nout; ++c) {
            vcol_merge(&results[c], &co[c]);
            vcol_free(&co[c]);
        }
        free(co);
    }
    free(chunk_out);
    pdh_free(&ix);
    return results;
}

/* ------------------------------------------------------------------ page writers */
/* Table::to_columnar / table_to_columnar INT32 rules (build_table.cpp:668-720,
 * execute.cpp:447-491): a page holds rows while 4 + 4*(values+1) + rows/8 + 1 <= 8192. */
typedef struct {
    uint8_t* pages;
    uint32_t npages, cap;
} pagebuf;

static uint8_t* pagebuf_new(pagebuf* b) {
    if (b->npages == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 4;
        b->pages = (uint8_t*)xrealloc(b->pages, (size_t)b->cap * PAGE);
    }
    uint8_t* p = b->pages + (size_t)b->npages * PAGE;
    memset(p, 0, PAGE);
    b->npages++;
    return p;
}

typedef struct {
    pagebuf pb;
    uint16_t num_rows;
    int32_t data[2048];
    uint32_t nd;
    uint8_t bitmap[8192];
} i32writer;

static void i32w_save(i32writer* w) {
    uint8_t* pg = pagebuf_new(&w->pb);
    uint16_t nr = w->num_rows, nd = (uint16_t)w->nd;
    memcpy(pg, &nr, 2);
    memcpy(pg + 2, &nd, 2);
    memcpy(pg + 4, w->data, (size_t)w->nd * 4);
    uint32_t bms = (w->num_rows + 7) / 8;
    memcpy(pg + PAGE - bms, w->bitmap, bms);
    w->num_rows = 0;
    w->nd = 0;
    memset(w->bitmap, 0, sizeof w->bitmap);
}
static void i32w_put(i32writer* w, int valid, int32_t v) {
    if (valid) {
        if (4 + (w->nd + 1) * 4 + (w->num_rows / 8 + 1) > PAGE) i32w_save(w);
        w->bitmap[w->num_rows >> 3] |= (uint8_t)(1u << (w->num_rows & 7));
        w->data[w->nd++] = v;
    } else {
        if (4 + w->nd * 4 + (w->num_rows / 8 + 1) > PAGE) i32w_save(w);
        w->bitmap[w->num_rows >> 3] &= (uint8_t) ~(1u << (w->num_rows & 7));
    }
    w->num_rows++;
}
static void i32w_finish(i32writer* w, hj_out_column* out) {
    if (w->num_rows != 0) i32w_save(w);
    out->type = HJ_INT32;
    out->num_pages = w->pb.npages;
    out->pages = w->pb.pages;
}

/* VARCHAR rules (build_table.cpp:752-833, execute.cpp:492-569): strings longer than
 * 8192-7 bytes go to 0xffff/0xfffe long-string pages of at most 8188 bytes each. */
typedef struct {
    pagebuf pb;
    uint16_t num_rows;
    uint8_t data[8192];
    uint32_t nd;
    uint16_t offsets[4096];
    uint32_t no;
    uint8_t bitmap[8192];
} strwriter;

static void sw_save(strwriter* w) {
    uint8_t* pg = pagebuf_new(&w->pb);
    uint16_t nr = w->num_rows, no = (uint16_t)w->no;
    memcpy(pg, &nr, 2);
    memcpy(pg + 2, &no, 2);
    memcpy(pg + 4, w->offsets, (size_t)w->no * 2);
    memcpy(pg + 4 + (size_t)w->no * 2, w->data, w->nd);
    uint32_t bms = (w->num_rows + 7) / 8;
    memcpy(pg + PAGE - bms, w->bitmap, bms);
    w->num_rows = 0;
    w->nd = 0;
    w->no = 0;
    memset(w->bitmap, 0, sizeof w->bitmap);
}
static void sw_long(strwriter* w, const char* s, uint64_t len) {
    uint64_t off = 0;
    int first = 1;
    while (off < len) {
        uint8_t* pg = pagebuf_new(&w->pb);
        uint16_t tag = first ? 0xffff : 0xfffe;
        first = 0;
        memcpy(pg, &tag, 2);
        uint64_t l = len - off < PAGE - 4 ? len - off : PAGE - 4;
        uint16_t l16 = (uint16_t)l;
        memcpy(pg + 2, &l16, 2);
        memcpy(pg + 4, s + off, l);
        off += l;
    }
}
static void sw_put(strwriter* w, int valid, const char* s, uint64_t len) {
    if (valid) {
        if (len > PAGE - 7) {
            if (w->num_rows > 0) sw_save(w);
            sw_long(w, s, len);
        } else {
            if (4 + (w->no + 1) * 2 + (w->nd + len) + (w->num_rows / 8 + 1) > PAGE) sw_save(w);
            w->bitmap[w->num_rows >> 3] |= (uint8_t)(1u << (w->num_rows & 7));
            memcpy(w->data + w->nd, s, len);
            w->nd += (uint32_t)len;
            w->offsets[w->no++] = (uint16_t)w->nd;
            w->num_rows++;
        }
    } else {
        if (4 + w->no * 2 + w->nd + (w->num_rows / 8 + 1) > PAGE) sw_save(w);
        w->bitmap[w->num_rows >> 3] &= (uint8_t) ~(1u << (w->num_rows & 7));
        w->num_rows++;
    }
}
static void sw_finish(strwriter* w, hj_out_column* out) {
    if (w->num_rows != 0) sw_save(w);
    out->type = HJ_VARCHAR;
    out->num_pages = w->pb.npages;
    out->pages = w->pb.pages;
}

int oracle_encode_int32(const int32_t* values, const uint8_t* validity, uint64_t n,
                        hj_out_column* out) {
    i32writer* w = (i32writer*)xcalloc(1, sizeof(i32writer));
    for (uint64_t i = 0; i < n; ++i) i32w_put(w, validity ? validity[i] != 0 : 1, values[i]);
    i32w_finish(w, out);
    free(w);
    return 0;
}

int oracle_encode_varchar(const uint64_t* offsets, const char* chars, const uint8_t* validity,
                          uint64_t n, hj_out_column* out) {
    strwriter* w = (strwriter*)xcalloc(1, sizeof(strwriter));
    for (uint64_t i = 0; i < n; ++i)
        sw_put(w, validity ? validity[i] != 0 : 1, chars + offsets[i], offsets[i + 1] - offsets[i]);
    sw_finish(w, out);
    free(w);
    return 0;
}

void oracle_out_column_free(hj_out_column* c) {
    free(c->pages);
    c->pages = NULL;
    c->num_pages = 0;
}

/* ------------------------------------------------------------------ page readers */
/* Table::from_columnar (build_table.cpp:513-637) */
uint64_t oracle_column_rows(const hj_column* col) {
    uint64_t rows = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        uint16_t nr = rd16(col->pages[p]);
        if (col->type == HJ_VARCHAR && nr == 0xffff)
            rows += 1;
        else if (col->type == HJ_VARCHAR && nr == 0xfffe)
            rows += 0;
        else
            rows += nr;
    }
    return rows;
}

uint64_t oracle_column_chars(const hj_column* col) {
    uint64_t chars = 0;
    if (col->type != HJ_VARCHAR) return 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        uint16_t nr = rd16(pg);
        if (nr == 0xffff || nr == 0xfffe) {
            chars += rd16(pg + 2);
        } else {
            uint16_t nn = rd16(pg + 2);
            if (nn) chars += rd16(pg + 4 + 2 * (nn - 1));
        }
    }
    return chars;
}

int oracle_decode_int32(const hj_column* col, int32_t* values, uint8_t* validity) {
    if (col->type != HJ_INT32) return HJ_E_TYPE;
    uint64_t row = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        uint16_t nr = rd16(pg);
        const uint8_t* bm = pg + PAGE - (nr + 7) / 8;
        uint32_t di = 0;
        for (uint32_t i = 0; i < nr; ++i, ++row) {
            if (bit_get(bm, i)) {
                values[row] = rd32(pg + 4 + 4 * (di++));
                validity[row] = 1;
            } else {
                values[row] = 0;
                validity[row] = 0;
            }
        }
    }
    return 0;
}

int oracle_decode_varchar(const hj_column* col, uint64_t* offsets, char* chars,
                          uint8_t* validity) {
    if (col->type != HJ_VARCHAR) return HJ_E_TYPE;
    uint64_t row = 0, pos = 0;
    offsets[0] = 0;
    for (uint32_t p = 0; p < col->num_pages; ++p) {
        const uint8_t* pg = col->pages[p];
        uint16_t nr = rd16(pg);
        if (nr == 0xffff) {
            uint16_t n = rd16(pg + 2);
            memcpy(chars + pos, pg + 4, n);
            pos += n;
            validity[row] = 1;
            offsets[++row] = pos;
        } else if (nr == 0xfffe) {
            if (row == 0) return HJ_E_INVALID; /* "long string page 0xfffe must follows a string" */
            uint16_t n = rd16(pg + 2);
            memcpy(chars + pos, pg + 4, n);
            pos += n;
            offsets[row] = pos;
        } else {
            uint16_t nn = rd16(pg + 2);
            const uint8_t* offs = pg + 4;
            const uint8_t* data = pg + 4 + 2 * (size_t)nn;
            const uint8_t* bm = pg + PAGE - (nr + 7) / 8;
            uint32_t di = 0;
            uint16_t start = 0;
            for (uint32_t i = 0; i < nr; ++i) {
                if (bit_get(bm, i)) {
                    uint16_t end = rd16(offs + 2 * (di++));
                    memcpy(chars + pos, data + start, (size_t)(end - start));
                    pos += (uint64_t)(end - start);
                    start = end;
                    validity[row] = 1;
                } else {
                    validity[row] = 0;
                }
                offsets[++row] = pos;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ deref_str_ref */
typedef struct {
    char* s;
    uint64_t n, cap;
} sbuf;
static void sbuf_add(sbuf* b, const void* p, uint64_t n) {
    if (b->n + n > b->cap) {
        b->cap = (b->n + n) * 2 + 64;
        b->s = (char*)xrealloc(b->s, b->cap);
    }
    memcpy(b->s + b->n, p, n);
    b->n += n;
}

/* src/execute.cpp:371-428 */
static int deref_str_ref(uint64_t ref, const hj_table* inputs, uint64_t ninputs, sbuf* out) {
    out->n = 0;
    uint64_t ti = sr_table(ref), ci = sr_column(ref), pi = sr_page(ref);
    if (ti >= ninputs || ci >= inputs[ti].num_columns) return HJ_E_INVALID;
    const hj_column* col = &inputs[ti].columns[ci];
    if (pi >= col->num_pages) return HJ_E_INVALID;
    const uint8_t* raw = col->pages[pi];
    uint16_t nr = rd16(raw);
    if (ref & SR_LONG) {
        if (nr != 0xffff) return HJ_E_INVALID; /* "Page doesn't refer to long string" */
        sbuf_add(out, raw + 4, rd16(raw + 2));
        for (++pi; pi < col->num_pages; ++pi) {
            raw = col->pages[pi];
            if (rd16(raw) != 0xfffe) break;
            sbuf_add(out, raw + 4, rd16(raw + 2));
        }
    } else {
        uint64_t idx = sr_offset(ref);
        uint16_t nn = rd16(raw + 2);
        const uint8_t* offs = raw + 4;
        const uint8_t* data = raw + 4 + 2 * (size_t)nn;
        uint16_t start = idx == 0 ? 0 : rd16(offs + 2 * (idx - 1));
        uint16_t end = rd16(offs + 2 * idx);
        sbuf_add(out, data + start, (uint64_t)(end - start));
    }
    return 0;
}

/* src/execute.cpp:431-573 */
static int table_to_columnar(vcol* cols, const hj_attr* attrs, uint64_t ncols,
                             const hj_table* inputs, uint64_t ninputs, hj_result* out) {
    out->num_rows = ncols ? cols[0].rows : 0;
    out->num_columns = (uint32_t)ncols;
    out->columns = (hj_out_column*)xcalloc(ncols ? ncols : 1, sizeof(hj_out_column));
    sbuf sb = {0};
    for (uint64_t c = 0; c < ncols; ++c) {
        out->columns[c].type = attrs[c].type;
        if (attrs[c].type == HJ_INT32) {
            i32writer* w = (i32writer*)xcalloc(1, sizeof(i32writer));
            for (uint64_t r = 0; r < cols[c].rows; ++r) {
                uint64_t v = vcol_get(&cols[c], r);
                if (vt_is_int32(v))
                    i32w_put(w, 1, (int32_t)(uint32_t)v);
                else if (vt_is_null(v))
                    i32w_put(w, 0, 0);
                else {
                    free(w->pb.pages);
                    free(w);
                    free(sb.s);
                    return HJ_E_TYPE;
                }
            }
            i32w_finish(w, &out->columns[c]);
            free(w);
        } else if (attrs[c].type == HJ_VARCHAR) {
            strwriter* w = (strwriter*)xcalloc(1, sizeof(strwriter));
            for (uint64_t r = 0; r < cols[c].rows; ++r) {
                uint64_t v = vcol_get(&cols[c], r);
                if (vt_is_strref(v)) {
                    int rc = deref_str_ref(v & ~(3ull << 62), inputs, ninputs, &sb);
                    if (rc) {
                        free(w->pb.pages);
                        free(w);
                        free(sb.s);
                        return rc;
                    }
                    sw_put(w, 1, sb.s, sb.n);
                } else if (vt_is_null(v)) {
                    sw_put(w, 0, NULL, 0);
                } else {
                    free(w->pb.pages);
                    free(w);
                    free(sb.s);
                    return HJ_E_TYPE; /* "not string or null" */
                }
            }
            sw_finish(w, &out->columns[c]);
            free(w);
        }
        /* INT64 / FP64: no case in the reference's switch -> a column with no pages */
    }
    free(sb.s);
    return 0;
}

void oracle_result_free(hj_result* r) {
    if (!r || !r->columns) return;
    for (uint32_t c = 0; c < r->num_columns; ++c) free(r->columns[c].pages);
    free(r->columns);
    r->columns = NULL;
    r->num_columns = 0;
    r->num_rows = 0;
}

/* ------------------------------------------------------------------ plan walk */
typedef struct {
    const hj_plan* plan;
    const hj_table* inputs;
    uint64_t ninputs;
    const uint64_t* scan_table; /* per node: input id used by the scan (after filtering) */
    int nthreads;
    errstate* err;
} exec_ctx;

static void free_cols(vcol* c, uint64_t n) {
    if (!c) return;
    for (uint64_t i = 0; i < n; ++i) vcol_free(&c[i]);
    free(c);
}

/* execute_impl, src/execute.cpp:584-596; execute_hash_join, :188-225 */
static vcol* execute_impl(exec_ctx* x, uint64_t node_idx) {
    if (node_idx >= x->plan->num_nodes) {
        err_set(x->err, HJ_E_INVALID, "node index out of range");
        return NULL;
    }
    const hj_node* node = &x->plan->nodes[node_idx];
    if (node->kind == HJ_NODE_SCAN) {
        uint64_t tid = x->scan_table[node_idx];
        if (tid >= x->ninputs) {
            err_set(x->err, HJ_E_INVALID, "base table out of range");
            return NULL;
        }
        /* a filtered scan reads its filtered copy, which is its own input (as after the
         * reference's load-time filter), so StrRefs name that copy */
        return copy_scan(&x->inputs[tid], tid, node->output_attrs, node->num_output_attrs,
                         x->err);
    }
    if (node->left >= x->plan->num_nodes || node->right >= x->plan->num_nodes) {
        err_set(x->err, HJ_E_INVALID, "child index out of range");
        return NULL;
    }
    uint64_t nl = x->plan->nodes[node->left].num_output_attrs;
    uint64_t nr = x->plan->nodes[node->right].num_output_attrs;
    vcol* left = execute_impl(x, node->left);
    if (!left) return NULL;
    vcol* right = execute_impl(x, node->right);
    if (!right) {
        free_cols(left, nl);
        return NULL;
    }
    vcol* res = NULL;
    if (!atomic_load(&x->err->failed))
        res = join_run(node->build_left, left, nl, right, nr, node->left_attr, node->right_attr,
                       node->output_attrs, node->num_output_attrs, x->nthreads, x->err);
    free_cols(left, nl);
    free_cols(right, nr);
    return res;
}

/* ------------------------------------------------------------------ scan predicates */
/* one character of s[i, n) in UTF-8: its code point and byte length; a byte that does not start
 * a complete sequence (lead 0xC2-0xF4 followed by its continuation bytes) is a character of its
 * own, 0x110000 + the byte */
static uint32_t utf8_char(const unsigned char* s, uint64_t n, uint64_t i, uint32_t* len) {
    const unsigned b = s[i];
    uint32_t need = 0, cp = b;
    if (b >= 0xC2 && b <= 0xDF) { need = 1; cp = b & 0x1F; }
    else if (b >= 0xE0 && b <= 0xEF) { need = 2; cp = b & 0x0F; }
    else if (b >= 0xF0 && b <= 0xF4) { need = 3; cp = b & 0x07; }
    else if (b >= 0x80) { *len = 1; return 0x110000u + b; }
    for (uint32_t k = 1; k <= need; ++k) {
        if (i + k >= n || (s[i + k] & 0xC0) != 0x80) { *len = 1; return 0x110000u + b; }
        cp = (cp << 6) | (s[i + k] & 0x3F);
    }
    *len = need + 1;
    return cp;
}

#define LIKE_ANY_RUN 0xFFFFFFFFu /* '%' */
#define LIKE_ANY_ONE 0xFFFFFFFEu /* '_' */

/* Comparison::like_match (include/statement.h:118-161): '%' -> ".*", '_' -> ".", other pattern
 * characters literal, the whole string matched by RE2, whose '.' is one UTF-8 character other
 * than '\n'. Restated as a dynamic program over (text characters, pattern characters). */
int oracle_like(const char* str, uint64_t slen, const char* pat, uint64_t plen) {
    const unsigned char* p = (const unsigned char*)pat;
    const unsigned char* t = (const unsigned char*)str;
    uint32_t* tok = (uint32_t*)xmalloc((plen + 1) * 4);
    uint64_t m = 0;
    for (uint64_t i = 0; i < plen;) {
        uint32_t l;
        const uint32_t cp = utf8_char(p, plen, i, &l);
        tok[m++] = p[i] == '%' ? LIKE_ANY_RUN : p[i] == '_' ? LIKE_ANY_ONE : cp;
        i += l;
    }
    uint8_t* cur = (uint8_t*)xcalloc(m + 1, 1);
    uint8_t* nxt = (uint8_t*)xcalloc(m + 1, 1);
    cur[0] = 1;
    for (uint64_t j = 1; j <= m; ++j) cur[j] = cur[j - 1] && tok[j - 1] == LIKE_ANY_RUN;
    for (uint64_t i = 0; i < slen;) {
        uint32_t l;
        const uint32_t c = utf8_char(t, slen, i, &l);
        i += l;
        nxt[0] = 0;
        for (uint64_t j = 1; j <= m; ++j) {
            const uint32_t k = tok[j - 1];
            if (k == LIKE_ANY_RUN)
                nxt[j] = nxt[j - 1] || (cur[j] && c != '\n');
            else if (k == LIKE_ANY_ONE)
                nxt[j] = cur[j - 1] && c != '\n';
            else
                nxt[j] = cur[j - 1] && c == k;
        }
        uint8_t* sw = cur;
        cur = nxt;
        nxt = sw;
    }
    const int r = cur[m];
    free(tok);
    free(cur);
    free(nxt);
    return r;
}

typedef struct {
    int32_t* v;      /* INT32 */
    uint64_t* off;   /* VARCHAR */
    char* ch;
    uint8_t* ok;
} pred_col;

/* keep[r] &= the program's result for row r (src/statement.cpp:46-230: a comparison of a NULL
 * row is 0, IS_NULL / IS_NOT_NULL read the validity, AND / OR / NOT combine the row's bits) */
static int apply_preds(const hj_table* in, const hj_pred* pr, uint64_t np, uint8_t* keep) {
    const uint64_t n = in->num_rows;
    if (np > 32) return HJ_E_INVALID;
    pred_col cols[32];
    memset(cols, 0, sizeof cols);
    int depth = 0, rc = 0;
    for (uint64_t k = 0; k < np && rc == 0; ++k) {
        const int op = pr[k].op;
        if (op == HJ_P_AND || op == HJ_P_OR) {
            if (depth < 2) rc = HJ_E_INVALID;
            --depth;
            continue;
        }
        if (op == HJ_P_NOT) {
            if (depth < 1) rc = HJ_E_INVALID;
            continue;
        }
        if (op < HJ_P_EQ || op > HJ_P_IS_NOT_NULL || pr[k].column >= in->num_columns) {
            rc = HJ_E_INVALID;
            break;
        }
        const hj_column* col = &in->columns[pr[k].column];
        if (col->type != HJ_INT32 && col->type != HJ_VARCHAR) rc = HJ_E_TYPE;
        if ((op == HJ_P_LIKE || op == HJ_P_NOT_LIKE) && col->type != HJ_VARCHAR) rc = HJ_E_TYPE;
        if (col->type == HJ_VARCHAR && op < HJ_P_IS_NULL && !pr[k].sval && pr[k].slen) rc = HJ_E_INVALID;
        if (++depth > 32) rc = HJ_E_INVALID;
        if (rc) break;
        pred_col* pc = &cols[k];
        pc->ok = (uint8_t*)xmalloc(n + 1);
        if (col->type == HJ_INT32) {
            pc->v = (int32_t*)xmalloc(n * 4 + 4);
            oracle_decode_int32(col, pc->v, pc->ok);
        } else {
            pc->off = (uint64_t*)xmalloc((n + 1) * 8);
            pc->ch = (char*)xmalloc(oracle_column_chars(col) + 1);
            oracle_decode_varchar(col, pc->off, pc->ch, pc->ok);
        }
    }
    if (rc == 0 && depth != 1 && np) rc = HJ_E_INVALID;
    for (uint64_t r = 0; rc == 0 && np && r < n; ++r) {
        uint8_t st[32];
        int sp = 0;
        for (uint64_t k = 0; k < np; ++k) {
            const int op = pr[k].op;
            if (op == HJ_P_AND) { --sp; st[sp - 1] = st[sp - 1] && st[sp]; continue; }
            if (op == HJ_P_OR) { --sp; st[sp - 1] = st[sp - 1] || st[sp]; continue; }
            if (op == HJ_P_NOT) { st[sp - 1] = !st[sp - 1]; continue; }
            const pred_col* pc = &cols[k];
            const int nn = pc->ok[r] != 0;
            int b = 0;
            if (op == HJ_P_IS_NULL) b = !nn;
            else if (op == HJ_P_IS_NOT_NULL) b = nn;
            else if (!nn) b = 0;
            else if (pc->v) {
                const int32_t x = pc->v[r], y = (int32_t)pr[k].ival;
                b = op == HJ_P_EQ ? x == y : op == HJ_P_NEQ ? x != y : op == HJ_P_LT ? x < y
                  : op == HJ_P_GT ? x > y : op == HJ_P_LEQ ? x <= y : x >= y;
            } else {
                const char* x = pc->ch + pc->off[r];
                const uint64_t xl = pc->off[r + 1] - pc->off[r], yl = pr[k].slen;
                if (op == HJ_P_LIKE || op == HJ_P_NOT_LIKE) {
                    const int mt = oracle_like(x, xl, pr[k].sval ? pr[k].sval : "", yl);
                    b = op == HJ_P_LIKE ? mt : !mt;
                } else {
                    int c = memcmp(x, pr[k].sval ? pr[k].sval : "", xl < yl ? xl : yl);
                    if (c == 0) c = xl < yl ? -1 : xl > yl ? 1 : 0;
                    b = op == HJ_P_EQ ? c == 0 : op == HJ_P_NEQ ? c != 0 : op == HJ_P_LT ? c < 0
                      : op == HJ_P_GT ? c > 0 : op == HJ_P_LEQ ? c <= 0 : c >= 0;
                }
            }
            st[sp++] = (uint8_t)b;
        }
        keep[r] &= st[0];
    }
    for (uint64_t k = 0; k < np && k < 32; ++k) {
        free(cols[k].v);
        free(cols[k].off);
        free(cols[k].ch);
        free(cols[k].ok);
    }
    return rc;
}

/* ------------------------------------------------------------------ load-time filter */
/* Rows pass iff every filter column is non-NULL and lo <= v <= hi (include/inner_column.h:
 * 258-259 ANDs the validity bitmap into every comparison; BETWEEN = GEQ AND LEQ,
 * tests/read_sql.cpp:585-591). Passing rows are re-encoded with the writer rules, keeping row
 * order and NULLs of the other columns (src/build_table.cpp:104-129, 319-371). */
static int filter_table(const hj_table* in, const hj_filter* f, uint64_t nf, const hj_pred* pr,
                        uint64_t np, hj_table* out, hj_column** out_cols,
                        const uint8_t*** out_ptrs, uint8_t** out_blocks) {
    uint64_t n = in->num_rows;
    uint8_t* keep = (uint8_t*)xmalloc(n ? n : 1);
    memset(keep, 1, n);
    {
        const int prc = apply_preds(in, pr, np, keep);
        if (prc) {
            free(keep);
            return prc;
        }
    }
    for (uint64_t k = 0; k < nf; ++k) {
        if (f[k].column >= in->num_columns || in->columns[f[k].column].type != HJ_INT32) {
            free(keep);
            return HJ_E_TYPE;
        }
        const hj_column* col = &in->columns[f[k].column];
        int32_t* v = (int32_t*)xmalloc(n * 4 + 4);
        uint8_t* ok = (uint8_t*)xmalloc(n + 1);
        oracle_decode_int32(col, v, ok);
        for (uint64_t r = 0; r < n; ++r)
            keep[r] &= (uint8_t)(ok[r] && v[r] >= f[k].lo && v[r] <= f[k].hi);
        free(v);
        free(ok);
    }
    uint64_t m = 0;
    for (uint64_t r = 0; r < n; ++r) m += keep[r];
    hj_column* cols = (hj_column*)xcalloc(in->num_columns ? in->num_columns : 1, sizeof(hj_column));
    const uint8_t** ptrs_all = NULL;
    uint64_t nptrs = 0;
    uint8_t** blocks = (uint8_t**)xcalloc(in->num_columns ? in->num_columns : 1, sizeof(uint8_t*));
    for (uint32_t c = 0; c < in->num_columns; ++c) {
        const hj_column* col = &in->columns[c];
        hj_out_column oc = {0};
        if (col->type == HJ_INT32) {
            int32_t* v = (int32_t*)xmalloc(n * 4 + 4);
            uint8_t* ok = (uint8_t*)xmalloc(n + 1);
            oracle_decode_int32(col, v, ok);
            uint64_t j = 0;
            for (uint64_t r = 0; r < n; ++r)
                if (keep[r]) {
                    v[j] = v[r];
                    ok[j] = ok[r];
                    ++j;
                }
            oracle_encode_int32(v, ok, m, &oc);
            free(v);
            free(ok);
        } else if (col->type == HJ_VARCHAR) {
            uint64_t nc = oracle_column_chars(col);
            uint64_t* off = (uint64_t*)xmalloc((n + 1) * 8);
            char* ch = (char*)xmalloc(nc + 1);
            uint8_t* ok = (uint8_t*)xmalloc(n + 1);
            oracle_decode_varchar(col, off, ch, ok);
            uint64_t* off2 = (uint64_t*)xmalloc((m + 1) * 8);
            char* ch2 = (char*)xmalloc(nc + 1);
            uint8_t* ok2 = (uint8_t*)xmalloc(m + 1);
            uint64_t j = 0, pos = 0;
            off2[0] = 0;
            for (uint64_t r = 0; r < n; ++r)
                if (keep[r]) {
                    uint64_t l = off[r + 1] - off[r];
                    memcpy(ch2 + pos, ch + off[r], l);
                    pos += l;
                    ok2[j] = ok[r];
                    off2[++j] = pos;
                }
            oracle_encode_varchar(off2, ch2, ok2, m, &oc);
            free(off);
            free(ch);
            free(ok);
            free(off2);
            free(ch2);
            free(ok2);
        } else {
            oc.type = col->type; /* other types are not produced by the filter path */
        }
        blocks[c] = oc.pages;
        cols[c].type = col->type;
        cols[c].num_pages = oc.num_pages;
        ptrs_all = (const uint8_t**)xrealloc(ptrs_all, (nptrs + oc.num_pages + 1) * sizeof(void*));
        for (uint32_t p = 0; p < oc.num_pages; ++p)
            ptrs_all[nptrs + p] = oc.pages + (size_t)p * PAGE;
        cols[c].pages = (const uint8_t* const*)(uintptr_t)nptrs; /* fixed up below */
        nptrs += oc.num_pages;
    }
    for (uint32_t c = 0; c < in->num_columns; ++c)
        cols[c].pages = ptrs_all + (uintptr_t)cols[c].pages;
    out->num_rows = m;
    out->num_columns = in->num_columns;
    out->columns = cols;
    *out_cols = cols;
    *out_ptrs = ptrs_all;
    *out_blocks = (uint8_t*)blocks;
    free(keep);
    return 0;
}

static uint64_t now_ns(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

int oracle_execute(const hj_plan* plan, int threads, hj_result* out, uint64_t* exec_ns,
                   char* err, size_t errlen) {
    memset(out, 0, sizeof *out);
    errstate es;
    memset(&es, 0, sizeof es);
    pthread_mutex_init(&es.mu, NULL);
    if (plan->root >= plan->num_nodes) {
        snprintf(err, errlen, "root out of range");
        return HJ_E_INVALID;
    }
    /* load-time filtering: filtered tables are appended as extra inputs (untimed) */
    uint64_t nin = plan->num_inputs;
    hj_table* inputs = (hj_table*)xcalloc(nin + plan->num_nodes + 1, sizeof(hj_table));
    memcpy(inputs, plan->inputs, nin * sizeof(hj_table));
    uint64_t* scan_table = (uint64_t*)xcalloc(plan->num_nodes, sizeof(uint64_t));
    void** owned = (void**)xcalloc(3 * (plan->num_nodes + 1), sizeof(void*));
    uint64_t nowned = 0;
    uint64_t* owned_ncols = (uint64_t*)xcalloc(plan->num_nodes + 1, sizeof(uint64_t));
    int rc = 0;
    for (uint64_t i = 0; i < plan->num_nodes && rc == 0; ++i) {
        const hj_node* nd = &plan->nodes[i];
        if (nd->kind != HJ_NODE_SCAN) continue;
        scan_table[i] = nd->base_table_id;
        if (nd->num_filters == 0 && nd->num_preds == 0) continue;
        if (nd->base_table_id >= nin) {
            rc = HJ_E_INVALID;
            break;
        }
        hj_column* cols;
        const uint8_t** ptrs;
        uint8_t* blocks;
        rc = filter_table(&plan->inputs[nd->base_table_id], nd->filters, nd->num_filters,
                          nd->preds, nd->num_preds,
                          &inputs[nin], &cols, &ptrs, &blocks);
        if (rc) break;
        owned_ncols[nowned / 3] = plan->inputs[nd->base_table_id].num_columns;
        owned[nowned++] = cols;
        owned[nowned++] = ptrs;
        owned[nowned++] = blocks;
        scan_table[i] = nin++;
    }
    if (rc == 0) {
        exec_ctx x = {plan, inputs, nin, scan_table, threads > 0 ? threads : default_threads(),
                      &es};
        uint64_t t0 = now_ns();
        vcol* res = execute_impl(&x, plan->root);
        const hj_node* root = &plan->nodes[plan->root];
        if (res && !atomic_load(&es.failed))
            rc = table_to_columnar(res, root->output_attrs, root->num_output_attrs, inputs, nin,
                                   out);
        uint64_t t1 = now_ns();
        if (exec_ns) *exec_ns = t1 - t0;
        free_cols(res, root->num_output_attrs);
        if (atomic_load(&es.failed)) rc = es.code ? es.code : HJ_E_INVALID;
    }
    if (rc && err) snprintf(err, errlen, "%s", atomic_load(&es.failed) ? es.msg : "oracle error");
    for (uint64_t k = 0; k < nowned; k += 3) {
        uint8_t** blocks = (uint8_t**)owned[k + 2];
        for (uint64_t c = 0; c < owned_ncols[k / 3]; ++c) free(blocks[c]);
        free(owned[k]);
        free(owned[k + 1]);
        free(owned[k + 2]);
    }
    free(owned);
    free(owned_ncols);
    free(inputs);
    free(scan_table);
    pthread_mutex_destroy(&es.mu);
    if (rc) oracle_result_free(out);
    return rc;
}

/* ------------------------------------------------------------------ nested-loop join */
/* Same plan semantics (post-order, output index into concat(left, right), NULL keys never
 * match), computed with decoded values and a quadratic loop; shares no code with the hash path
 * above beyond copy_scan's value decoding and the page writer. */
static vcol* nl_impl(exec_ctx* x, uint64_t node_idx) {
    const hj_node* node = &x->plan->nodes[node_idx];
    if (node->kind == HJ_NODE_SCAN)
        return copy_scan(&x->inputs[x->scan_table[node_idx]], x->scan_table[node_idx],
                         node->output_attrs, node->num_output_attrs, x->err);
    uint64_t nl = x->plan->nodes[node->left].num_output_attrs;
    uint64_t nr = x->plan->nodes[node->right].num_output_attrs;
    vcol* L = nl_impl(x, node->left);
    vcol* R = nl_impl(x, node->right);
    uint64_t nout = node->num_output_attrs;
    vcol* res = (vcol*)xcalloc(nout ? nout : 1, sizeof(vcol));
    for (uint64_t c = 0; c < nout; ++c) vcol_init(&res[c], node->output_attrs[c].type);
    if (L && R && node->left_attr < nl && node->right_attr < nr) {
        const vcol* lk = &L[node->left_attr];
        const vcol* rk = &R[node->right_attr];
        for (uint64_t i = 0; i < lk->rows; ++i) {
            uint64_t a = vcol_get(lk, i);
            if (!vt_is_int32(a)) continue;
            for (uint64_t j = 0; j < rk->rows; ++j) {
                uint64_t b = vcol_get(rk, j);
                if (!vt_is_int32(b) || (uint32_t)a != (uint32_t)b) continue;
                for (uint64_t c = 0; c < nout; ++c) {
                    uint64_t ci = node->output_attrs[c].index;
                    vcol_append(&res[c], ci < nl ? vcol_get(&L[ci], i) : vcol_get(&R[ci - nl], j));
                }
            }
        }
    }
    free_cols(L, nl);
    free_cols(R, nr);
    return res;
}

int oracle_nested_loop_execute(const hj_plan* plan, hj_result* out, char* err, size_t errlen) {
    memset(out, 0, sizeof *out);
    /* reuse the filter staging of oracle_execute by running it with the loop join */
    errstate es;
    memset(&es, 0, sizeof es);
    pthread_mutex_init(&es.mu, NULL);
    uint64_t nin = plan->num_inputs;
    hj_table* inputs = (hj_table*)xcalloc(nin + plan->num_nodes + 1, sizeof(hj_table));
    memcpy(inputs, plan->inputs, nin * sizeof(hj_table));
    uint64_t* scan_table = (uint64_t*)xcalloc(plan->num_nodes, sizeof(uint64_t));
    void* owned[3 * 64];
    uint64_t owned_nc[64], nowned = 0;
    int rc = 0;
    for (uint64_t i = 0; i < plan->num_nodes && rc == 0; ++i) {
        const hj_node* nd = &plan->nodes[i];
        if (nd->kind != HJ_NODE_SCAN) continue;
        scan_table[i] = nd->base_table_id;
        if (!nd->num_filters && !nd->num_preds) continue;
        if (nowned >= 3 * 64) {
            rc = HJ_E_INVALID;
            break;
        }
        hj_column* cols;
        const uint8_t** ptrs;
        uint8_t* blocks;
        rc = filter_table(&plan->inputs[nd->base_table_id], nd->filters, nd->num_filters,
                          nd->preds, nd->num_preds,
                          &inputs[nin], &cols, &ptrs, &blocks);
        if (rc) break;
        owned_nc[nowned / 3] = plan->inputs[nd->base_table_id].num_columns;
        owned[nowned++] = cols;
        owned[nowned++] = ptrs;
        owned[nowned++] = blocks;
        scan_table[i] = nin++;
    }
    if (rc == 0) {
        exec_ctx x = {plan, inputs, nin, scan_table, 1, &es};
        vcol* res = nl_impl(&x, plan->root);
        const hj_node* root = &plan->nodes[plan->root];
        if (!atomic_load(&es.failed))
            rc = table_to_columnar(res, root->output_attrs, root->num_output_attrs, inputs, nin,
                                   out);
        free_cols(res, root->num_output_attrs);
        if (atomic_load(&es.failed)) rc = es.code ? es.code : HJ_E_INVALID;
    }
    if (rc && err) snprintf(err, errlen, "%s", atomic_load(&es.failed) ? es.msg : "oracle error");
    for (uint64_t k = 0; k < nowned; k += 3) {
        uint8_t** blocks = (uint8_t**)owned[k + 2];
        for (uint64_t c = 0; c < owned_nc[k / 3]; ++c) free(blocks[c]);
        free(owned[k]);
        free(owned[k + 1]);
        free(owned[k + 2]);
    }
    free(inputs);
    free(scan_table);
    pthread_mutex_destroy(&es.mu);
    if (rc) oracle_result_free(out);
    return rc;
}

/* ------------------------------------------------------------------ raw-key PDH join */
uint64_t oracle_pdh_join(const uint32_t* build_keys, const uint8_t* build_valid, uint64_t nb,
                         const uint32_t* probe_keys, const uint8_t* probe_valid, uint64_t np,
                         int threads, uint64_t* pairs, uint64_t cap) {
    errstate es;
    memset(&es, 0, sizeof es);
    pthread_mutex_init(&es.mu, NULL);
    pdh ix;
    memset(&ix, 0, sizeof ix);
    keysrc src = {NULL, build_keys, build_valid, nb};
    pdh_build(&ix, &src, (uint64_t)(threads > 0 ? threads : default_threads()), &es);
    uint64_t cnt = 0;
    int overflow = 0;
    for (uint64_t r = 0; r < np; ++r) {
        if (probe_valid && !probe_valid[r]) continue;
        uint64_t b, e;
        pdh_range(&ix, probe_keys[r], &b, &e);
        for (uint64_t i = b; i < e; ++i) {
            if (ix.tuples[i].key != probe_keys[r]) continue;
            if (cnt < cap) {
                pairs[2 * cnt] = r;
                pairs[2 * cnt + 1] = ix.tuples[i].row;
            } else {
                overflow = 1;
            }
            ++cnt;
        }
    }
    pdh_free(&ix);
    pthread_mutex_destroy(&es.mu);
    return overflow ? (uint64_t)-1 : cnt;
}

/* ------------------------------------------------------------------ UnchainedHash */
/* include/unchained.h:14-247 for Key = Payload = uint64_t; std::hash<uint64_t> is the identity
 * in libstdc++, so hash = key * 11400714819323198485 (2^64 / golden ratio). */
typedef struct {
    uint64_t key, payload;
    uint16_t suffix;
} uc_tuple;

struct oracle_unchained {
    uc_tuple* t;
    uint64_t* hashes;
    uint32_t n, cap;
    uint64_t* dir;
    uint32_t dir_size;
    uint8_t prefix_bits;
};

static inline uint64_t uc_hash(uint64_t key) { return key * 11400714819323198485ull; }
static inline uint32_t uc_prefix(const oracle_unchained* h, uint64_t hv) {
    if (h->prefix_bits == 0) return 0;
    return (uint32_t)(hv >> (64 - h->prefix_bits));
}
static inline uint16_t uc_bloom_mask(uint16_t s) {
    return (uint16_t)((1u << (s & 0xF)) | (1u << ((s >> 4) & 0xF)) | (1u << ((s >> 8) & 0xF)) |
                      (1u << ((s >> 12) & 0xF)));
}
static uint32_t uc_next_pow2(uint32_t x) {
    if (x == 0) return 1;
    --x;
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    return x + 1;
}

oracle_unchained* oracle_unchained_new(void) {
    return (oracle_unchained*)xcalloc(1, sizeof(oracle_unchained));
}
void oracle_unchained_clear(oracle_unchained* h) {
    free(h->t);
    free(h->hashes);
    free(h->dir);
    memset(h, 0, sizeof *h);
}
void oracle_unchained_free(oracle_unchained* h) {
    if (!h) return;
    oracle_unchained_clear(h);
    free(h);
}
void oracle_unchained_insert(oracle_unchained* h, uint64_t key, uint64_t payload) {
    if (h->n == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 16;
        h->t = (uc_tuple*)xrealloc(h->t, h->cap * sizeof(uc_tuple));
        h->hashes = (uint64_t*)xrealloc(h->hashes, h->cap * sizeof(uint64_t));
    }
    uint64_t hv = uc_hash(key);
    h->t[h->n] = (uc_tuple){key, payload, (uint16_t)(hv & 0xFFFFu)};
    h->hashes[h->n] = hv;
    h->n++;
}
/* finalize_build, unchained.h:163-217 */
void oracle_unchained_finalize(oracle_unchained* h) {
    uint32_t n = h->n;
    if (n == 0) {
        oracle_unchained_clear(h);
        return;
    }
    h->dir_size = uc_next_pow2(n);
    h->prefix_bits = 0;
    for (uint32_t tmp = h->dir_size; tmp > 1; tmp >>= 1) h->prefix_bits++;
    free(h->dir);
    h->dir = (uint64_t*)xcalloc((size_t)h->dir_size + 1, sizeof(uint64_t));
    uint32_t* counts = (uint32_t*)xcalloc(h->dir_size, sizeof(uint32_t));
    uint32_t* pc = (uint32_t*)xmalloc((size_t)n * 4);
    for (uint32_t i = 0; i < n; ++i) {
        pc[i] = uc_prefix(h, h->hashes[i]);
        counts[pc[i]]++;
    }
    uint32_t* starts = (uint32_t*)xcalloc((size_t)h->dir_size + 1, sizeof(uint32_t));
    uint32_t off = 0;
    for (uint32_t p = 0; p < h->dir_size; ++p) {
        starts[p] = off;
        off += counts[p];
    }
    starts[h->dir_size] = off;
    uc_tuple* nt = (uc_tuple*)xmalloc((size_t)n * sizeof(uc_tuple));
    uint32_t* wp = (uint32_t*)xmalloc((size_t)h->dir_size * 4);
    uint16_t* blooms = (uint16_t*)xcalloc(h->dir_size, 2);
    memcpy(wp, starts, (size_t)h->dir_size * 4);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t p = pc[i];
        nt[wp[p]++] = h->t[i];
        blooms[p] |= uc_bloom_mask(h->t[i].suffix);
    }
    free(h->t);
    h->t = nt;
    h->cap = n;
    free(h->hashes);
    h->hashes = NULL;
    for (uint32_t p = 0; p < h->dir_size; ++p)
        h->dir[p] = ((uint64_t)starts[p] << 16) | (uint64_t)blooms[p];
    h->dir[h->dir_size] = (uint64_t)starts[h->dir_size] << 16;
    free(counts);
    free(pc);
    free(starts);
    free(wp);
    free(blooms);
}
/* probe, unchained.h:220-246 */
uint64_t oracle_unchained_probe(const oracle_unchained* h, uint64_t key, uint64_t* payloads,
                                uint64_t cap) {
    if (h->dir_size == 0 || h->dir == NULL) return 0;
    uint64_t hv = uc_hash(key);
    uint32_t p = uc_prefix(h, hv);
    uint16_t suffix = (uint16_t)(hv & 0xFFFFu);
    if (p >= h->dir_size) return 0;
    uint64_t e = h->dir[p], ne = h->dir[p + 1];
    uint16_t bloom = (uint16_t)(e & 0xFFFFu), mask = uc_bloom_mask(suffix);
    if ((bloom & mask) != mask) return 0;
    uint64_t cnt = 0;
    for (uint32_t i = (uint32_t)(e >> 16); i < (uint32_t)(ne >> 16); ++i) {
        if (h->t[i].suffix != suffix) continue;
        if (h->t[i].key == key) {
            if (cnt < cap) payloads[cnt] = h->t[i].payload;
            ++cnt;
        }
    }
    return cnt;
}
uint32_t oracle_unchained_dir_size(const oracle_unchained* h) { return h->dir_size; }

/* ------------------------------------------------------------------ .tbl cache files */
/* DumpTable::dump (include/table.h:200-211): TableMeta (include/table.h:30-35) written field by
 * field at its C layout offsets, zeros to the end of the 8192-byte header page, then each
 * column's pages. */
int oracle_cache_write(const hj_table* t, const char* path) {
    if (t->num_columns > 16) return -1;
    unsigned char head[HJ_PAGE_SIZE];
    memset(head, 0, sizeof head);
    uint64_t v = t->num_rows;
    memcpy(head + 0, &v, 8);
    v = t->num_columns;
    memcpy(head + 8, &v, 8);
    for (uint32_t i = 0; i < t->num_columns; ++i) {
        int32_t ty = t->columns[i].type;
        uint64_t np = t->columns[i].num_pages;
        memcpy(head + 16 + 4 * i, &ty, 4);
        memcpy(head + 80 + 8 * i, &np, 8);
    }
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int ok = fwrite(head, 1, sizeof head, f) == sizeof head;
    for (uint32_t i = 0; ok && i < t->num_columns; ++i)
        for (uint32_t j = 0; ok && j < t->columns[i].num_pages; ++j)
            ok = fwrite(t->columns[i].pages[j], 1, HJ_PAGE_SIZE, f) == HJ_PAGE_SIZE;
    if (fclose(f) != 0) ok = 0;
    return ok ? 0 : -1;
}

/* Table::from_cache (src/build_table.cpp:147-203), reading instead of mapping: the header's
 * counts, then num_pages[i] pages per column in column order. */
int oracle_cache_read(const char* path, hj_result* out) {
    memset(out, 0, sizeof *out);
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    unsigned char head[HJ_PAGE_SIZE];
    if (fread(head, 1, sizeof head, f) != sizeof head) {
        fclose(f);
        return -1;
    }
    uint64_t rows, ncols, pages[16], total = 0;
    int32_t types[16];
    memcpy(&rows, head + 0, 8);
    memcpy(&ncols, head + 8, 8);
    if (ncols > 16) {
        fclose(f);
        return -1;
    }
    for (uint64_t i = 0; i < ncols; ++i) {
        memcpy(&types[i], head + 16 + 4 * i, 4);
        memcpy(&pages[i], head + 80 + 8 * i, 8);
        total += pages[i];
    }
    fseek(f, 0, SEEK_END);
    const long size = ftell(f);
    if (size < 0 || (uint64_t)size != (total + 1) * HJ_PAGE_SIZE) {
        fclose(f);
        return -1;
    }
    fseek(f, HJ_PAGE_SIZE, SEEK_SET);
    out->num_rows = rows;
    out->num_columns = (uint32_t)ncols;
    out->columns = (hj_out_column*)calloc(ncols ? ncols : 1, sizeof(hj_out_column));
    int ok = out->columns != NULL;
    for (uint64_t i = 0; ok && i < ncols; ++i) {
        hj_out_column* c = &out->columns[i];
        c->type = types[i];
        c->num_pages = (uint32_t)pages[i];
        if (!pages[i]) continue;
        c->pages = (uint8_t*)malloc(pages[i] * HJ_PAGE_SIZE);
        ok = c->pages && fread(c->pages, 1, pages[i] * HJ_PAGE_SIZE, f) == pages[i] * HJ_PAGE_SIZE;
    }
    fclose(f);
    if (!ok) {
        oracle_result_free(out);
        return -1;
    }
    return 0;
}
