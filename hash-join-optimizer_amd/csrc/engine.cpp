// engine.cpp - plan execution on one GPU and the extern "C" layer of include/hjoin.h.
//
// execute() (src/execute.cpp:584-608) walks the plan post-order like the reference, but
// intermediate results are late-materialized row-id lists instead of value_t columns:
//   scan node  -> a lazy relation over the resident base table (identity row ids + filters)
//   join node  -> build side: key extraction (ballot compaction) -> LSD radix sort by the
//                 Fibonacci hash -> unchained directory (slot starts + Bloom bits);
//                 probe side: per-block match counts -> device scan -> write of one row-id
//                 array per base table the node's output_attrs still reference
//   root       -> gather of the output columns (INT32 values, VARCHAR bytes) into flat
//                 HBM buffers; page encoding happens on the host afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/hjoin.h"
#include "hj_internal.hpp"
#include "hostpool.hpp"
#include "kernels.hpp"
#include "pagecodec.hpp"
#include "pagewrite.hpp"
#include "tblcache.hpp"
#include "pipeline.hpp"
#include "predicate.hpp"
#include "materialize.hpp"

namespace hj {

// string bytes of an output column are allocated from rows x longest string (no host round trip)
// up to this size; beyond it they are counted first
constexpr uint64_t kCharsBound = 1ull << 30;
// counters of the pipeline (4 words) and of each of kMaxStages build jobs (4 words each)
constexpr uint32_t kMetaWords = 4 * (kMaxStages + 1);
// totals of the direct output: rows, string bytes per column, overflow
constexpr uint32_t kEmitOut = 2 + kEmitCols;
// blocks of the fused pass per block of the direct output. Each output block first sums the
// counts of every pass block before its own, so fewer, larger groups suit a small output; when no
// scan of the pipeline is filtered, nearly every source row of a key join finds its match and
// the output is as large as the source, which wants more blocks.
constexpr uint32_t kEmitGroup = 8;
constexpr uint32_t kEmitGroupUnfiltered = 4;
static_assert(kMaxOutStr >= kEmitCols, "every output column may be a string column");

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string& m) {
    std::lock_guard<std::mutex> g(g_err_mu);
    g_err = m;
}

// Wait for a query's round trip: poll the stream for up to 2 ms (a short query's result is
// ready within microseconds, and a sleeping host thread takes longer than that to wake), then
// block.
void wait_stream(hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady)
            throw Error(HJ_E_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(2000)) break;
    }
    HJ_HIP(hipStreamSynchronize(s));
}

template <class T>
T fetch(hj_context* ctx, const T* dptr) {
    HJ_HIP(hipMemcpyAsync(ctx->pinned, dptr, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    wait_stream(ctx->stream);
    T v;
    std::memcpy(&v, ctx->pinned, sizeof(T));
    return v;
}

// host-side timing of a query's phases (environment HJ_TRACE_HOST=1): microseconds since the
// call began, printed to stderr when it returns
struct HostTrace {
    bool on = std::getenv("HJ_TRACE_HOST") != nullptr;
    std::chrono::steady_clock::time_point t0;
    std::vector<std::pair<const char*, double>> marks;
    void begin() {
        if (!on) return;
        marks.clear();
        t0 = std::chrono::steady_clock::now();
    }
    void mark(const char* what) {
        if (!on) return;
        marks.push_back({what, std::chrono::duration<double, std::micro>(
                                   std::chrono::steady_clock::now() - t0).count()});
    }
    void print() {
        if (!on) return;
        std::string line = "[host]";
        char b[64];
        for (auto& m : marks) {
            std::snprintf(b, sizeof b, " %s %.1f", m.first, m.second);
            line += b;
        }
        std::fprintf(stderr, "%s\n", line.c_str());
    }
};
HostTrace& host_trace() {
    static thread_local HostTrace t;
    return t;
}

uint32_t log2_pow2_ceil(uint64_t n) {
    uint32_t k = 0;
    while ((1ull << k) < n) ++k;
    return k;
}

// field widths of the reference's StrRef (include/value_t.h:14-17): the reference's copy_scan
// encodes every string of a scanned VARCHAR column as (table, column, page, index) and throws
// std::out_of_range when a field does not fit (include/value_t.h:36-47)
constexpr uint64_t kStrRefTableMask = (1ull << 10) - 1;
constexpr uint64_t kStrRefColumnMask = (1ull << 10) - 1;
constexpr uint64_t kStrRefPageMask = (1ull << 18) - 1;

// ------------------------------------------------------------------ relations
struct Slot {
    uint32_t table;
    const uint32_t* rid;  // device row ids into the base table; nullptr = identity
};
struct AttrRef {
    uint32_t slot;
    uint32_t col;
    int32_t type;
};
struct Rel {
    bool is_scan = false;
    uint64_t n = 0;
    std::vector<Slot> slots;
    std::vector<AttrRef> attrs;
    uint32_t table = 0;                // scan only
    std::vector<hj_filter> filters;    // scan only
    DCol pred_view{};                  // scan only: the predicate program's 0 / 1 per row
};

// filter column standing for a scan's predicate program (its result, filtered on = 1)
constexpr uint32_t kPredColumn = 0xffffffffu;

// a LIKE pattern as tokens: UTF-8 characters (the decoding of oracle_like and the device
// matcher: a byte that does not start a complete sequence is a character of its own), with '%'
// and '_' as kLikeAnyRun / kLikeAnyOne
std::vector<uint32_t> like_tokens(const uint8_t* p, uint64_t n) {
    std::vector<uint32_t> t;
    for (uint64_t i = 0; i < n;) {
        const uint32_t b = p[i];
        uint32_t need = 0, cp = b;
        if (b >= 0xC2 && b <= 0xDF) { need = 1; cp = b & 0x1F; }
        else if (b >= 0xE0 && b <= 0xEF) { need = 2; cp = b & 0x0F; }
        else if (b >= 0xF0 && b <= 0xF4) { need = 3; cp = b & 0x07; }
        else if (b >= 0x80) { need = 0; cp = 0x110000u + b; }
        uint32_t len = 1;
        if (need) {
            bool ok = true;
            for (uint32_t k = 1; k <= need && ok; ++k)
                ok = i + k < n && (p[i + k] & 0xC0) == 0x80;
            if (ok) {
                for (uint32_t k = 1; k <= need; ++k) cp = (cp << 6) | (p[i + k] & 0x3F);
                len = need + 1;
            } else {
                cp = 0x110000u + b;
            }
        }
        t.push_back(b == '%' ? kLikeAnyRun : b == '_' ? kLikeAnyOne : cp);
        i += len;
    }
    return t;
}

class Executor {
  public:
    Executor(hj_context* ctx, const hj_plan* plan, const hj_dtable* const* tables)
        : ctx_(ctx), plan_(plan), tables_(tables), s_(ctx->stream), A_(ctx->arena) {}
    // cached tables read by this query's kernels are released once those kernels are done
    ~Executor() {
        if (cache_refs_.empty()) return;
        (void)hipStreamSynchronize(s_);
        for (auto& k : cache_refs_) ctx_->dcache->release(k);
    }


    hj_dresult* run() {
        if (!plan_->nodes || plan_->root >= plan_->num_nodes)
            throw Error(HJ_E_INVALID, "plan root out of range");
        depth_ = 0;
        check_strrefs(plan_->root);
        depth_ = 0;
        host_trace().mark("checked");
        prepare();
        host_trace().mark("prepared");
        if (ctx_->use_pipeline) {
            for (int attempt = 0; attempt < 2; ++attempt) {
                Rel pr;
                hj_dresult* dr = nullptr;
                const Pipe st = try_pipeline(pr, &dr, attempt == 0);
                if (st == Pipe::kResult) return dr;
                if (st == Pipe::kRel) return materialize(pr, plan_->nodes[plan_->root]);
                A_.clear_to(prep_mark_);  // the query's column index stays
                if (st != Pipe::kRetry) break;  // not a chain, or a duplicate build key
            }
        }
        flush_pending();
        Rel r = node(plan_->root);
        if (direct_) return direct_;  // the root join wrote its output columns itself
        return materialize(r, plan_->nodes[plan_->root]);
    }

  private:
    hj_context* ctx_;
    const hj_plan* plan_;
    const hj_dtable* const* tables_;
    hipStream_t s_;
    Arena& A_;
    int depth_ = 0;
    hj_dresult* direct_ = nullptr;  // set when the root join wrote the result columns
    // the query's own row-space access to the columns its scans read (built by prepare(), in
    // the arena below prep_mark_): a group index, or for an INT32 column without one a decoded
    // row-space array
    struct QCol {
        uint64_t* gidx = nullptr;
        int32_t* flat = nullptr;
        uint64_t* vbits = nullptr;
    };
    std::map<uint64_t, QCol> qcols_;
    size_t prep_mark_ = 0;

    static uint64_t qkey(uint32_t t, uint32_t c) { return ((uint64_t)t << 32) | c; }

    // the view of column c of input t that this query's kernels read
    DCol qview(uint32_t t, uint32_t c) const {
        DCol d = column(t, c).view();
        auto it = qcols_.find(qkey(t, c));
        if (it != qcols_.end()) {
            d.gidx = it->second.gidx;
            d.flat = it->second.flat;
            d.vbits = it->second.vbits;
        }
        return d;
    }

    // Index (or, where a page holds fewer than 64 rows before the last one, decode) every INT32
    // and VARCHAR column that a scan of the plan reads: outputs, filters and predicates. This is
    // the query's first device work; the reference's copy_scan does the same job in its execute()
    // (src/execute.cpp:265-369).
    //
    // A left-deep chain's builds read only the build scans' INT32 keys and filters, so the
    // probe side's columns (the source scan's, and strings that are only output) are indexed by
    // extra blocks of the build's filter launch (pending_), after it on the join-at-a-time path.
    // Dense INT32 columns of uniform pages need no index (arithmetic) unless the fused probe
    // streams them (the source table's).
    void prepare() {
        qcols_.clear();
        std::memset(&pending_, 0, sizeof pending_);
        uint64_t source = 0;
        std::vector<uint64_t> joins;
        int64_t src_table = -1;
        if (ctx_->use_pipeline && chain(plan_->root, source, joins, 0) && !joins.empty() &&
            plan_->nodes[source].base_table_id < plan_->num_inputs)
            src_table = (int64_t)plan_->nodes[source].base_table_id;
        struct Want {
            uint32_t t, c;
            bool late;  // may wait for the build launch
        };
        std::vector<Want> want;
        for (uint64_t i = 0; i < plan_->num_nodes; ++i) {
            const hj_node& nd = plan_->nodes[i];
            if (nd.kind != HJ_NODE_SCAN || nd.base_table_id >= plan_->num_inputs ||
                !tables_[nd.base_table_id])
                continue;
            const uint32_t t = (uint32_t)nd.base_table_id;
            const size_t nc = tables_[t]->cols.size();
            const bool src = (int64_t)t == src_table;
            auto add = [&](uint64_t c, bool late) {
                if (c < nc) want.push_back({t, (uint32_t)c, src_table >= 0 && late});
            };
            for (uint64_t a = 0; a < nd.num_output_attrs && nd.output_attrs; ++a) {
                const uint64_t c = nd.output_attrs[a].index;
                add(c, src || (c < nc && tables_[t]->cols[c].type == HJ_VARCHAR));
            }
            for (uint64_t f = 0; f < nd.num_filters && nd.filters; ++f) add(nd.filters[f].column, src);
            // predicate programs run on the scan's rows before the builds
            for (uint64_t e = 0; e < nd.num_preds && nd.preds; ++e)
                if (nd.preds[e].op < HJ_P_AND) add(nd.preds[e].column, false);
        }
        for (auto& w : want)  // wanted early anywhere: early
            for (auto& o : want)
                if (!o.late && o.t == w.t && o.c == w.c) w.late = false;
        GroupIndexBatch now;
        std::memset(&now, 0, sizeof now);
        auto launch_now = [&](GroupIndexBatch& b) {
            if (!b.n) return;
            { Prof p_(ctx_, "group_index"); launch_group_index(b, s_); }
            std::memset(&b, 0, sizeof b);
        };
        for (auto& w : want) {
            const uint64_t key = qkey(w.t, w.c);
            if (qcols_.count(key)) continue;
            const DevColumn& dc = column(w.t, w.c);
            if (dc.type != HJ_INT32 && dc.type != HJ_VARCHAR) continue;
            QCol q;
            const uint64_t G = (dc.num_rows + 63) / 64;
            const bool dense = dc.view().umagic != 0;
            if (dense && (int64_t)w.t != src_table) continue;  // addressed by arithmetic
            if (dc.regular) {
                q.gidx = scratch<uint64_t>(2 * G + 2);
                if (G) {
                    GroupIndexBatch& b = w.late ? pending_ : now;
                    if (b.n == kIndexCols) launch_now(b);  // a full batch runs at once
                    b.col[b.n] = dc.view();
                    b.out[b.n] = q.gidx;
                    b.blk0[b.n + 1] = b.blk0[b.n] + (uint32_t)((G + 255) / 256);
                    if (&b == &pending_) pending_src_[b.n] = (int64_t)w.t == src_table;
                    ++b.n;
                }
            } else if (dc.type == HJ_INT32) {
                q.flat = scratch<int32_t>(dc.num_rows + 1);
                if (!dc.all_dense) {
                    q.vbits = scratch<uint64_t>(G + 1);
                    HJ_HIP(hipMemsetAsync(q.vbits, 0, (G + 1) * 8, s_));
                }
                Prof p_(ctx_, "decode_int32");
                launch_decode_int32(dc.view(), q.flat, q.vbits, s_);
            } else {
                continue;  // a VARCHAR column with long strings: rows found by page search
            }
            qcols_[key] = q;
        }
        launch_now(now);
        prep_mark_ = A_.mark();
    }

    // the probe side's index work not yet run (a pipeline takes the source's columns into its
    // fused probe and the rest into its build's filter launch)
    GroupIndexBatch pending_;
    bool pending_src_[kIndexCols] = {};  // entry i is a column of the pipeline's source
    void flush_pending() {
        if (!pending_.n) return;
        { Prof p_(ctx_, "group_index"); launch_group_index(pending_, s_); }
        std::memset(&pending_, 0, sizeof pending_);
    }

    template <class T>
    T* scratch(uint64_t n) {
        return (T*)A_.scratch(n * sizeof(T));
    }
    template <class T>
    T* result(uint64_t n) {
        return (T*)A_.result(n * sizeof(T));
    }

    const hj_dtable* table(uint64_t id) const {
        if (id >= plan_->num_inputs || !tables_[id])
            throw Error(HJ_E_INVALID, "scan of an input table that does not exist");
        return tables_[id];
    }
    const DevColumn& column(uint32_t t, uint32_t c) const {
        const hj_dtable* tb = table(t);
        if (c >= tb->cols.size()) throw Error(HJ_E_INVALID, "column index out of range");
        return tb->cols[c];
    }

    // The reference's scans encode each string of an output VARCHAR column as a StrRef and
    // throw where a field overflows (src/execute.cpp:323-358, include/value_t.h:36-47); the
    // engine keeps no StrRefs but reports the same error, with the same message, for the first
    // scan in the reference's post-order (left subtree, right subtree) whose output overflows.
    void check_strrefs(uint64_t idx) {
        if (idx >= plan_->num_nodes || ++depth_ > 4096) return;  // reported by the plan walk
        const hj_node& nd = plan_->nodes[idx];
        if (nd.kind == HJ_NODE_JOIN) {
            check_strrefs(nd.left);
            check_strrefs(nd.right);
        } else if (nd.kind == HJ_NODE_SCAN && nd.base_table_id < plan_->num_inputs &&
                   tables_[nd.base_table_id]) {
            const hj_dtable* tb = tables_[nd.base_table_id];
            for (uint64_t i = 0; i < nd.num_output_attrs && nd.output_attrs; ++i) {
                const uint64_t c = nd.output_attrs[i].index;
                if (c >= tb->cols.size() || tb->cols[c].type != HJ_VARCHAR) continue;
                const auto over = [](const char* field, uint64_t v, uint64_t mask) {
                    return Error(HJ_E_RANGE, std::string("StrRef ") + field + " overflow: " +
                                                 std::to_string(v) + " > " + std::to_string(mask));
                };
                const DevColumn& dc = tb->cols[c];
                if (!dc.has_strings) continue;
                if (nd.base_table_id > kStrRefTableMask)
                    throw over("table", nd.base_table_id, kStrRefTableMask);
                if (c > kStrRefColumnMask) throw over("column", c, kStrRefColumnMask);
                if (dc.strref_page_over != ~0ull)
                    throw over("page", dc.strref_page_over, kStrRefPageMask);
            }
        }
        --depth_;
    }

    Rel node(uint64_t idx) {
        if (idx >= plan_->num_nodes) throw Error(HJ_E_INVALID, "plan node index out of range");
        if (++depth_ > 4096) throw Error(HJ_E_INVALID, "plan is not a tree");
        const hj_node& nd = plan_->nodes[idx];
        Rel r;
        if (nd.kind == HJ_NODE_SCAN) {
            r = scan(nd);
        } else if (nd.kind == HJ_NODE_JOIN) {
            Rel L = node(nd.left);
            Rel R = node(nd.right);
            r = join(nd, L, R);
        } else {
            throw Error(HJ_E_INVALID, "unknown plan node kind");
        }
        --depth_;
        return r;
    }

    Rel scan(const hj_node& nd, bool run_preds = true) {
        const hj_dtable* tb = table(nd.base_table_id);
        Rel r;
        r.is_scan = true;
        r.table = (uint32_t)nd.base_table_id;
        r.n = tb->num_rows;
        r.slots.push_back({r.table, nullptr});
        for (uint64_t i = 0; i < nd.num_output_attrs; ++i) {
            const hj_attr& a = nd.output_attrs[i];
            if (a.index >= tb->cols.size()) throw Error(HJ_E_INVALID, "scan attribute out of range");
            r.attrs.push_back({0, (uint32_t)a.index, tb->cols[a.index].type});
        }
        if (nd.num_filters + (nd.num_preds ? 1 : 0) > (uint64_t)kMaxFilters)
            throw Error(HJ_E_INVALID, "too many filters on one scan");
        for (uint64_t f = 0; f < nd.num_filters; ++f) {
            const hj_filter& fl = nd.filters[f];
            if (fl.column >= tb->cols.size() || tb->cols[fl.column].type != HJ_INT32)
                throw Error(HJ_E_TYPE, "filter column must be an INT32 column of the scanned table");
            r.filters.push_back(fl);
        }
        if (nd.num_preds && run_preds) predicate(nd, tb, r);
        return r;
    }

    // the scan's predicate program, run once over the table's rows into an INT32 0 / 1 array
    // that the scan filters on like any range filter
    void predicate(const hj_node& nd, const hj_dtable* tb, Rel& r) {
        if (nd.num_preds > (uint64_t)kMaxPredOps)
            throw Error(HJ_E_INVALID, "a scan predicate has more than 32 entries");
        if (!nd.preds) throw Error(HJ_E_INVALID, "scan predicate missing");
        if (tb->cols.empty()) throw Error(HJ_E_INVALID, "predicate on a table without columns");
        PredProg pp;
        std::memset(&pp, 0, sizeof pp);
        std::vector<uint8_t> lit;
        std::vector<uint32_t> used;
        int depth = 0;
        for (uint64_t k = 0; k < nd.num_preds; ++k) {
            const hj_pred& e = nd.preds[k];
            PredOp& o = pp.op[k];
            o.op = e.op;
            if (e.op == HJ_P_AND || e.op == HJ_P_OR || e.op == HJ_P_NOT) {
                const int need = e.op == HJ_P_NOT ? 1 : 2;
                if (depth < need) throw Error(HJ_E_INVALID, "scan predicate: operator without operands");
                depth -= need - 1;
                continue;
            }
            if (e.op < HJ_P_EQ || e.op > HJ_P_IS_NOT_NULL)
                throw Error(HJ_E_INVALID, "scan predicate: unknown operator");
            if (e.column >= tb->cols.size())
                throw Error(HJ_E_INVALID, "scan predicate: column out of range");
            const int32_t ty = tb->cols[e.column].type;
            if (ty != HJ_INT32 && ty != HJ_VARCHAR)
                throw Error(HJ_E_TYPE, "scan predicates take INT32 and VARCHAR columns");
            const bool like = e.op == HJ_P_LIKE || e.op == HJ_P_NOT_LIKE;
            if (like && ty != HJ_VARCHAR) throw Error(HJ_E_TYPE, "LIKE needs a VARCHAR column");
            size_t slot = 0;
            while (slot < used.size() && used[slot] != e.column) ++slot;
            if (slot == used.size()) {
                if (used.size() == (size_t)kMaxPredCols)
                    throw Error(HJ_E_INVALID, "a scan predicate reads more than 8 columns");
                used.push_back(e.column);
            }
            o.col = (int32_t)slot;
            if (ty == HJ_INT32) {
                o.ival = (int32_t)e.ival;  // the reference's static_cast<int32_t> of the literal
            } else if (e.op < HJ_P_IS_NULL) {
                if (e.slen && !e.sval) throw Error(HJ_E_INVALID, "scan predicate: literal missing");
                while (lit.size() % 4) lit.push_back(0);
                o.loff = (uint32_t)lit.size();
                if (like) {
                    const std::vector<uint32_t> t = like_tokens((const uint8_t*)e.sval, e.slen);
                    if (t.size() > (size_t)kMaxLikeTokens)
                        throw Error(HJ_E_INVALID, "LIKE patterns longer than 63 characters are not supported");
                    o.llen = (uint32_t)t.size();
                    uint32_t run = 0;
                    bool bytes = t.size() <= (size_t)kMaxByteTokens;
                    for (size_t j = 0; j < t.size(); ++j) {
                        if (t[j] == kLikeAnyRun) o.pct |= 1ull << (j + 1);
                        if (t[j] == kLikeAnyOne) o.any |= 1ull << (j + 1);
                        run = t[j] == kLikeAnyRun ? run + 1 : 0;
                        o.pad = std::max(o.pad, run);
                        bytes = bytes && (t[j] == kLikeAnyRun || t[j] < 128);
                    }
                    o.ival = bytes ? 1 : 0;
                    const uint8_t* tb8 = (const uint8_t*)t.data();
                    lit.insert(lit.end(), tb8, tb8 + t.size() * 4);
                } else {
                    if (e.slen >= (1ull << 32)) throw Error(HJ_E_INVALID, "scan predicate: literal too long");
                    o.llen = (uint32_t)e.slen;
                    lit.insert(lit.end(), (const uint8_t*)e.sval, (const uint8_t*)e.sval + e.slen);
                }
            }
            if (++depth > 32) throw Error(HJ_E_INVALID, "scan predicate deeper than 32");
        }
        if (depth != 1) throw Error(HJ_E_INVALID, "a scan predicate must leave exactly one result");
        pp.nops = (int32_t)nd.num_preds;
        pp.ncol = (int32_t)used.size();
        for (size_t i = 0; i < used.size(); ++i) pp.col[i] = qview(r.table, used[i]);
        uint8_t* dlit = result<uint8_t>(lit.size() + 4);
        if (!lit.empty()) {
            HJ_HIP(hipMemcpyAsync(dlit, lit.data(), lit.size(), hipMemcpyHostToDevice, s_));
            HJ_HIP(hipStreamSynchronize(s_));
        }
        pp.lit = dlit;
        pp.n = tb->num_rows;
        pp.out = result<int32_t>(pp.n ? pp.n : 1);
        { Prof p_(ctx_, "predicate"); eval_predicate(pp, s_); }
        // row structure (pages, row index) of column 0; values from the 0 / 1 array
        r.pred_view = tb->cols[0].view();
        r.pred_view.type = HJ_INT32;
        r.pred_view.flat = pp.out;
        r.pred_view.vbits = nullptr;
        r.filters.push_back({kPredColumn, 1, 1, 0});
    }

    DCol filter_view(const Rel& r, uint32_t col) const {
        return col == kPredColumn ? r.pred_view : qview(r.table, col);
    }

    ScanSrc scan_src(const Rel& r, const AttrRef& key) const {
        ScanSrc s;
        std::memset(&s, 0, sizeof s);
        s.key = qview(r.table, key.col);
        s.nf = (int32_t)r.filters.size();
        for (size_t f = 0; f < r.filters.size(); ++f) {
            s.fcol[f] = filter_view(r, r.filters[f].column);
            s.flo[f] = r.filters[f].lo;
            s.fhi[f] = r.filters[f].hi;
        }
        return s;
    }

    RelSrc rel_src(const Rel& r, const AttrRef& key) const {
        RelSrc s;
        s.key = qview(r.slots[key.slot].table, key.col);
        s.rid = r.slots[key.slot].rid;
        s.n = r.n;
        return s;
    }

    const AttrRef& key_attr(const Rel& r, uint64_t idx) const {
        if (idx >= r.attrs.size()) throw Error(HJ_E_INVALID, "join attribute out of range");
        const AttrRef& a = r.attrs[idx];
        const Slot& sl = r.slots[a.slot];
        if (column(sl.table, a.col).type != HJ_INT32)
            throw Error(HJ_E_TYPE, "Join key isn't of type int32_t");
        return a;
    }

    Rel join(const hj_node& nd, Rel& L, Rel& R) {
        const bool bl = nd.build_left != 0;
        Rel& B = bl ? L : R;
        Rel& P = bl ? R : L;
        const AttrRef bkey = key_attr(B, bl ? nd.left_attr : nd.right_attr);
        const AttrRef pkey = key_attr(P, bl ? nd.right_attr : nd.left_attr);

        // output slots: one per (relation, slot) pair the output attributes reference
        Rel out;
        std::vector<std::pair<int, uint32_t>> used;  // (0 = L, 1 = R), slot
        for (uint64_t i = 0; i < nd.num_output_attrs; ++i) {
            uint64_t idx = nd.output_attrs[i].index;
            int side;
            AttrRef a;
            if (idx < L.attrs.size()) {
                side = 0;
                a = L.attrs[idx];
            } else if (idx < L.attrs.size() + R.attrs.size()) {
                side = 1;
                a = R.attrs[idx - L.attrs.size()];
            } else {
                throw Error(HJ_E_INVALID, "output attribute out of range");
            }
            auto key = std::make_pair(side, a.slot);
            auto it = std::find(used.begin(), used.end(), key);
            uint32_t ns;
            if (it == used.end()) {
                ns = (uint32_t)used.size();
                used.push_back(key);
            } else {
                ns = (uint32_t)(it - used.begin());
            }
            out.attrs.push_back({ns, a.col, a.type});
        }
        if (used.size() > (size_t)kMaxSlots)
            throw Error(HJ_E_INVALID, "a join output references more than 16 base tables");

        const size_t mark = A_.mark();
        // ---------------- build
        uint32_t *bh = nullptr, *brow = nullptr;
        uint64_t nb = 0;
        {
            const bool sc = B.is_scan;
            ScanSrc ss;
            RelSrc rs;
            uint32_t nblk;
            if (sc) {
                ss = scan_src(B, bkey);
                nblk = extract_scan_blocks(ss);
            } else {
                rs = rel_src(B, bkey);
                nblk = rel_blocks(rs.n);
            }
            if (nblk) {
                uint32_t* counts = scratch<uint32_t>(nblk);
                uint32_t* offs = scratch<uint32_t>(nblk);
                uint32_t* part = scratch<uint32_t>(scan_partials_needed(nblk));
                uint32_t* tot = scratch<uint32_t>(1);
                if (sc)
                    { Prof p_(ctx_, "extract_scan_count"); extract_scan_count(ss, counts, s_); }
                else
                    { Prof p_(ctx_, "extract_rel_count"); extract_rel_count(rs, counts, s_); }
                { Prof p_(ctx_, "scan"); scan_excl_sum_u32(counts, offs, nblk, part, tot, s_); }
                nb = fetch(ctx_, tot);
                if (nb) {
                    bh = scratch<uint32_t>(nb);
                    brow = scratch<uint32_t>(nb);
                    if (sc)
                        { Prof p_(ctx_, "extract_scan_write"); extract_scan_write(ss, offs, bh, brow, s_); }
                    else
                        { Prof p_(ctx_, "extract_rel_write"); extract_rel_write(rs, offs, bh, brow, s_); }
                }
            }
        }
        if (nb == 0) {
            A_.reset(mark);
            return empty_like(out, used, L, R);
        }
        {
            uint32_t* th = scratch<uint32_t>(nb);
            uint32_t* tr = scratch<uint32_t>(nb);
            void* sortscr = A_.scratch(radix_sort_scratch_bytes(nb));
            { Prof p_(ctx_, "radix_sort_pairs"); radix_sort_pairs(bh, brow, th, tr, nb, sortscr, s_); }
        }
        const uint32_t k = log2_pow2_ceil(nb);
        const uint64_t D = 1ull << k;
        uint64_t* dir = scratch<uint64_t>(D + 1);
        {
            uint32_t* e = scratch<uint32_t>(D);
            uint32_t* bloom = scratch<uint32_t>(D);
            uint32_t* part = scratch<uint32_t>(scan_partials_needed(D));
            { Prof p_(ctx_, "build_directory"); build_directory(bh, nb, k, dir, e, bloom, part, s_); }
        }
        DirView dv;
        dv.dir = dir;
        dv.sh = bh;
        dv.srow = brow;
        dv.k = k;
        dv.pad = 0;
        dv.n = nb;

        // ---------------- probe
        const bool psc = P.is_scan;
        ScanSrc pss;
        RelSrc prs;
        uint32_t pblk;
        if (psc) {
            pss = scan_src(P, pkey);
            pblk = pss.key.num_pages;
        } else {
            prs = rel_src(P, pkey);
            pblk = rel_blocks(prs.n);
        }
        uint64_t total = 0;
        uint64_t* poffs = nullptr;
        uint64_t* match = nullptr;
        if (pblk) {
            uint64_t* counts = scratch<uint64_t>(pblk);
            poffs = scratch<uint64_t>(pblk);
            uint64_t* part = scratch<uint64_t>(scan_partials_needed(pblk));
            uint64_t* tot = scratch<uint64_t>(1);
            // each probe row's lookup kept for the write pass (8 bytes a row), when the arena
            // has room for it with the output to spare
            const uint64_t prows = psc ? pss.key.num_rows : prs.n;
            if (ctx_->match_stash && A_.free_bytes() / 8 >= prows * sizeof(uint64_t))
                match = scratch<uint64_t>(prows);
            if (psc)
                { Prof p_(ctx_, "probe_scan_count"); probe_scan_count(pss, dv, counts, match, s_); }
            else
                { Prof p_(ctx_, "probe_rel_count"); probe_rel_count(prs, dv, counts, match, s_); }
            { Prof p_(ctx_, "scan"); scan_excl_sum_u64(counts, poffs, pblk, part, tot, s_); }
            total = fetch(ctx_, tot);
        }
        if (total >= (1ull << 32))
            throw Error(HJ_E_INVALID, "a join produced 2^32 or more rows (row ids are 32-bit)");
        if (total == 0) {
            A_.reset(mark);
            return empty_like(out, used, L, R);
        }
        OutSlots o;
        std::memset(&o, 0, sizeof o);
        // the root join writes INT32 output columns straight from the probe (no row ids, no
        // gather pass) when each has a row-space mirror and its declared type matches
        const hj_node& root = plan_->nodes[plan_->root];
        bool direct = &nd == &root && nd.num_output_attrs <= (uint64_t)kMaxSlots &&
                      nd.num_output_attrs > 0;
        for (uint64_t i = 0; direct && i < nd.num_output_attrs; ++i) {
            const AttrRef& a = out.attrs[i];
            const Rel& src = used[a.slot].first == 0 ? L : R;
            const DCol dc = qview(src.slots[used[a.slot].second].table, a.col);
            direct = dc.type == HJ_INT32 && root.output_attrs[i].type == HJ_INT32;
        }
        std::unique_ptr<hj_dresult> dres;
        if (direct) {
            dres.reset(new hj_dresult());
            dres->num_rows = total;
            dres->live = true;
            const uint64_t vbytes = ((total + 63) / 64) * 8;
            o.values = 1;
            o.n = (int32_t)nd.num_output_attrs;
            for (uint64_t i = 0; i < nd.num_output_attrs; ++i) {
                const AttrRef& a = out.attrs[i];
                const Rel& src = used[a.slot].first == 0 ? L : R;
                const DCol dc = qview(src.slots[used[a.slot].second].table, a.col);
                hj_dresult::Col c;
                c.type = HJ_INT32;
                c.values = result<int32_t>(total);
                c.validity = result<uint8_t>(vbytes);
                o.dst[i] = (uint32_t*)c.values;
                o.src[i] = src.slots[used[a.slot].second].rid;
                o.side[i] = (used[a.slot].first == 0) == bl ? 1 : 0;
                o.col[i] = dc;
                // validity bytes unless the column has no NULL row
                o.valid8[i] = (dc.flat ? dc.vbits != nullptr : !dc.all_dense)
                                  ? scratch<uint8_t>(total + 8)
                                  : nullptr;
                dres->cols.push_back(c);
            }
        } else {
            o.n = (int32_t)used.size();
            for (size_t i = 0; i < used.size(); ++i) {
                const Rel& src = used[i].first == 0 ? L : R;
                const bool from_build = (used[i].first == 0) == bl;
                uint32_t* dst = result<uint32_t>(total);
                o.dst[i] = dst;
                o.src[i] = src.slots[used[i].second].rid;
                o.side[i] = from_build ? 1 : 0;
                out.slots.push_back({src.slots[used[i].second].table, dst});
            }
        }
        if (o.values) {
            // build-side output values in the directory's row order: the matches of a probe
            // row then read them consecutively instead of one scattered read each (worth it
            // when there are more matches than build rows)
            for (int i = 0; i < o.n; ++i) {
                if (!o.side[i] || total < nb) continue;
                int32_t* bv = scratch<int32_t>(nb);
                uint8_t* bvv = o.valid8[i] ? scratch<uint8_t>(nb) : nullptr;
                { Prof p_(ctx_, "dir_gather"); dir_gather_values(dv, o, i, bv, bvv, s_); }
                o.bval[i] = bv;
                o.bvalid[i] = bvv;
            }
        }
        out.n = total;
        HeavyEntry* heavy = scratch<HeavyEntry>(total / kLightMax + 1);
        uint32_t* hcount = scratch<uint32_t>(1);
        HJ_HIP(hipMemsetAsync(hcount, 0, sizeof(uint32_t), s_));
        if (psc)
            { Prof p_(ctx_, "probe_scan_write"); probe_scan_write(pss, dv, poffs, o, heavy, hcount, match, s_); }
        else
            { Prof p_(ctx_, "probe_rel_write"); probe_rel_write(prs, dv, poffs, o, heavy, hcount, match, s_); }
        { Prof p_(ctx_, "expand_heavy"); expand_heavy(dv, o, heavy, hcount, s_); }
        if (dres) {
            for (int i = 0; i < o.n; ++i) {
                Prof p_(ctx_, "pack_validity");
                pack_valid8(o.valid8[i], total, dres->cols[i].validity, s_);
            }
            direct_ = dres.release();
            out.slots.clear();
            for (size_t i = 0; i < used.size(); ++i)
                out.slots.push_back({(used[i].first == 0 ? L : R).slots[used[i].second].table, nullptr});
        }
        // the scratch (build side, directory, counts) is reused by the next operator; the
        // stream orders the reuse after these kernels
        A_.reset(mark);
        return out;
    }

    // ---------------------------------------------------------------- fused pipeline
    // A left-deep chain: every join's build child is a base scan, the probe side descends to a
    // base scan (the source). All build sides are built in one batch; one kernel streams the
    // source pages and runs every join per row. Requires unique build keys: a duplicate hit or
    // an over-long slot is reported by the kernels and the plan is re-run join at a time.
    struct AttrSlot {
        int slot;
        uint32_t col;
    };
    std::vector<AttrSlot> pipe_keys_;  // the pipeline's probe keys (slot, column)
    int pay_out_ = -1;   // the output column whose values the pipeline's table carries
    int ppay_out_ = -1;  // the output column whose values the partition pass carries
    std::vector<std::string> cache_refs_;  // directory cache entries this query holds
    std::vector<uint32_t> job_tuples_;     // tuples of each built pipeline job, after the pass

    // directory cache key of a build scan keyed by column c: the table's upload id, the column,
    // the scan's filters and its predicate program
    std::string cache_key(const hj_node& bn, uint32_t c) const {
        std::string k = std::to_string(table(bn.base_table_id)->uid) + ":" + std::to_string(c);
        for (uint64_t f = 0; f < bn.num_filters; ++f)
            k += "|f" + std::to_string(bn.filters[f].column) + "," +
                 std::to_string(bn.filters[f].lo) + "," + std::to_string(bn.filters[f].hi);
        for (uint64_t e = 0; e < bn.num_preds && bn.preds; ++e) {
            const hj_pred& p = bn.preds[e];
            k += "|p" + std::to_string(p.op) + "," + std::to_string(p.column) + "," +
                 std::to_string(p.ival) + "," + std::to_string(p.slen) + ":";
            if (p.sval && p.slen) k.append(p.sval, p.slen);
        }
        return k;
    }

    bool chain(uint64_t idx, uint64_t& source, std::vector<uint64_t>& joins, int depth) {
        if (idx >= plan_->num_nodes || depth > kMaxStages + 1) return false;
        const hj_node& nd = plan_->nodes[idx];
        if (nd.kind == HJ_NODE_SCAN) {
            source = idx;
            return true;
        }
        if (nd.kind != HJ_NODE_JOIN) return false;
        const uint64_t b = nd.build_left ? nd.left : nd.right;
        const uint64_t pr = nd.build_left ? nd.right : nd.left;
        if (b >= plan_->num_nodes || plan_->nodes[b].kind != HJ_NODE_SCAN) return false;
        if (!chain(pr, source, joins, depth + 1)) return false;
        joins.push_back(idx);
        return true;
    }

    // the pipeline's kernels stream keys in 64-row groups from their pages through the group
    // index, with 32-bit buffer offsets; a filter column may also be a row-space array
    static bool idx_ok(const DCol& d) {
        const uint64_t G = (d.num_rows + 63) / 64;
        return !d.flat && d.gidx && (uint64_t)d.num_pages * HJ_PAGE_SIZE < (1ull << 32) &&
               G * 16 < (1ull << 32);
    }
    // a build scan's key or filter column: also a dense column of uniform pages (arithmetic)
    static bool build_ok(const DCol& d) {
        if (d.flat) return false;
        return (d.gidx ? idx_ok(d) : d.umagic != 0) &&
               (uint64_t)d.num_pages * HJ_PAGE_SIZE < (1ull << 32);
    }
    static bool filter_ok(const DCol& d) {
        return d.flat ? d.num_rows * 4 < (1ull << 32) : build_ok(d);
    }

    static uint32_t lg2ceil(uint64_t n) {
        uint32_t k = 0;
        while ((1ull << k) < n) ++k;
        return k;
    }

    // outcome of the fused pipeline: not applicable (or a build side with a duplicate key: join
    // at a time instead), a relation of output row ids, a materialized result, or a string total
    // beyond the capacity of the direct output (run again with the materializer)
    enum class Pipe { kNo, kRel, kResult, kRetry };

    Pipe try_pipeline(Rel& out, hj_dresult** dres, bool allow_emit) {
        pipe_keys_.clear();
        pay_out_ = -1;
        ppay_out_ = -1;
        uint64_t source = 0;
        std::vector<uint64_t> joins;
        if (!chain(plan_->root, source, joins, 0) || joins.empty() ||
            joins.size() > (size_t)kMaxStages)
            return Pipe::kNo;
        const Rel srel = scan(plan_->nodes[source]);
        std::vector<AttrSlot> cur;
        for (auto& a : srel.attrs) cur.push_back({0, a.col});
        std::vector<uint32_t> slot_table{srel.table};

        BuildBatch hb;
        PipeDesc hp;
        bool any_filter = false;  // a build scan with filters or predicates
        std::memset(&hb, 0, sizeof hb);
        std::memset(&hp, 0, sizeof hp);
        hb.njobs = 0;  // build jobs: the joins whose table is not in the directory cache
        hp.nst = (int32_t)joins.size();
        std::vector<const DirCache::Entry*> cached(joins.size(), nullptr);
        std::vector<std::string> ckeys(joins.size());
        std::vector<uint64_t> stage_rows(joins.size(), 0);
        // descriptors (build batch, pipeline) and the counters (pipeline, jobs, direct output) in
        // one block, uploaded together before the first kernel (the counters as zeros)
        const size_t bb_bytes = sizeof(BuildBatch);
        const size_t pd_off = (bb_bytes + 255) & ~(size_t)255;
        const size_t meta_off = (pd_off + sizeof(PipeDesc) + 255) & ~(size_t)255;
        const size_t desc_bytes = meta_off + (kMetaWords + 2 * kEmitOut + kCountWays * kMaxStages) * 4;
        if (desc_bytes > hj_context::kPinnedBytes - hj_context::kDescOffset)
            throw Error(HJ_E_INVALID, "pipeline descriptors exceed the staging buffer");
        uint8_t* ddesc = (uint8_t*)A_.scratch(desc_bytes);
        uint32_t* metas = (uint32_t*)(ddesc + meta_off);
        hp.meta = metas;
        for (size_t i = 0; i < joins.size(); ++i) {
            const hj_node& jn = plan_->nodes[joins[i]];
            const bool bl = jn.build_left != 0;
            const hj_node& bn = plan_->nodes[bl ? jn.left : jn.right];
            // a build side keyed by the source may come from the directory cache
            {
                const uint64_t ba = bl ? jn.left_attr : jn.right_attr;
                const uint64_t pa = bl ? jn.right_attr : jn.left_attr;
                if (ctx_->dcache && pa < cur.size() && cur[pa].slot == 0 &&
                    ba < bn.num_output_attrs && bn.output_attrs) {
                    ckeys[i] = cache_key(bn, (uint32_t)bn.output_attrs[ba].index);
                    cached[i] = ctx_->dcache->acquire(ckeys[i]);
                    if (cached[i]) cache_refs_.push_back(ckeys[i]);
                }
            }
            // a cached build side needs only its shape: its filters and predicate program ran
            // when its table was built
            const Rel brel = scan(bn, !cached[i]);
            if (bn.num_filters || bn.num_preds) any_filter = true;
            std::vector<AttrSlot> bat;
            for (auto& a : brel.attrs) bat.push_back({(int)i + 1, a.col});
            slot_table.push_back(brel.table);
            const uint64_t battr = bl ? jn.left_attr : jn.right_attr;
            const uint64_t pattr = bl ? jn.right_attr : jn.left_attr;
            if (battr >= bat.size() || pattr >= cur.size())
                throw Error(HJ_E_INVALID, "join attribute out of range");
            const AttrSlot bk = bat[battr], pk = cur[pattr];
            pipe_keys_.push_back(pk);  // never NULL in the output: NULL keys do not match
            const DevColumn& bcol = column(slot_table[bk.slot], bk.col);
            const DevColumn& pcol = column(slot_table[pk.slot], pk.col);
            if (bcol.type != HJ_INT32 || pcol.type != HJ_INT32)
                throw Error(HJ_E_TYPE, "Join key isn't of type int32_t");
            if (i == 0 && pk.slot != 0) return Pipe::kNo;
            if (brel.n >= (1ull << 28)) return Pipe::kNo;  // key slots of the table

            PipeStage& st = hp.st[i];
            st.key_slot = pk.slot;
            st.keycol = qview(slot_table[pk.slot], pk.col);
            if (pk.slot == 0 && !idx_ok(st.keycol)) return Pipe::kNo;
            if (cached[i]) {  // no build job: the stream join reads the cached table
                st.job = -1;
                stage_rows[i] = cached[i]->tuples;
            } else {
                stage_rows[i] = brel.n;
                st.job = hb.njobs;
                BuildJob& jb = hb.job[hb.njobs++];
                jb.key = qview(slot_table[bk.slot], bk.col);
                if (!build_ok(jb.key)) return Pipe::kNo;
                jb.nf = (int32_t)brel.filters.size();
                for (size_t f = 0; f < brel.filters.size(); ++f) {
                    jb.fcol[f] = filter_view(brel, brel.filters[f].column);
                    if (!filter_ok(jb.fcol[f])) return Pipe::kNo;
                    jb.flo[f] = brel.filters[f].lo;
                    jb.fhi[f] = brel.filters[f].hi;
                }
                const uint32_t rows = (uint32_t)brel.n;
                jb.rows = rows;
                jb.kmax = table_bits(rows);
                jb.meta = metas + 4 * (st.job + 1);
                jb.nblk = std::max<uint32_t>(1, (rows + kBuildRows - 1) / kBuildRows);
                jb.blk0 = hb.grid;
                hb.grid += jb.nblk;
                const uint64_t scap = (uint64_t)jb.nblk * kBuildRows;
                jb.st_key = scratch<uint32_t>(scap);
                jb.st_row = scratch<uint32_t>(scap);
                jb.blk = scratch<uint32_t>(jb.nblk);
                // clear pass: 16-byte stores, 4 per thread and round, over the largest table
                const uint64_t stores = (1ull << jb.kmax) / 2;
                hb.cgrid = (uint32_t)std::max<uint64_t>(
                    hb.cgrid, std::min<uint64_t>(1024, (stores + 4 * 256 - 1) / (4 * 256)));
                jb.table = scratch<uint32_t>(2ull * (1ull << jb.kmax) + 8);
                jb.bf = scratch<uint32_t>(kBloomWords);
            }

            const std::vector<AttrSlot>& L = bl ? bat : cur;
            const std::vector<AttrSlot>& R = bl ? cur : bat;
            std::vector<AttrSlot> nxt;
            for (uint64_t a = 0; a < jn.num_output_attrs; ++a) {
                const uint64_t idx = jn.output_attrs[a].index;
                if (idx < L.size())
                    nxt.push_back(L[idx]);
                else if (idx < L.size() + R.size())
                    nxt.push_back(R[idx - L.size()]);
                else
                    throw Error(HJ_E_INVALID, "output attribute out of range");
            }
            cur.swap(nxt);
        }
        stage_job_.assign(joins.size(), -1);
        for (size_t i = 0; i < joins.size(); ++i) stage_job_[i] = hp.st[i].job;
        for (int j = 0; j < kMaxStages; ++j) hb.blk0[j] = j < hb.njobs ? hb.job[j].blk0 : ~0u;
        // partitioned build (segments built in LDS) when every job has at most 2^kMaxFineBits
        // fine partitions and kMaxPartBlocks filter blocks; else clear + compare-and-swap inserts
        hb.part = ctx_->part_build ? 1 : 0;
        for (int j = 0; j < hb.njobs; ++j)
            if (fine_bits(hb.job[j].kmax) > kMaxFineBits || hb.job[j].nblk > kMaxPartBlocks)
                hb.part = 0;
        if (hb.part)
            for (int j = 0; j < hb.njobs; ++j) {
                BuildJob& jb = hb.job[j];
                jb.pbits = std::min(fine_bits(jb.kmax), kPartBits);
                const uint64_t P = 1ull << jb.pbits;
                jb.poff = scratch<uint16_t>((uint64_t)jb.nblk * (P + 1));
            }
        hb.sgrid = (uint32_t)ctx_->num_cus * 4;  // segment pass: a grid-stride loop
        hb.pcnt = metas + kMetaWords + 2 * kEmitOut;

        hp.nf = (int32_t)srel.filters.size();
        for (size_t f = 0; f < srel.filters.size(); ++f) {
            hp.fcol[f] = filter_view(srel, srel.filters[f].column);
            hp.flo[f] = srel.filters[f].lo;
            hp.fhi[f] = srel.filters[f].hi;
        }
        // output slots referenced by the root's attributes
        std::vector<int> oslots;
        Rel res;
        for (auto& a : cur) {
            auto it = std::find(oslots.begin(), oslots.end(), a.slot);
            uint32_t q;
            if (it == oslots.end()) {
                q = (uint32_t)oslots.size();
                oslots.push_back(a.slot);
            } else {
                q = (uint32_t)(it - oslots.begin());
            }
            res.attrs.push_back({q, a.col, column(slot_table[a.slot], a.col).type});
        }
        const uint64_t nsrc = srel.n;
        // the fused pass addresses the source's key columns with 32-bit byte offsets
        if (nsrc >= (1ull << 29)) return Pipe::kNo;
        hp.src_rows = (uint32_t)nsrc;
        hp.nout = (int32_t)oslots.size();
        const uint64_t ncap = nsrc ? nsrc : 1;
        // joins keyed by the source run fused, smallest build table first (its Bloom filter
        // is the one held in LDS); the others run afterwards over the survivors
        for (int st = 0; st < hp.nst; ++st) {
            hp.qpos[st] = -1;
            if (hp.st[st].key_slot == 0)
                hp.order[hp.nstream++] = st;
            else
                ++hp.ndep;
        }
        std::stable_sort(hp.order, hp.order + hp.nstream, [&](int a, int b) {
            return stage_rows[a] < stage_rows[b];
        });
        for (int q = 0; q < hp.nstream; ++q) {
            const int st = hp.order[q];
            hp.qpos[st] = q;
            if (cached[st]) {
                hp.sj[q] = {hp.st[st].keycol, cached[st]->table, cached[st]->meta, cached[st]->bf};
            } else {
                const BuildJob& jb = hb.job[hp.st[st].job];
                hp.sj[q] = {hp.st[st].keycol, jb.table, jb.meta, jb.bf};
            }
        }
        // output: written directly from the fused pass's staged survivors when every join is
        // fused (at most one output row per source row: cap rows), else compacted row ids
        const hj_node& root = plan_->nodes[plan_->root];
        bool emit = allow_emit && hp.ndep == 0 && root.num_output_attrs == cur.size() &&
                    cur.size() <= (size_t)kEmitCols;
        if (emit) {
            // the direct output's buffers at capacity must leave the arena room for the pass
            uint64_t need = 0;
            for (const AttrSlot& a : cur) {
                const DevColumn& dc = column(slot_table[a.slot], a.col);
                need += ncap + ncap / 8 + 1024;
                need += dc.type == HJ_VARCHAR
                            ? 8 * (ncap + 1) + std::min<uint64_t>(ncap * (uint64_t)dc.max_len,
                                                                  kCharsBound)
                            : 4 * ncap;
            }
            emit = need <= A_.free_bytes() / 2;
        }
        if (!emit)
            for (size_t q = 0; q < oslots.size(); ++q) {
                hp.out_slot[q] = oslots[q];
                hp.final_out[q] = result<uint32_t>(ncap);
                res.slots.push_back({slot_table[oslots[q]], hp.final_out[q]});
            }
        {
            // fused pass: contiguous ranges of 64-row groups per block
            const int U = probe_unroll(hp.nstream);
            const uint64_t groups = (nsrc + 63) / 64;
            // at most the blocks resident at once, and at most what the direct output's
            // prefix sums cover
            const uint64_t maxb = std::min<uint64_t>(
                probe_resident_blocks(hp.nstream, (uint32_t)ctx_->num_cus), kEmitMaxBlocks);
            const uint64_t per = (uint64_t)U * 4;  // groups per block and iteration
            uint64_t gpb = std::max<uint64_t>(per, (groups + maxb - 1) / maxb);
            gpb = (gpb + per - 1) / per * per;
            hp.probe_gpb = (uint32_t)gpb;
            hp.probe_blocks = (uint32_t)std::max<uint64_t>(1, (groups + gpb - 1) / gpb);
            uint64_t cap = (uint64_t)hp.probe_blocks * gpb * 64;
            // one join over a table larger than an XCD's L2 (more than 2^18 build rows: at
            // least 2^20 key slots, 8 MiB; a cached table or an unfiltered build, whose tuple
            // count is known here): the sliced probe, one table slice per XCD
            const uint64_t nbp = (nsrc + kPartRows - 1) / kPartRows;
            const bool exact_rows = hp.nstream == 1 && (cached[hp.order[0]] || !any_filter);
            if (ctx_->xcd_slices && hp.nstream == 1 && hp.ndep == 0 && emit && exact_rows &&
                stage_rows[hp.order[0]] > (1u << 18) && nbp >= 8 &&
                nbp <= (uint64_t)(kEmitMaxBlocks / 8) * kSliceMaxK) {
                const uint64_t chunks0 = std::min<uint64_t>(kEmitMaxBlocks / 8, nbp);
                const uint64_t K = (nbp + chunks0 - 1) / chunks0;
                const uint64_t chunks = (nbp + K - 1) / K;
                hp.xslices = 8;
                hp.part_blocks = (uint32_t)nbp;
                hp.part_k = (uint32_t)K;
                hp.probe_blocks = (uint32_t)(8 * chunks);
                cap = nbp * kPartRows;
                hp.ptk = scratch<uint32_t>(cap);
                hp.ptr = scratch<uint32_t>(cap);
                hp.pcnt = scratch<uint32_t>(nbp * 8);
                hp.blk_origin = scratch<uint32_t>(hp.probe_blocks);
                // The join's only build-side output, an INT32 column without NULL, rides in
                // the table's row words (a table built by this query, not published to the
                // directory cache): the sliced probe stages the value from the matching
                // bucket, and the output copies it instead of reading it at a random row.
                const int st0 = hp.order[0], j0 = hp.st[st0].job;
                int nb = 0, bi = -1;
                for (size_t i = 0; i < cur.size(); ++i)
                    if (cur[i].slot == st0 + 1) {
                        ++nb;
                        bi = (int)i;
                    }
                if (j0 >= 0 && nb == 1) {
                    const DevColumn& pc = column(slot_table[st0 + 1], cur[bi].col);
                    if (pc.type == HJ_INT32 && pc.all_dense) {
                        hb.job[j0].pay = qview(slot_table[st0 + 1], cur[bi].col);
                        hb.job[j0].pay_on = 1;
                        pay_out_ = bi;
                    }
                }
                // likewise the source's only output: the partition pass carries its value
                // with the key, and the probe stages that value instead of the row
                int ns = 0, si = -1;
                for (size_t i = 0; i < cur.size(); ++i)
                    if (cur[i].slot == 0) {
                        ++ns;
                        si = (int)i;
                    }
                if (ns == 1) {
                    const DevColumn& sc = column(slot_table[0], cur[si].col);
                    if (sc.type == HJ_INT32 && sc.all_dense) {
                        hp.ppay = qview(slot_table[0], cur[si].col);
                        hp.ppay_on = 1;
                        ppay_out_ = si;
                    }
                }
            }
            for (int q = 0; q <= hp.nstream; ++q) hp.stage[q] = scratch<uint32_t>(cap);
            hp.blk_cnt = scratch<uint32_t>(hp.probe_blocks);
            hp.blk_off = scratch<uint32_t>(hp.probe_blocks);
        }
        if (emit) {
            // output VARCHAR columns: the fused pass counts their bytes per block
            for (size_t i = 0; i < cur.size(); ++i) {
                const int ps = cur[i].slot;
                const DevColumn& dc = column(slot_table[ps], cur[i].col);
                if (dc.type != HJ_VARCHAR) continue;
                hp.ostr[hp.nostr] = qview(slot_table[ps], cur[i].col);
                hp.ostr_slot[hp.nostr] = ps ? 1 + hp.qpos[ps - 1] : 0;
                ++hp.nostr;
            }
            hp.blk_bytes = scratch<uint64_t>((uint64_t)hp.probe_blocks * kMaxOutStr);
        }
        if (hp.ndep) {
            hp.l_row = scratch<uint32_t>(ncap);
            for (int st = 0; st < hp.nst; ++st) hp.l_b[st] = scratch<uint32_t>(ncap);
            hp.tile_cnt = scratch<uint32_t>(ncap / 4096 + 2);
            hp.tile_off = scratch<uint32_t>(ncap / 4096 + 2);
        }
        hp.grid = (uint32_t)ctx_->num_cus * 8;
        // The probe side's index, taken here, after the last return without a launch: the
        // plain fused probe's blocks index the source's columns (up to kProbeIdxCols) for their
        // own row ranges as they start; every other column (all of them before a sliced probe)
        // goes to extra blocks of the build's filter launch, or a launch of its own when every
        // table is cached.
        if (pending_.n) {
            GroupIndexBatch rest;
            std::memset(&rest, 0, sizeof rest);
            for (int i = 0; i < pending_.n; ++i) {
                if (!hp.xslices && pending_src_[i] && hp.nsidx < kProbeIdxCols) {
                    hp.sidx_table = ctx_->index_tables ? ~0u : 0u;  // ~0: the whole table
                    hp.sidx_col[hp.nsidx] = pending_.col[i];
                    hp.sidx_out[hp.nsidx] = pending_.out[i];
                    ++hp.nsidx;
                    continue;
                }
                rest.col[rest.n] = pending_.col[i];
                rest.out[rest.n] = pending_.out[i];
                rest.blk0[rest.n + 1] = rest.blk0[rest.n] + (pending_.blk0[i + 1] - pending_.blk0[i]);
                ++rest.n;
            }
            hb.gib = rest;
            hb.igrid = rest.blk0[rest.n];
            std::memset(&pending_, 0, sizeof pending_);
        }

        // descriptors and zeroed counters -> pinned staging -> device
        uint8_t* pin = (uint8_t*)ctx_->pinned + hj_context::kDescOffset;
        hp.jobs = (const BuildJob*)(ddesc + offsetof(BuildBatch, job));
        std::memset(pin, 0, desc_bytes);
        std::memcpy(pin, &hb, bb_bytes);
        std::memcpy(pin + pd_off, &hp, sizeof(PipeDesc));
        host_trace().mark("planned");
        HJ_HIP(hipMemcpyAsync(ddesc, pin, desc_bytes, hipMemcpyHostToDevice, s_));
        host_trace().mark("desc_copy");
        const BuildBatch* dbb = (const BuildBatch*)ddesc;
        const PipeDesc* dpd = (const PipeDesc*)(ddesc + pd_off);
        if (hb.njobs) {
            Prof p_(ctx_, "pipe_build");
            pipe_build(hb, dbb, s_);
        } else if (hb.igrid) {  // every table cached: the probe side's index on its own
            Prof p_(ctx_, "group_index");
            launch_group_index(hb.gib, s_);
        }
        host_trace().mark("build_q");
        {
            Prof p_(ctx_, "probe_fused", true);
            pipe_probe_fused(hp, dpd, s_, p_.a, p_.b);
        }
        if (emit) {
            const Pipe r = emit_result(hp, oslots, slot_table, cur, metas, any_filter, dres);
            if (r == Pipe::kResult) publish(hb, ckeys, cached);
            return r;
        }
        { Prof p_(ctx_, "probe_tail"); pipe_probe_tail(hp, dpd, s_); }
        struct F4 {
            uint32_t v[4];
        };
        const F4 fl = fetch(ctx_, (const F4*)hp.meta);
        if (std::getenv("HJ_DEBUG_PIPE")) {
            for (int j = 0; j < hb.njobs; ++j) {
                const F4 jm = fetch(ctx_, (const F4*)hb.job[j].meta);
                std::fprintf(stderr, "[pipe] job %d: tuples %u duplicate %u\n", j, jm.v[0],
                             jm.v[1]);
            }
            std::fprintf(stderr, "[pipe] fallback %u survivors %u out %u\n", fl.v[0], fl.v[2],
                         fl.v[3]);
        }
        if (fl.v[0]) return Pipe::kNo;
        res.n = fl.v[3];
        out = std::move(res);
        if (ctx_->dcache && hb.njobs) {
            job_tuples_.assign(hb.njobs, 0);
            for (int j = 0; j < hb.njobs; ++j)
                job_tuples_[j] = fetch(ctx_, (const F4*)hb.job[j].meta).v[0];
            publish(hb, ckeys, cached);
        }
        return Pipe::kRel;
    }

    // the tables this query built for stream joins, copied into the directory cache
    void publish(const BuildBatch& hb, const std::vector<std::string>& ckeys,
                 const std::vector<const DirCache::Entry*>& cached) {
        if (!ctx_->dcache) return;
        for (size_t i = 0; i < ckeys.size(); ++i) {
            if (ckeys[i].empty() || cached[i]) continue;
            const int j = plan_stage_job(i);
            if (j < 0 || j >= (int)job_tuples_.size()) continue;
            const BuildJob& jb = hb.job[j];
            if (jb.pay_on) continue;  // row words hold values, not rows
            ctx_->dcache->insert(ckeys[i], jb.table, table_bits(job_tuples_[j]), jb.meta, jb.bf,
                                 job_tuples_[j], s_);
        }
    }
    std::vector<int32_t> stage_job_;  // pipeline stage -> its build job (-1: cached)
    int plan_stage_job(size_t i) const { return i < stage_job_.size() ? stage_job_[i] : -1; }

    // the root's columns straight from the fused pass (emit_rows), then one round trip for the
    // counters: fallback flag, rows, string totals, overflow
    Pipe emit_result(const PipeDesc& hp, const std::vector<int>& oslots,
                     const std::vector<uint32_t>& slot_table, const std::vector<AttrSlot>& cur,
                     uint32_t* metas, bool any_filter, hj_dresult** dres) {
        const hj_node& root = plan_->nodes[plan_->root];
        const uint64_t cap = hp.src_rows ? hp.src_rows : 1;
        const uint64_t vbytes = ((cap + 63) / 64) * 8;
        std::unique_ptr<hj_dresult> res(new hj_dresult());
        EmitDesc ed;
        std::memset(&ed, 0, sizeof ed);
        ed.ncol = (int32_t)cur.size();
        bool into = ctx_->emit_into.size() == cur.size();
        for (size_t i = 0; into && i < cur.size(); ++i)
            into = column(slot_table[cur[i].slot], cur[i].col).type == HJ_INT32;
        int nstr = 0;  // VARCHAR columns in output order: the fused pass's byte count entries
        for (size_t i = 0; i < cur.size(); ++i) {
            const int ps = cur[i].slot;
            const DevColumn& dc = column(slot_table[ps], cur[i].col);
            if (root.output_attrs[i].type != dc.type)
                throw Error(HJ_E_TYPE, "output attribute type does not match its column");
            if (dc.type != HJ_INT32 && dc.type != HJ_VARCHAR)
                throw Error(HJ_E_TYPE, "unsupported output column type (only INT32 and VARCHAR)");
            EmitCol& ec = ed.c[i];
            ec.col = qview(slot_table[ps], cur[i].col);
            // an output column with no NULL row: a column without NULL, or a probe key
            ec.nonull = dc.all_dense ? 1 : 0;
            for (const AttrSlot& k : pipe_keys_)
                if (k.slot == cur[i].slot && k.col == cur[i].col) ec.nonull = 1;
            ec.slot = ps ? 1 + hp.qpos[ps - 1] : 0;
            ec.staged = ((int)i == pay_out_ || (int)i == ppay_out_) ? 1 : 0;
            ec.type = dc.type;
            if (dc.type == HJ_VARCHAR) ec.sidx = nstr++;
            ec.valid8 = scratch<uint8_t>(cap);
            hj_dresult::Col c;
            c.type = dc.type;
            if (into) {  // the caller's buffers (every output INT32)
                c.validity = ctx_->emit_into[i].second;
                c.values = ctx_->emit_into[i].first;
                ec.validity = c.validity;
                ec.values = c.values;
                res->cols.push_back(c);
                continue;
            }
            c.validity = result<uint8_t>(vbytes);
            ec.validity = c.validity;
            if (dc.type == HJ_INT32) {
                c.values = result<int32_t>(cap);
                ec.values = c.values;
            } else {
                c.offsets = result<uint64_t>(cap + 1);
                ec.offsets = c.offsets;
                ec.chars_cap = std::min<uint64_t>(cap * (uint64_t)dc.max_len, kCharsBound);
                c.chars = result<char>(ec.chars_cap ? ec.chars_cap : 1);
                ec.chars = c.chars;
            }
            res->cols.push_back(c);
        }
        (void)oslots;
        ed.nstr = nstr;
        ed.blocks = hp.probe_blocks;
        ed.rows_per_block = hp.probe_gpb * 64;
        ed.blk_origin = hp.xslices ? hp.blk_origin : nullptr;
        ed.blk_cnt = hp.blk_cnt;
        for (int q = 0; q <= hp.nstream; ++q) ed.stage[q] = hp.stage[q];
        ed.meta = hp.meta;
        ed.cap_rows = cap;
        ed.group = any_filter || hp.nf ? kEmitGroup : kEmitGroupUnfiltered;
        ed.blk_bytes = hp.blk_bytes;
        ed.bytes_stride = kMaxOutStr;
        ed.out = (uint64_t*)(metas + kMetaWords);
        {
            Prof p_(ctx_, "emit");
            if (!emit_rows(ed, s_))
                throw Error(HJ_E_INVALID, "direct output: " + std::to_string(ed.blocks) +
                                              " fused-pass blocks exceed " +
                                              std::to_string(kEmitMaxBlocks));
        }
        // one round trip: pipeline counters and the direct output's totals
        host_trace().mark("emit_q");
        uint8_t* pin = (uint8_t*)ctx_->pinned;
        const size_t bytes = kMetaWords * 4 + kEmitOut * 8;
        HJ_HIP(hipMemcpyAsync(pin, metas, bytes, hipMemcpyDeviceToHost, s_));
        wait_stream(s_);
        host_trace().mark("synced");
        uint32_t m[kMetaWords];
        uint64_t o[kEmitOut];
        std::memcpy(m, pin, sizeof m);
        std::memcpy(o, pin + kMetaWords * 4, sizeof o);
        job_tuples_.assign(kMaxStages, 0);
        for (int j = 0; j < kMaxStages; ++j) job_tuples_[j] = m[4 + 4 * j];
        if (std::getenv("HJ_DEBUG_PIPE"))
            std::fprintf(stderr, "[pipe] fallback %u rows %llu overflow %llu\n", m[0],
                         (unsigned long long)o[0], (unsigned long long)o[1 + kEmitCols]);
        if (m[0]) return Pipe::kNo;
        if (o[1 + kEmitCols]) return Pipe::kRetry;
        res->num_rows = o[0];
        for (size_t i = 0; i < cur.size(); ++i)
            if (res->cols[i].type == HJ_VARCHAR) res->cols[i].num_chars = o[1 + i];
        res->live = true;
        *dres = res.release();
        return Pipe::kResult;
    }

    Rel empty_like(Rel out, const std::vector<std::pair<int, uint32_t>>& used, const Rel& L,
                   const Rel& R) {
        out.n = 0;
        for (auto& u : used) {
            const Rel& src = u.first == 0 ? L : R;
            out.slots.push_back({src.slots[u.second].table, nullptr});
        }
        return out;
    }

    hj_dresult* materialize(Rel& r, const hj_node& root) {
        // a filtered scan at the root is compacted to a row-id list first
        if (r.is_scan && !r.filters.empty()) {
            AttrRef any{0, 0, HJ_INT32};  // column 0: the row structure of the pages
            ScanSrc ss = scan_src(r, any);
            uint32_t nblk = ss.key.num_pages;
            uint64_t n = 0;
            uint32_t* rows = nullptr;
            if (nblk) {
                uint32_t* counts = scratch<uint32_t>(nblk);
                uint32_t* offs = scratch<uint32_t>(nblk);
                uint32_t* part = scratch<uint32_t>(scan_partials_needed(nblk));
                uint32_t* tot = scratch<uint32_t>(1);
                { Prof p_(ctx_, "select_scan_count"); select_scan_count(ss, counts, s_); }
                { Prof p_(ctx_, "scan"); scan_excl_sum_u32(counts, offs, nblk, part, tot, s_); }
                n = fetch(ctx_, tot);
                rows = result<uint32_t>(n ? n : 1);
                if (n) { Prof p_(ctx_, "select_scan_write"); select_scan_write(ss, offs, rows, s_); }
            }
            r.n = n;
            r.slots[0].rid = rows;
            r.is_scan = false;
        }
        auto* res = new hj_dresult();
        res->num_rows = r.n;
        const uint64_t n = r.n;
        const uint64_t vbytes = ((n + 63) / 64) * 8;
        const uint32_t tiles = (uint32_t)((n + kMatTile - 1) / kMatTile);
        std::vector<MatCol> mcols;
        for (size_t i = 0; i < root.num_output_attrs; ++i) {
            const AttrRef& a = r.attrs[i];
            const Slot& sl = r.slots[a.slot];
            const DevColumn& dc = column(sl.table, a.col);
            if (root.output_attrs[i].type != dc.type)
                throw Error(HJ_E_TYPE, "output attribute type does not match its column");
            if (dc.type != HJ_INT32 && dc.type != HJ_VARCHAR)
                throw Error(HJ_E_TYPE, "unsupported output column type (only INT32 and VARCHAR)");
            hj_dresult::Col c;
            c.type = dc.type;
            c.validity = result<uint8_t>(vbytes ? vbytes : 8);
            MatCol m;
            std::memset(&m, 0, sizeof m);
            m.col = qview(sl.table, a.col);
            m.rid = sl.rid;
            m.type = dc.type;
            m.validity = c.validity;
            if (dc.type == HJ_INT32) {
                c.values = result<int32_t>(n ? n : 1);
                m.values = c.values;
            } else {
                c.offsets = result<uint64_t>(n + 1);
                m.offsets = c.offsets;
                m.tile_sum = scratch<uint64_t>((uint64_t)tiles + 1);
                m.src = scratch<uint64_t>(n ? n : 1);
                if (!n) HJ_HIP(hipMemsetAsync(c.offsets, 0, sizeof(uint64_t), s_));
            }
            res->cols.push_back(c);
            mcols.push_back(m);
        }
        if (!n) {
            res->live = true;
            return res;
        }
        // string bytes: bounded by rows x longest string when that is small, else counted first
        for (size_t b0 = 0; b0 < mcols.size(); b0 += kMatCols) {
            MatBatch mb;
            std::memset(&mb, 0, sizeof mb);
            mb.n = n;
            mb.tiles = tiles;
            mb.ncol = (int32_t)std::min<size_t>(kMatCols, mcols.size() - b0);
            for (int j = 0; j < mb.ncol; ++j) mb.c[j] = mcols[b0 + j];
            { Prof p_(ctx_, "mat_gather"); mat_gather(mb, s_); }
            bool any_str = false;
            for (int j = 0; j < mb.ncol; ++j) {
                if (mb.c[j].type != HJ_VARCHAR) continue;
                any_str = true;
                const DevColumn& dc = column(r.slots[r.attrs[b0 + j].slot].table,
                                             r.attrs[b0 + j].col);
                const uint64_t bound = n * (uint64_t)dc.max_len;
                uint64_t cap = bound;
                if (bound > kCharsBound) cap = fetch(ctx_, mb.c[j].tile_sum + tiles);
                mb.c[j].chars = result<char>(cap ? cap : 1);
                res->cols[b0 + j].chars = mb.c[j].chars;
                res->cols[b0 + j].num_chars = cap;
            }
            if (any_str) { Prof p_(ctx_, "mat_strings"); mat_strings(mb, s_); }
        }
        res->live = true;
        return res;
    }
};

// ------------------------------------------------------------------ upload
uint32_t rd16h(const uint8_t* p) {
    uint16_t v;
    std::memcpy(&v, p, 2);
    return v;
}

// copies segments (dst, src, bytes) with up to 8 threads for large totals
struct CopySeg {
    uint8_t* dst;
    const uint8_t* src;
    size_t bytes;
};
void parallel_copy(const std::vector<CopySeg>& segs) {
    size_t total = 0;
    for (const auto& g : segs) total += g.bytes;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<size_t>({8, hw, total / (4u << 20) + 1});
    auto run = [&segs](size_t lo, size_t hi) {  // bytes [lo, hi) of the concatenated segments
        size_t base = 0;
        for (const auto& g : segs) {
            const size_t a = std::max(lo, base), e = std::min(hi, base + g.bytes);
            if (a < e) std::memcpy(g.dst + (a - base), g.src + (a - base), e - a);
            base += g.bytes;
        }
    };
    if (nt <= 1) {
        run(0, total);
        return;
    }
    std::vector<std::thread> th;
    const size_t step = (total + nt - 1) / nt;
    for (unsigned t = 1; t < nt; ++t)
        th.emplace_back(run, std::min(total, t * step), std::min(total, (t + 1) * step));
    run(0, std::min(total, step));
    for (auto& x : th) x.join();
}

// the context's pinned double buffer for pages on their way between host and HBM
void ensure_page_stage(hj_context* ctx) {
    if (ctx->page_stage) return;
    HJ_HIP(hipHostMalloc(&ctx->page_stage, 2 * hj_context::kPageStageHalf, hipHostMallocDefault));
    for (auto& e : ctx->page_ev) HJ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// host pages -> contiguous device pages, in batches of 32 MiB through the pinned double buffer:
// the host fills one half (up to 8 threads) while the other half's copy to HBM runs
void copy_pages_h2d(hj_context* ctx, uint8_t* dst, const uint8_t* const* pages, uint32_t np) {
    ensure_page_stage(ctx);
    constexpr uint64_t kHalfPages = hj_context::kPageStageHalf / HJ_PAGE_SIZE;
    const uint64_t nb = (np + kHalfPages - 1) / kHalfPages;
    for (uint64_t k = 0; k < nb; ++k) {
        uint8_t* half = ctx->page_stage + (k & 1) * hj_context::kPageStageHalf;
        // the copy out of this half two batches ago must be done before it is refilled
        if (k >= 2) HJ_HIP(hipEventSynchronize(ctx->page_ev[k & 1]));
        const uint64_t p0 = k * kHalfPages, cnt = std::min<uint64_t>(kHalfPages, np - p0);
        std::vector<CopySeg> segs;
        for (uint64_t j = 0; j < cnt; ++j) {
            const uint8_t* src = pages[p0 + j];
            if (!src) throw Error(HJ_E_INVALID, "null page pointer");
            if (!segs.empty() && segs.back().src + segs.back().bytes == src)
                segs.back().bytes += HJ_PAGE_SIZE;  // pages adjacent in host memory: one run
            else
                segs.push_back({half + j * HJ_PAGE_SIZE, src, HJ_PAGE_SIZE});
        }
        parallel_copy(segs);
        HJ_HIP(hipMemcpyAsync(dst + p0 * HJ_PAGE_SIZE, half, cnt * HJ_PAGE_SIZE,
                              hipMemcpyHostToDevice, ctx->stream));
        HJ_HIP(hipEventRecord(ctx->page_ev[k & 1], ctx->stream));
    }
    HJ_HIP(hipStreamSynchronize(ctx->stream));
}

void upload_column(hj_context* ctx, const hj_column& hc, uint64_t table_rows, hj_dtable* t,
                   DevColumn& dc) {
    dc.type = hc.type;
    dc.num_pages = hc.num_pages;
    if (hc.num_pages && !hc.pages) throw Error(HJ_E_INVALID, "column pages missing");
    dc.h_row_start.assign((size_t)hc.num_pages + 1, 0);
    std::vector<uint32_t> group_start(hc.num_pages ? hc.num_pages : 1, 0);
    uint64_t rows = 0, groups = 0, long_len = 0;
    uint32_t first_rows = 0;
    bool uniform = true;
    dc.all_dense = 1;
    for (uint32_t p = 0; p < hc.num_pages; ++p) {
        const uint8_t* pg = hc.pages[p];
        if (!pg) throw Error(HJ_E_INVALID, "null page pointer");
        const uint32_t nr = rd16h(pg), nn = rd16h(pg + 2);
        dc.h_row_start[p] = (uint32_t)rows;
        group_start[p] = (uint32_t)groups;
        uint32_t prow = nr;
        if (hc.type == HJ_VARCHAR && p > kStrRefPageMask && dc.strref_page_over == ~0ull &&
            (nr == 0xffffu || (nr != 0xfffeu && nn > 0)))
            dc.strref_page_over = p;
        if (hc.type == HJ_VARCHAR && (nr == 0xffffu || (nr != 0xfffeu && nn > 0)))
            dc.has_strings = true;
        if (hc.type == HJ_VARCHAR && (nr == 0xffffu || nr == 0xfffeu)) {
            if (nr == 0xfffeu && (p == 0)) throw Error(HJ_E_INVALID, "0xfffe page without a long string");
            if (nn > HJ_PAGE_SIZE - 4) throw Error(HJ_E_INVALID, "corrupt long-string page");
            prow = nr == 0xffffu ? 1 : 0;
            dc.num_chars += nn;
            long_len = (nr == 0xffffu ? 0 : long_len) + nn;
            dc.max_len = (uint32_t)std::max<uint64_t>(dc.max_len, long_len);
            dc.all_dense = 0;  // long-string pages have no bitmap but break uniformity
            uniform = false;
        } else if (hc.type == HJ_INT32) {
            if (nn > nr || 4ull + 4ull * nn + (nr + 7) / 8 > HJ_PAGE_SIZE)
                throw Error(HJ_E_INVALID, "corrupt INT32 page");
            if (nn != nr) {
                dc.all_dense = 0;
                groups += (nr + 63) / 64;
            }
        } else if (hc.type == HJ_VARCHAR) {
            if (nn > nr || 4ull + 2ull * nn + (nr + 7) / 8 > HJ_PAGE_SIZE)
                throw Error(HJ_E_INVALID, "corrupt VARCHAR page");
            uint32_t prev = 0;
            for (uint32_t i = 0; i < nn; ++i) {
                uint32_t e = rd16h(pg + 4 + 2 * i);
                if (e < prev) throw Error(HJ_E_INVALID, "corrupt VARCHAR offsets");
                dc.max_len = std::max(dc.max_len, e - prev);
                prev = e;
            }
            if (4ull + 2ull * nn + prev + (nr + 7) / 8 > HJ_PAGE_SIZE)
                throw Error(HJ_E_INVALID, "corrupt VARCHAR page");
            dc.num_chars += prev;
            if (nn != nr) {
                dc.all_dense = 0;
                groups += (nr + 63) / 64;
            }
        } else {
            prow = nr;  // INT64/FP64 are kept resident but never decoded
        }
        if (p == 0) first_rows = prow;
        if (p + 1 < hc.num_pages && prow != first_rows) uniform = false;
        if (p + 1 < hc.num_pages && prow < 64) dc.regular = false;
        if (hc.type == HJ_VARCHAR && (nr == 0xffffu || nr == 0xfffeu)) dc.regular = false;
        rows += prow;
        if (rows >= (1ull << 32)) throw Error(HJ_E_INVALID, "a table must have fewer than 2^32 rows");
    }
    dc.h_row_start[hc.num_pages] = (uint32_t)rows;
    if (rows != table_rows && !(hc.num_pages == 0 && table_rows == 0))
        throw Error(HJ_E_INVALID, "column row count (" + std::to_string(rows) +
                                      ") differs from the table's num_rows (" +
                                      std::to_string(table_rows) + ")");
    dc.num_rows = rows;
    dc.uniform_rows = (uniform && hc.num_pages && first_rows) ? first_rows : 0;
    if (hc.type == HJ_VARCHAR && dc.uniform_rows) {
        // VARCHAR pages hold a varying number of rows; only use the division shortcut when
        // every page really is the same size (checked above)
    }
    const size_t bytes = (size_t)hc.num_pages * HJ_PAGE_SIZE;
    auto dev_alloc = [&](size_t n) -> void* {
        void* p = nullptr;
        HJ_HIP(hipMalloc(&p, n ? n : 256));
        t->allocations.push_back(p);
        return p;
    };
    dc.pages = (uint8_t*)dev_alloc(bytes + 256);  // slack for wide loads near the end
    if (bytes) copy_pages_h2d(ctx, dc.pages, hc.pages, hc.num_pages);
    dc.row_start = (uint32_t*)dev_alloc(dc.h_row_start.size() * 4);
    HJ_HIP(hipMemcpyAsync(dc.row_start, dc.h_row_start.data(), dc.h_row_start.size() * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    dc.group_start = (uint32_t*)dev_alloc(group_start.size() * 4);
    HJ_HIP(hipMemcpyAsync(dc.group_start, group_start.data(), group_start.size() * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    {
        // chunk index: page of row c * 2048 for every chunk, plus one extra entry
        const uint64_t nch = (rows >> HJ_CHUNK_SHIFT) + 2;
        std::vector<uint32_t> cp(nch, hc.num_pages ? hc.num_pages - 1 : 0);
        uint32_t p = 0;
        for (uint64_t c = 0; c < nch; ++c) {
            const uint64_t r = c << HJ_CHUNK_SHIFT;
            if (r >= rows) break;
            while (p + 1 < hc.num_pages && dc.h_row_start[p + 1] <= r) ++p;
            cp[c] = p;
        }
        dc.chunk_page = (uint32_t*)dev_alloc(nch * 4);
        HJ_HIP(hipMemcpyAsync(dc.chunk_page, cp.data(), nch * 4, hipMemcpyHostToDevice,
                              ctx->stream));
        HJ_HIP(hipStreamSynchronize(ctx->stream));
    }
    // page-level index only (the pages themselves stay as the caller gave them): each query
    // indexes or decodes the rows it reads on its own
    dc.rank = (uint16_t*)dev_alloc(groups * 2);
    if (!dc.all_dense && groups) launch_rank_index(dc.view(), dc.rank, ctx->stream);
    HJ_HIP(hipStreamSynchronize(ctx->stream));
}

// ------------------------------------------------------------------ error plumbing
template <class F>
int guarded(hj_context* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return HJ_OK;
    } catch (const Error& e) {
        if (ctx) ctx->last_error = e.what();
        set_global_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host out of memory";
        set_global_error("host out of memory");
        return HJ_E_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        set_global_error(e.what());
        return HJ_E_INVALID;
    }
}

void free_dresult(hj_context* ctx, hj_dresult* r) {
    if (!r) return;
    if (ctx && ctx->live_result == r) {
        ctx->live_result = nullptr;
        ctx->arena.clear();
    }
    delete r;
}

// Device page encoder: plan every column's pages (one wave per column), read the page counts,
// then write the pages in HBM in batches of up to 32 MiB, each batch copied by DMA straight into
// the result columns' page blocks (pinned host memory recycled across results, hostpool.cpp);
// stream order keeps a batch's stage until its copies are done. Byte-identical to the host
// writer (pagecodec.cpp) on the same columns.
void to_host_pages(hj_context* ctx, const hj_dresult* r, hj_result* out) {
    constexpr uint64_t kStagePages = hj_context::kPageStageHalf / kPageBytes;
    const uint64_t n = r->num_rows;
    const size_t ncols = r->cols.size();
    out->num_rows = n;
    out->num_columns = (uint32_t)ncols;
    out->columns = (hj_out_column*)std::calloc(ncols ? ncols : 1, sizeof(hj_out_column));
    if (!out->columns) throw std::bad_alloc();
    Arena& A = ctx->arena;
    const size_t mark = A.mark();
    try {
        for (size_t c0 = 0; c0 < ncols; c0 += kPageCols) {
            const uint32_t nc = (uint32_t)std::min<size_t>(kPageCols, ncols - c0);
            PageBatch pb{};
            pb.ncol = nc;
            pb.counts = (uint64_t*)A.scratch(2 * kPageCols * sizeof(uint64_t));
            pb.valid_rows = pb.counts + kPageCols;
            HJ_HIP(hipMemsetAsync(pb.valid_rows, 0, kPageCols * sizeof(uint64_t), ctx->stream));
            for (uint32_t i = 0; i < nc; ++i) {
                const auto& col = r->cols[c0 + i];
                PageCol& pc = pb.c[i];
                pc.type = col.type;
                pc.n = n;
                pc.validity = col.validity;
                pc.values = col.values;
                pc.offsets = col.offsets;
                pc.chars = col.chars;
                // every page holds at least one row or one 8188-byte piece of a long string
                pc.cap = col.type == HJ_INT32 ? int32_page_bound(n)
                                              : n + col.num_chars / kLongPiece + 2;
                pc.refs = (PageRef*)A.scratch(pc.cap * sizeof(PageRef));
                if (col.type == HJ_VARCHAR && n > kJumpMinRows && n < (1ull << 31)) {
                    // planned by jumps (pagewrite.hip): the planner's scratch
                    const uint64_t coarse = (n + 1) / 1024 + 3;
                    pc.jump = 1;
                    pc.jlp = (uint32_t*)A.scratch((n + 1) * sizeof(uint32_t));
                    pc.jbp = (uint64_t*)A.scratch((n / 4096 + 2) * sizeof(uint64_t));
                    pc.jnx = (uint32_t*)A.scratch((n + 1) * sizeof(uint32_t));
                    pc.ja = (uint32_t*)A.scratch((n + 1) * sizeof(uint32_t));
                    pc.jb = (uint32_t*)A.scratch((n + 1) * sizeof(uint32_t));
                    pc.jc = (uint32_t*)A.scratch(coarse * sizeof(uint32_t));
                    pc.jm = (uint32_t*)A.scratch(coarse * 32 * sizeof(uint32_t));
                    pc.jnode = (uint32_t*)A.scratch(coarse * 1024 * sizeof(uint32_t));
                    pc.jcnt = (uint32_t*)A.scratch(64);
                    pb.jump_rows = std::max<uint64_t>(pb.jump_rows, n);
                }
            }
            {
                Prof p_(ctx, "page_plan");
                page_plan(pb, n, ctx->stream);
            }
            HJ_HIP(hipMemcpyAsync(ctx->pinned, pb.counts, nc * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
            HJ_HIP(hipStreamSynchronize(ctx->stream));
            uint64_t total = 0;
            for (uint32_t i = 0; i < nc; ++i) {
                const uint64_t cnt = ctx->pinned[i];
                if (cnt > pb.c[i].cap) throw Error(HJ_E_INVALID, "page planner exceeded its bound");
                pb.c[i].first = total;
                total += cnt;
                hj_out_column& oc = out->columns[c0 + i];
                oc.type = pb.c[i].type;
                oc.num_pages = (uint32_t)cnt;
                if (cnt) {
                    oc.pages = host_pages_alloc(cnt * HJ_PAGE_SIZE);
                    if (!oc.pages) throw std::bad_alloc();
                }
            }
            if (!total) continue;
            const uint64_t stage = std::min(total, kStagePages);
            const uint64_t nbatch = (total + stage - 1) / stage;
            pb.pages = (uint8_t*)A.scratch(stage * kPageBytes);
            for (uint64_t k = 0; k < nbatch; ++k) {
                const uint64_t p0 = k * stage, cnt = std::min(stage, total - p0);
                {
                    Prof p_(ctx, "page_write");
                    page_write(pb, p0, cnt, ctx->stream);
                }
                for (uint32_t i = 0; i < nc; ++i) {
                    const uint64_t a = std::max(p0, pb.c[i].first);
                    const uint64_t e =
                        std::min(p0 + cnt, pb.c[i].first + out->columns[c0 + i].num_pages);
                    if (a < e)
                        HJ_HIP(hipMemcpyAsync(
                            out->columns[c0 + i].pages + (a - pb.c[i].first) * kPageBytes,
                            pb.pages + (a - p0) * kPageBytes, (e - a) * kPageBytes,
                            hipMemcpyDeviceToHost, ctx->stream));
                }
            }
            HJ_HIP(hipStreamSynchronize(ctx->stream));
            A.reset(mark);
        }
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        A.reset(mark);
        throw;
    }
    A.reset(mark);
}

void to_host(hj_context* ctx, const hj_dresult* r, hj_result* out) {
    HJ_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->gpu_pages) {
        to_host_pages(ctx, r, out);
        return;
    }
    const uint64_t n = r->num_rows;
    out->num_rows = n;
    out->num_columns = (uint32_t)r->cols.size();
    out->columns = (hj_out_column*)std::calloc(r->cols.size() ? r->cols.size() : 1,
                                               sizeof(hj_out_column));
    if (!out->columns) throw std::bad_alloc();
    const uint64_t vbytes = ((n + 63) / 64) * 8;
    for (size_t c = 0; c < r->cols.size(); ++c) {
        const auto& col = r->cols[c];
        std::vector<uint8_t> valid(vbytes ? vbytes : 8);
        if (vbytes) HJ_HIP(hipMemcpy(valid.data(), col.validity, vbytes, hipMemcpyDeviceToHost));
        auto vf = [&valid](uint64_t i) { return ((valid[i >> 3] >> (i & 7)) & 1) != 0; };
        if (col.type == HJ_INT32) {
            std::vector<int32_t> vals(n ? n : 1);
            if (n) HJ_HIP(hipMemcpy(vals.data(), col.values, n * 4, hipMemcpyDeviceToHost));
            encode_int32(vals.data(), vf, n, &out->columns[c]);
        } else {
            std::vector<uint64_t> offs(n + 1);
            HJ_HIP(hipMemcpy(offs.data(), col.offsets, (n + 1) * 8, hipMemcpyDeviceToHost));
            const uint64_t nc = offs[n];
            std::vector<char> chars(nc ? nc : 1);
            if (nc) HJ_HIP(hipMemcpy(chars.data(), col.chars, nc, hipMemcpyDeviceToHost));
            encode_varchar(offs.data(), chars.data(), vf, n, &out->columns[c]);
        }
    }
}

}  // namespace
}  // namespace hj

// ====================================================================== directory object
// ---------------------------------------------------------------- fixed device staging
// hj_directory_probe_noalloc stages its probe keys, counts and first rows in this array instead
// of allocating device memory: every device the library runs on holds its own copy until the
// process ends. Host code never uses the array's name as an address (on the host it is only a
// stand-in); the address each device gives for it is fetched once per device, with that device
// current, when the first context is created there, and kept in g_scratch_ptr.
constexpr uint64_t kNoAllocScratchBytes = HJ_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch` (the caller made the device current)
void init_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw hj::Error(HJ_E_NODEV, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw hj::Error(HJ_E_HIP,
                        std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (uint8_t*)ptr;
}
}  // namespace

struct hj_directory {
    uint64_t n = 0;
    uint32_t k = 0;
    uint64_t max_slot = 0;
    std::vector<void*> allocs;
    hj::DirView view{};
};

using namespace hj;

extern "C" {

const char* hj_global_error(void) {
    std::lock_guard<std::mutex> g(g_err_mu);
    static thread_local std::string copy;
    copy = g_err;
    return copy.c_str();
}

#if !defined(HJ_HASH_ALGO_HIP)
#error "the engine is the HASH_ALGO=hip table; build it through build.py (-DHJ_HASH_ALGO_HIP)"
#endif
const char* hj_hash_algo(void) { return "hip"; }

const char* hj_last_error(const hj_context* ctx) {
    return ctx ? ctx->last_error.c_str() : hj_global_error();
}

int hj_context_create(uint64_t arena_bytes, hj_context** out) {
    if (!out) return HJ_E_INVALID;
    *out = nullptr;
    auto* ctx = new (std::nothrow) hj_context();
    if (!ctx) return HJ_E_OOM;
    int rc = guarded(nullptr, [&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error(HJ_E_NODEV, "no HIP device available");
        HJ_HIP(hipGetDevice(&ctx->device));
        hipDeviceProp_t prop;
        HJ_HIP(hipGetDeviceProperties(&prop, ctx->device));
        ctx->num_cus = prop.multiProcessorCount;
        HJ_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        uint64_t bytes = arena_bytes;
        if (!bytes) {
            const char* e = std::getenv("HJ_ARENA_BYTES");
            bytes = e ? std::strtoull(e, nullptr, 10) : 0;
        }
        if (!bytes) bytes = 16ull << 30;
        HJ_HIP(hipMalloc(&ctx->arena_mem, bytes));
        ctx->arena.init(ctx->arena_mem, bytes);
        HJ_HIP(hipHostMalloc(&ctx->pinned, hj_context::kPinnedBytes, hipHostMallocDefault));
        if (const char* e = std::getenv("HJ_PIPELINE")) ctx->use_pipeline = std::atoi(e) != 0;
        init_noalloc_scratch(ctx->device);
    });
    if (rc) {
        if (ctx->arena_mem) (void)hipFree(ctx->arena_mem);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return HJ_OK;
}

int hj_context_destroy(hj_context* ctx) {
    if (!ctx) return HJ_E_INVALID;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->live_result) delete ctx->live_result;
    for (auto& r : ctx->prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->page_stage) (void)hipHostFree(ctx->page_stage);
    for (auto e : ctx->page_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->arena_mem) (void)hipFree(ctx->arena_mem);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HJ_OK;
}

void* hj_context_stream(hj_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int hj_set_option(hj_context* ctx, const char* name, int64_t value) {
    if (!ctx || !name) return HJ_E_INVALID;
    if (std::strcmp(name, "pipeline") == 0) {
        ctx->use_pipeline = value != 0;
        return HJ_OK;
    }
    if (std::strcmp(name, "partitioned_build") == 0) {
        ctx->part_build = value != 0;
        return HJ_OK;
    }
    if (std::strcmp(name, "gpu_pages") == 0) {
        ctx->gpu_pages = value != 0;
        return HJ_OK;
    }
    if (std::strcmp(name, "xcd_slices") == 0) {
        ctx->xcd_slices = value != 0;
        return HJ_OK;
    }
    if (std::strcmp(name, "index_tables") == 0) {
        ctx->index_tables = value != 0;
        return HJ_OK;
    }
    if (std::strcmp(name, "match_stash") == 0) {
        ctx->match_stash = value != 0;
        return HJ_OK;
    }
    if (std::strcmp(name, "dir_cache_bytes") == 0) {
        if (value < 0) return HJ_E_INVALID;
        (void)hipStreamSynchronize(ctx->stream);
        if (value == 0) {
            ctx->dcache.reset();
        } else {
            ctx->dcache = std::make_shared<DirCache>();
            ctx->dcache->budget = (uint64_t)value;
        }
        return HJ_OK;
    }
    ctx->last_error = std::string("unknown option ") + name;
    return HJ_E_INVALID;
}

int hj_dir_cache_stats(hj_context* ctx, uint64_t* hits, uint64_t* misses, uint64_t* entries,
                       uint64_t* bytes) {
    if (!ctx) return HJ_E_INVALID;
    uint64_t h = 0, m = 0, e = 0, b = 0;
    if (ctx->dcache) {
        std::lock_guard<std::mutex> lk(ctx->dcache->mu);
        h = ctx->dcache->hits;
        m = ctx->dcache->misses;
        e = ctx->dcache->map.size();
        b = ctx->dcache->used;
    }
    if (hits) *hits = h;
    if (misses) *misses = m;
    if (entries) *entries = e;
    if (bytes) *bytes = b;
    return HJ_OK;
}

int hj_profile_enable(hj_context* ctx, const char* prefix) {
    if (!ctx) return HJ_E_INVALID;
    ctx->prof_on = prefix != nullptr;
    ctx->prof_prefix = prefix ? prefix : "";
    return HJ_OK;
}

int hj_profile_report(hj_context* ctx, char* buf, uint64_t len) {
    if (!ctx || !buf || !len) return HJ_E_INVALID;
    return guarded(ctx, [&] {
        HJ_HIP(hipStreamSynchronize(ctx->stream));
        std::map<std::string, std::pair<uint64_t, double>> agg;
        std::vector<std::string> order;
        for (auto& r : ctx->prof) {
            float ms = 0.f;
            HJ_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            auto it = agg.find(r.name);
            if (it == agg.end()) {
                order.push_back(r.name);
                agg[r.name] = {1, ms};
            } else {
                it->second.first += 1;
                it->second.second += ms;
            }
            ctx->event_pool.push_back(r.a);
            ctx->event_pool.push_back(r.b);
        }
        ctx->prof.clear();
        std::string out;
        char line[256];
        for (auto& n : order) {
            std::snprintf(line, sizeof line, "%s %llu %.6f\n", n.c_str(),
                          (unsigned long long)agg[n].first, agg[n].second);
            out += line;
        }
        std::snprintf(buf, len, "%s", out.c_str());
    });
}

int hj_synchronize(hj_context* ctx) {
    if (!ctx) return HJ_E_INVALID;
    return guarded(ctx, [&] { HJ_HIP(hipStreamSynchronize(ctx->stream)); });
}

int hj_table_upload(hj_context* ctx, const hj_table* table, hj_dtable** out) {
    if (!ctx || !table || !out) return HJ_E_INVALID;
    *out = nullptr;
    auto* t = new hj_dtable();
    static std::atomic<uint64_t> next_uid{1};
    t->uid = next_uid.fetch_add(1);
    int rc = guarded(ctx, [&] {
        if (table->num_columns && !table->columns) throw Error(HJ_E_INVALID, "columns missing");
        t->num_rows = table->num_rows;
        t->cols.resize(table->num_columns);
        for (uint32_t c = 0; c < table->num_columns; ++c)
            upload_column(ctx, table->columns[c], table->num_rows, t, t->cols[c]);
    });
    if (rc) {
        for (void* p : t->allocations) (void)hipFree(p);
        delete t;
        return rc;
    }
    *out = t;
    return HJ_OK;
}

int hj_table_upload_cache(hj_context* ctx, const char* path, hj_dtable** out) {
    if (!ctx || !path || !out) return HJ_E_INVALID;
    *out = nullptr;
    hj_cache* c = nullptr;
    int rc = guarded(ctx, [&] { c = open_cache(path); });
    if (rc) return rc;
    rc = hj_table_upload(ctx, &c->table, out);
    close_cache(c);
    return rc;
}

int hj_cache_open(const char* path, hj_cache** out) {
    if (!path || !out) return HJ_E_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] { *out = open_cache(path); });
}

const hj_table* hj_cache_table(const hj_cache* c) { return c ? &c->table : nullptr; }

void hj_cache_close(hj_cache* c) { close_cache(c); }

int hj_cache_write(const hj_table* table, const char* path) {
    if (!table || !path) return HJ_E_INVALID;
    return guarded(nullptr, [&] { write_cache(*table, path); });
}

int hj_table_free(hj_context* ctx, hj_dtable* t) {
    if (!t) return HJ_E_INVALID;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : t->allocations) (void)hipFree(p);
    delete t;
    return HJ_OK;
}

uint64_t hj_table_num_rows(const hj_dtable* t) { return t ? t->num_rows : 0; }

int hj_execute_device(hj_context* ctx, const hj_plan* plan, const hj_dtable* const* tables,
                      hj_dresult** out) {
    if (!ctx || !plan || !out || (!tables && plan->num_inputs)) return HJ_E_INVALID;
    *out = nullptr;
    return guarded(ctx, [&] {
        if (ctx->live_result)
            throw Error(HJ_E_BUSY, "free the previous device result before the next execute");
        host_trace().begin();
        ctx->arena.clear();
        Executor ex(ctx, plan, tables);
        hj_dresult* r = nullptr;
        try {
            r = ex.run();
            host_trace().mark("done");
            host_trace().print();
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->arena.clear();
            throw;
        }
        ctx->live_result = r;
        *out = r;
    });
}

uint64_t hj_dresult_num_rows(const hj_dresult* r) { return r ? r->num_rows : 0; }

int hj_dresult_column(const hj_dresult* r, uint32_t c, int32_t* type, const void** values,
                      const void** chars, const uint8_t** validity) {
    if (!r || c >= r->cols.size()) return HJ_E_INVALID;
    const auto& col = r->cols[c];
    if (type) *type = col.type;
    if (values) *values = col.type == HJ_INT32 ? (const void*)col.values : (const void*)col.offsets;
    if (chars) *chars = col.chars;
    if (validity) *validity = col.validity;
    return HJ_OK;
}

int hj_dresult_to_host(hj_context* ctx, const hj_dresult* r, hj_result* result) {
    if (!ctx || !r || !result) return HJ_E_INVALID;
    std::memset(result, 0, sizeof *result);
    int rc = guarded(ctx, [&] { to_host(ctx, r, result); });
    if (rc) hj_result_free(result);
    return rc;
}

int hj_dresult_free(hj_context* ctx, hj_dresult* r) {
    if (!r || r->batch_view) return HJ_E_INVALID;  // a batch's results go with the batch
    // no wait: the result lives in the context's arena, which the next query reuses in stream
    // order, after this query's kernels
    free_dresult(ctx, r);
    return HJ_OK;
}

void hj_result_free(hj_result* r) {
    if (!r || !r->columns) return;
    for (uint32_t c = 0; c < r->num_columns; ++c) host_pages_free(r->columns[c].pages);
    std::free(r->columns);
    r->columns = nullptr;
    r->num_columns = 0;
    r->num_rows = 0;
}

int hj_execute(hj_context* ctx, const hj_plan* plan, hj_result* result) {
    if (!ctx || !plan || !result) return HJ_E_INVALID;
    std::memset(result, 0, sizeof *result);
    std::vector<hj_dtable*> dts(plan->num_inputs, nullptr);
    int rc = HJ_OK;
    for (uint64_t i = 0; i < plan->num_inputs && rc == HJ_OK; ++i)
        rc = hj_table_upload(ctx, &plan->inputs[i], &dts[i]);
    hj_dresult* dr = nullptr;
    if (rc == HJ_OK) rc = hj_execute_device(ctx, plan, dts.data(), &dr);
    if (rc == HJ_OK) rc = hj_dresult_to_host(ctx, dr, result);
    std::string err = ctx->last_error;
    if (dr) hj_dresult_free(ctx, dr);
    for (auto* t : dts)
        if (t) hj_table_free(ctx, t);
    ctx->last_error = err;
    return rc;
}

// ---------------------------------------------------------------- page codec
int hj_encode_int32(const int32_t* values, const uint8_t* validity, uint64_t n, hj_out_column* out) {
    if (!out || (n && !values)) return HJ_E_INVALID;
    return guarded(nullptr, [&] {
        encode_int32(values, [validity](uint64_t i) { return !validity || validity[i] != 0; }, n,
                     out);
    });
}

int hj_encode_varchar(const uint64_t* offsets, const char* chars, const uint8_t* validity,
                      uint64_t n, hj_out_column* out) {
    if (!out || !offsets) return HJ_E_INVALID;
    return guarded(nullptr, [&] {
        encode_varchar(offsets, chars, [validity](uint64_t i) { return !validity || validity[i] != 0; },
                       n, out);
    });
}

void hj_out_column_free(hj_out_column* col) {
    if (!col) return;
    host_pages_free(col->pages);
    col->pages = nullptr;
    col->num_pages = 0;
}

uint64_t hj_column_num_rows(const hj_column* col) { return col ? column_rows(col) : 0; }
uint64_t hj_column_num_chars(const hj_column* col) { return col ? column_chars(col) : 0; }
int hj_decode_int32(const hj_column* col, int32_t* values, uint8_t* validity) {
    if (!col) return HJ_E_INVALID;
    return decode_int32(col, values, validity);
}
int hj_decode_varchar(const hj_column* col, uint64_t* offsets, char* chars, uint8_t* validity) {
    if (!col) return HJ_E_INVALID;
    return decode_varchar(col, offsets, chars, validity);
}

// ---------------------------------------------------------------- directory test surface
int hj_directory_build(hj_context* ctx, const uint32_t* keys, uint64_t n, hj_directory** out) {
    if (!ctx || !out || (n && !keys)) return HJ_E_INVALID;
    *out = nullptr;
    auto* d = new hj_directory();
    int rc = guarded(ctx, [&] {
        if (n >= (1ull << 32)) throw Error(HJ_E_INVALID, "too many keys");
        auto alloc = [&](size_t b) -> void* {
            void* p = nullptr;
            HJ_HIP(hipMalloc(&p, b ? b : 256));
            d->allocs.push_back(p);
            return p;
        };
        d->n = n;
        d->k = log2_pow2_ceil(n ? n : 1);
        const uint64_t D = 1ull << d->k;
        uint32_t* dk = (uint32_t*)alloc(n * 4);
        uint32_t* h = (uint32_t*)alloc(n * 4);
        uint32_t* row = (uint32_t*)alloc(n * 4);
        uint32_t* th = (uint32_t*)alloc(n * 4);
        uint32_t* tr = (uint32_t*)alloc(n * 4);
        void* scr = alloc(radix_sort_scratch_bytes(n ? n : 1));
        uint64_t* dir = (uint64_t*)alloc((D + 1) * 8);
        uint32_t* e = (uint32_t*)alloc(D * 4);
        uint32_t* bloom = (uint32_t*)alloc(D * 4);
        uint32_t* part = (uint32_t*)alloc(scan_partials_needed(D) * 4);
        if (n) HJ_HIP(hipMemcpyAsync(dk, keys, n * 4, hipMemcpyHostToDevice, ctx->stream));
        hash_keys(dk, n, h, row, ctx->stream);
        radix_sort_pairs(h, row, th, tr, n, scr, ctx->stream);
        build_directory(h, n, d->k, dir, e, bloom, part, ctx->stream);
        std::vector<uint64_t> hd(D + 1);
        HJ_HIP(hipMemcpyAsync(hd.data(), dir, (D + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
        HJ_HIP(hipStreamSynchronize(ctx->stream));
        for (uint64_t s = 0; s < D; ++s)
            d->max_slot = std::max<uint64_t>(d->max_slot, (hd[s + 1] >> 16) - (hd[s] >> 16));
        d->view.dir = dir;
        d->view.sh = h;
        d->view.srow = row;
        d->view.k = d->k;
        d->view.pad = 0;
        d->view.n = n;
    });
    if (rc) {
        for (void* p : d->allocs) (void)hipFree(p);
        delete d;
        return rc;
    }
    *out = d;
    return HJ_OK;
}

int hj_directory_probe(hj_context* ctx, const hj_directory* dir, const uint32_t* probe_keys,
                       uint64_t m, uint64_t* counts, uint64_t* first_row) {
    if (!ctx || !dir || (m && (!probe_keys || !counts))) return HJ_E_INVALID;
    return guarded(ctx, [&] {
        if (!m) return;
        uint32_t* dk = nullptr;
        uint64_t *dc = nullptr, *df = nullptr;
        HJ_HIP(hipMalloc(&dk, m * 4));
        HJ_HIP(hipMalloc(&dc, m * 8));
        HJ_HIP(hipMalloc(&df, m * 8));
        HJ_HIP(hipMemcpyAsync(dk, probe_keys, m * 4, hipMemcpyHostToDevice, ctx->stream));
        dir_probe_keys(dir->view, dk, m, dc, df, ctx->stream);
        HJ_HIP(hipMemcpyAsync(counts, dc, m * 8, hipMemcpyDeviceToHost, ctx->stream));
        if (first_row)
            HJ_HIP(hipMemcpyAsync(first_row, df, m * 8, hipMemcpyDeviceToHost, ctx->stream));
        HJ_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(dk);
        (void)hipFree(dc);
        (void)hipFree(df);
    });
}

int hj_directory_probe_noalloc(hj_context* ctx, const hj_directory* dir,
                               const uint32_t* probe_keys, uint64_t m, uint64_t* counts,
                               uint64_t* first_row) {
    if (!ctx || !dir || (m && (!probe_keys || !counts))) return HJ_E_INVALID;
    return guarded(ctx, [&] {
        if (m > HJ_NOALLOC_MAX_PROBE_KEYS)
            throw Error(HJ_E_INVALID, "hj_directory_probe_noalloc: " + std::to_string(m) +
                                          " probe keys exceed the limit of " +
                                          std::to_string((uint64_t)HJ_NOALLOC_MAX_PROBE_KEYS));
        if (!m) return;
        uint8_t* base = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device]
                                                                      : nullptr;
        if (!base) throw Error(HJ_E_HIP, "no staging address for this device");
        // [keys: 4 m B][counts: 8 m B][first rows: 8 m B], each from a 256-byte boundary; the
        // kernel writes every count (and first row) it stages, so nothing needs clearing
        const uint64_t okeys = 0, ocnt = Arena::align(4 * m), ofirst = ocnt + Arena::align(8 * m);
        static_assert(HJ_NOALLOC_MAX_PROBE_KEYS * 20ull + 512 <= HJ_NOALLOC_SCRATCH_BYTES,
                      "the staging holds the largest accepted probe");
        uint32_t* dk = (uint32_t*)(base + okeys);
        uint64_t* dc = (uint64_t*)(base + ocnt);
        uint64_t* df = first_row ? (uint64_t*)(base + ofirst) : nullptr;
        HJ_HIP(hipMemcpyAsync(dk, probe_keys, m * 4, hipMemcpyHostToDevice, ctx->stream));
        dir_probe_keys(dir->view, dk, m, dc, df, ctx->stream);
        HJ_HIP(hipGetLastError());
        HJ_HIP(hipMemcpyAsync(counts, dc, m * 8, hipMemcpyDeviceToHost, ctx->stream));
        if (first_row)
            HJ_HIP(hipMemcpyAsync(first_row, df, m * 8, hipMemcpyDeviceToHost, ctx->stream));
        HJ_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int hj_directory_probe_pairs(hj_context* ctx, const hj_directory* dir, const uint32_t* probe_keys,
                             uint64_t m, uint64_t** pairs, uint64_t* num_pairs) {
    if (!ctx || !dir || !pairs || !num_pairs || (m && !probe_keys)) return HJ_E_INVALID;
    *pairs = nullptr;
    *num_pairs = 0;
    return guarded(ctx, [&] {
        std::vector<uint64_t> counts(m);
        if (m) {
            int rc = hj_directory_probe(ctx, dir, probe_keys, m, counts.data(), nullptr);
            if (rc) throw Error(rc, ctx->last_error);
        }
        std::vector<uint64_t> offs(m + 1, 0);
        for (uint64_t i = 0; i < m; ++i) offs[i + 1] = offs[i] + counts[i];
        const uint64_t total = offs[m];
        uint64_t* host = (uint64_t*)std::malloc((total ? total : 1) * 16);
        if (!host) throw std::bad_alloc();
        if (total) {
            uint32_t* dk = nullptr;
            uint64_t *doff = nullptr, *dp = nullptr;
            HJ_HIP(hipMalloc(&dk, m * 4));
            HJ_HIP(hipMalloc(&doff, (m + 1) * 8));
            HJ_HIP(hipMalloc(&dp, total * 16));
            HJ_HIP(hipMemcpyAsync(dk, probe_keys, m * 4, hipMemcpyHostToDevice, ctx->stream));
            HJ_HIP(hipMemcpyAsync(doff, offs.data(), (m + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
            dir_probe_pairs(dir->view, dk, m, doff, dp, ctx->stream);
            HJ_HIP(hipMemcpyAsync(host, dp, total * 16, hipMemcpyDeviceToHost, ctx->stream));
            HJ_HIP(hipStreamSynchronize(ctx->stream));
            (void)hipFree(dk);
            (void)hipFree(doff);
            (void)hipFree(dp);
        }
        *pairs = host;
        *num_pairs = total;
    });
}

int hj_directory_info(const hj_directory* dir, uint64_t* num_slots, uint64_t* num_tuples,
                      uint64_t* max_slot_len) {
    if (!dir) return HJ_E_INVALID;
    if (num_slots) *num_slots = 1ull << dir->k;
    if (num_tuples) *num_tuples = dir->n;
    if (max_slot_len) *max_slot_len = dir->max_slot;
    return HJ_OK;
}

int hj_directory_free(hj_context* ctx, hj_directory* dir) {
    if (!dir) return HJ_E_INVALID;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : dir->allocs) (void)hipFree(p);
    delete dir;
    return HJ_OK;
}

void hj_free(void* p) { std::free(p); }

}  // extern "C"
