// batch.cpp - host side of the batched mode: plans many independent queries into one work list,
// allocates their tables, indexes and outputs in one device allocation, and runs all of them in
// four launches (batch.hip): clear, index, insert, probe.
//
// No counterpart in the reference (its execute() runs one plan at a time). The launches are
// captured once as a HIP graph (use_graph) and replayed on every run. A query is batched
// when it is a left-deep chain whose build sides are base scans and whose join keys are INT32
// columns of the chain's source scan (a star join), with range filters at most in its scans and
// INT32 outputs only; any other query, and any batched query whose build keys repeat, runs one
// at a time through the ordinary engine after the batch's launch. Results are the same either
// way (bit-exact rows; the row order is not part of the contract).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hjoin.h"
#include "batch.hpp"
#include "hj_internal.hpp"
#include "pipeline.hpp"

using namespace hj;

namespace {

constexpr uint32_t kMaxQueries = 1u << 20;

struct QueryPlan {
    bool batched = false;
    bool launch = false;  // runs in the launches (else on its own, copied into its outputs)
    BatchQuery q{};
    // host-side sizes for the allocation
    uint64_t src_rows = 0;
    uint32_t k[kBatchJoins] = {};
    uint32_t brows[kBatchJoins] = {};
    std::vector<const DevColumn*> idx_cols;  // columns read through a group index
};

struct Slot {
    int slot;  // 0 source, 1 + j build table of join j
    uint32_t col;
};

// Resolve a plan into a batched query: a chain of joins whose build sides are scans and whose
// probe keys are columns of the chain's source scan. Returns false for any other plan.
bool plan_query(const hj_plan* p, const hj_dtable* const* tables, QueryPlan& qp) {
    if (!p || !p->nodes || p->root >= p->num_nodes) return false;
    std::vector<uint64_t> joins;
    uint64_t idx = p->root;
    for (int depth = 0;; ++depth) {
        if (depth > kBatchJoins || idx >= p->num_nodes) return false;
        const hj_node& nd = p->nodes[idx];
        if (nd.kind == HJ_NODE_SCAN) break;
        if (nd.kind != HJ_NODE_JOIN) return false;
        const uint64_t b = nd.build_left ? nd.left : nd.right;
        if (b >= p->num_nodes || p->nodes[b].kind != HJ_NODE_SCAN) return false;
        joins.push_back(idx);
        idx = nd.build_left ? nd.right : nd.left;
    }
    std::reverse(joins.begin(), joins.end());
    if (joins.empty() || joins.size() > (size_t)kBatchJoins) return false;
    const uint64_t source = idx;
    // a scan node -> its table, checked (no predicate programs; range filters on INT32 columns)
    auto scan_table = [&](const hj_node& nd) -> const hj_dtable* {
        if (nd.base_table_id >= p->num_inputs || !tables[nd.base_table_id]) return nullptr;
        if (nd.num_preds || nd.num_filters > (uint64_t)kMaxFilters) return nullptr;
        const hj_dtable* t = tables[nd.base_table_id];
        for (uint64_t f = 0; f < nd.num_filters; ++f)
            if (nd.filters[f].column >= t->cols.size() ||
                t->cols[nd.filters[f].column].type != HJ_INT32)
                return nullptr;
        for (uint64_t a = 0; a < nd.num_output_attrs; ++a)
            if (!nd.output_attrs || nd.output_attrs[a].index >= t->cols.size()) return nullptr;
        return t;
    };
    const hj_node& sn = p->nodes[source];
    const hj_dtable* st = scan_table(sn);
    if (!st || st->num_rows >= (1ull << 29)) return false;
    std::vector<const hj_dtable*> slot_table{st};
    std::vector<Slot> cur;
    for (uint64_t a = 0; a < sn.num_output_attrs; ++a) cur.push_back({0, (uint32_t)sn.output_attrs[a].index});

    // a column the kernel streams in 64-row groups: indexed (regular pages) or dense uniform
    auto streamable = [](const DevColumn& c) {
        return c.type == HJ_INT32 && (c.view().umagic != 0 || c.regular) &&
               (uint64_t)c.num_pages * HJ_PAGE_SIZE < (1ull << 32);
    };
    auto want_index = [&](const DevColumn& c) {
        if (c.view().umagic == 0 && c.regular && c.num_rows)
            if (std::find(qp.idx_cols.begin(), qp.idx_cols.end(), &c) == qp.idx_cols.end())
                qp.idx_cols.push_back(&c);
    };
    BatchQuery& q = qp.q;
    std::memset(&q, 0, sizeof q);
    q.nsf = (int32_t)sn.num_filters;
    for (uint64_t f = 0; f < sn.num_filters; ++f) {
        const DevColumn& c = st->cols[sn.filters[f].column];
        if (!streamable(c)) return false;
        want_index(c);
        q.sfcol[f] = c.view();
        q.sflo[f] = sn.filters[f].lo;
        q.sfhi[f] = sn.filters[f].hi;
    }
    q.nj = (int32_t)joins.size();
    std::vector<Slot> keys;
    for (size_t i = 0; i < joins.size(); ++i) {
        const hj_node& jn = p->nodes[joins[i]];
        const bool bl = jn.build_left != 0;
        const hj_node& bn = p->nodes[bl ? jn.left : jn.right];
        const hj_dtable* bt = scan_table(bn);
        if (!bt || bt->num_rows >= (1ull << 28)) return false;
        std::vector<Slot> bat;
        for (uint64_t a = 0; a < bn.num_output_attrs; ++a)
            bat.push_back({(int)i + 1, (uint32_t)bn.output_attrs[a].index});
        slot_table.push_back(bt);
        const uint64_t battr = bl ? jn.left_attr : jn.right_attr;
        const uint64_t pattr = bl ? jn.right_attr : jn.left_attr;
        if (battr >= bat.size() || pattr >= cur.size()) return false;
        const Slot bk = bat[battr], pk = cur[pattr];
        if (pk.slot != 0) return false;  // keyed by a build table: not a star join
        keys.push_back(bk);  // a join key is never NULL in the output (NULL does not match)
        keys.push_back(pk);
        const DevColumn& bc = bt->cols[bk.col];
        const DevColumn& pc = st->cols[pk.col];
        if (!streamable(bc) || !streamable(pc)) return false;
        BatchJoin& J = q.j[i];
        J.bkey = bc.view();
        J.pkey = pc.view();
        want_index(bc);
        want_index(pc);
        J.nbf = (int32_t)bn.num_filters;
        for (uint64_t f = 0; f < bn.num_filters; ++f) {
            const DevColumn& c = bt->cols[bn.filters[f].column];
            if (!streamable(c)) return false;
            want_index(c);
            J.bfcol[f] = c.view();
            J.bflo[f] = bn.filters[f].lo;
            J.bfhi[f] = bn.filters[f].hi;
        }
        J.brows = (uint32_t)bt->num_rows;
        qp.brows[i] = J.brows;
        const std::vector<Slot>& L = bl ? bat : cur;
        const std::vector<Slot>& R = bl ? cur : bat;
        std::vector<Slot> nxt;
        for (uint64_t a = 0; a < jn.num_output_attrs; ++a) {
            if (!jn.output_attrs) return false;
            const uint64_t x = jn.output_attrs[a].index;
            if (x < L.size())
                nxt.push_back(L[x]);
            else if (x < L.size() + R.size())
                nxt.push_back(R[x - L.size()]);
            else
                return false;
        }
        cur.swap(nxt);
    }
    const hj_node& root = p->nodes[p->root];
    if (cur.size() > (size_t)kBatchOut || root.num_output_attrs != cur.size()) return false;
    q.nout = (int32_t)cur.size();
    for (size_t c = 0; c < cur.size(); ++c) {
        const DevColumn& dc = slot_table[cur[c].slot]->cols[cur[c].col];
        if (dc.type != HJ_INT32 || root.output_attrs[c].type != HJ_INT32) return false;
        q.oslot[c] = cur[c].slot;
        q.ocol[c] = dc.view();
        want_index(dc);
        const bool key = std::any_of(keys.begin(), keys.end(), [&](const Slot& k) {
            return k.slot == cur[c].slot && k.col == cur[c].col;
        });
        if (!dc.all_dense && !key) q.onull |= 1u << c;
    }
    q.src_rows = (uint32_t)st->num_rows;
    qp.src_rows = st->num_rows;
    qp.batched = true;
    return true;
}

uint64_t pad64(uint64_t words) { return (words + 63) & ~63ull; }  // 256-byte units of u32

}  // namespace

struct hj_batch {
    hj_context* ctx = nullptr;
    uint32_t n = 0;
    hj_batch_params prm{};
    std::vector<const hj_plan*> plans;
    std::vector<std::vector<const hj_dtable*>> tables;
    std::vector<QueryPlan> qp;
    std::vector<int32_t> qidx;   // query -> its batched index, -1 when it runs on its own
    uint32_t nbatched = 0;
    uint8_t* dmem = nullptr;
    uint64_t dbytes = 0;
    BatchDesc* d_desc = nullptr;
    uint32_t* d_ctr = nullptr;   // head words, then kQCtrWords per batched query
    uint32_t* h_ctr = nullptr;   // pinned copy
    uint64_t ctr_words = 0;
    uint32_t nfill = 0;
    uint64_t max_fill = 0;
    uint32_t nitems = 0, nidx = 0, probe0 = 0, wide0 = 0;
    BatchLaunch launch{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hj_dresult> views;   // per batched query, valid after a run
    std::vector<hj_result> own;      // results of the queries run on their own (host pages)
    std::vector<int> own_rc;
    std::vector<std::string> own_err;
    std::vector<uint8_t> fallback;   // per query: ran on its own in the last run
    bool ran = false;
    // contexts (own stream, arena and host thread during a run) that run the star joins routed
    // past the launches beside them; created by the first run that has such queries
    std::vector<hj_context*> helpers;
};

namespace {

void batch_release(hj_batch* b) {
    if (!b) return;
    for (auto* h : b->helpers) hj_context_destroy(h);
    for (auto& r : b->own) hj_result_free(&r);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->h_ctr) (void)hipHostFree(b->h_ctr);
    if (b->dmem) (void)hipFree(b->dmem);
    delete b;
}

template <class F>
int batch_guard(hj_context* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return HJ_OK;
    } catch (const Error& e) {
        if (ctx) ctx->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host out of memory";
        return HJ_E_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return HJ_E_INVALID;
    }
}

void check_params(hj_batch_params& p) {
    if (!p.probe_item_groups) p.probe_item_groups = 256;
    if (!p.insert_item_rows) p.insert_item_rows = 8192;
    if (!p.load_shift) p.load_shift = 1;
    if (!p.large_load_shift) p.large_load_shift = p.load_shift;
    if (!p.large_build_rows) p.large_build_rows = ~0u;
    if (!p.unroll) p.unroll = 4;
    if (p.probe_item_groups > (1u << 20) || p.insert_item_rows % 1024 || p.radix_bits > 12 ||
        p.load_shift > 3 || p.large_load_shift > 3 || (p.unroll != 2 && p.unroll != 4) ||
        p.coop > 1)
        throw Error(HJ_E_INVALID, "batch parameters out of range");
}

// plan every query, lay out the device allocation, write the descriptors
void batch_build(hj_batch* b) {
    hj_context* ctx = b->ctx;
    const uint32_t n = b->n;
    b->qp.resize(n);
    b->qidx.assign(n, -1);
    b->fallback.assign(n, 0);
    b->own.assign(n, hj_result{});
    b->own_rc.assign(n, HJ_OK);
    b->own_err.assign(n, std::string());
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < n; ++i) {
        if (plan_query(b->plans[i], b->tables[i].data(), b->qp[i])) {
            b->qidx[i] = (int32_t)b->nbatched++;
            uint64_t brows = 0;
            for (int j = 0; j < b->qp[i].q.nj; ++j) brows += b->qp[i].brows[j];
            b->qp[i].launch =
                (!b->prm.plain_probe_rows || b->qp[i].src_rows < b->prm.plain_probe_rows) &&
                (!b->prm.plain_build_rows || brows < b->prm.plain_build_rows);
            order.push_back(i);
        }
    }
    // largest sources first: the small queries' items fill the end of the launch
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return b->qp[x].src_rows > b->qp[y].src_rows;
    });
    const uint32_t nb = b->nbatched;
    // the columns indexed by the batch (shared between queries that read the same column)
    std::vector<const DevColumn*> icols;
    std::map<const DevColumn*, uint32_t> icol_id;
    std::map<const uint8_t*, uint32_t> icol_pages;
    std::vector<uint32_t> in_launch;
    for (uint32_t i : order)
        if (b->qp[i].launch) in_launch.push_back(i);
    for (uint32_t i : in_launch)
        for (const DevColumn* c : b->qp[i].idx_cols)
            if (!icol_id.count(c)) {
                icol_id[c] = (uint32_t)icols.size();
                icol_pages[c->pages] = (uint32_t)icols.size();
                icols.push_back(c);
            }
    // work items: index, then inserts, then probes, each in query order
    std::vector<BatchItem> items;
    for (uint32_t c = 0; c < icols.size(); ++c) {
        const uint64_t G = (icols[c]->num_rows + 63) / 64;
        for (uint64_t g = 0; g < G; g += kIndexItemGroups)
            items.push_back({kItemIndex << 24, c, (uint32_t)g,
                             (uint32_t)std::min<uint64_t>(G, g + kIndexItemGroups)});
    }
    b->nidx = (uint32_t)items.size();
    const hj_batch_params& P = b->prm;
    for (uint32_t i : order) {
        QueryPlan& qp = b->qp[i];
        const uint32_t qi = (uint32_t)b->qidx[i];
        for (int j = 0; j < qp.q.nj; ++j) {
            if (!qp.launch) {
                qp.k[j] = kMinTableBits;  // no table: the query runs on its own
                continue;
            }
            const uint32_t rows = qp.brows[j];
            const uint32_t shift = rows >= P.large_build_rows ? P.large_load_shift : P.load_shift;
            uint32_t k = kMinTableBits;
            while ((1ull << k) < ((uint64_t)rows << shift)) ++k;
            qp.k[j] = k;
            qp.q.j[j].k = k;
            for (uint64_t r = 0; r < rows; r += P.insert_item_rows) {
                items.push_back({(kItemInsert << 24) | qi, (uint32_t)j, (uint32_t)r,
                                 (uint32_t)std::min<uint64_t>(rows, r + P.insert_item_rows)});
            }
        }
    }
    b->probe0 = (uint32_t)items.size();
    // probe items: queries of one or two joins, then of three or four; within each, item k of
    // every query before item k + 1 of any, so that the workgroups running at once reserve
    // their output ranges from many queries' counters rather than all from the largest one's
    for (int wide = 0; wide < 2; ++wide) {
        if (wide) b->wide0 = (uint32_t)items.size();
        for (uint64_t k = 0;; ++k) {
            bool any = false;
            for (uint32_t i : in_launch) {
                if ((b->qp[i].q.nj > 2) != (wide != 0)) continue;
                const uint64_t G = (b->qp[i].src_rows + 63) / 64;
                const uint64_t g = k * P.probe_item_groups;
                if (g >= G) continue;
                any = true;
                items.push_back({(kItemProbe << 24) | (uint32_t)b->qidx[i], 0, (uint32_t)g,
                                 (uint32_t)std::min<uint64_t>(G, g + P.probe_item_groups)});
            }
            if (!any) break;
        }
    }
    b->nitems = (uint32_t)items.size();

    // layout (256-byte aligned pieces): descriptors | counters | tables | group indexes |
    // bitmaps set to zero | bitmaps set to ones | values
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = off;
        off += (bytes + 255) & ~255ull;
        return at;
    };
    const uint64_t o_desc = take(sizeof(BatchDesc));
    const uint64_t o_q = take(sizeof(BatchQuery) * std::max<uint32_t>(1, nb));
    const uint64_t o_items = take(sizeof(BatchItem) * std::max<size_t>(1, items.size()));
    const uint64_t o_icol = take(sizeof(DCol) * std::max<size_t>(1, icols.size()));
    const uint64_t o_iout = take(sizeof(uint64_t*) * std::max<size_t>(1, icols.size()));
    const uint64_t o_fill = take(sizeof(FillSeg) * 4);
    b->ctr_words = pad64(4 + (uint64_t)kQCtrWords * nb);
    const uint64_t o_ctr = take(b->ctr_words * 4);
    const uint64_t o_tab = off;
    std::vector<std::vector<uint64_t>> tab_off(n);
    for (uint32_t i : order) {
        for (int j = 0; j < b->qp[i].q.nj; ++j)
            tab_off[i].push_back(take(pad64(2ull * (1ull << b->qp[i].k[j]) + 8) * 4));
    }
    const uint64_t tab_words = (off - o_tab) / 4;
    std::vector<uint64_t> gidx_off(icols.size());
    for (size_t c = 0; c < icols.size(); ++c)
        gidx_off[c] = take((2 * ((icols[c]->num_rows + 63) / 64) + 2) * 8);
    // bitmaps: 64-bit words of validity per output column (bit i of byte i / 8)
    std::vector<std::vector<uint64_t>> vb_off(n, std::vector<uint64_t>(kBatchOut, 0));
    const uint64_t o_vz = off;
    for (uint32_t i : order)
        for (int c = 0; c < b->qp[i].q.nout; ++c)
            if ((b->qp[i].q.onull >> c) & 1u)
                vb_off[i][c] = take(pad64(2 * ((b->qp[i].src_rows + 63) / 64) + 2) * 4);
    const uint64_t vz_words = (off - o_vz) / 4;
    const uint64_t o_vo = off;
    for (uint32_t i : order)
        for (int c = 0; c < b->qp[i].q.nout; ++c)
            if (!((b->qp[i].q.onull >> c) & 1u))
                vb_off[i][c] = take(pad64(2 * ((b->qp[i].src_rows + 63) / 64) + 2) * 4);
    const uint64_t vo_words = (off - o_vo) / 4;
    std::vector<std::vector<uint64_t>> val_off(n, std::vector<uint64_t>(kBatchOut, 0));
    for (uint32_t i : order)
        for (int c = 0; c < b->qp[i].q.nout; ++c)
            val_off[i][c] = take(std::max<uint64_t>(1, b->qp[i].src_rows) * 4);
    b->dbytes = off;
    HJ_HIP(hipMalloc(&b->dmem, b->dbytes));
    uint8_t* D = b->dmem;

    // descriptors
    std::vector<uint8_t> host(o_tab, 0);
    std::vector<BatchQuery> hq(std::max<uint32_t>(1, nb));
    for (uint32_t i : order) {
        QueryPlan& qp = b->qp[i];
        const uint32_t qi = (uint32_t)b->qidx[i];
        BatchQuery& q = qp.q;
        q.ctr = (uint32_t*)(D + o_ctr) + 4 + (uint64_t)kQCtrWords * qi;
        // the views carry the batch's group index of their column (a column is known by its
        // device pages)
        auto fix = [&](DCol& v) {
            auto it = icol_pages.find(v.pages);
            if (it != icol_pages.end()) v.gidx = (const uint64_t*)(D + gidx_off[it->second]);
        };
        for (int f = 0; f < q.nsf; ++f) fix(q.sfcol[f]);
        for (int j = 0; j < q.nj; ++j) {
            BatchJoin& J = q.j[j];
            fix(J.bkey);
            fix(J.pkey);
            for (int f = 0; f < J.nbf; ++f) fix(J.bfcol[f]);
            J.table = (uint32_t*)(D + tab_off[i][j]);
        }
        for (int c = 0; c < q.nout; ++c) {
            fix(q.ocol[c]);
            q.out[c] = (int32_t*)(D + val_off[i][c]);
            q.ovalid[c] = (uint32_t*)(D + vb_off[i][c]);
        }
        hq[qi] = q;
    }
    std::memcpy(host.data() + o_q, hq.data(), sizeof(BatchQuery) * nb);
    std::memcpy(host.data() + o_items, items.data(), sizeof(BatchItem) * items.size());
    std::vector<DCol> hic(icols.size());
    std::vector<uint64_t*> hio(icols.size());
    for (size_t c = 0; c < icols.size(); ++c) {
        hic[c] = icols[c]->view();
        hio[c] = (uint64_t*)(D + gidx_off[c]);
    }
    std::memcpy(host.data() + o_icol, hic.data(), sizeof(DCol) * hic.size());
    std::memcpy(host.data() + o_iout, hio.data(), sizeof(uint64_t*) * hio.size());
    FillSeg fills[4];
    uint32_t nfill = 0;
    b->max_fill = 0;
    auto add_fill = [&](uint64_t o, uint64_t words, uint32_t v) {
        if (!words) return;
        fills[nfill++] = {(uint32_t*)(D + o), words, v, 0};
        b->max_fill = std::max(b->max_fill, words);
    };
    add_fill(o_ctr, b->ctr_words, 0);
    add_fill(o_tab, tab_words, kEmptyKey);
    add_fill(o_vz, vz_words, 0);
    add_fill(o_vo, vo_words, ~0u);
    b->nfill = nfill;
    std::memcpy(host.data() + o_fill, fills, sizeof fills);
    BatchDesc d{};
    d.q = (const BatchQuery*)(D + o_q);
    d.items = (const BatchItem*)(D + o_items);
    d.nitems = b->nitems;
    d.nidx = b->nidx;
    d.icol = (const DCol*)(D + o_icol);
    d.iout = (uint64_t* const*)(D + o_iout);
    d.head = (uint32_t*)(D + o_ctr);
    d.probe0 = b->probe0;
    d.wide0 = b->wide0;
    d.radix_bits = P.radix_bits;
    d.coop = P.coop;
    d.fill = (const FillSeg*)(D + o_fill);
    d.nfill = nfill;
    std::memcpy(host.data() + o_desc, &d, sizeof d);
    HJ_HIP(hipMemcpy(D, host.data(), host.size(), hipMemcpyHostToDevice));
    b->d_desc = (BatchDesc*)(D + o_desc);
    b->d_ctr = (uint32_t*)(D + o_ctr);
    HJ_HIP(hipHostMalloc((void**)&b->h_ctr, b->ctr_words * 4, hipHostMallocDefault));
    BatchLaunch& L = b->launch;
    L.nfill = nfill;
    L.max_fill_words = b->max_fill;
    L.nidx = b->nidx;
    L.nins = b->probe0 - b->nidx;
    L.nprobe = b->wide0 - b->probe0;
    L.nwide = b->nitems - b->wide0;
    L.unroll = P.unroll;
    L.coop = P.coop;
    L.grids = batch_grids(P.unroll, P.coop, (uint32_t)ctx->num_cus);
    b->views.resize(nb);
    for (uint32_t i : order) {
        hj_dresult& v = b->views[(uint32_t)b->qidx[i]];
        v.batch_view = true;
        v.cols.resize(b->qp[i].q.nout);
        for (int c = 0; c < b->qp[i].q.nout; ++c) {
            v.cols[c].type = HJ_INT32;
            v.cols[c].values = (int32_t*)(D + val_off[i][c]);
            v.cols[c].validity = D + vb_off[i][c];
        }
    }
}

// HJ_BATCH_TRACE=1: each step of a run on stderr (capture, launch, wait); 2: without a graph,
// also each of the four launches waited for on its own. For diagnosis.
int batch_trace() {
    static const int level = [] {
        const char* e = std::getenv("HJ_BATCH_TRACE");
        return e ? std::atoi(e) : 0;
    }();
    return level;
}
void trace(const char* what, const hj_batch* b) {
    if (batch_trace())
        std::fprintf(stderr, "[hj_batch] %s: %u queries (%u batched), items %u index, %u insert, %u + %u probe\n",
                     what, b->n, b->nbatched, b->launch.nidx, b->launch.nins, b->launch.nprobe,
                     b->launch.nwide);
}

void batch_enqueue(hj_batch* b) {
    hipStream_t s = b->ctx->stream;
    if (!b->nitems) {
        HJ_HIP(hipMemsetAsync(b->d_ctr, 0, b->ctr_words * 4, s));
        return;
    }
    if (!b->prm.use_graph) {
        trace("launch", b);
        if (batch_trace() >= 2) {
            static const char* name[4] = {"clear", "index", "insert", "probe"};
            for (int ph = 0; ph < 4; ++ph) {
                batch_launch(b->d_desc, b->launch, s, ph, ph);
                HJ_HIP(hipGetLastError());
                HJ_HIP(hipStreamSynchronize(s));
                trace(name[ph], b);
            }
            return;
        }
        Prof p_(b->ctx, "batch_run");
        batch_launch(b->d_desc, b->launch, s);
        HJ_HIP(hipGetLastError());
        return;
    }
    if (!b->exec) {
        trace("capture", b);
        HJ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        batch_launch(b->d_desc, b->launch, s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) throw Error(HJ_E_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        HJ_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    trace("graph launch", b);
    Prof p_(b->ctx, "batch_run");  // the four launches, when kernel timing is on
    HJ_HIP(hipGraphLaunch(b->exec, s));
}

}  // namespace

extern "C" {

int hj_batch_create(hj_context* ctx, uint32_t n, const hj_plan* const* plans,
                    const hj_dtable* const* const* tables, const hj_batch_params* params,
                    hj_batch** out) {
    if (!ctx || !out || (n && (!plans || !tables)) || n > kMaxQueries) return HJ_E_INVALID;
    *out = nullptr;
    auto* b = new (std::nothrow) hj_batch();
    if (!b) return HJ_E_OOM;
    const int rc = batch_guard(ctx, [&] {
        b->ctx = ctx;
        b->n = n;
        if (params) b->prm = *params;
        if (!params) b->prm.use_graph = 1;
        check_params(b->prm);
        for (uint32_t i = 0; i < n; ++i) {
            const hj_plan* p = plans[i];
            if (!p || (!tables[i] && p->num_inputs)) throw Error(HJ_E_INVALID, "batch query without plan or tables");
            b->plans.push_back(p);
            b->tables.emplace_back(tables[i], tables[i] + p->num_inputs);
        }
        batch_build(b);
    });
    if (rc) {
        batch_release(b);
        return rc;
    }
    *out = b;
    return HJ_OK;
}

namespace {
constexpr uint32_t kBatchHelpers = 4;           // helper contexts of a batch
constexpr uint64_t kHelperArena = 8ull << 30;   // at most this much arena each

// the helper contexts (at most kBatchHelpers, at most one per routed query), with the batch
// context's engine options and directory cache; none where creating one fails
void ensure_helpers(hj_batch* b, size_t routed) {
    const size_t want = std::min<size_t>(kBatchHelpers, routed);
    const uint64_t arena = std::min<uint64_t>(b->ctx->arena.cap, kHelperArena);
    while (b->helpers.size() < want) {
        hj_context* h = nullptr;
        if (hj_context_create(arena, &h) != HJ_OK) break;
        h->use_pipeline = b->ctx->use_pipeline;
        h->part_build = b->ctx->part_build;
        h->gpu_pages = b->ctx->gpu_pages;
        h->xcd_slices = b->ctx->xcd_slices;
        h->match_stash = b->ctx->match_stash;
        h->index_tables = b->ctx->index_tables;
        h->dcache = b->ctx->dcache;
        b->helpers.push_back(h);
    }
}

// query i on its own on context c: a star join's columns go into the batch's outputs (written
// there by the fused pass, else copied in stream order), any other query's result to host pages
int run_own(hj_batch* b, hj_context* c, uint32_t i) {
    const int32_t qi = b->qidx[i];
    b->fallback[i] = 1;
    hj_dresult* r = nullptr;
    if (qi >= 0) {
        const hj_dresult& v = b->views[qi];
        for (const auto& col : v.cols) c->emit_into.push_back({col.values, col.validity});
    }
    int rc = hj_execute_device(c, b->plans[i], b->tables[i].data(), &r);
    c->emit_into.clear();
    if (rc == HJ_OK && qi >= 0) {
        rc = batch_guard(c, [&] {
            hj_dresult& v = b->views[qi];
            const uint64_t rows = r->num_rows;
            if (r->cols.size() != v.cols.size() || rows > b->qp[i].src_rows)
                throw Error(HJ_E_INVALID, "batch result shape differs from the plan's");
            for (size_t k = 0; k < v.cols.size(); ++k) {
                if (r->cols[k].values == v.cols[k].values &&
                    r->cols[k].validity == v.cols[k].validity)
                    continue;  // written in place
                if (rows)
                    HJ_HIP(hipMemcpyAsync(v.cols[k].values, r->cols[k].values, rows * 4,
                                          hipMemcpyDeviceToDevice, c->stream));
                const uint64_t vb = (rows + 63) / 64 * 8;
                if (vb && r->cols[k].validity)
                    HJ_HIP(hipMemcpyAsync(v.cols[k].validity, r->cols[k].validity, vb,
                                          hipMemcpyDeviceToDevice, c->stream));
                else if (vb)
                    HJ_HIP(hipMemsetAsync(v.cols[k].validity, 0xff, vb, c->stream));
            }
            HJ_HIP(hipStreamSynchronize(c->stream));
            v.num_rows = rows;
        });
        hj_dresult_free(c, r);
    } else if (rc == HJ_OK) {
        rc = hj_dresult_to_host(c, r, &b->own[i]);
        hj_dresult_free(c, r);
    }
    b->own_rc[i] = rc;
    b->own_err[i] = rc ? c->last_error : std::string();
    return rc;
}
}  // namespace

int hj_batch_run(hj_context* ctx, hj_batch* b) {
    if (!ctx || !b || ctx != b->ctx) return HJ_E_INVALID;
    return batch_guard(ctx, [&] {
        if (ctx->live_result)
            throw Error(HJ_E_BUSY, "free the previous device result before a batch runs");
        for (auto& r : b->own) hj_result_free(&r);
        std::fill(b->fallback.begin(), b->fallback.end(), 0);
        std::fill(b->own_rc.begin(), b->own_rc.end(), HJ_OK);
        b->ran = false;
        // star joins routed past the launches run on the helper contexts, one host thread each,
        // pulling queries from a shared counter, while this thread runs the launches and the
        // queries that run on their own here
        std::vector<uint32_t> routed;
        for (uint32_t i = 0; i < b->n; ++i)
            if (b->qidx[i] >= 0 && !b->qp[i].launch) routed.push_back(i);
        ensure_helpers(b, routed.size());
        std::atomic<size_t> next{0};
        std::mutex mu;
        std::vector<uint32_t> retry;
        std::vector<std::thread> th;
        struct Joiner {  // the helpers are joined however this scope ends
            std::vector<std::thread>& t;
            ~Joiner() {
                for (auto& x : t)
                    if (x.joinable()) x.join();
            }
        } joiner{th};
        for (hj_context* h : b->helpers)
            th.emplace_back([&, h] {
                for (size_t k; (k = next.fetch_add(1)) < routed.size();)
                    if (run_own(b, h, routed[k]) == HJ_E_OOM) {
                        std::lock_guard<std::mutex> lk(mu);
                        retry.push_back(routed[k]);
                    }
            });
        if (b->helpers.empty())
            for (uint32_t i : routed) run_own(b, ctx, i);
        batch_enqueue(b);
        HJ_HIP(hipMemcpyAsync(b->h_ctr, b->d_ctr, b->ctr_words * 4, hipMemcpyDeviceToHost,
                              ctx->stream));
        HJ_HIP(hipStreamSynchronize(ctx->stream));
        trace("done", b);
        for (uint32_t i = 0; i < b->n; ++i) {
            const int32_t qi = b->qidx[i];
            if (qi >= 0 && b->qp[i].launch) {
                const uint32_t* c = b->h_ctr + 4 + (uint64_t)kQCtrWords * qi;
                b->views[qi].num_rows = c[2];
                if (c[1] == 0) continue;  // the launches' result is complete
            }
            if (qi >= 0 && !b->qp[i].launch) continue;  // routed: run by the helpers
            // on its own: not a star join, or a repeated build key
            run_own(b, ctx, i);
        }
        for (auto& x : th) x.join();
        for (uint32_t i : retry) run_own(b, ctx, i);  // a helper's arena was too small
        HJ_HIP(hipStreamSynchronize(ctx->stream));
        b->ran = true;
    });
}

uint32_t hj_batch_size(const hj_batch* b) { return b ? b->n : 0; }

int hj_batch_info(const hj_batch* b, uint32_t i, int32_t* batched, uint64_t* rows) {
    if (!b || i >= b->n) return HJ_E_INVALID;
    const int32_t qi = b->qidx[i];
    if (batched) *batched = (qi >= 0 && b->qp[i].launch && !(b->ran && b->fallback[i])) ? 1 : 0;
    if (rows) *rows = !b->ran ? 0 : qi >= 0 ? b->views[qi].num_rows : b->own[i].num_rows;
    return HJ_OK;
}

int hj_batch_result(const hj_batch* b, uint32_t i, const hj_dresult** out) {
    if (!b || !out || i >= b->n || !b->ran) return HJ_E_INVALID;
    *out = nullptr;
    if (b->qidx[i] < 0 || b->own_rc[i]) return HJ_E_INVALID;
    *out = &b->views[b->qidx[i]];
    return HJ_OK;
}

int hj_batch_result_to_host(hj_context* ctx, hj_batch* b, uint32_t i, hj_result* out) {
    if (!ctx || !b || !out || i >= b->n || !b->ran || ctx != b->ctx) return HJ_E_INVALID;
    std::memset(out, 0, sizeof *out);
    if (b->own_rc[i]) {
        ctx->last_error = b->own_err[i];
        return b->own_rc[i];
    }
    if (b->qidx[i] >= 0) return hj_dresult_to_host(ctx, &b->views[b->qidx[i]], out);
    if (!b->own[i].columns && b->own[i].num_rows == 0 && b->own[i].num_columns == 0)
        return HJ_E_INVALID;  // handed over already
    *out = b->own[i];
    std::memset(&b->own[i], 0, sizeof b->own[i]);
    return HJ_OK;
}

uint64_t hj_batch_device_bytes(const hj_batch* b) { return b ? b->dbytes : 0; }

int hj_batch_free(hj_context* ctx, hj_batch* b) {
    if (!b || (ctx && ctx != b->ctx)) return HJ_E_INVALID;
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_release(b);
    return HJ_OK;
}

}  // extern "C"
