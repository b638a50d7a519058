// hjoin_contest.hpp - the reference's C++ entry points over the C ABI of include/hjoin.h.
//
//   namespace Contest { void* build_context(); void destroy_context(void*);
//                       ColumnarTable execute(const Plan& plan, void* context); }
//   (include/plan.h:398-405, src/execute.cpp:598-614)
//
// Header-only and written against the shape of the reference's plan types (include/plan.h:71-211:
// Plan{nodes, inputs, root}, PlanNode{variant<ScanNode, JoinNode> data,
// vector<tuple<size_t, DataType>> output_attrs}, ColumnarTable{num_rows, columns},
// Column{type, pages, new_page()}, Page{byte data[8192]}), so the same code compiles against the
// reference's plan.h or against any types of that shape. A build that includes plan.h adds, in
// one translation unit:
//
//   #include "plan.h"
//   #include "hjoin_contest.hpp"
//   namespace Contest {
//   void* build_context() { return hj_contest::build_context(); }
//   void destroy_context(void* c) { hj_contest::destroy_context(c); }
//   ColumnarTable execute(const Plan& p, void* c) { return hj_contest::execute<ColumnarTable>(p, c); }
//   }
//
// Errors from the engine are thrown as std::runtime_error with the engine's message (the
// reference's messages where one exists, e.g. "Join key isn't of type int32_t",
// src/execute.cpp:81), as the reference's execute() throws.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

#include "hjoin.h"

// The adapter stands for the reference's execute.cpp under HASH_ALGO=hip; a translation unit
// configured for one of the reference's CPU tables (include/hash_config.h) is a different build.
#if defined(USE_ROBINHOOD_HASH) || defined(USE_HOPSCOTCH_HASH) || defined(USE_CUCKOO_HASH) || \
    defined(USE_UNCHAINED_HASH) || defined(USE_STD_HASH)
#error "hjoin_contest.hpp replaces execute.cpp only under HASH_ALGO=hip"
#endif

namespace hj_contest {

namespace detail {
template <class T, class = void>
struct is_join : std::false_type {};
template <class T>
struct is_join<T, std::void_t<decltype(std::declval<T>().build_left)>> : std::true_type {};

inline void check(hj_context* ctx, int rc) {
    if (rc != HJ_OK) throw std::runtime_error(ctx ? hj_last_error(ctx) : hj_global_error());
}
}  // namespace detail

// Contest::build_context: a HIP stream and an HBM arena sized by the engine (0 = default)
inline void* build_context() {
    hj_context* ctx = nullptr;
    detail::check(nullptr, hj_context_create(0, &ctx));
    return ctx;
}

inline void destroy_context(void* ctx) { hj_context_destroy(static_cast<hj_context*>(ctx)); }

// Contest::execute: Plan -> hj_plan (borrowing the input pages), engine call, result pages copied
// into Column::new_page() pages of the returned table
template <class ColumnarTable, class Plan>
ColumnarTable execute(const Plan& plan, void* context) {
    hj_context* ctx = static_cast<hj_context*>(context);
    bool own = false;
    if (!ctx) {  // the reference's build_context returns nullptr; accept it
        ctx = static_cast<hj_context*>(build_context());
        own = true;
    }
    struct Guard {
        hj_context* c;
        bool own;
        ~Guard() {
            if (own) hj_context_destroy(c);
        }
    } guard{ctx, own};

    // inputs: page pointer arrays per column
    std::vector<std::vector<std::vector<const uint8_t*>>> page_ptrs(plan.inputs.size());
    std::vector<std::vector<hj_column>> cols(plan.inputs.size());
    std::vector<hj_table> tables(plan.inputs.size());
    for (size_t t = 0; t < plan.inputs.size(); ++t) {
        const auto& in = plan.inputs[t];
        page_ptrs[t].resize(in.columns.size());
        cols[t].resize(in.columns.size());
        for (size_t c = 0; c < in.columns.size(); ++c) {
            const auto& col = in.columns[c];
            for (const auto* pg : col.pages)
                page_ptrs[t][c].push_back(reinterpret_cast<const uint8_t*>(pg));
            cols[t][c].type = static_cast<int32_t>(col.type);
            cols[t][c].num_pages = static_cast<uint32_t>(col.pages.size());
            cols[t][c].pages = page_ptrs[t][c].data();
        }
        tables[t].num_rows = in.num_rows;
        tables[t].num_columns = static_cast<uint32_t>(in.columns.size());
        tables[t].reserved = 0;
        tables[t].columns = cols[t].data();
    }
    // nodes
    std::vector<std::vector<hj_attr>> attrs(plan.nodes.size());
    std::vector<hj_node> nodes(plan.nodes.size());
    for (size_t i = 0; i < plan.nodes.size(); ++i) {
        const auto& pn = plan.nodes[i];
        for (const auto& oa : pn.output_attrs) {
            hj_attr a;
            a.index = static_cast<uint64_t>(std::get<0>(oa));
            a.type = static_cast<int32_t>(std::get<1>(oa));
            a.reserved = 0;
            attrs[i].push_back(a);
        }
        hj_node& n = nodes[i];
        std::memset(&n, 0, sizeof n);
        std::visit(
            [&](const auto& d) {
                using D = std::decay_t<decltype(d)>;
                if constexpr (detail::is_join<D>::value) {
                    n.kind = HJ_NODE_JOIN;
                    n.build_left = d.build_left ? 1 : 0;
                    n.left = d.left;
                    n.right = d.right;
                    n.left_attr = d.left_attr;
                    n.right_attr = d.right_attr;
                } else {
                    n.kind = HJ_NODE_SCAN;
                    n.base_table_id = d.base_table_id;
                }
            },
            pn.data);
        n.num_output_attrs = attrs[i].size();
        n.output_attrs = attrs[i].data();
    }
    hj_plan hp;
    hp.num_nodes = nodes.size();
    hp.nodes = nodes.data();
    hp.num_inputs = tables.size();
    hp.inputs = tables.data();
    hp.root = plan.root;

    hj_result res;
    detail::check(ctx, hj_execute(ctx, &hp, &res));
    ColumnarTable out;
    out.num_rows = res.num_rows;
    try {
        using ColumnT = typename decltype(out.columns)::value_type;
        using DataTypeT = decltype(std::declval<ColumnT>().type);
        for (uint32_t c = 0; c < res.num_columns; ++c) {
            const hj_out_column& oc = res.columns[c];
            out.columns.emplace_back(static_cast<DataTypeT>(oc.type));
            auto& col = out.columns.back();
            for (uint32_t p = 0; p < oc.num_pages; ++p)
                std::memcpy(col.new_page(), oc.pages + (size_t)p * HJ_PAGE_SIZE, HJ_PAGE_SIZE);
        }
    } catch (...) {
        hj_result_free(&res);
        throw;
    }
    hj_result_free(&res);
    return out;
}

}  // namespace hj_contest
