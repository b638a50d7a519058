// contest_drop_in.cpp - Contest::execute through include/hjoin_contest.hpp, checked against a
// nested-loop join in this file.
//
// The plan and table types below have the shape of the reference's (include/plan.h:71-211,
// include/attribute.h:8-13): the adapter is written against that shape, and these types stand
// in for plan.h, which needs libraries this image lacks. Inputs are written with the engine's
// page writer (hj_encode_*); results are read back with its page reader (hj_decode_*), both
// checked byte for byte against the oracle's in tests/test_pages.py.
//
// Built by __graft_entry__.build() (g++, linked against libhjoin.so); run by
// tests/test_contest_drop_in.py on the GPU. Prints one line per case and "contest ok" at the end.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <optional>
#include <random>
#include <string>
#include <tuple>
#include <variant>
#include <vector>

// ---------------------------------------------------------------- plan.h-shaped types
enum class DataType { INT32, INT64, FP64, VARCHAR };
struct ScanNode {
    size_t base_table_id;
};
struct JoinNode {
    bool build_left;
    size_t left, right, left_attr, right_attr;
};
struct PlanNode {
    std::variant<ScanNode, JoinNode> data;
    std::vector<std::tuple<size_t, DataType>> output_attrs;
    PlanNode(std::variant<ScanNode, JoinNode> d, std::vector<std::tuple<size_t, DataType>> o)
        : data(std::move(d)), output_attrs(std::move(o)) {}
};
struct alignas(8) Page {
    std::byte data[8192];
};
struct Column {
    DataType type;
    std::vector<Page*> pages;
    explicit Column(DataType t) : type(t) {}
    Column(Column&& o) noexcept : type(o.type), pages(std::move(o.pages)) { o.pages.clear(); }
    Column(const Column&) = delete;
    ~Column() {
        for (Page* p : pages) delete p;
    }
    Page* new_page() {
        pages.push_back(new Page);
        return pages.back();
    }
};
struct ColumnarTable {
    size_t num_rows{0};
    std::vector<Column> columns;
};
struct Plan {
    std::vector<PlanNode> nodes;
    std::vector<ColumnarTable> inputs;
    size_t root = 0;
    size_t new_scan_node(size_t t, std::vector<std::tuple<size_t, DataType>> o) {
        nodes.emplace_back(ScanNode{t}, std::move(o));
        return nodes.size() - 1;
    }
    size_t new_join_node(bool bl, size_t l, size_t r, size_t la, size_t ra,
                         std::vector<std::tuple<size_t, DataType>> o) {
        nodes.emplace_back(JoinNode{bl, l, r, la, ra}, std::move(o));
        return nodes.size() - 1;
    }
};

#include "../../include/hjoin_contest.hpp"

namespace Contest {
void* build_context() { return hj_contest::build_context(); }
void destroy_context(void* c) { hj_contest::destroy_context(c); }
ColumnarTable execute(const Plan& p, void* c) { return hj_contest::execute<ColumnarTable>(p, c); }
}  // namespace Contest

// ---------------------------------------------------------------- rows <-> pages
using Cell = std::optional<std::variant<int32_t, std::string>>;
using Rows = std::vector<std::vector<Cell>>;

static Column to_column(const Rows& rows, size_t c, DataType t) {
    const size_t n = rows.size();
    std::vector<uint8_t> ok(n);
    hj_out_column oc;
    if (t == DataType::INT32) {
        std::vector<int32_t> v(n);
        for (size_t i = 0; i < n; ++i) {
            ok[i] = rows[i][c].has_value();
            v[i] = ok[i] ? std::get<int32_t>(*rows[i][c]) : 0;
        }
        hj_encode_int32(v.data(), ok.data(), n, &oc);
    } else {
        std::vector<uint64_t> off(n + 1, 0);
        std::string chars;
        for (size_t i = 0; i < n; ++i) {
            ok[i] = rows[i][c].has_value();
            if (ok[i]) chars += std::get<std::string>(*rows[i][c]);
            off[i + 1] = chars.size();
        }
        hj_encode_varchar(off.data(), chars.data(), ok.data(), n, &oc);
    }
    Column col(t);
    for (uint32_t p = 0; p < oc.num_pages; ++p)
        std::memcpy(col.new_page(), oc.pages + (size_t)p * HJ_PAGE_SIZE, HJ_PAGE_SIZE);
    hj_out_column_free(&oc);
    return col;
}

static ColumnarTable to_table(const Rows& rows, const std::vector<DataType>& types) {
    ColumnarTable t;
    t.num_rows = rows.size();
    for (size_t c = 0; c < types.size(); ++c) t.columns.push_back(to_column(rows, c, types[c]));
    return t;
}

static Rows from_table(const ColumnarTable& t) {
    Rows rows(t.num_rows, std::vector<Cell>(t.columns.size()));
    for (size_t c = 0; c < t.columns.size(); ++c) {
        std::vector<const uint8_t*> ptrs;
        for (const Page* p : t.columns[c].pages) ptrs.push_back((const uint8_t*)p);
        hj_column hc{(int32_t)t.columns[c].type, (uint32_t)ptrs.size(), ptrs.data()};
        const uint64_t n = hj_column_num_rows(&hc);
        if (n != t.num_rows) throw std::runtime_error("column row count differs from num_rows");
        std::vector<uint8_t> ok(n + 1);
        if (t.columns[c].type == DataType::INT32) {
            std::vector<int32_t> v(n + 1);
            hj_decode_int32(&hc, v.data(), ok.data());
            for (size_t i = 0; i < n; ++i)
                if (ok[i]) rows[i][c] = v[i];
        } else {
            std::vector<uint64_t> off(n + 1);
            std::vector<char> chars(hj_column_num_chars(&hc) + 1);
            hj_decode_varchar(&hc, off.data(), chars.data(), ok.data());
            for (size_t i = 0; i < n; ++i)
                if (ok[i]) rows[i][c] = std::string(chars.data() + off[i], off[i + 1] - off[i]);
        }
    }
    return rows;
}

// ---------------------------------------------------------------- nested-loop reference
static Rows nested_loop(const Plan& plan, const std::vector<Rows>& data, size_t node) {
    const PlanNode& pn = plan.nodes[node];
    if (const auto* s = std::get_if<ScanNode>(&pn.data)) {
        Rows out;
        for (const auto& r : data[s->base_table_id]) {
            std::vector<Cell> o;
            for (const auto& a : pn.output_attrs) o.push_back(r[std::get<0>(a)]);
            out.push_back(o);
        }
        return out;
    }
    const auto& j = std::get<JoinNode>(pn.data);
    const Rows L = nested_loop(plan, data, j.left), R = nested_loop(plan, data, j.right);
    Rows out;
    for (const auto& l : L)
        for (const auto& r : R) {
            const Cell& a = l[j.left_attr];
            const Cell& b = r[j.right_attr];
            if (!a || !b || std::get<int32_t>(*a) != std::get<int32_t>(*b)) continue;
            std::vector<Cell> o;
            for (const auto& at : pn.output_attrs) {
                const size_t k = std::get<0>(at);
                o.push_back(k < l.size() ? l[k] : r[k - l.size()]);
            }
            out.push_back(o);
        }
    return out;
}

static std::string key_of(const std::vector<Cell>& r) {
    std::string s;
    for (const auto& c : r) {
        if (!c)
            s += "N|";
        else if (std::holds_alternative<int32_t>(*c))
            s += "i" + std::to_string(std::get<int32_t>(*c)) + "|";
        else
            s += "s" + std::to_string(std::get<std::string>(*c).size()) + ":" +
                 std::get<std::string>(*c) + "|";
    }
    return s;
}

static std::map<std::string, int> multiset(const Rows& rows) {
    std::map<std::string, int> m;
    for (const auto& r : rows) ++m[key_of(r)];
    return m;
}

// table: INT32 key (domain dom, NULL with probability pn), INT32 payload, VARCHAR payload
static Rows random_rows(std::mt19937_64& g, size_t n, int dom, double pn, bool long_strings) {
    Rows rows;
    std::uniform_int_distribution<int> key(0, dom - 1), len(0, 40), ch(33, 126);
    std::uniform_real_distribution<double> u(0, 1);
    for (size_t i = 0; i < n; ++i) {
        std::vector<Cell> r(3);
        if (u(g) >= pn) r[0] = key(g);
        r[1] = (int32_t)(g() & 0x7fffffff);
        if (u(g) >= 0.1) {
            std::string s(long_strings && i % 97 == 3 ? 9000 + i % 13 : len(g), 'x');
            for (auto& c : s) c = (char)ch(g);
            r[2] = s;
        }
        rows.push_back(r);
    }
    return rows;
}

int main() {
    using DT = DataType;
    void* ctx = Contest::build_context();
    std::mt19937_64 g(7);
    int cases = 0;
    for (int it = 0; it < 6; ++it) {
        const size_t na = 50 + (g() % 3000), nb = 50 + (g() % 3000), nc = 20 + (g() % 500);
        const int dom = 10 + (int)(g() % 2000);
        std::vector<Rows> data = {random_rows(g, na, dom, 0.05, it == 2),
                                  random_rows(g, nb, dom, 0.1, false),
                                  random_rows(g, nc, dom, 0.0, false)};
        Plan plan;
        const std::vector<DT> types = {DT::INT32, DT::INT32, DT::VARCHAR};
        for (const auto& d : data) plan.inputs.push_back(to_table(d, types));
        const size_t s0 = plan.new_scan_node(0, {{0, DT::INT32}, {1, DT::INT32}, {2, DT::VARCHAR}});
        const size_t s1 = plan.new_scan_node(1, {{0, DT::INT32}, {2, DT::VARCHAR}});
        const size_t s2 = plan.new_scan_node(2, {{0, DT::INT32}, {1, DT::INT32}});
        const size_t j1 = plan.new_join_node(it % 2 == 0, s0, s1, 0, 0,
                                             {{0, DT::INT32}, {2, DT::VARCHAR}, {4, DT::VARCHAR}, {1, DT::INT32}});
        plan.root = plan.new_join_node(it % 3 == 0, j1, s2, 0, 0,
                                       {{1, DT::VARCHAR}, {2, DT::VARCHAR}, {5, DT::INT32}, {3, DT::INT32}});
        const ColumnarTable res = Contest::execute(plan, ctx);
        const auto got = multiset(from_table(res));
        const auto want = multiset(nested_loop(plan, data, plan.root));
        size_t n = 0;
        for (const auto& kv : want) n += kv.second;
        std::printf("case %d: %zu rows, %s\n", it, n, got == want ? "equal" : "DIFFERENT");
        if (got != want) return 1;
        ++cases;
    }
    // a non-INT32 join key is the reference's error (src/execute.cpp:81)
    {
        Plan plan;
        std::vector<Rows> data = {random_rows(g, 10, 5, 0.0, false)};
        plan.inputs.push_back(to_table(data[0], {DT::INT32, DT::INT32, DT::VARCHAR}));
        const size_t a = plan.new_scan_node(0, {{2, DT::VARCHAR}});
        const size_t b = plan.new_scan_node(0, {{2, DT::VARCHAR}});
        plan.root = plan.new_join_node(true, a, b, 0, 0, {{0, DT::VARCHAR}});
        bool thrown = false;
        try {
            Contest::execute(plan, ctx);
        } catch (const std::runtime_error& e) {
            thrown = std::string(e.what()).find("Join key isn't of type int32_t") != std::string::npos;
        }
        std::printf("error case: %s\n", thrown ? "thrown" : "NOT THROWN");
        if (!thrown) return 1;
    }
    Contest::destroy_context(ctx);
    std::printf("contest ok: %d join cases\n", cases);
    return 0;
}
