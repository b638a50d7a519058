"""Load-time scan predicates (the reference's Comparison / LogicalOperation trees, src/statement.cpp:
46-230, include/inner_column.h:250-560, include/statement.h:118-161).

CPU: the oracle's LIKE against Python's re on the reference's own pattern translation ('%' ->
".*", '_' -> ".", the listed specials escaped, full match, '.' not matching a newline, as RE2's
default); the oracle's predicate evaluation against a row-by-row evaluation of the same trees.
GPU (marked): the engine's device evaluation against the oracle on random programs, in every
engine mode, and on long, multi-byte and newline strings.
"""
import ctypes as C
import random
import re

import numpy as np
import pytest

import hjtest

I32, STR = hjtest.DataType.INT32, hjtest.DataType.VARCHAR
SPECIAL = set("\\.^$|?*+()[]{}")


def re_like(s, p):
    rx = "".join(".*" if c == "%" else "." if c == "_" else ("\\" + c if c in SPECIAL else c)
                 for c in p)
    return re.fullmatch(rx, s) is not None


def oracle_like(oracle, s, p):
    f = oracle.lib.oracle_like
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint64]
    bs, bp = s.encode(), p.encode()
    return bool(f(bs, len(bs), bp, len(bp)))


def test_oracle_like_matches_python_re(oracle):
    alpha = ["a", "b", "c", "\n", "é", "中", "😀", ".", "*", "(", "[", "%", "_", "\\", "$", "^"]
    rng = random.Random(7)
    for _ in range(20000):
        s = "".join(rng.choice(alpha) for _ in range(rng.randint(0, 9)))
        p = "".join(rng.choice(alpha + ["%", "%", "_", "_"]) for _ in range(rng.randint(0, 7)))
        assert oracle_like(oracle, s, p) == re_like(s, p), (s, p)


@pytest.mark.parametrize("s,p,want", [
    ("abc", "abc", True), ("abc", "a%", True), ("abc", "%c", True), ("abc", "_b_", True),
    ("abc", "ab", False), ("", "%", True), ("", "_", False), ("a\nb", "a%b", False),
    ("a\nb", "a_b", False), ("é", "_", True), ("中文", "__", True), ("中文", "_", False),
    ("x.y", "x.y", True), ("xay", "x.y", False), ("50%", "50%", True), ("a%b", "a\\%b", False),
])
def test_oracle_like_known_answers(oracle, s, p, want):
    assert oracle_like(oracle, s, p) == want
    assert re_like(s, p) == want


# ---------------------------------------------------------------- program evaluation
def eval_row(prog, row):
    """the reference's column-level semantics (src/statement.cpp:46-230) for one row"""
    st = []
    for e in prog:
        op = e[0]
        if op == "and":
            b = st.pop()
            st.append(st.pop() and b)
        elif op == "or":
            b = st.pop()
            st.append(st.pop() or b)
        elif op == "not":
            st.append(not st.pop())
        else:
            v = row[e[1]]
            if op == "is_null":
                st.append(v is None)
            elif op == "is_not_null":
                st.append(v is not None)
            elif v is None:
                st.append(False)
            elif op in ("like", "not_like"):
                m = re_like(v, e[2])
                st.append(m if op == "like" else not m)
            else:
                a, b = (v.encode(), e[2].encode()) if isinstance(v, str) else (v, e[2])
                st.append({"eq": a == b, "neq": a != b, "lt": a < b, "gt": a > b,
                           "leq": a <= b, "geq": a >= b}[op])
    assert len(st) == 1
    return st[0]


def random_program(rng, table, depth=3):
    """a random postfix program over the table's INT32 and VARCHAR columns"""
    types = table.types

    def leaf():
        c = int(rng.integers(0, len(types)))
        vals = [r[c] for r in table.rows if r[c] is not None]
        k = int(rng.integers(0, 10))
        if k == 0:
            return [("is_null", c)]
        if k == 1:
            return [("is_not_null", c)]
        if types[c] == I32:
            lit = int(rng.choice(vals)) if vals and rng.random() < 0.7 else int(rng.integers(-5, 50))
            return [(str(rng.choice(["eq", "neq", "lt", "gt", "leq", "geq"])), c, lit)]
        base = str(rng.choice(vals)) if vals else "abc"
        if k < 6:
            chars = list(base)
            for i in range(len(chars)):
                x = rng.random()
                if x < 0.2:
                    chars[i] = "_"
                elif x < 0.35:
                    chars[i] = "%"
            if rng.random() < 0.3:
                chars = ["%"] + chars[len(chars) // 2:]
            pat = "".join(chars)[:60]
            return [(str(rng.choice(["like", "not_like"])), c, pat)]
        lit = base if rng.random() < 0.5 else base[: int(rng.integers(0, len(base) + 1))]
        return [(str(rng.choice(["eq", "neq", "lt", "gt", "leq", "geq"])), c, lit)]

    def node(d):
        if d == 0 or rng.random() < 0.35:
            return leaf()
        k = rng.random()
        if k < 0.2:
            return node(d - 1) + [("not",)]
        return node(d - 1) + node(d - 1) + [("and",) if k < 0.6 else ("or",)]

    return node(depth)


def predicate_table(rng, n, null_p=0.2, unicode=False, long_every=0):
    ints = rng.integers(0, 40, size=n)
    alpha = list("abcde%_.") + (["é", "中", "\n", "😀"] if unicode else [])
    rows = []
    for i in range(n):
        L = int(rng.integers(0, 8))
        if long_every and i % long_every == long_every - 1:
            L = int(rng.integers(8186, 12000))
        s = "".join(rng.choice(alpha) for _ in range(L))
        rows.append([None if rng.random() < null_p else int(ints[i]),
                     None if rng.random() < null_p else s])
    return hjtest.Table(rows, [I32, STR])


def scan_plan(oracle, table, prog):
    plan = hjtest.Plan()
    plan.inputs.append(table.to_columnar(oracle.codec))
    plan.new_scan_node(0, [(0, I32), (1, STR)], preds=prog)
    plan.root = 0
    return plan


@pytest.mark.parametrize("seed", range(12))
def test_oracle_programs_match_row_evaluation(oracle, seed):
    rng = np.random.default_rng(500 + seed)
    t = predicate_table(rng, 300, unicode=seed % 2 == 1)
    prog = random_program(rng, t)
    kept = hjtest.Table([r for r in t.rows if eval_row(prog, r)], t.types)
    want = oracle.rows(kept.to_columnar(oracle.codec))  # decoded as the results are
    assert oracle.rows(oracle.execute(scan_plan(oracle, t, prog))) == want


# ---------------------------------------------------------------- the engine (GPU)
@pytest.fixture(params=["pipeline", "join_at_a_time"])
def mode(request, gpu_ctx):
    gpu_ctx.set_option("pipeline", 0 if request.param == "join_at_a_time" else 1)
    yield request.param
    gpu_ctx.set_option("pipeline", 1)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(16))
def test_scan_predicates_vs_oracle(gpu_ctx, oracle, seed):
    rng = np.random.default_rng(700 + seed)
    t = predicate_table(rng, 2000, unicode=seed % 2 == 1, long_every=97 if seed % 4 == 3 else 0)
    prog = random_program(rng, t)
    plan = scan_plan(oracle, t, prog)
    assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(12))
def test_join_with_predicates_vs_oracle(gpu_ctx, oracle, mode, seed):
    # a star join whose dimension scans (build sides) and fact scan carry predicate programs
    rng = np.random.default_rng(900 + seed)
    nd = 3
    dims = [predicate_table(rng, 400, unicode=True) for _ in range(nd)]
    for d in dims:  # unique keys in column 0
        for i, r in enumerate(d.rows):
            r[0] = i
    fact_rows = [[int(rng.integers(0, 450)) if rng.random() > 0.05 else None for _ in range(nd)]
                 for _ in range(3000)]
    fact = hjtest.Table(fact_rows, [I32] * nd)
    plan = hjtest.Plan()
    plan.inputs.append(fact.to_columnar(oracle.codec))
    for d in dims:
        plan.inputs.append(d.to_columnar(oracle.codec))
    cur = plan.new_scan_node(0, [(j, I32) for j in range(nd)],
                             preds=[("is_not_null", 0), ("lt", 1, 400), ("and",)]
                             if seed % 2 else ())
    types = [I32] * nd  # the left side's columns
    for j in range(nd):
        prog = random_program(rng, dims[j], depth=2)
        ds = plan.new_scan_node(1 + j, [(0, I32), (1, STR)], preds=prog)
        attrs = [(i, ty) for i, ty in enumerate(types)] + [(len(types) + 1, STR)]
        cur = plan.new_join_node(False, cur, ds, j, 0, attrs)
        types.append(STR)
    plan.root = cur
    assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan))


# patterns of '%' and ASCII characters only, at most 31 tokens: the kernel runs these over
# bytes; the last ones (a non-ASCII literal, '_', 32 tokens) take the character automaton
BYTE_PATTERNS = ["%ab%", "a%", "%b", "ab", "%a%b%", "%%c", "", "%", "%%", "a.c%", "%\n%", "%.%",
                 "%e%d", "ab%%", "%c%c%c", "%" + "a" * 30, "a%b%c%d%e%a%b%c%d%e%a%b%c%d%e%",
                 "%é%", "%a_%", "%" + "b" * 31]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_scan_like_byte_patterns_vs_oracle(gpu_ctx, oracle, seed):
    # texts with multi-byte characters, '\n' and (odd seeds) long strings over several pages
    rng = np.random.default_rng(1100 + seed)
    t = predicate_table(rng, 3000, unicode=True, long_every=211 if seed % 2 else 0)
    for pat in BYTE_PATTERNS:
        for op in ("like", "not_like"):
            plan = scan_plan(oracle, t, [(op, 1, pat)])
            assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan)), (op, pat)
    # six LIKE entries in one program: the fifth and sixth have no table in LDS
    prog = [("like", 1, BYTE_PATTERNS[0])]
    for pat in BYTE_PATTERNS[1:6]:
        prog += [("like", 1, pat), ("or",)]
    plan = scan_plan(oracle, t, prog)
    assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan))


@pytest.mark.gpu
def test_predicate_errors(gpu_ctx, hjo, oracle):
    t = predicate_table(np.random.default_rng(1), 50)
    for prog, code in [([("like", 0, "%1")], -2),              # LIKE on INT32
                       ([("eq", 0, 1), ("and",)], -1),          # AND without two operands
                       ([("eq", 0, 1), ("eq", 0, 2)], -1),      # two results left
                       ([("like", 1, "a" * 64)], -1)]:          # pattern too long
        with pytest.raises(hjo.EngineError) as ei:
            gpu_ctx.execute(scan_plan(oracle, t, prog))
        assert ei.value.code == code
