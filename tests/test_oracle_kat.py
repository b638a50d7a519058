"""The oracle against the reference's own known-answer tests and against a nested-loop join.

Pins oracle_execute (restating src/execute.cpp:35-608) with:
  - the 10 execute() cases of tests/unit_tests.cpp (tests/golden/unit_tests_kat.json);
  - a quadratic nested-loop join over decoded values (no hashing, no partitions) on random
    left-deep / right-deep / bushy plans with NULLs, duplicate keys, strings and filters;
  - the closed form |out| = sum_k b_k * p_k for a two-way join on random key histograms.
"""
import numpy as np
import pytest

import hjtest


@pytest.mark.parametrize("case", hjtest.load_kat_cases(), ids=lambda c: c["name"])
def test_reference_unit_cases(oracle, case):
    plan = hjtest.plan_from_case(case, oracle.codec)
    res = oracle.execute(plan)
    assert res.num_rows == len(case["expected"])
    assert [int(c.type) for c in res.columns] == [int(hjtest.DataType[t]) for t in case["types"]]
    assert oracle.rows(res) == hjtest.expected_rows(case)


@pytest.mark.parametrize("case", hjtest.load_kat_cases(), ids=lambda c: c["name"])
def test_nested_loop_matches_reference_cases(oracle, case):
    plan = hjtest.plan_from_case(case, oracle.codec)
    assert oracle.rows(oracle.nested_loop(plan)) == hjtest.expected_rows(case)


@pytest.mark.parametrize("seed", range(40))
def test_random_plans_vs_nested_loop(oracle, seed):
    rng = np.random.default_rng(1000 + seed)
    plan = hjtest.random_plan(rng, oracle.codec, max_rows=40, key_domain=6)
    a = oracle.rows(oracle.execute(plan, threads=int(rng.integers(2, 9))))
    b = oracle.rows(oracle.nested_loop(plan))
    assert a == b


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_thread_count_does_not_change_multiset(oracle, threads):
    rng = np.random.default_rng(7)
    plan = hjtest.random_plan(rng, oracle.codec, ntables=3, max_rows=300, key_domain=20)
    base = oracle.rows(oracle.execute(plan, threads=4))
    assert oracle.rows(oracle.execute(plan, threads=threads)) == base


def test_two_way_count_closed_form(oracle):
    """|R join S| = sum over non-NULL keys k of count_R(k) * count_S(k)"""
    rng = np.random.default_rng(3)
    nR, nS, dom = 5000, 20000, 700
    rk = rng.integers(0, dom, nR).astype(np.int32)
    sk = rng.integers(0, dom, nS).astype(np.int32)
    rv = (rng.random(nR) > 0.05).astype(np.uint8)
    sv = (rng.random(nS) > 0.1).astype(np.uint8)
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(nR, [oracle.codec.encode_int32(rk, rv)]))
    plan.inputs.append(hjtest.ColumnarTable(nS, [oracle.codec.encode_int32(sk, sv)]))
    plan.new_scan_node(0, [(0, hjtest.DataType.INT32)])
    plan.new_scan_node(1, [(0, hjtest.DataType.INT32)])
    plan.new_join_node(True, 0, 1, 0, 0, [(0, hjtest.DataType.INT32), (1, hjtest.DataType.INT32)])
    plan.root = 2
    res = oracle.execute(plan)
    br = np.bincount(rk[rv == 1], minlength=dom).astype(np.int64)
    bs = np.bincount(sk[sv == 1], minlength=dom).astype(np.int64)
    assert res.num_rows == int((br * bs).sum())


def test_filter_semantics(oracle):
    """a filtered scan keeps rows with a non-NULL attribute inside [lo, hi], in order"""
    rng = np.random.default_rng(11)
    n = 3000
    key = np.arange(n, dtype=np.int32)
    attr = rng.integers(0, 100, n).astype(np.int32)
    av = (rng.random(n) > 0.2).astype(np.uint8)
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(n, [oracle.codec.encode_int32(key),
                                                oracle.codec.encode_int32(attr, av)]))
    plan.new_scan_node(0, [(0, hjtest.DataType.INT32)], filters=[(1, 10, 19)])
    plan.root = 0
    got = [r[0] for r in oracle.rows(oracle.execute(plan))]
    want = [int(k) for k, a, v in zip(key, attr, av) if v and 10 <= a <= 19]
    assert got == sorted(want)
