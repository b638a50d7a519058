"""Result pages written on the device (option "gpu_pages", default on) against the oracle's page
writer, byte for byte.

A root scan without filters returns its input rows in order, so its output columns must be the
very pages the oracle's writer made for the input (src/execute.cpp:431-573 applied to the same
rows). Join results are compared with the host writer (gpu_pages = 0) and with the oracle's
writer re-encoding the rows that the oracle's reader decodes from them.
"""
import zlib

import numpy as np
import pytest

import hjtest

pytestmark = pytest.mark.gpu

PAGE = 8192
I32, STR = hjtest.DataType.INT32, hjtest.DataType.VARCHAR


@pytest.fixture
def paged(gpu_ctx):
    gpu_ctx.set_option("gpu_pages", 1)
    yield gpu_ctx
    gpu_ctx.set_option("gpu_pages", 1)


def _execute_both(ctx, plan):
    ctx.set_option("gpu_pages", 1)
    dev = ctx.execute(plan)
    ctx.set_option("gpu_pages", 0)
    host = ctx.execute(plan)
    ctx.set_option("gpu_pages", 1)
    return dev, host


def _same_pages(a, b):
    assert a.num_rows == b.num_rows
    assert len(a.columns) == len(b.columns)
    for j, (x, y) in enumerate(zip(a.columns, b.columns)):
        assert x.type == y.type, j
        assert x.num_pages == y.num_pages, (j, x.num_pages, y.num_pages)
        assert np.array_equal(x.pages, y.pages), j


def _scan_plan(table):
    plan = hjtest.Plan()
    plan.inputs.append(table)
    plan.new_scan_node(0, [(j, c.type) for j, c in enumerate(table.columns)])
    plan.root = 0
    return plan


def _int32_col(codec, n, null_p, rng):
    v = rng.integers(-2**31, 2**31, size=n, dtype=np.int64).astype(np.int32)
    ok = (rng.random(n) >= null_p).astype(np.uint8)
    return codec.encode_int32(v, ok)


def _str_col(codec, lens, null_p, rng):
    strs = []
    for L in lens:
        if rng.random() < null_p:
            strs.append(None)
        else:
            strs.append(bytes(rng.integers(33, 127, size=int(L), dtype=np.uint8)))
    return codec.encode_strings(strs)


INT32_CASES = {
    "dense_3_pages_plus_5": (1984 * 3 + 5, 0.0),
    "one_row": (1, 0.0),
    "one_null": (1, 1.0),
    "all_null_two_pages": (65496 + 10, 1.0),
    "sparse_nulls": (100_000, 0.03),
    "half_null": (50_000, 0.5),
    "mostly_null": (300_000, 0.97),
}


@pytest.mark.parametrize("name", sorted(INT32_CASES))
def test_scan_int32_pages_match_oracle_writer(paged, oracle, name):
    n, null_p = INT32_CASES[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    col = _int32_col(oracle.codec, n, null_p, rng)
    res = paged.execute(_scan_plan(hjtest.ColumnarTable(n, [col])))
    assert res.num_rows == n
    assert res.columns[0].num_pages == col.num_pages
    assert np.array_equal(res.columns[0].pages, col.pages)


STR_CASES = {
    "short": lambda rng, n: rng.integers(0, 70, n),
    "empty_strings": lambda rng, n: np.zeros(n, np.int64),
    "wide": lambda rng, n: rng.integers(0, 3000, n),
    "near_long": lambda rng, n: rng.integers(8180, 8192, n),
    "long_mix": lambda rng, n: np.where(rng.random(n) < 0.03, rng.integers(8186, 30000, n),
                                        rng.integers(0, 200, n)),
    "piece_edges": lambda rng, n: np.resize(
        np.array([8185, 8186, 8188, 8189, 16376, 16377, 0, 3]), n),
}


@pytest.mark.parametrize("null_p", [0.0, 0.3, 0.95])
@pytest.mark.parametrize("name", sorted(STR_CASES))
def test_scan_varchar_pages_match_oracle_writer(paged, oracle, name, null_p):
    rng = np.random.default_rng(zlib.crc32(name.encode()) + int(null_p * 100))
    n = 3000 if name != "short" else 40_000
    col = _str_col(oracle.codec, STR_CASES[name](rng, n), null_p, rng)
    res = paged.execute(_scan_plan(hjtest.ColumnarTable(n, [col])))
    assert res.num_rows == n
    assert res.columns[0].num_pages == col.num_pages
    assert np.array_equal(res.columns[0].pages, col.pages)


def test_scan_many_columns(paged, oracle):
    # more output columns than one planner launch takes (16): two column groups
    rng = np.random.default_rng(77)
    n = 20_000
    cols = []
    for j in range(21):
        if j % 3 == 2:
            cols.append(_str_col(oracle.codec, rng.integers(0, 40, n), 0.2, rng))
        else:
            cols.append(_int32_col(oracle.codec, n, [0.0, 0.1, 0.6][j % 3], rng))
    res = paged.execute(_scan_plan(hjtest.ColumnarTable(n, cols)))
    for j, c in enumerate(cols):
        assert np.array_equal(res.columns[j].pages, c.pages), j


def test_scan_staged_batches(paged, oracle):
    # more pages than two writer batches (4096 pages each): the copy-out cycles through both
    # halves of the pinned double buffer
    n = 1984 * 9000 + 77
    rng = np.random.default_rng(8)
    v = rng.integers(-2**31, 2**31, size=n, dtype=np.int64).astype(np.int32)
    ok = np.ones(n, np.uint8)
    ok[rng.integers(0, n, 1000)] = 0
    col = oracle.codec.encode_int32(v, ok)
    assert col.num_pages > 2 * 4096
    res = paged.execute(_scan_plan(hjtest.ColumnarTable(n, [col])))
    assert res.columns[0].num_pages == col.num_pages
    assert np.array_equal(res.columns[0].pages, col.pages)


def test_empty_result_has_no_pages(paged, oracle):
    t = hjtest.Table([[1, "a"], [2, None]], [I32, STR]).to_columnar(oracle.codec)
    plan = hjtest.Plan()
    plan.inputs.append(t)
    plan.new_scan_node(0, [(0, I32), (1, STR)], filters=[(0, 10, 20)])
    plan.root = 0
    res = paged.execute(plan)
    assert res.num_rows == 0
    assert [c.num_pages for c in res.columns] == [0, 0]


@pytest.mark.parametrize("seed", range(12))
def test_join_result_pages_device_vs_host_and_oracle(gpu_ctx, oracle, seed):
    rng = np.random.default_rng(9100 + seed)
    rows, dom = [(300, 60), (3000, 3000), (6000, 3000)][seed % 3]
    plan = hjtest.random_plan(rng, oracle.codec, max_rows=rows, key_domain=dom,
                              long_every=int(rng.choice([0, 97])))
    dev, host = _execute_both(gpu_ctx, plan)
    _same_pages(dev, host)
    # the oracle's writer on the decoded rows gives the same pages
    for c in dev.columns:
        vals, ok = oracle.codec.decode(c)
        if c.type == I32:
            again = oracle.codec.encode_int32(vals, ok.astype(np.uint8))
        else:
            off, ch = vals
            again = oracle.codec.encode_varchar(off, ch, ok.astype(np.uint8))
        assert again.num_pages == c.num_pages
        assert np.array_equal(again.pages, c.pages)
    assert oracle.rows(dev) == oracle.rows(oracle.execute(plan))
