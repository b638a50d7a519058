"""Long VARCHAR result columns (more than 65536 rows) are planned by jumps (every row's next page
start found in parallel, squared into jumps of 32 and 1024 pages, the chain from row 0 walked
and expanded): the pages must still be byte for byte the oracle writer's pages of the same rows.
Cases: random lengths, one fixed length, long strings that take pages of their own (also at
4096-row block edges), mostly-NULL rows, all-NULL stretches longer than a page, and several such
columns (with an INT32 column) in one result."""
import zlib

import numpy as np
import pytest

import hjtest

pytestmark = pytest.mark.gpu

SEG = 4096


def _strings(rng, lens, valid):
    out = []
    for L, ok in zip(lens, valid):
        out.append(bytes(rng.integers(33, 127, size=int(L), dtype=np.uint8)) if ok else None)
    return out


def _case(name, rng):
    if name == "random_short":
        n = 300_000
        return rng.integers(0, 70, n), rng.random(n) >= 0.05
    if name == "fixed_length":
        n = 140_000
        return np.full(n, 40), np.ones(n, bool)
    if name == "long_mix":
        n = 100_000
        lens = np.where(rng.random(n) < 0.01, rng.integers(8186, 30000, n), rng.integers(0, 120, n))
        return lens, rng.random(n) >= 0.1
    if name == "boundaries":
        n = 20 * SEG + 123
        lens = rng.integers(0, 60, n)
        valid = np.ones(n, bool)
        for k in range(1, 20):  # a long string, an empty string and a NULL at block edges
            lens[k * SEG] = 9000 if k % 2 else 0
            valid[k * SEG + 1] = False
            lens[k * SEG - 1] = 20000
        return lens, valid
    if name == "mostly_null":
        n = 600_000
        return rng.integers(0, 50, n), rng.random(n) >= 0.97
    if name == "null_stretches":
        n = 400_000
        valid = rng.random(n) >= 0.1
        valid[50_000:200_000] = False  # all-NULL pages across several segments
        return rng.integers(1, 30, n), valid
    raise KeyError(name)


@pytest.mark.parametrize("name", ["random_short", "fixed_length", "long_mix", "boundaries",
                                  "mostly_null", "null_stretches"])
def test_long_varchar_pages_match_oracle_writer(gpu_ctx, oracle, name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    lens, valid = _case(name, rng)
    n = len(lens)
    col = oracle.codec.encode_strings(_strings(rng, lens, valid))
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(n, [col]))
    plan.new_scan_node(0, [(0, hjtest.DataType.VARCHAR)])
    plan.root = 0
    print("long column pages", name, n, col.num_pages, flush=True)
    res = gpu_ctx.execute(plan)
    assert res.num_rows == n
    assert res.columns[0].num_pages == col.num_pages
    assert np.array_equal(res.columns[0].pages, col.pages)


def test_long_columns_together(gpu_ctx, oracle):
    rng = np.random.default_rng(31)
    n = 150_000
    cols = [oracle.codec.encode_strings(_strings(rng, rng.integers(0, 90, n), rng.random(n) >= 0.2)),
            oracle.codec.encode_int32(rng.integers(0, 1000, n).astype(np.int32),
                                      (rng.random(n) >= 0.1).astype(np.uint8)),
            oracle.codec.encode_strings(_strings(rng, np.full(n, 7), np.ones(n, bool))),
            oracle.codec.encode_strings(_strings(rng, rng.integers(0, 3000, n), rng.random(n) >= 0.5))]
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(n, cols))
    plan.new_scan_node(0, [(j, c.type) for j, c in enumerate(cols)])
    plan.root = 0
    res = gpu_ctx.execute(plan)
    for j, c in enumerate(cols):
        assert res.columns[j].num_pages == c.num_pages, j
        assert np.array_equal(res.columns[j].pages, c.pages), j
