"""Inputs whose pages are cut where the reference's writer would not cut them (include/plan.h
allows any row count per page): the engine reads every row straight from the pages, through
the query's 64-row group index where every page but the last holds >= 64 rows, by arithmetic
for dense pages of one row count, and otherwise through a decoded row-space array (INT32) or a
page search (VARCHAR). Each layout is joined in both engine modes and compared with the
oracle's execute() on the same rows re-paged by the reference's writer rules: the reference reads
a column with no NULL in place assuming 1984 rows per page (include/column_t.h:14,48-58), so on
dense pages of other sizes its execute() would read other rows; the rows the pages hold are the
meaning of the input, and that is what the engine must join."""
import numpy as np
import pytest

import hjtest

pytestmark = pytest.mark.gpu
I32, STR = hjtest.DataType.INT32, hjtest.DataType.VARCHAR


def layout_plan(rng, layout, n_build=3000, n_probe=40000):
    """build (unique key, INT32 payload, VARCHAR) JOIN probe (FK with NULLs, INT32 payload with
    NULLs, VARCHAR with NULLs), both cut into pages by `layout`"""
    def sizes(n):
        if layout == "tiny":       # pages of 1 .. 200 rows: some below 64
            return hjtest.split_sizes(rng, n, 1, 200)
        if layout == "regular":    # 64 .. 300 rows per page, varying
            return hjtest.split_sizes(rng, n, 64, 300)
        u = {"uniform64": 64, "uniform100": 100, "uniform1000": 1000}[layout]
        return [u] * (n // u) + ([n % u] if n % u else [])
    dense = layout.startswith("uniform")

    def str_sizes(n):  # strings of up to 20 bytes: at most 300 rows fit a page
        return hjtest.split_sizes(rng, n, 64, 300) if dense or layout == "regular" else sizes(n)
    bk = rng.permutation(n_build).astype(np.int32)
    bp = rng.integers(-2**31, 2**31, n_build).astype(np.int32)
    bs = ["b%d" % x for x in rng.integers(0, 10**6, n_build)]
    pk = rng.integers(0, 2 * n_build, n_probe).astype(np.int32)
    pv = np.ones(n_probe, np.uint8) if dense else (rng.random(n_probe) > 0.2).astype(np.uint8)
    pp = rng.integers(-2**31, 2**31, n_probe).astype(np.int32)
    ppv = np.ones(n_probe, np.uint8) if dense else (rng.random(n_probe) > 0.3).astype(np.uint8)
    ps = [None if (not dense and rng.random() < 0.25) else "p" * int(rng.integers(0, 20))
          for _ in range(n_probe)]
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(n_build, [
        hjtest.int32_pages(bk, np.ones(n_build, np.uint8), sizes(n_build)),
        hjtest.int32_pages(bp, np.ones(n_build, np.uint8), sizes(n_build)),
        hjtest.varchar_pages(bs, str_sizes(n_build))]))
    plan.inputs.append(hjtest.ColumnarTable(n_probe, [
        hjtest.int32_pages(pk, pv, sizes(n_probe)),
        hjtest.int32_pages(pp, ppv, sizes(n_probe)),
        hjtest.varchar_pages(ps, str_sizes(n_probe))]))
    plan.new_scan_node(0, [(0, I32), (1, I32), (2, STR)])
    plan.new_scan_node(1, [(0, I32), (1, I32), (2, STR)])
    plan.new_join_node(False, 1, 0, 0, 0, [(1, I32), (2, STR), (4, I32), (5, STR), (0, I32)])
    plan.root = 2
    return plan


def writer_pages(plan, codec):
    """the plan over its inputs' rows re-encoded with the writer rules (src/build_table.cpp)"""
    q = hjtest.Plan()
    q.inputs = [hjtest.Table.from_columnar(t, codec).to_columnar(codec) for t in plan.inputs]
    q.nodes = plan.nodes
    q.root = plan.root
    return q


@pytest.mark.parametrize("pipeline", [1, 0])
@pytest.mark.parametrize("layout", ["tiny", "regular", "uniform64", "uniform100", "uniform1000"])
def test_page_layouts_vs_oracle(gpu_ctx, oracle, layout, pipeline):
    rng = np.random.default_rng(sum(map(ord, layout)))
    plan = layout_plan(rng, layout)
    want = oracle.rows(oracle.execute(writer_pages(plan, oracle.codec)))
    assert len(want) > 1000
    gpu_ctx.set_option("pipeline", pipeline)
    try:
        got = oracle.rows(gpu_ctx.execute(plan))
    finally:
        gpu_ctx.set_option("pipeline", 1)
    assert got == want


@pytest.mark.parametrize("layout", ["regular", "uniform64", "uniform100", "uniform1000"])
def test_page_layouts_probe_index_without_tables(gpu_ctx, oracle, layout):
    """the fused probe's blocks index their source rows through LDS page tables; with the
    tables off (option index_tables 0) every group is indexed by its own page search, the
    path a block takes when its rows span more pages than a table holds"""
    rng = np.random.default_rng(sum(map(ord, layout)) + 7)
    plan = layout_plan(rng, layout)
    want = oracle.rows(oracle.execute(writer_pages(plan, oracle.codec)))
    for tables in (0, 1):
        gpu_ctx.set_option("index_tables", tables)
        try:
            got = oracle.rows(gpu_ctx.execute(plan))
        finally:
            gpu_ctx.set_option("index_tables", 1)
        assert got == want, tables


@pytest.mark.parametrize("layout", ["tiny", "regular", "uniform100"])
def test_page_layouts_root_scan_filtered(gpu_ctx, oracle, layout):
    """a filtered root scan reads its filter column and outputs through the same access paths"""
    rng = np.random.default_rng(7 + len(layout))
    plan = layout_plan(rng, layout)
    p = hjtest.Plan()
    p.inputs = plan.inputs
    p.new_scan_node(1, [(2, STR), (1, I32), (0, I32)], filters=[(0, 100, 2500)])
    p.root = 0
    want = oracle.rows(oracle.execute(writer_pages(p, oracle.codec)))
    assert len(want) > 1000
    assert oracle.rows(gpu_ctx.execute(p)) == want
