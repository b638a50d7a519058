"""One-join pipelines over tables of more than 2^18 build rows run the sliced probe (a partition
pass orders each 4096-row block's live (key, row) pairs by the top 3 bits of the key hash; probe
block c * 8 + x, dispatched to XCD x, looks up slice x of a chunk of blocks): the same rows as
with the option off and as the oracle's, with NULL keys, INT32_MIN / INT32_MAX keys, a source
filter, VARCHAR output and the build on either side. Where the join's only build-side output is
an INT32 column without NULL, the table carries its values in place of rows (payload tables):
also checked with that column holding NULLs (rows gathered as usual), with the build key as the
output, and with the directory cache on (payload tables are not published to it). Likewise the
partition pass carries the source's only output, an INT32 column without NULL, with the key
(checked with and without NULLs in that column)."""
import numpy as np
import pytest

import hjtest

pytestmark = pytest.mark.gpu
I32, STR = hjtest.DataType.INT32, hjtest.DataType.VARCHAR


def _plan(oracle, seed, nb, npr, build_left, src_filter, strings, bnull=0.0, out_key=False,
          pnull=0.0):
    rng = np.random.default_rng(seed)
    ext = np.array([-2**31, 2**31 - 1, 0, -1], np.int32)
    body = rng.choice(np.arange(1, 4 * nb, dtype=np.int64), nb - 4, replace=False)
    bk = np.concatenate([ext, body.astype(np.int32)])
    bk = bk[rng.permutation(nb)]
    dom = np.concatenate([bk, rng.integers(4 * nb, 8 * nb, nb).astype(np.int32)])
    pk = dom[rng.integers(0, len(dom), npr)]
    pv = (rng.random(npr) > 0.1).astype(np.uint8)
    pp = rng.integers(0, 1000, npr).astype(np.int32)
    bval = None if not bnull else (rng.random(nb) > bnull).astype(np.uint8)
    bcols = [oracle.codec.encode_int32(bk),
             oracle.codec.encode_int32(np.arange(nb, dtype=np.int32), bval)]
    btypes = [(0, I32), (1, I32)]
    if strings:
        words = [f"s{i % 977}-{'x' * (i % 23)}" for i in range(nb)]
        bcols.append(oracle.codec.encode_strings(words))
        btypes.append((2, STR))
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(nb, bcols))
    ppv = None if not pnull else (rng.random(npr) > pnull).astype(np.uint8)
    plan.inputs.append(hjtest.ColumnarTable(npr, [oracle.codec.encode_int32(pk, pv),
                                                  oracle.codec.encode_int32(pp, ppv)]))
    b = plan.new_scan_node(0, btypes)
    p = plan.new_scan_node(1, [(0, I32), (1, I32)], [(1, 0, 899)] if src_filter else ())
    nbc = len(btypes)
    bo = 0 if out_key else 1  # the build-side INT32 output: its key or its row column
    if build_left:  # output: probe payload, build row, (build string)
        out = [(nbc + 1, I32), (bo, I32)] + ([(2, STR)] if strings else [])
        plan.root = plan.new_join_node(True, b, p, 0, 0, out)
    else:
        out = [(1, I32), (2 + bo, I32)] + ([(4, STR)] if strings else [])
        plan.root = plan.new_join_node(False, p, b, 0, 0, out)
    return plan


@pytest.mark.parametrize("case", [
    dict(seed=1, nb=300_000, npr=1_500_000, build_left=True, src_filter=False, strings=False),
    dict(seed=2, nb=600_000, npr=3_000_000, build_left=False, src_filter=True, strings=False),
    dict(seed=3, nb=400_000, npr=1_000_000, build_left=True, src_filter=True, strings=True),
    # 512 partition blocks of 2048 rows: 256 chunks of 2, so 2048 probe blocks, exactly as
    # many as the output pass's prefix sums cover
    dict(seed=4, nb=300_000, npr=1 << 20, build_left=False, src_filter=False, strings=True),
    dict(seed=5, nb=350_000, npr=1_200_000, build_left=True, src_filter=False, strings=False,
         bnull=0.2),
    dict(seed=6, nb=350_000, npr=1_200_000, build_left=False, src_filter=True, strings=False,
         out_key=True),
    dict(seed=8, nb=350_000, npr=1_200_000, build_left=False, src_filter=False, strings=False,
         pnull=0.3),
], ids=["plain", "filtered-probe-left", "strings", "output-block-cap", "payload-null",
        "key-output", "probe-payload-null"])
def test_xcd_slices_match_oracle(gpu_ctx, oracle, case):
    plan = _plan(oracle, **case)
    want = oracle.rows(oracle.execute(plan))
    assert len(want) > 100_000
    print("xcd slices case", case, flush=True)
    for on in (1, 0):
        gpu_ctx.set_option("xcd_slices", on)
        try:
            got = oracle.rows(gpu_ctx.execute(plan))
        finally:
            gpu_ctx.set_option("xcd_slices", 1)
        assert got == want, on


def test_payload_tables_with_directory_cache(hjo, oracle):
    """with the directory cache on, a payload table is not published: the same join twice, then
    one whose build output is a string (row words needed), all equal to the oracle"""
    plain = _plan(oracle, seed=7, nb=300_000, npr=1_000_000, build_left=True, src_filter=False,
                  strings=False)
    strs = _plan(oracle, seed=7, nb=300_000, npr=1_000_000, build_left=True, src_filter=False,
                 strings=True)
    want_plain = oracle.rows(oracle.execute(plain))
    want_strs = oracle.rows(oracle.execute(strs))
    with hjo.Context(arena_bytes=4 << 30) as ctx:
        ctx.set_option("dir_cache_bytes", 1 << 30)
        for plan, want in ((plain, want_plain), (plain, want_plain), (strs, want_strs),
                           (plain, want_plain)):
            assert oracle.rows(ctx.execute(plan)) == want
