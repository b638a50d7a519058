"""Engine (HIP, gfx950) against the oracle: sorted multisets of output rows must be identical.

Every test here runs the HIP path through the C ABI (libhjoin.so); results are decoded with the
oracle's page reader so the engine's own reader is not trusted for the comparison.
"""
import numpy as np
import pytest

import hjtest

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["pipeline", "pipeline_cas", "join_at_a_time", "jat_lookup_twice"])
def mode(request, gpu_ctx):
    """run each test with the fused probe pipeline (tables built segment by segment in LDS, or
    by compare-and-swap inserts into cleared tables) and with join-at-a-time (its write pass
    reading the count pass's stored lookups, or looking every key up again)"""
    jat = request.param in ("join_at_a_time", "jat_lookup_twice")
    gpu_ctx.set_option("pipeline", 0 if jat else 1)
    gpu_ctx.set_option("partitioned_build", 0 if request.param == "pipeline_cas" else 1)
    gpu_ctx.set_option("match_stash", 0 if request.param == "jat_lookup_twice" else 1)
    yield request.param
    gpu_ctx.set_option("pipeline", 1)
    gpu_ctx.set_option("partitioned_build", 1)
    gpu_ctx.set_option("match_stash", 1)


@pytest.mark.parametrize("case", hjtest.load_kat_cases(), ids=lambda c: c["name"])
def test_reference_unit_cases_on_gpu(gpu_ctx, oracle, case, mode):
    plan = hjtest.plan_from_case(case, oracle.codec)
    res = gpu_ctx.execute(plan)
    assert res.num_rows == len(case["expected"])
    assert [int(c.type) for c in res.columns] == [int(hjtest.DataType[t]) for t in case["types"]]
    assert oracle.rows(res) == hjtest.expected_rows(case)


@pytest.mark.parametrize("seed", range(30))
def test_random_plans_vs_oracle(gpu_ctx, oracle, seed, mode):
    rng = np.random.default_rng(2000 + seed)
    rows, dom = [(30, 5), (300, 60), (3000, 4000)][seed % 3]
    plan = hjtest.random_plan(rng, oracle.codec, max_rows=rows, key_domain=dom,
                              long_every=int(rng.choice([0, 0, 211])))
    want = oracle.rows(oracle.execute(plan))
    got = oracle.rows(gpu_ctx.execute(plan))
    assert got == want


def test_root_scan_with_filter(gpu_ctx, oracle):
    rng = np.random.default_rng(4)
    t = hjtest.random_table(rng, 5000, 2, 100, null_p=0.2, with_str=True)
    plan = hjtest.Plan()
    plan.inputs.append(t.to_columnar(oracle.codec))
    plan.new_scan_node(0, [(2, hjtest.DataType.VARCHAR), (0, hjtest.DataType.INT32)],
                       filters=[(1, 20, 60)])
    plan.root = 0
    assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan))


def test_root_scan_plain(gpu_ctx, oracle):
    rng = np.random.default_rng(5)
    t = hjtest.random_table(rng, 7000, 1, 100, null_p=0.3, with_str=True, long_every=1500)
    plan = hjtest.Plan()
    plan.inputs.append(t.to_columnar(oracle.codec))
    plan.new_scan_node(0, [(1, hjtest.DataType.VARCHAR), (0, hjtest.DataType.INT32)])
    plan.root = 0
    assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan))


def test_non_int32_key_is_an_error(gpu_ctx, hjo, oracle):
    t = hjtest.Table([[1, "a"], [2, "b"]], [hjtest.DataType.INT32, hjtest.DataType.VARCHAR])
    plan = hjtest.Plan()
    plan.inputs += [t.to_columnar(oracle.codec), t.to_columnar(oracle.codec)]
    plan.new_scan_node(0, [(1, hjtest.DataType.VARCHAR)])
    plan.new_scan_node(1, [(1, hjtest.DataType.VARCHAR)])
    plan.new_join_node(True, 0, 1, 0, 0, [(0, hjtest.DataType.VARCHAR)])
    plan.root = 2
    with pytest.raises(hjo.EngineError) as e:
        gpu_ctx.execute(plan)
    assert e.value.code == hjo.abi.HJ_E_TYPE
    assert "Join key isn't of type int32_t" in e.value.msg
    # the context stays usable after a failed query
    case = hjtest.load_kat_cases()[2]
    p2 = hjtest.plan_from_case(case, oracle.codec)
    assert oracle.rows(gpu_ctx.execute(p2)) == hjtest.expected_rows(case)


def two_way(codec, bk, bp, bv, pk, pp, pv):
    I = hjtest.DataType.INT32
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(len(bk), [codec.encode_int32(bk, bv),
                                                      codec.encode_int32(bp)]))
    plan.inputs.append(hjtest.ColumnarTable(len(pk), [codec.encode_int32(pk, pv),
                                                      codec.encode_int32(pp)]))
    plan.new_scan_node(0, [(0, I), (1, I)])
    plan.new_scan_node(1, [(0, I), (1, I)])
    plan.new_join_node(True, 0, 1, 0, 0, [(1, I), (3, I)])
    plan.root = 2
    return plan


def pairs_of(codec, res):
    a, oka = codec.decode(res.columns[0])
    b, okb = codec.decode(res.columns[1])
    assert oka.all() and okb.all()
    p = np.stack([a.astype(np.int64), b.astype(np.int64)], 1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def expected_pairs(bk, bp, bv, pk, pp, pv):
    """numpy equi-join (sort + searchsorted), independent of the engine and the oracle"""
    bk, bp = bk[bv == 1], bp[bv == 1]
    pk, pp = pk[pv == 1], pp[pv == 1]
    o = np.argsort(bk, kind="stable")
    sk, sp = bk[o], bp[o]
    lo = np.searchsorted(sk, pk, "left")
    hi = np.searchsorted(sk, pk, "right")
    cnt = hi - lo
    prow = np.repeat(np.arange(len(pk)), cnt)
    start = np.repeat(lo - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
    bidx = start + np.arange(len(prow))
    p = np.stack([sp[bidx].astype(np.int64), pp[prow].astype(np.int64)], 1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def test_resident_int32_mirror(gpu_ctx, oracle):
    """the row-space INT32 mirror decoded at upload (values + validity) over 1M rows with 5 %
    NULLs, runs of NULLs and an all-NULL stretch, read back through a root scan's gathers"""
    rng = np.random.default_rng(17)
    n = 1_000_000
    v = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    ok = (rng.random(n) >= 0.05).astype(np.uint8)
    ok[300_000:420_000] = 0                       # all-NULL pages
    ok[700_000:700_100] = 0
    I = hjtest.DataType.INT32
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(n, [oracle.codec.encode_int32(v, ok)]))
    plan.new_scan_node(0, [(0, I)])
    plan.root = 0
    dt = gpu_ctx.upload(plan.inputs[0])
    try:
        for _ in range(2):
            r = gpu_ctx.execute_device(plan, [dt])
            res = r.to_host()
            r.free()
            got, gok = oracle.codec.decode(res.columns[0])
            assert np.array_equal(gok.astype(np.uint8), ok)
            assert np.array_equal(got[ok == 1], v[ok == 1])
    finally:
        dt.free()


def test_extreme_keys(gpu_ctx, oracle, mode):
    """INT32_MIN (the pipeline table's free-slot marker, kept aside), INT32_MAX, 0 and -1 as
    unique build keys and as probe keys; then INT32_MIN twice in the build (duplicate key)"""
    rng = np.random.default_rng(5)
    ext = np.array([-2**31, 2**31 - 1, 0, -1], np.int32)
    bk = np.concatenate([ext, rng.choice(np.arange(1, 10**6), 5000, replace=False)]).astype(np.int32)
    bk = bk[rng.permutation(len(bk))]
    bp = rng.integers(0, 2**31, len(bk)).astype(np.int32)
    bv = np.ones(len(bk), np.uint8)
    pk = np.concatenate([np.repeat(ext, 3), rng.choice(bk, 20000)]).astype(np.int32)
    pp = rng.integers(0, 2**31, len(pk)).astype(np.int32)
    pv = (rng.random(len(pk)) > 0.1).astype(np.uint8)
    plan = two_way(oracle.codec, bk, bp, bv, pk, pp, pv)
    assert np.array_equal(pairs_of(oracle.codec, gpu_ctx.execute(plan)),
                          expected_pairs(bk, bp, bv, pk, pp, pv))
    bk2 = np.concatenate([bk, [-2**31]]).astype(np.int32)
    bp2 = np.concatenate([bp, [12345]]).astype(np.int32)
    bv2 = np.ones(len(bk2), np.uint8)
    plan = two_way(oracle.codec, bk2, bp2, bv2, pk, pp, pv)
    assert np.array_equal(pairs_of(oracle.codec, gpu_ctx.execute(plan)),
                          expected_pairs(bk2, bp2, bv2, pk, pp, pv))


def test_two_way_unique_build_1m(gpu_ctx, oracle, mode):
    """the shape of the 2-way config at 1/10 size: unique shuffled build keys, 5 % NULL FKs"""
    rng = np.random.default_rng(42)
    nb, np_ = 100_000, 1_000_000
    bk = rng.permutation(nb).astype(np.int32)
    bp = rng.integers(0, 2**31, nb).astype(np.int32)
    pk = rng.integers(0, 2 * nb, np_).astype(np.int32)
    pv = (rng.random(np_) >= 0.05).astype(np.uint8)
    pp = rng.integers(0, 2**31, np_).astype(np.int32)
    bv = np.ones(nb, np.uint8)
    plan = two_way(oracle.codec, bk, bp, bv, pk, pp, pv)
    got = pairs_of(oracle.codec, gpu_ctx.execute(plan))
    assert np.array_equal(got, expected_pairs(bk, bp, bv, pk, pp, pv))


def test_many_to_many_heavy_keys(gpu_ctx, oracle, mode):
    """a heavy key (> 64K matches per probe row) goes through the chunked heavy expansion; long
    slots go through the binary search; result = numpy join"""
    rng = np.random.default_rng(8)
    bk = np.concatenate([np.full(150_000, 7), rng.integers(0, 50_000, 100_000),
                         np.full(3000, 123456)]).astype(np.int32)
    bk = bk[rng.permutation(len(bk))]
    bp = rng.integers(0, 2**31, len(bk)).astype(np.int32)
    bv = (rng.random(len(bk)) > 0.02).astype(np.uint8)
    pk = np.concatenate([[7, 7, 123456], rng.integers(0, 60_000, 40_000)]).astype(np.int32)
    pp = rng.integers(0, 2**31, len(pk)).astype(np.int32)
    pv = np.ones(len(pk), np.uint8)
    plan = two_way(oracle.codec, bk, bp, bv, pk, pp, pv)
    want = expected_pairs(bk, bp, bv, pk, pp, pv)
    got = pairs_of(oracle.codec, gpu_ctx.execute(plan))
    assert len(got) == len(want)
    assert np.array_equal(got, want)


# ------------------------------------------------------------------ directory (unchained_tests)
def test_directory_reference_cases(gpu_ctx, hjo):
    """tests/unchained_tests.cpp cases with integer keys; payload = build row id"""
    d = hjo.Directory(gpu_ctx, [])
    c, f = d.probe([42, 1])
    assert c.tolist() == [0, 0]                       # :10, :19 empty
    d.free()
    d = hjo.Directory(gpu_ctx, [7])
    c, f = d.probe([7, 8])
    assert c.tolist() == [1, 0] and f[0] == 0         # :38, :52, :63
    d.free()
    d = hjo.Directory(gpu_ctx, np.arange(10000))
    keys = [0, 1, 2, 17, 999, 4096, 9999]
    c, f = d.probe(keys)
    assert c.tolist() == [1] * 7 and f.tolist() == keys   # :74
    d.free()
    d = hjo.Directory(gpu_ctx, [5, 5, 5])
    pr = d.probe_pairs([5])
    assert sorted(pr[:, 1].tolist()) == [0, 1, 2]     # :93 duplicates
    d.free()
    for n in (1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17):    # :159 power-of-two boundaries
        d = hjo.Directory(gpu_ctx, np.arange(n))
        assert d.info()["num_slots"] == 1 << (n - 1).bit_length()
        c, f = d.probe([0, n // 2, n - 1])
        assert c.tolist() == [1, 1, 1] and f.tolist() == [0, n // 2, n - 1]
        d.free()


def test_directory_clustered_keys(gpu_ctx, hjo):
    """unchained_tests.cpp:200-225 analogue: keys sharing high bits still resolve exactly"""
    keys = (0x12340000 + np.arange(2000)).astype(np.uint32)
    d = hjo.Directory(gpu_ctx, keys)
    c, f = d.probe(keys[[0, 1, 2, 123, 999, 1999]])
    assert c.tolist() == [1] * 6 and f.tolist() == [0, 1, 2, 123, 999, 1999]
    c, _ = d.probe([0x12340000 + 777777])
    assert c.tolist() == [0]
    d.free()


@pytest.mark.parametrize("seed", range(4))
def test_directory_random_vs_oracle_pdh(gpu_ctx, hjo, oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 200_000))
    keys = rng.integers(0, max(1, n // int(rng.choice([1, 3, 50]))), n).astype(np.uint32)
    probes = rng.integers(0, n + 10, 50_000).astype(np.uint32)
    d = hjo.Directory(gpu_ctx, keys)
    got = d.probe_pairs(probes)
    d.free()
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    ones_b, ones_p = np.ones(n, np.uint8), np.ones(len(probes), np.uint8)
    cap = 4 * (len(got) + 16)
    buf = np.zeros(2 * cap, np.uint64)
    m = oracle.lib.oracle_pdh_join(keys.ctypes.data, ones_b.ctypes.data, n, probes.ctypes.data,
                                   ones_p.ctypes.data, len(probes), 4, buf.ctypes.data, cap)
    want = buf[:2 * m].reshape(-1, 2)
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", range(6))
def test_left_deep_chains_vs_oracle(gpu_ctx, oracle, seed, mode):
    """left-deep chains with base-scan build sides (the fused pipeline's shape): unique or
    duplicated build keys (duplicates make the pipeline fall back), filters, snowflake keys"""
    rng = np.random.default_rng(3000 + seed)
    I, S = hjtest.DataType.INT32, hjtest.DataType.VARCHAR
    nfact = int(rng.integers(0, 20000))
    ndims = int(rng.integers(1, 5))
    plan = hjtest.Plan()
    dims = []
    for d in range(ndims):
        n = int(rng.integers(1, 3000))
        keys = rng.permutation(n) if seed % 3 else rng.integers(0, n, n)  # duplicates when seed%3==0
        attr = rng.integers(0, 100, n)
        s = hjtest.random_strings(rng, n, 0, 10)
        t = hjtest.Table([[int(k), int(a), x] for k, a, x in zip(keys, attr, s)], [I, I, S])
        dims.append(n)
        plan.inputs.append(t.to_columnar(oracle.codec))
    fk = [rng.integers(0, dims[d] + 5, nfact) for d in range(ndims)]
    nul = [rng.random(nfact) < 0.1 for d in range(ndims)]
    fact = hjtest.Table([[None if nul[d][i] else int(fk[d][i]) for d in range(ndims)]
                         for i in range(nfact)], [I] * ndims)
    plan.inputs.append(fact.to_columnar(oracle.codec))
    ftab = ndims
    cur = plan.new_scan_node(ftab, [(d, I) for d in range(ndims)])
    cur_attrs = [("f", d) for d in range(ndims)]
    for d in range(ndims):
        flt = [(1, 10, 10 + int(rng.integers(5, 90)))] if rng.random() < 0.5 else []
        sc = plan.new_scan_node(d, [(0, I), (2, S), (1, I)], flt)
        bl = bool(rng.random() < 0.5)
        pkey = cur_attrs.index(("f", d))
        if bl:
            left_attrs, right_attrs = [("d", d, 0), ("d", d, 2), ("d", d, 1)], cur_attrs
            la, ra = 0, pkey
        else:
            left_attrs, right_attrs = cur_attrs, [("d", d, 0), ("d", d, 2), ("d", d, 1)]
            la, ra = pkey, 0
        allattrs = left_attrs + right_attrs
        keep = [i for i, a in enumerate(allattrs) if a[0] == "f" or rng.random() < 0.5]
        ty = [S if a[0] == "d" and a[2] == 2 else I for a in allattrs]
        cur = plan.new_join_node(bl, sc if bl else cur, cur if bl else sc, la, ra,
                                 [(i, ty[i]) for i in keep])
        cur_attrs = [allattrs[i] for i in keep]
    plan.root = cur
    want = oracle.rows(oracle.execute(plan))
    assert oracle.rows(gpu_ctx.execute(plan)) == want


def test_star4_reduced_scale_vs_numpy(gpu_ctx, hjo, mode):
    """the 4-way star configuration at 1/20 scale: output rows equal numpy lookups exactly"""
    dg = hjtest.import_module("hash-join-optimizer_amd.datagen")
    d = dg.star4(seed=43, n_fact=1_800_000, dims=(125_000, 200_000, 5_000),
                 sel=(0.3, 0.3, 0.5))
    codec = hjo.Codec()
    plan = dg.star4_plan(d, dg.star4_tables(d, codec))
    got = hjtest.sorted_rows(hjtest.Table.from_columnar(gpu_ctx.execute(plan), codec))
    rows, drows = dg.star4_expected(d)
    exp = []
    for i, fr in enumerate(rows.tolist()):
        r = [int(d["fact"][j][fr]) for j in range(3)]
        for j, dm in enumerate(d["dims"]):
            rr = int(drows[j][i])
            r.append(bytes(dm["schars"][int(dm["soff"][rr]):int(dm["soff"][rr + 1])]).decode())
        exp.append(tuple(r))
    assert len(got) == len(exp) > 1000
    assert got == hjtest.sorted_rows(hjtest.Table(exp, [0, 0, 0, 3, 3, 3]))


def test_cache_file_upload(gpu_ctx, hjo, oracle, tmp_path):
    # Table::from_cache straight into HBM: the same rows as uploading the ColumnarTables
    rng = np.random.default_rng(31)
    plan = hjtest.random_plan(rng, oracle.codec, ntables=3, max_rows=3000, key_domain=400,
                              long_every=97)
    paths = []
    for i, t in enumerate(plan.inputs):
        p = tmp_path / f"t{i}.tbl"
        hjo.write_cache(t, p)
        paths.append(p)
    dts = [gpu_ctx.upload_cache(p) for p in paths]
    try:
        r = gpu_ctx.execute_device(plan, dts)
        got = oracle.rows(r.to_host())
        r.free()
    finally:
        for d in dts:
            d.free()
    assert got == oracle.rows(oracle.execute(plan))


def test_cache_file_upload_many_pages(gpu_ctx, hjo, oracle, tmp_path):
    # 9001 pages: three batches through the pinned double buffer; a root scan returns them
    n = 1984 * 9000 + 11
    v = np.arange(n, dtype=np.int64).astype(np.int32)
    col = oracle.codec.encode_int32(v)
    t = hjtest.ColumnarTable(n, [col])
    p = tmp_path / "big.tbl"
    hjo.write_cache(t, p)
    dt = gpu_ctx.upload_cache(p)
    plan = hjtest.Plan()
    plan.inputs.append(t)
    plan.new_scan_node(0, [(0, hjtest.DataType.INT32)])
    plan.root = 0
    try:
        r = gpu_ctx.execute_device(plan, [dt])
        res = r.to_host()
        r.free()
    finally:
        dt.free()
    assert res.num_rows == n
    assert np.array_equal(res.columns[0].pages, col.pages)


@pytest.mark.parametrize("nulls", [True, False])
def test_star_join_takes_the_fused_pipeline(gpu_ctx, oracle, nulls):
    """a star join over pages the writer cut (with and without NULL keys) runs as the fused
    probe pipeline (its kernels appear in the engine's kernel timing), not join at a time"""
    import importlib
    dg = importlib.import_module("hash-join-optimizer_amd.datagen")
    d = dg.star4(seed=44, n_fact=200_000, dims=(20_000, 30_000, 3_000), sel=(0.3, 0.3, 0.3),
                 null_p=0.03 if nulls else 0.0)
    plan = dg.star4_plan(d, dg.star4_tables(d, oracle.codec))
    gpu_ctx.profile("")
    try:
        got = oracle.rows(gpu_ctx.execute(plan))
        prof = gpu_ctx.profile_report()
    finally:
        gpu_ctx.profile(None)
    assert "probe_fused" in prof and "pipe_build" in prof, sorted(prof)
    assert "probe_scan_write" not in prof
    assert got == oracle.rows(oracle.execute(plan))
