"""Benchmark: one step = one execute of the 4-way star join (fact 36M rows x 3 Zipf(1.1) FKs,
dimensions 2.5M / 4M / 100K rows with range filters pushed into their scans, VARCHAR output)
on one MI355X, inputs resident in HBM, output materialized in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload star4|two_way|m2m|batch|serve|job]

Prints one JSON line. value = join tuples/s = sum over joins of (|build_j| + |probe_j|) / step
time; pct_hbm_peak = algorithmic bytes / step time / 8 TB/s (keys and filter attributes read once,
validity bits, output columns gathered and written once). The roofline object is the dominant
kernel timed with HIP events on the engine's stream over as many steps again (the timed steps
run without events). execute_ms is the drop-in hj_execute end to end (host pages uploaded,
result pages copied out), median of 3. cpu_baseline is the
oracle (a CPU restatement of the reference's execute()) on the host, timed around execute() with
pre-filtered dimensions, as the reference filters before execute(). result_pages times the copy of
the result into the reference's page format (outside the timed steps): pages written on the
device against flat columns paged on the host.
"""
import argparse
import ctypes as C
import importlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
hjo = importlib.import_module("hash-join-optimizer_amd")
dg = importlib.import_module("hash-join-optimizer_amd.datagen")

HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ workloads
def star4_workload(codec, scale):
    n_fact = int(36_000_000 * scale)
    dims = tuple(max(1000, int(x * scale)) for x in dg.STAR_DIMS)
    t = time.time()
    d = dg.star4(seed=43, n_fact=n_fact, dims=dims)
    tables = dg.star4_tables(d, codec)
    plan = dg.star4_plan(d, tables)
    log(f"star4 generated+encoded in {time.time() - t:.1f}s")
    # per-join sizes (exact, from numpy lookups)
    n = n_fact
    ok = np.ones(n, dtype=bool)
    builds, probes, outs = [], [n], []
    strlens = []
    for fk, v, dm in zip(d["fact"], d["fact_valid"], d["dims"]):
        inv = np.empty(len(dm["key"]), dtype=np.int64)
        inv[dm["key"]] = np.arange(len(dm["key"]))
        r = inv[fk]
        passed = (dm["attr"][r] >= dm["lo"]) & (dm["attr"][r] <= dm["hi"])
        ok &= (v == 1) & passed
        builds.append(int(((dm["attr"] >= dm["lo"]) & (dm["attr"] <= dm["hi"])).sum()))
        outs.append(int(ok.sum()))
        strlens.append((r, dm))
    probes += outs[:-1]
    rows = np.nonzero(ok)[0]
    out_rows = len(rows)
    str_bytes = 0
    for r, dm in strlens:
        rr = r[rows]
        str_bytes += int((dm["soff"][rr + 1] - dm["soff"][rr]).sum()) + 2 * out_rows
    key_bytes = sum(4 * int(v.sum()) + (n + 7) // 8 for v in d["fact_valid"])  # packed values+bits
    dim_bytes = sum(8 * len(dm["key"]) for dm in d["dims"])                      # key + attr
    gathered = out_rows * 3 * 4 + str_bytes
    b_alg = key_bytes + dim_bytes + 2 * gathered
    info = dict(builds=builds, probes=probes, outs=outs, out_rows=out_rows, b_alg=b_alg,
                fk_bytes=[4 * int(v.sum()) + (n + 7) // 8 for v in d["fact_valid"]], out_slots=4)
    workload = {"workload": "4-way star join ((F x sD1) x sD2) x sD3",
                "fact_rows": n_fact, "dim_rows": list(dims), "fk_dist": "zipf(1.1)",
                "fk_null": 0.03, "filters": [[dm["lo"], dm["hi"]] for dm in d["dims"]],
                "filtered_dim_rows": builds, "output_rows": out_rows,
                "output": "3 fact INT32 FKs + 3 dimension VARCHARs", "seed": 43}
    return d, tables, plan, info, workload


def two_way_workload(codec, scale):
    d = dg.two_way(seed=42, n_build=int(1_000_000 * scale), n_probe=int(10_000_000 * scale))
    plan = dg.two_way_plan(d, codec)
    out_rows = dg.two_way_expected_count(d)
    nb, np_ = len(d["bk"]), len(d["pk"])
    b_alg = 4 * nb + 4 * int(d["pv"].sum()) + (np_ + 7) // 8 + 2 * out_rows * 8
    info = dict(builds=[nb], probes=[np_], outs=[out_rows], out_rows=out_rows, b_alg=b_alg,
                fk_bytes=[4 * int(d["pv"].sum()) + (np_ + 7) // 8], out_slots=2)
    workload = {"workload": "2-way INT32 join 1M unique x 10M FK (5% NULL)", "seed": 42,
                "output_rows": out_rows}
    return d, plan.inputs, plan, info, workload


def m2m_workload(codec, scale):
    """many-to-many skew stress: 50M build keys ~ Zipf(1.3) over [0, 5M) joined with 200M probe
    keys ~ Zipf(1.3) under an independent key permutation (10 % NULL), INT32 payloads out"""
    t = time.time()
    # seed 61: 702,588,819 output rows, 3.5x the probe rows (seeds 45-69 give 0.4x-7.0x; the
    # independent key permutations decide how the two sides' heavy keys meet)
    d = dg.many_to_many(seed=61, n_build=int(50_000_000 * scale), n_probe=int(200_000_000 * scale))
    plan = dg.two_way_plan(d, codec)
    out_rows = dg.many_to_many_expected_count(d)
    log(f"m2m generated+encoded in {time.time() - t:.1f}s, {out_rows} output rows")
    nb, np_ = len(d["bk"]), len(d["pk"])
    b_alg = 4 * nb + 4 * int(d["pv"].sum()) + (np_ + 7) // 8 + 2 * out_rows * 8
    info = dict(builds=[nb], probes=[np_], outs=[out_rows], out_rows=out_rows, b_alg=b_alg,
                fk_bytes=[4 * int(d["pv"].sum()) + (np_ + 7) // 8], out_slots=2)
    workload = {"workload": "many-to-many INT32 join 50M x 200M, keys Zipf(1.3) over [0, 5M) with "
                            "independent permutations, 10 % NULL probe keys",
                "build_rows": nb, "probe_rows": np_, "output_rows": out_rows, "seed": 61,
                "engine": "join-at-a-time (option pipeline = 0: the build keys repeat)"}
    return d, plan.inputs, plan, info, workload


def serve_workload(codec, args):
    """Serving: star joins over resident tables (datagen.serve_star / serve_queries): a fact
    table whose prefixes of 100K, 1M, 4M, 16M and 36M rows are resident, with five FK columns
    (Zipf(1.1), 3 % NULL) over the three dimensions of the 4-way configuration (2.5M / 4M / 100K
    rows); each query joins 1-5 of the FKs (2- to 6-way) with range filters of 1-20 %
    selectivity drawn from 16 per dimension (Zipf over the pool), fact prefix log-uniform in
    [100K, 36M] rows."""
    fact_sizes = [100_000, 1_000_000, 4_000_000, 16_000_000, 36_000_000]
    t = time.time()
    d = dg.serve_star(seed=47, n_fact=max(fact_sizes), dims=dg.STAR_DIMS)
    tables = dg.serve_tables(d, fact_sizes, codec)
    queries = dg.serve_queries(np.random.default_rng(49), args.queries, fact_sizes)
    log(f"serve tables generated+encoded in {time.time() - t:.1f}s")
    return tables, queries


def batch_workload(codec, args):
    """the batch of independent 2- and 3-way INT32 joins (datagen.batch_joins, seed 48): builds
    log-uniform in [1K, 4M] x scale rows, probes log-uniform in [10K, 40M] x scale"""
    t = time.time()
    plans = dg.batch_joins(seed=48, n=args.queries, scale=args.scale, codec=codec)
    log(f"batch of {len(plans)} joins generated+encoded in {time.time() - t:.1f}s")
    return plans


def batch_main(args):
    """the batch of independent joins: 256 sequential plain GPU runs (hj_execute_device, one
    query at a time) against one hj_batch_run of all of them (four graph-captured launches);
    results stay in HBM in both, and every batched result's row count is checked against the
    plain run's"""
    codec = hjo.Codec()
    plans = batch_workload(codec, args)
    ctx = hjo.Context(arena_bytes=24 << 30)
    t = time.time()
    dts = [[ctx.upload(t_) for t_ in p.inputs] for p in plans]
    log(f"uploaded in {time.time() - t:.1f}s")
    probe_rows = int(sum(p.inputs[0].num_rows for p in plans))
    build_rows = int(sum(sum(t_.num_rows for t_ in p.inputs[1:]) for p in plans))
    # plain: one query at a time (the drop-in path), results left in HBM
    plain_rows = []
    for p, d in zip(plans[:4], dts[:4]):  # warm-up
        ctx.execute_device(p, d).free()
    ctx.synchronize()
    ts = []
    for rep in range(args.reps):
        t0 = time.perf_counter()
        rows = []
        for p, d in zip(plans, dts):
            r = ctx.execute_device(p, d)
            rows.append(r.num_rows)
            r.free()
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
        plain_rows = rows
    plain_s = statistics.median(ts)
    log(f"plain: {len(plans) / plain_s:.1f} joins/s ({plain_s * 1e3:.2f} ms for {len(plans)})")
    params = "tuned" if not args.untuned else None
    t = time.time()
    with hjo.Batch(ctx, list(zip(plans, dts)), params) as b:
        log(f"batch created in {time.time() - t:.2f}s, {b.device_bytes() / 2**30:.2f} GiB, "
            f"params {b.params or 'defaults'}")
        b.run()  # graph capture and first launch
        for i in range(len(plans)):
            batched, rows = b.info(i)
            assert rows == plain_rows[i], f"query {i}: {rows} rows batched, {plain_rows[i]} plain"
        nb = sum(b.info(i)[0] for i in range(len(plans)))
        ts = []
        for rep in range(max(args.reps, 3)):
            t0 = time.perf_counter()
            b.run()
            ts.append(time.perf_counter() - t0)
        batch_s = statistics.median(ts)
        ctx.profile("batch")
        b.run()
        prof = ctx.profile_report()
        ctx.profile(None)
    log(f"batched: {len(plans) / batch_s:.1f} joins/s ({batch_s * 1e3:.2f} ms), {nb} in the launches")
    tuples = probe_rows + build_rows
    out = {"metric": "batched queries/s", "value": len(plans) / batch_s, "unit": "queries/s",
           "n_gpus": 1, "steps": max(args.reps, 3), "warmup": 1,
           "ms_per_step": batch_s * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32", "data": "synthetic, seeded",
           "config": {"workload": "batch of independent 2- and 3-way INT32 joins "
                                  "(datagen.batch_joins, seed 48): builds log-uniform [1K, 4M] x "
                                  "scale rows, probes log-uniform [10K, 40M] x scale, uniform / "
                                  "Zipf(0.8, 1.2) FKs, match rate U[0.1, 1.0], 5 % NULL FKs",
                      "queries": len(plans), "scale": args.scale, "probe_rows": probe_rows,
                      "build_rows": build_rows,
                      "output_rows": int(sum(plain_rows))},
           "batched_queries": nb,
           "plain_qps": len(plans) / plain_s, "plain_ms": plain_s * 1e3,
           "speedup_vs_plain": plain_s / batch_s,
           "join_tuples_per_s": tuples / batch_s,
           "plain_join_tuples_per_s": tuples / plain_s,
           "batch_kernel_ms": prof, "params": b.params or "defaults",
           "results": "left in HBM in both paths; batched row counts equal the plain run's"}
    for d in dts:
        for t_ in d:
            t_.free()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_served(srv, dts, queries, rate=None, seed=0):
    """submit every query (at once, or at Poisson arrival times of `rate` per second), wait for
    all; returns (wall seconds, latencies in ms)"""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(queries)) if rate else np.zeros(len(queries))
    per_query = bool(dts) and isinstance(dts[0], list)  # each query's own tables
    t0 = time.perf_counter()
    due = t0
    tickets = []
    for i, (q, g) in enumerate(zip(queries, gaps)):
        due += g
        while rate and time.perf_counter() < due:
            pass
        tickets.append(srv.submit(q, dts[i] if per_query else dts))
    lat = []
    nbytes = 0
    for tk in tickets:
        b, q_ms, r_ms = srv.wait(tk, keep=False)  # the result's pages are released unread
        lat.append(q_ms + r_ms)
        nbytes += b
    run_served.result_bytes = nbytes
    return time.perf_counter() - t0, lat


def serve_main(args):
    """the plain path (one slot, one query at a time, no directory cache) against the served
    path (args.slots slots, shared directory cache, reserved-memory admission): a burst of all
    queries at once (queries/s), then Poisson arrivals at each rate (p50 / p99 latency from
    submission to the result in host memory)"""
    codec = hjo.Codec()
    tables, queries = serve_workload(codec, args)
    ctx = hjo.Context()
    dts = [ctx.upload(t) for t in tables]
    rates = [float(r) for r in args.rates.split(",")]
    out = {"metric": "batched queries/s", "unit": "queries/s", "n_gpus": 1,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic, seeded",
           "config": {"workload": "served 2- to 6-way star joins over resident dimensions "
                                  "(2.5M / 4M / 100K rows, 1-20 % range filters from a pool of 16 "
                                  "per dimension), fact prefix 100K-36M rows log-uniform",
                      "queries": len(queries), "slots": args.slots, "poisson_rates": rates,
                      "results": "pages in host memory"}}
    for name, slots, cache in (("plain", 1, 0), ("served", args.slots, 1 << 30)):
        with hjo.Server(slots=slots, arena_bytes=16 << 30, dir_cache_bytes=cache) as srv:
            run_served(srv, dts, queries[:8])  # warm-up (and, served, the first cached tables)
            wall, lat = run_served(srv, dts, queries)
            out[f"{name}_qps"] = len(queries) / wall
            out[f"{name}_batch_p50_ms"] = float(np.percentile(lat, 50))
            out[f"{name}_batch_p99_ms"] = float(np.percentile(lat, 99))
            out["result_bytes"] = run_served.result_bytes
            for rate in rates:
                wall, lat = run_served(srv, dts, queries, rate=rate, seed=1)
                key = f"{name}_poisson_{int(rate)}"
                out[key] = {"p50_ms": float(np.percentile(lat, 50)),
                            "p99_ms": float(np.percentile(lat, 99)),
                            "qps": len(queries) / wall}
                log(f"{name} ({slots} slots) Poisson {rate:.0f}/s: p50 {out[key]['p50_ms']:.2f} ms "
                    f"p99 {out[key]['p99_ms']:.2f} ms, {out[key]['qps']:.1f} q/s")
            if cache:
                out[f"{name}_dir_cache"] = srv.cache_stats()
        log(f"{name} ({slots} slots): {out[name + '_qps']:.1f} q/s for a burst of {len(queries)}")
    out["value"] = out["served_qps"]
    out["speedup_vs_plain"] = out["served_qps"] / out["plain_qps"]
    for t_ in dts:
        t_.free()
    ctx.close()
    print(json.dumps(out), flush=True)


def job_main(args):
    """the Join Order Benchmark's 113 queries (SQL + PostgreSQL plans, tests/golden) through the
    front-end on synthetic tables of the JOB schema (datagen.job_tables, one data set for all
    queries, JOB_ROWS x scale rows): the total of the per-query times, resident tables with
    results left in HBM (hj_execute_device, the steps) and the drop-in hj_execute (host pages in,
    result pages out); filters run inside each call as the scans' predicate programs"""
    jf = importlib.import_module("hash-join-optimizer_amd.jobfront")
    jr = importlib.import_module("hash-join-optimizer_amd.jobrun")
    scale = args.scale if args.scale != 1.0 else 200.0
    schema_text, queries = jf.load_fixture(os.path.join(ROOT, "tests", "golden", "job_queries.json"))
    schema = jf.parse_schema(schema_text)
    codec = hjo.Codec()
    t = time.time()
    raw = dg.job_tables(schema_text, queries, seed=70, scale=scale)
    tables = dg.job_columnar(raw, codec)
    nrows = sum(n for n, _ in raw.values())
    log(f"JOB tables: {nrows} rows at scale {scale} generated in {time.time() - t:.1f}s")
    t0 = time.perf_counter()
    ctx = hjo.Context()
    ctx_ms = (time.perf_counter() - t0) * 1e3
    dts = {k: ctx.upload(c) for k, c in tables.items()}
    quiet = lambda m: None  # noqa: E731
    for _ in range(args.warmup):
        jr.run(ctx, queries, schema, tables, dts, log=quiet)
    totals, per_query = [], None
    for _ in range(args.steps):
        ms, _, per_query = jr.run(ctx, queries, schema, tables, dts, log=quiet)
        totals.append(ms)
    step_ms = statistics.median(totals)
    dropin = []
    for _ in range(max(1, min(3, args.steps))):
        ms, _, dq = jr.run(ctx, queries, schema, tables, host_pages=True, log=quiet)
        dropin.append(ms)
    gpu_rows = {name: rows for name, _, rows, _ in per_query}
    assert all(gpu_rows[n] == r for n, _, r, _ in dq), "hj_execute and hj_execute_device differ"
    slow = sorted(per_query, key=lambda x: -x[1])[:5]
    log(f"JOB resident: {step_ms:.2f} ms total; slowest {[(n, round(m, 3)) for n, m, _, _ in slow]}")
    log(f"JOB drop-in hj_execute: {statistics.median(dropin):.2f} ms total")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import hjtest  # test infrastructure: loads oracle/liboracle.so
        o = hjtest.Oracle()
        threads = cpu_threads()
        ns_tot = 0
        for name, sql, pg in queries:
            plan, order, _ = jf.translate(sql, pg, schema, tables)
            ct, ns = o.execute(plan, threads=threads, timed=True)
            ns_tot += ns
            assert ct.num_rows == gpu_rows[name], f"{name}: {gpu_rows[name]} rows, oracle {ct.num_rows}"
        cpu = {"value": len(queries) / (ns_tot / 1e9), "unit": "queries/s", "cores": threads,
               "kind": "port", "ms": ns_tot / 1e6,
               "sample": "all 113 queries once, sum of the oracle's execute() times (scan filters "
                         "applied before its timer, as the reference filters at load time); "
                         "every query's row count equals the GPU's",
               "host": host_cpu()}
        log(f"JOB oracle execute(): {ns_tot / 1e6:.1f} ms total on {threads} threads")
    out = {"metric": "JOB queries/s", "value": len(queries) / (step_ms / 1e3), "unit": "queries/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic JOB-schema tables (datagen.job_tables, seed 70); the queries and "
                   "plans are the benchmark's own",
           "config": {"workload": "Join Order Benchmark, 113 queries, one step = all of them "
                                  "one after another over resident tables",
                      "scale": scale, "table_rows": nrows, "queries": len(queries),
                      "output_rows": int(sum(gpu_rows.values()))},
           "total_ms_resident": step_ms, "total_ms_hj_execute": statistics.median(dropin),
           "context_ms": ctx_ms,
           "slowest_queries_ms": {n: m for n, m, _, _ in slow},
           "roofline": None, "cpu_baseline": cpu}
    for d in dts.values():
        d.free()
    ctx.close()
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ kernel byte model
def kernel_bytes(name, info, launches):
    """algorithmic HBM bytes of all launches of one step of kernel `name` (None: no model)"""
    fk1 = info["fk_bytes"][0]
    outs, probes = info["outs"], info["probes"]
    if name == "probe_fused":
        # the fused probe pass (one launch): every probe-side key column read once in full
        # (packed values + validity bits), plus the surviving rows' ids written once per slot
        return sum(info["fk_bytes"]) + info["out_rows"] * info["out_slots"] * 4
    if name == "probe_scan_count":
        return fk1
    if name == "probe_scan_write":
        return fk1 + outs[0] * 2 * 4
    if name == "probe_rel_count":
        return sum(p * 8 for p in probes[1:])          # row id + gathered key per probe row
    if name == "probe_rel_write":
        return sum(p * 8 for p in probes[1:]) + sum(o * 4 * (j + 2) for j, o in enumerate(outs[1:]))
    return None


# ------------------------------------------------------------------ measured HBM traffic
TRAFFIC_FILES = {("probe_fused", "star4"): "k_sp_probe_star4_traffic.md"}


def measured_traffic(kernel, workload_name, scale):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None when there is none"""
    f = TRAFFIC_FILES.get((kernel, workload_name))
    if f is None or scale != 1.0:
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", f)):
            if line.strip().startswith("bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


# ------------------------------------------------------------------ result pages
def result_pages(ctx, plan, dts, reps=3):
    """wall time of hj_dresult_to_host (the result's 8 KiB pages into host memory) with the
    pages written on the device (gpu_pages = 1) and with flat columns paged on the host (0), best
    of reps after one untimed call; the page kernels' own time from HIP events (mean of reps)"""
    lib = ctx.lib
    r = ctx.execute_device(plan, dts)
    ctx.synchronize()
    out = {"rows": r.num_rows}
    for mode in (1, 0):
        ctx.set_option("gpu_pages", mode)
        res = hjo.abi.hj_result()  # one untimed call: first launches and pinned staging
        ctx._check(lib.hj_dresult_to_host(ctx.handle, r.handle, C.byref(res)))
        lib.hj_result_free(C.byref(res))
        if mode:
            ctx.profile("page_")
        ts = []
        for _ in range(reps):
            res = hjo.abi.hj_result()
            t0 = time.perf_counter()
            ctx._check(lib.hj_dresult_to_host(ctx.handle, r.handle, C.byref(res)))
            ts.append(time.perf_counter() - t0)
            pages = sum(res.columns[j].num_pages for j in range(res.num_columns))
            lib.hj_result_free(C.byref(res))
        if mode:
            prof = ctx.profile_report()
            ctx.profile(None)
            out["kernel_ms"] = {k: ms / n for k, (n, ms) in prof.items()}
        out["pages"] = pages
        out["gpu_pages_ms" if mode else "host_pages_ms"] = min(ts) * 1e3
    ctx.set_option("gpu_pages", 1)
    r.free()
    out["bytes"] = out["pages"] * 8192
    return out


# ------------------------------------------------------------------ scan predicates
def scan_predicate(ctx, dts, d, reps=3):
    """a LIKE '%ab%' scan predicate over the 4M-row dimension's VARCHAR column (star4 only):
    the predicate kernel's event time (mean of reps after one untimed call) and its algorithmic
    bytes (8-byte string reference + the string's bytes per row)"""
    dm = d["dims"][1]
    n = len(dm["key"])
    plan = hjo.Plan()
    plan.inputs = [None, None, None, None]
    plan.new_scan_node(2, [(0, hjo.DataType.INT32)], preds=[("like", 2, "%ab%")])
    plan.root = 0
    ctx.execute_device(plan, dts).free()  # one untimed call: the kernel's first launch
    ctx.profile("predicate")
    rows = 0
    for _ in range(reps):
        r = ctx.execute_device(plan, dts)
        rows = r.num_rows
        r.free()
    prof = ctx.profile_report()
    ctx.profile(None)
    launches, ms = prof.get("predicate", (0, 0.0))
    byts = 8 * n + int(dm["soff"][-1])
    us = ms / max(1, launches) * 1e3
    return {"rows": n, "pattern": "%ab%", "kept": rows, "kernel_us": us,
            "alg_bytes": byts, "GB/s": byts / (us * 1e-6) / 1e9 if us else None}


# ------------------------------------------------------------------ CPU baseline (oracle)
def cpu_threads():
    """threads of the CPU baseline: the host's share of cores (OMP_NUM_THREADS where the host
    sets it, else every core)"""
    n = os.cpu_count() or 1
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(2, n)  # the reference's partition shift needs >= 2 threads


def host_cpu():
    """nproc, the CPU model and the socket count (lscpu)"""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = v.strip()
    except Exception:
        pass
    return info


def cpu_baseline(d, workload_name, tuples, threads):
    """the oracle (CPU restatement of the reference's execute()), used only as the baseline"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import hjtest  # test infrastructure: loads oracle/liboracle.so
    o = hjtest.Oracle()
    if workload_name == "star4":
        tabs = dg.star4_prefiltered_tables(d, o.codec)
        plan = dg.star4_plan(d, tabs, prefiltered=True)
        sample = "full star4 workload; dimensions pre-filtered (untimed, as the reference filters at load time)"
    else:
        plan = dg.two_way_plan(d, o.codec)
        sample = "full two_way workload"
    times = []
    for _ in range(3):
        ct, ns = o.execute(plan, threads=threads, timed=True)
        times.append(ns / 1e9)
        log(f"cpu oracle execute: {ns / 1e6:.1f} ms, {ct.num_rows} rows")
    t = statistics.median(times)
    return {"value": tuples / t, "unit": "join tuples/s", "cores": threads, "kind": "port",
            "sample": sample + f"; median of 3 execute() calls, {t * 1e3:.1f} ms each; "
                      f"{threads} probe threads plus the reference's 12-thread scan pool",
            "ms": t * 1e3, "host": host_cpu()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="star4",
                    choices=["star4", "two_way", "m2m", "serve", "batch", "job"])
    ap.add_argument("--queries", type=int, default=256, help="serve: queries per run")
    ap.add_argument("--slots", type=int, default=4, help="serve: contexts of the served path")
    ap.add_argument("--rates", default="50,200,500", help="serve: Poisson arrival rates per second")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reps", type=int, default=3, help="batch: timed repetitions (median)")
    ap.add_argument("--untuned", action="store_true",
                    help="batch: the engine's default parameters instead of the tuned table")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="engine option for the measured context (hj_set_option), repeatable")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 3 times (a plain copy kernel of "
                         "known size, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 256 << 20, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the join engine has no "
              f"multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve":
        serve_main(args)
        return
    if args.workload == "batch":
        batch_main(args)
        return
    if args.workload == "job":
        job_main(args)
        return
    codec = hjo.Codec()
    if args.workload == "star4":
        d, tables, plan, info, workload = star4_workload(codec, args.scale)
    elif args.workload == "m2m":
        d, tables, plan, info, workload = m2m_workload(codec, args.scale)
    else:
        d, tables, plan, info, workload = two_way_workload(codec, args.scale)
    tuples = sum(info["builds"]) + sum(info["probes"])

    ctx = hjo.Context(arena_bytes=(64 << 30) if args.workload == "m2m" else 0)
    if args.workload == "m2m":
        ctx.set_option("pipeline", 0)
    for kv in args.set:
        name, _, value = kv.partition("=")
        ctx.set_option(name, int(value))
    t = time.time()
    dts = [ctx.upload(tb) for tb in tables]
    log(f"uploaded {len(dts)} tables in {time.time() - t:.1f}s")

    # untimed warmup, then a few steps with every engine kernel timed to find the dominant one
    for _ in range(args.warmup):
        r = ctx.execute_device(plan, dts)
        got = r.num_rows
        r.free()
    assert got == info["out_rows"], f"output rows {got} != expected {info['out_rows']}"
    ctx.synchronize()
    ctx.profile("")
    nprof = 3
    for _ in range(nprof):
        r = ctx.execute_device(plan, dts)
        r.free()
    prof = ctx.profile_report()
    ctx.profile(None)
    per_step = {k: (n / nprof, ms / nprof) for k, (n, ms) in prof.items()}
    for k, (n, ms) in sorted(per_step.items(), key=lambda x: -x[1][1]):
        log(f"  {k:22s} {n:5.1f} launches/step {ms:8.3f} ms/step")
    modeled = [k for k in per_step if kernel_bytes(k, info, 1) is not None]
    dom = max(modeled or per_step, key=lambda k: per_step[k][1])

    # timed region: no kernel timing at all (events add packets, and gaps, to the stream)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.execute_device(plan, dts)
        r.free()
    ctx.synchronize()
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    # the dominant kernel's launches, bracketed by events over as many steps again
    ctx.profile(dom)
    for _ in range(args.steps):
        r = ctx.execute_device(plan, dts)
        r.free()
    prof = ctx.profile_report()
    ctx.profile(None)

    launches, dom_ms = prof.get(dom, (0, 0.0))
    roofline = None
    kb = kernel_bytes(dom, info, launches / max(1, args.steps))
    if launches and kb is not None:
        per_launch_bytes = kb / (launches / args.steps)
        avg_s = dom_ms / launches / 1e3
        achieved = per_launch_bytes / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": achieved / (HBM_PEAK / 1e9),
                    "traffic": measured_traffic(dom, args.workload, args.scale),
                    "bytes_per_launch": per_launch_bytes, "avg_launch_ms": avg_s * 1e3}

    # the many-to-many result is ~600M rows: its pages are not copied out here
    pages = result_pages(ctx, plan, dts) if args.workload != "m2m" else None
    log(f"result pages: {pages}")
    pred = scan_predicate(ctx, dts, d) if args.workload == "star4" else None
    if pred:
        log(f"scan predicate: {pred}")

    # the drop-in call end to end (hj_execute on host pages: upload, execute, result pages
    # copied out), the reference's timed execute()
    execute_ms = None
    if args.workload != "m2m":
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            res = ctx.execute(plan)
            ts.append(time.perf_counter() - t0)
            assert res.num_rows == info["out_rows"]
            del res
        execute_ms = statistics.median(ts) * 1e3
        log(f"hj_execute (host pages in, result pages out): {execute_ms:.2f} ms")

    cpu = None
    if not args.no_cpu_baseline and args.workload != "m2m":  # the CPU path takes minutes there
        cpu = cpu_baseline(d, args.workload, tuples, cpu_threads())

    out = {
        "metric": "join tuples/s", "value": tuples / step_s, "unit": "tuples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic, seeded", "config": workload,
        "pct_hbm_peak": 100.0 * info["b_alg"] / step_s / HBM_PEAK,
        "alg_bytes_per_step": info["b_alg"], "output_rows_per_s": info["out_rows"] / step_s,
        "kernel_ms_per_step": {k: v[1] for k, v in per_step.items()},
        "roofline": roofline, "cpu_baseline": cpu, "result_pages": pages,
        "scan_predicate": pred,
        "execute_ms": execute_ms,
        "step_reads": "every key, filter and output value and validity bit is read from the "
                      "resident input pages inside the step; the per-query 64-row group index "
                      "is built inside the step (the source's columns by the fused probe's "
                      "blocks, each for its row range; other columns by extra blocks of the "
                      "build's filter launch)",
    }
    for t_ in dts:
        t_.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
