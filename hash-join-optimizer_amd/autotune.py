"""Autotuner of the batched mode (hj_batch_params), project-written, no counterpart in the reference.

    python -c "import sys; sys.path.insert(0, '.'); import importlib; \
               importlib.import_module('hash-join-optimizer_amd.autotune').main()" \
           --scale 0.1 --out tuning.json                                            (on a GPU)

It times hj_batch_run over a seeded batch of 2- and 3-way joins (datagen.batch_joins) and
chooses, one parameter at a time in this order (coordinate descent from the engine's defaults,
each parameter keeping the best value found so far for the others):
  lane mapping and unroll   coop 0 (one lane per probe row) / 1 (four lanes cooperate on a row),
                            x 2 or 4 64-row groups per wave and round
  probe item size           64-row groups per probe item: 64, 256, 1024
  directory sizing          key slots 2^k >= rows << load_shift (load <= 1/2 or 1/4), and the
                            same for builds of at least 256K rows only
  radix bits                0, 4, 8, 12: the insert rounds ordered by that many top hash bits
  plain route               sources of at least 1M / 4M / 16M rows (or none) run through the
                            fused pipeline one at a time, their results written into the batch
  plain build route         likewise queries whose builds hold at least 256K / 1M / 4M rows (or
                            none): the pipeline builds tables in LDS segments
Every timing is the median of `reps` runs of the same graph; every configuration's results are
checked against the first configuration's row counts. The chosen parameters and all timings are
written as JSON; tuned_params() reads the committed tables (tuning/batch_mi355x.json, one
per tuned batch size) and picks the one tuned on the batch nearest in total source rows.
"""
import argparse
import importlib
import json
import math
import os
import statistics
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "batch_mi355x.json")

DEFAULTS = {"coop": 0, "unroll": 4, "probe_item_groups": 256, "load_shift": 1,
            "large_build_rows": 0, "large_load_shift": 0, "radix_bits": 0, "insert_item_rows": 8192,
            "plain_probe_rows": 0, "plain_build_rows": 0}

STEPS = [
    ("lane mapping", [{"coop": c, "unroll": u} for c in (0, 1) for u in (2, 4)]),
    ("probe item size", [{"probe_item_groups": g} for g in (64, 256, 1024)]),
    ("directory sizing", [{"load_shift": 1, "large_build_rows": 0, "large_load_shift": 0},
                          {"load_shift": 2, "large_build_rows": 0, "large_load_shift": 0},
                          {"load_shift": 1, "large_build_rows": 1 << 18, "large_load_shift": 2},
                          {"load_shift": 2, "large_build_rows": 1 << 18, "large_load_shift": 1}]),
    ("radix bits", [{"radix_bits": r} for r in (0, 4, 8, 12)]),
    ("plain route", [{"plain_probe_rows": r} for r in (0, 1 << 20, 1 << 22, 1 << 24)]),
    ("plain build route", [{"plain_build_rows": r} for r in (0, 1 << 18, 1 << 20, 1 << 22)]),
]


def tuned_params(queries=None):
    """the committed table's chosen parameters for a batch like `queries` (list of (plan,
    tables)): the table tuned on the batch whose total source rows are nearest on a log scale
    (the engine's defaults when there is no table)"""
    try:
        with open(TABLE) as f:
            tables = json.load(f)["tables"]
    except (OSError, KeyError, ValueError):
        return dict(DEFAULTS)
    rows = sum(p.inputs[0].num_rows if p.inputs else d[0].num_rows for p, d in queries or [])
    if not rows or not tables:
        return dict(tables[0]["best"]) if tables else dict(DEFAULTS)
    best = min(tables, key=lambda t: abs(math.log(max(1, t["probe_rows"]) / rows)))
    return dict(best["best"])


def time_batch(hjo, ctx, queries, params, reps):
    """(median ms of hj_batch_run, rows per query)"""
    with hjo.Batch(ctx, queries, dict(params, use_graph=1)) as b:
        b.run()  # graph capture and first launch
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            b.run()
            ts.append((time.perf_counter() - t0) * 1e3)
        rows = [b.info(i)[1] for i in range(len(queries))]
        batched = sum(b.info(i)[0] for i in range(len(queries)))
    return statistics.median(ts), rows, batched


def tune(hjo, ctx, queries, reps=5, log=print):
    best = dict(DEFAULTS)
    best_ms, ref_rows, _ = time_batch(hjo, ctx, queries, best, reps)
    trials = [{"params": dict(best), "ms": best_ms}]
    log(f"defaults: {best_ms:.3f} ms")
    for name, options in STEPS:
        for opt in options:
            cand = dict(best, **opt)
            if cand == best:
                continue
            ms, rows, batched = time_batch(hjo, ctx, queries, cand, reps)
            if rows != ref_rows:
                raise AssertionError(f"{cand}: result rows differ from the defaults'")
            trials.append({"params": cand, "ms": ms})
            log(f"{name} {opt}: {ms:.3f} ms")
            if ms < best_ms:
                best, best_ms = cand, ms
        log(f"after {name}: {best} ({best_ms:.3f} ms)")
    return best, best_ms, trials


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=0.1)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    hjo = importlib.import_module("hash-join-optimizer_amd")
    dg = importlib.import_module("hash-join-optimizer_amd.datagen")
    ctx = hjo.Context()
    plans = dg.batch_joins(seed=48, n=args.queries, scale=args.scale, codec=hjo.Codec())
    dts = [[ctx.upload(t) for t in p.inputs] for p in plans]
    queries = list(zip(plans, dts))
    best, best_ms, trials = tune(hjo, ctx, queries, args.reps,
                                 log=lambda m: print("[autotune]", m, flush=True))
    out = {"device": "MI355X (gfx950)", "workload": "datagen.batch_joins(seed=48)",
           "queries": args.queries, "scale": args.scale,
           "probe_rows": int(sum(p.inputs[0].num_rows for p in plans)),
           "reps": args.reps, "best": best, "best_ms": best_ms, "trials": trials}
    text = json.dumps(out, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text)
    for d in dts:
        for t in d:
            t.free()
    ctx.close()


if __name__ == "__main__":
    main()
