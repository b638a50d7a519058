"""Seeded synthetic inputs for the benchmark configurations (numpy PCG64 generators).

  two_way      1M unique shuffled build keys + INT32 payload  JOIN  10M probe rows, FK ~ U[0, 2M),
               5 % NULL FKs, INT32 payload; output (build payload, probe payload). Seed 42.
  star4        fact of 36M rows with three INT32 FKs ~ Zipf(1.1) over dimensions of 2.5M / 4M /
               100K rows (3 % NULL each); each dimension: unique shuffled keys, an INT32 attribute
               ~ U[0, 10000) with a range filter of 1-10 % selectivity, a VARCHAR of 8-64 printable
               ASCII chars. Plan ((F x sD1) x sD2) x sD3, builds on the filtered dimensions,
               output = the 3 fact FKs + the 3 dimension strings.
  many_to_many 50M build keys ~ Zipf(1.3) over [0, 5M) and 200M probe keys ~ Zipf(1.3) with an
               independent rank->key permutation, 10 % NULL probe keys, INT32 payloads.

Zipf(s) over a finite domain [0, N) is sampled by inverse CDF over ranks 1..N (weights r^-s),
followed by a seeded Fisher-Yates rank->key permutation per column; NULLs are Bernoulli per row.
Every generator returns plain numpy arrays; build_*_plan turns them into a Plan.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .plan import ColumnarTable, Codec, DataType, Plan

I32 = DataType.INT32
STR = DataType.VARCHAR


def zipf_ranks(rng, n_domain, s, size, chunk=1 << 22):
    """ranks 0..n_domain-1 with weights (rank+1)^-s by inverse CDF; the uniforms are drawn in
    chunk order from rng, and the chunks' CDF searches (which release the GIL) run on a pool of
    host threads, so the draw is the same for any thread count"""
    w = np.arange(1, n_domain + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    out = np.empty(size, dtype=np.int64)

    def search(i, u):
        out[i:i + len(u)] = np.searchsorted(cdf, u, side="right")

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        pending = []
        for i in range(0, size, chunk):
            pending.append(pool.submit(search, i, rng.random(min(chunk, size - i))))
            if len(pending) >= 32:  # bounded: at most 32 chunks of uniforms in memory
                pending.pop(0).result()
        for f in pending:
            f.result()
    np.minimum(out, n_domain - 1, out=out)
    return out


def zipf_keys(rng, n_domain, s, size, perm=None):
    if perm is None:
        perm = rng.permutation(n_domain).astype(np.int32)
    return perm[zipf_ranks(rng, n_domain, s, size)], perm


def random_strings(rng, n, lo=8, hi=64):
    lens = rng.integers(lo, hi + 1, n).astype(np.uint64)
    off = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    chars = rng.integers(33, 127, int(off[-1]), dtype=np.uint8)
    return off, chars


# ------------------------------------------------------------------ two_way (2-way INT32 join)
def two_way(seed=42, n_build=1_000_000, n_probe=10_000_000, null_p=0.05):
    rng = np.random.default_rng(seed)
    bk = rng.permutation(n_build).astype(np.int32)
    bp = rng.integers(0, 2**31, n_build).astype(np.int32)
    pk = rng.integers(0, 2 * n_build, n_probe).astype(np.int32)
    pv = (rng.random(n_probe) >= null_p).astype(np.uint8)
    pp = rng.integers(0, 2**31, n_probe).astype(np.int32)
    return dict(bk=bk, bp=bp, pk=pk, pv=pv, pp=pp)


def two_way_plan(d, codec=None):
    codec = codec or Codec()
    plan = Plan()
    plan.inputs.append(ColumnarTable(len(d["bk"]), [codec.encode_int32(d["bk"]),
                                                    codec.encode_int32(d["bp"])]))
    plan.inputs.append(ColumnarTable(len(d["pk"]), [codec.encode_int32(d["pk"], d["pv"]),
                                                    codec.encode_int32(d["pp"])]))
    plan.new_scan_node(0, [(0, I32), (1, I32)])
    plan.new_scan_node(1, [(0, I32), (1, I32)])
    plan.new_join_node(True, 0, 1, 0, 0, [(1, I32), (3, I32)])
    plan.root = 2
    return plan


def two_way_expected_count(d):
    nb = len(d["bk"])
    hit = (d["pv"] == 1) & (d["pk"] < nb)  # build keys are exactly 0..nb-1
    return int(hit.sum())


# ------------------------------------------------------------------ star4 (4-way star join)
STAR_DIMS = (2_500_000, 4_000_000, 100_000)


def star4(seed=43, n_fact=36_000_000, dims=STAR_DIMS, s=1.1, null_p=0.03, sel=None):
    rng = np.random.default_rng(seed)
    out = {"fact": [], "fact_valid": [], "dims": []}
    for i, nd in enumerate(dims):
        key = rng.permutation(nd).astype(np.int32)
        attr = rng.integers(0, 10000, nd).astype(np.int32)
        soff, schars = random_strings(rng, nd)
        w = int(round(10000 * (sel[i] if sel else rng.uniform(0.01, 0.10))))
        lo = int(rng.integers(0, 10000 - w + 1))
        out["dims"].append(dict(key=key, attr=attr, soff=soff, schars=schars,
                                lo=lo, hi=lo + w - 1))
    for i, nd in enumerate(dims):
        fk, _ = zipf_keys(rng, nd, s, n_fact)
        out["fact"].append(fk.astype(np.int32))
        out["fact_valid"].append((rng.random(n_fact) >= null_p).astype(np.uint8))
    return out


def star4_tables(d, codec=None):
    codec = codec or Codec()
    n = len(d["fact"][0])
    fact = ColumnarTable(n, [codec.encode_int32(fk, v)
                             for fk, v in zip(d["fact"], d["fact_valid"])])
    dims = []
    for dm in d["dims"]:
        dims.append(ColumnarTable(len(dm["key"]), [codec.encode_int32(dm["key"]),
                                                   codec.encode_int32(dm["attr"]),
                                                   codec.encode_varchar(dm["soff"], dm["schars"])]))
    return [fact] + dims


def star4_plan(d, tables=None, codec=None, prefiltered=False):
    """((F x sD1) x sD2) x sD3 with the range filters pushed into the dimension scans (or, with
    prefiltered=True, over already-filtered dimension tables, the reference's load-time filter)."""
    plan = Plan()
    plan.inputs = list(tables if tables is not None else star4_tables(d, codec))
    f = plan.new_scan_node(0, [(0, I32), (1, I32), (2, I32)])
    scans = []
    for i, dm in enumerate(d["dims"]):
        flt = () if prefiltered else [(1, dm["lo"], dm["hi"])]
        scans.append(plan.new_scan_node(1 + i, [(0, I32), (2, STR)], flt))
    j1 = plan.new_join_node(False, f, scans[0], 0, 0, [(0, I32), (1, I32), (2, I32), (4, STR)])
    j2 = plan.new_join_node(False, j1, scans[1], 1, 0,
                            [(0, I32), (1, I32), (2, I32), (3, STR), (5, STR)])
    j3 = plan.new_join_node(False, j2, scans[2], 2, 0,
                            [(0, I32), (1, I32), (2, I32), (3, STR), (4, STR), (6, STR)])
    plan.root = j3
    return plan


def star4_prefiltered_tables(d, codec=None):
    """the dimension tables after the load-time filter (what the reference's execute() gets)"""
    codec = codec or Codec()
    tabs = star4_tables(dict(fact=d["fact"], fact_valid=d["fact_valid"], dims=[]), codec)
    for dm in d["dims"]:
        m = (dm["attr"] >= dm["lo"]) & (dm["attr"] <= dm["hi"])
        idx = np.nonzero(m)[0]
        lens = (dm["soff"][idx + 1] - dm["soff"][idx]).astype(np.uint64)
        off = np.zeros(len(idx) + 1, np.uint64)
        np.cumsum(lens, out=off[1:])
        starts = dm["soff"][idx].astype(np.int64)
        sel = np.concatenate([np.arange(s0, s0 + l) for s0, l in zip(starts, lens.astype(np.int64))]) \
            if len(idx) else np.zeros(0, np.int64)
        chars = dm["schars"][sel] if len(sel) else np.zeros(0, np.uint8)
        tabs.append(ColumnarTable(len(idx), [codec.encode_int32(dm["key"][idx]),
                                             codec.encode_int32(dm["attr"][idx]),
                                             codec.encode_varchar(off, chars)]))
    return tabs


def star4_expected(d):
    """exact output row count and the matched rows (fact row, dim rows) by numpy lookups"""
    n = len(d["fact"][0])
    ok = np.ones(n, dtype=bool)
    drow = []
    for fk, v, dm in zip(d["fact"], d["fact_valid"], d["dims"]):
        inv = np.empty(len(dm["key"]), dtype=np.int64)
        inv[dm["key"]] = np.arange(len(dm["key"]))
        r = inv[fk]
        passed = (dm["attr"][r] >= dm["lo"]) & (dm["attr"][r] <= dm["hi"])
        ok &= (v == 1) & passed
        drow.append(r)
    rows = np.nonzero(ok)[0]
    return rows, [r[rows] for r in drow]


# ------------------------------------------------------------------ many_to_many (skew stress)
def many_to_many(seed=44, n_build=50_000_000, n_probe=200_000_000, domain=5_000_000, s=1.3,
                 null_p=0.10):
    rng = np.random.default_rng(seed)
    bk, _ = zipf_keys(rng, domain, s, n_build)
    pk, _ = zipf_keys(rng, domain, s, n_probe)
    pv = (rng.random(n_probe) >= null_p).astype(np.uint8)
    bp = rng.integers(0, 2**31, n_build).astype(np.int32)
    pp = rng.integers(0, 2**31, n_probe).astype(np.int32)
    return dict(bk=bk.astype(np.int32), bp=bp, pk=pk.astype(np.int32), pv=pv, pp=pp,
                domain=domain)


def many_to_many_expected_count(d):
    dom = d["domain"]
    b = np.bincount(d["bk"], minlength=dom).astype(np.int64)
    p = np.bincount(d["pk"][d["pv"] == 1], minlength=dom).astype(np.int64)
    return int((b * p).sum())


def _batch_query(ss, scale, codec, null_p):
    """one query of batch_joins from its own seed sequence"""
    rng = np.random.default_rng(ss)
    ways = int(rng.integers(2, 4))
    nb = [int(np.exp(rng.uniform(np.log(1_000 * scale + 1), np.log(4_000_000 * scale + 1))))
          for _ in range(ways - 1)]
    npr = int(np.exp(rng.uniform(np.log(10_000 * scale + 1), np.log(40_000_000 * scale + 1))))
    plan = Plan()
    probe_cols = []
    for m_rows in nb:
        m = rng.uniform(0.1, 1.0)
        dist = rng.choice(["uniform", "zipf0.8", "zipf1.2"])
        if dist == "uniform":
            fk = rng.integers(0, m_rows, npr, dtype=np.int32)
        else:
            w = np.arange(1, m_rows + 1, dtype=np.float64) ** (-(0.8 if dist == "zipf0.8" else 1.2))
            cdf = np.cumsum(w)
            cdf /= cdf[-1]
            perm = rng.permutation(m_rows).astype(np.int32)
            ranks = np.minimum(np.searchsorted(cdf, rng.random(npr), side="right"), m_rows - 1)
            fk = perm[ranks]
        miss = rng.random(npr) >= m
        fk = np.where(miss, (m_rows + rng.integers(0, m_rows + 1, npr)).astype(np.int32), fk)
        valid = (rng.random(npr) >= null_p).astype(np.uint8)
        probe_cols.append(codec.encode_int32(fk.astype(np.int32), valid))
    probe_cols.append(codec.encode_int32(rng.integers(0, 2**31, npr, dtype=np.int32)))
    plan.inputs.append(ColumnarTable(npr, probe_cols))
    for m_rows in nb:
        plan.inputs.append(ColumnarTable(m_rows, [
            codec.encode_int32(rng.permutation(m_rows).astype(np.int32)),
            codec.encode_int32(rng.integers(0, 2**31, m_rows, dtype=np.int32))]))
    nk = len(nb)
    cur = plan.new_scan_node(0, [(j, I32) for j in range(nk + 1)])
    width = nk + 1
    for j in range(nk):
        b = plan.new_scan_node(1 + j, [(0, I32), (1, I32)])
        # the probe side's attributes: the j build payloads prepended so far, the FKs, its
        # payload; FK j is attribute 2j
        cur = plan.new_join_node(True, b, cur, 0, 2 * j,
                                 [(1, I32)] + [(2 + i, I32) for i in range(width)])
        width += 1
    plan.root = cur
    return plan


def batch_joins(seed=48, n=256, scale=1.0, codec=None, null_p=0.05, threads=8):
    """n independent 2- and 3-way INT32 joins, each over its own tables: build sides of unique
    shuffled keys + an INT32 payload, rows log-uniform in [1K, 4M] x scale; a probe side of
    rows log-uniform in [10K, 40M] x scale with one FK per join (uniform, Zipf(0.8) or Zipf(1.2)
    over the build keys, match rate ~ U[0.1, 1.0]: the other FKs miss every build key; NULL with
    probability null_p) + an INT32 payload. Join j is keyed by the probe side's FK j; the output
    is every build payload and every probe attribute. Query i is drawn from child i of the seed's
    sequence (numpy SeedSequence.spawn), so the queries are generated in parallel and the batch
    is the same for any thread count. Returns a list of Plans with inputs."""
    codec = codec or Codec()
    seqs = np.random.SeedSequence(seed).spawn(n)
    with ThreadPoolExecutor(max_workers=max(1, min(threads, n))) as pool:
        return list(pool.map(lambda ss: _batch_query(ss, scale, codec, null_p), seqs))


# the serving workload's fact table: five FK columns over the three dimensions of star4 (the
# fourth and fifth are second roles of dimensions 0 and 1, under their own Zipf permutations)
SERVE_FK_DIM = (0, 1, 2, 0, 1)


def serve_star(seed=47, n_fact=36_000_000, dims=STAR_DIMS, s=1.1, null_p=0.03):
    """dimension tables as star4 (unique shuffled keys, an INT32 attribute ~ U[0, 10000), a
    VARCHAR of 8-64 chars) and a fact table with the five FK columns of SERVE_FK_DIM, each
    Zipf(s) over its dimension's keys with NULL probability null_p"""
    rng = np.random.default_rng(seed)
    out = {"fact": [], "fact_valid": [], "dims": []}
    for nd in dims:
        soff, schars = random_strings(rng, nd)
        out["dims"].append(dict(key=rng.permutation(nd).astype(np.int32),
                                attr=rng.integers(0, 10000, nd).astype(np.int32),
                                soff=soff, schars=schars))
    for di in SERVE_FK_DIM:
        fk, _ = zipf_keys(rng, dims[di], s, n_fact)
        out["fact"].append(fk.astype(np.int32))
        out["fact_valid"].append((rng.random(n_fact) >= null_p).astype(np.uint8))
    return out


def serve_tables(d, fact_sizes, codec=None):
    """the resident tables of the serving workload: the fact table's prefixes of fact_sizes rows,
    then the dimensions (key, attribute, string)"""
    codec = codec or Codec()
    facts = [ColumnarTable(n, [codec.encode_int32(fk[:n], v[:n])
                               for fk, v in zip(d["fact"], d["fact_valid"])]) for n in fact_sizes]
    dims = [ColumnarTable(len(dm["key"]), [codec.encode_int32(dm["key"]),
                                           codec.encode_int32(dm["attr"]),
                                           codec.encode_varchar(dm["soff"], dm["schars"])])
            for dm in d["dims"]]
    return facts + dims


def serve_queries(rng, n, fact_sizes, n_dims=3, pool=16, max_joins=5):
    """n star joins over serve_tables' tables: a fact prefix of log-uniform [100K, 36M] rows
    (the nearest of fact_sizes), 1..max_joins of its FK columns (2- to 6-way), each joined with its
    dimension under a range filter on the attribute of 1-20 % selectivity drawn from a pool of
    `pool` filters per dimension (Zipf(1) over the pool: popular filters repeat, as parameterized
    queries do); output the FKs used and the joined dimensions' strings. Returns Plans without
    inputs (input i is table i of serve_tables)."""
    prng = np.random.default_rng(rng.integers(1 << 62))
    filters = []
    for _ in range(n_dims):
        f = []
        for _ in range(pool):
            w = int(prng.integers(100, 2001))
            lo = int(prng.integers(0, 10_000 - w))
            f.append((lo, lo + w - 1))
        filters.append(f)
    wpool = 1.0 / np.arange(1, pool + 1)
    wpool /= wpool.sum()
    plans = []
    nf = len(fact_sizes)
    for _ in range(n):
        size = float(np.exp(rng.uniform(np.log(100_000), np.log(36_000_000))))
        fi = int(np.argmin([abs(np.log(size / f)) for f in fact_sizes]))
        k = int(rng.integers(1, max_joins + 1))
        fks = sorted(rng.choice(len(SERVE_FK_DIM), size=k, replace=False).tolist())
        plan = Plan()
        plan.inputs = [None] * (nf + n_dims)
        cur = plan.new_scan_node(fi, [(j, I32) for j in fks])
        types = [I32] * len(fks)
        for pos, j in enumerate(fks):
            di = SERVE_FK_DIM[j]
            lo, hi = filters[di][int(rng.choice(pool, p=wpool))]
            ds = plan.new_scan_node(nf + di, [(0, I32), (2, STR)], [(1, lo, hi)])
            cur = plan.new_join_node(False, cur, ds, pos, 0,
                                     [(i, t) for i, t in enumerate(types)] + [(len(types) + 1, STR)])
            types.append(STR)
        plan.root = cur
        plans.append(plan)
    return plans


# ------------------------------------------------------------------ Join Order Benchmark schema
JOB_ROWS = {"title": 500, "name": 600, "char_name": 400, "company_name": 200, "keyword": 300,
            "aka_name": 400, "aka_title": 300, "cast_info": 2500, "movie_companies": 1200,
            "movie_info": 2000, "movie_info_idx": 1000, "movie_keyword": 1500,
            "movie_link": 300, "complete_cast": 400, "person_info": 1200, "comp_cast_type": 4,
            "company_type": 4, "info_type": 113, "kind_type": 7, "link_type": 18,
            "role_type": 12}
JOB_FK = {"movie_id": "title", "linked_movie_id": "title", "episode_of_id": "title",
          "person_id": "name", "person_role_id": "char_name", "company_id": "company_name",
          "keyword_id": "keyword", "info_type_id": "info_type", "kind_id": "kind_type",
          "role_id": "role_type", "company_type_id": "company_type",
          "link_type_id": "link_type", "subject_id": "comp_cast_type",
          "status_id": "comp_cast_type"}
JOB_INT_RANGE = {"production_year": (1900, 2021), "season_nr": (1, 30), "episode_nr": (1, 100),
                 "nr_order": (1, 50)}
_JOB_ALPHA = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz ()[]-'0123456789ABCDEFGHIJ", np.uint8)


def _job_strings(rng, n, lo=1, hi=12):
    lens = rng.integers(lo, hi + 1, n)
    chars = rng.choice(_JOB_ALPHA, int(lens.sum()))
    cuts = np.concatenate([[0], np.cumsum(lens)])
    blob = bytes(chars)
    return [blob[cuts[i]:cuts[i + 1]].decode() for i in range(n)]


def _job_instantiate(rng, pattern):
    """a string that LIKE `pattern` matches: '%' becomes 0-3 random characters, '_' one"""
    out = []
    for ch in pattern:
        if ch == "%":
            out.append(_job_strings(rng, 1, 0, 3)[0])
        elif ch == "_":
            out.append(_job_strings(rng, 1, 1, 1)[0])
        else:
            out.append(ch)
    return "".join(out)


def job_tables(schema_text, queries, seed=70, scale=1.0, null_p=0.1, hot_p=0.85, cold_p=0.2,
               fk_hot_p=0.6):
    """Synthetic tables of the Join Order Benchmark schema (CREATE TABLE text) for the queries
    [(name, sql, plan)]: {table: (num_rows, [column])} for the tables they read, an INT32 column
    a pair (int32 values, validity bool array), a VARCHAR column a list of str / None.

    id columns are 1..n. A table's first rows are its hot rows (a fifth of a large table; in a
    type table of at most 200 rows, as many rows as the queries name literals for it): a hot row
    takes, column by column with probability hot_p (cold_p for other rows), one of the literals
    the queries compare that column with (LIKE patterns instantiated), else a random value (an
    INT32 uniform over JOB_INT_RANGE or [0, 10^6), a string of 1-12 characters); a type table's
    row i takes the i-th literal. Foreign keys (JOB_FK by column name) point at a hot row with
    probability fk_hot_p, else uniformly at any row, so that filtered rows join. Columns declared
    without NOT NULL are NULL with probability null_p. Row counts are JOB_ROWS x scale (the type
    tables keep theirs)."""
    from . import jobfront
    schema = jobfront.parse_schema(schema_text)
    nullable = jobfront.nullable_columns(schema_text)
    pools, used = {}, []
    for _, sql, _ in queries:
        parsed = jobfront.ParsedSQL(sql, schema)
        for t, _ in parsed.entities:
            if t not in used:
                used.append(t)
        for k, v in jobfront.literals(parsed).items():
            pools.setdefault(k, []).extend(v)
    rows = {t: (n if n <= 200 else max(1, int(n * scale))) for t, n in JOB_ROWS.items()}
    rng = np.random.default_rng(seed)
    bank = _job_strings(rng, 4096)
    choices = {}  # (table, column) -> literal values a row may take
    for (t, c), pool in pools.items():
        ch = []
        for x in dict.fromkeys(pool):
            if isinstance(x, tuple):
                ch.extend(_job_instantiate(rng, x[1]) for _ in range(4))
            else:
                ch.append(x)
        choices[(t, c)] = ch
    hot = {}
    for t, n in rows.items():
        if n <= 200:
            k = max([len(v) for (tt, _), v in choices.items() if tt == t] or [1])
            hot[t] = max(1, min(n, k))
        else:
            hot[t] = max(1, n // 5)
    out = {}
    for t in sorted(used, key=list(schema).index):
        n = rows[t]
        is_hot = np.arange(n) < hot[t]
        data = []
        for c, ty in schema[t]:
            ch = choices.get((t, c), [])
            nul = (rng.random(n) < null_p) if (t, c) in nullable else np.zeros(n, bool)
            pick = rng.random(n) < np.where(is_hot, hot_p, cold_p)
            if ty == I32:
                if c == "id":
                    v = np.arange(1, n + 1, dtype=np.int64)
                elif c in JOB_FK:
                    tgt = JOB_FK[c]
                    v = np.where(rng.random(n) < fk_hot_p, rng.integers(1, hot[tgt] + 1, n),
                                 rng.integers(1, rows[tgt] + 1, n))
                else:
                    lo, hi = JOB_INT_RANGE.get(c, (0, 1_000_000))
                    v = rng.integers(lo, hi, n)
                    ints = [x for x in ch if isinstance(x, int)]
                    if ints:
                        v[pick] = np.asarray(ints)[rng.integers(0, len(ints), int(pick.sum()))]
                data.append((v.astype(np.int32), ~nul))
                continue
            strs = [x if isinstance(x, str) else str(x) for x in ch]
            if n <= 200:
                vals = [strs[i] if i < len(strs) else bank[(i * 7919) % len(bank)]
                        for i in range(n)]
            else:
                ci = rng.integers(0, max(1, len(strs)), n)
                bi = rng.integers(0, len(bank), n)
                vals = [strs[ci[i]] if (pick[i] and strs) else bank[bi[i]] for i in range(n)]
            data.append([None if nul[i] else vals[i] for i in range(n)])
        out[t] = (n, data)
    return out


def job_columnar(tables, codec=None):
    """{table: ColumnarTable} of job_tables' output"""
    codec = codec or Codec()
    out = {}
    for t, (n, cols) in tables.items():
        enc = []
        for c in cols:
            if isinstance(c, tuple):
                enc.append(codec.encode_int32(c[0], c[1].astype(np.uint8)))
            else:
                enc.append(codec.encode_strings(c))
        out[t] = ColumnarTable(n, enc)
    return out
