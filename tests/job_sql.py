"""SQLite as an independent evaluator of the Join Order Benchmark queries on synthetic tables
(test infrastructure for the SQL / plan front-end).

Two SQL texts are evaluated per query:
  * the query itself, with its select list written as plain columns (no MIN), as the reference's
    executed_sql does for its DuckDB comparison (tests/read_sql.cpp:700-735);
  * the query re-written from the front-end's parse tree: every join condition the parse found,
    and each table's filter expression. With two_valued=False it is plain SQL (three-valued
    logic), which must give the rows of the query itself if the parse has the query's structure;
    with two_valued=True each comparison, LIKE, IN, BETWEEN and IS NULL is wrapped as
    COALESCE(atom, 0), which is the reference's two-valued filter semantics (a comparison of a
    NULL row is false, NOT negates it), so its rows are what the engine must return.
LIKE is case-sensitive (PRAGMA case_sensitive_like) and text compares bytewise, as the
reference's comparisons do.
"""
import importlib
import re
import sqlite3

plan_mod = importlib.import_module("hash-join-optimizer_amd.plan")
jobfront = importlib.import_module("hash-join-optimizer_amd.jobfront")
sort_key = plan_mod.sort_key


def database(schema, tables):
    """in-memory SQLite database holding datagen.job_tables' tables"""
    db = sqlite3.connect(":memory:")
    db.execute("PRAGMA case_sensitive_like = ON")
    for t, cols in schema.items():
        if t not in tables:
            continue
        n, data = tables[t]
        decl = ", ".join(f"{c} {'INTEGER' if ty == jobfront.I32 else 'TEXT'}" for c, ty in cols)
        db.execute(f"CREATE TABLE {t} ({decl})")
        pycols = []
        for c in data:
            if isinstance(c, tuple):
                v, ok = c
                pycols.append([int(x) if o else None for x, o in zip(v.tolist(), ok.tolist())])
            else:
                pycols.append(c)
        marks = ", ".join("?" * len(cols))
        db.executemany(f"INSERT INTO {t} VALUES ({marks})", list(zip(*pycols)))
    db.commit()
    return db


def _names(parsed):
    rev = {ent: a for a, ent in parsed.alias.items()}
    return {ent: rev.get(ent, ent[0]) for ent in parsed.entities}


def _lit(v):
    if isinstance(v, str):
        return "'" + v.replace("'", "''") + "'"
    return str(int(v))


def _select(parsed, names):
    return "SELECT " + ", ".join(f"{names[e]}.{c}" for e, c in parsed.output)


def original_sql(sql, parsed):
    """the query with its select list as plain columns"""
    names = _names(parsed)
    i = re.search(r"\bFROM\b", sql, re.I).start()
    return _select(parsed, names) + " " + sql[i:].strip().rstrip(";")


def tree_sql(parsed, two_valued):
    names = _names(parsed)

    def col(ent, c):
        return f"{names[ent]}.{c}"

    def atom(text):
        return f"COALESCE(({text}), 0)" if two_valued else f"({text})"

    def expr(ent, x):
        k = x[0]
        if k in ("and", "or"):
            return f"({expr(ent, x[1])} {k.upper()} {expr(ent, x[2])})"
        if k == "not":
            return f"(NOT {expr(ent, x[1])})"
        if k == "cmp":
            op = {"eq": "=", "neq": "<>", "lt": "<", "gt": ">", "leq": "<=", "geq": ">="}[x[1]]
            return atom(f"{col(ent, x[2])} {op} {_lit(x[3])}")
        if k == "like":
            return atom(f"{col(ent, x[2])} {'NOT ' if x[1] else ''}LIKE {_lit(x[3])}")
        if k == "in":
            return atom(f"{col(ent, x[1])} IN ({', '.join(_lit(v) for v in x[2])})")
        if k == "between":
            return atom(f"{col(ent, x[1])} BETWEEN {_lit(x[2])} AND {_lit(x[3])}")
        if k == "isnull":
            return atom(f"{col(ent, x[1])} IS NULL")
        raise ValueError(k)

    conds = []
    seen = set()
    for a, adj in parsed.join_graph.items():
        for b, (ca, cb) in adj.items():
            if (b, a) in seen:
                continue
            seen.add((a, b))
            conds.append(f"{col(a, ca)} = {col(b, cb)}")
    for ent, e in parsed.filters.items():
        conds.append(expr(ent, e))
    frm = ", ".join(f"{e[0]} AS {names[e]}" for e in parsed.entities)
    where = (" WHERE " + " AND ".join(conds)) if conds else ""
    return f"{_select(parsed, names)} FROM {frm}{where}"


def rows(db, text):
    return sorted((tuple(r) for r in db.execute(text).fetchall()), key=sort_key)
