"""SQL and PostgreSQL-plan front-end of the Join Order Benchmark driver: a JOB query (job/*.sql)
and its PostgreSQL EXPLAIN plan (an entry of plans.json) -> a Plan for the engine.

This restates the reference's tests/read_sql.cpp:738-1192 (ParsedSQL::parse_sql, parse_expr,
load_join_pipeline) and its schema table (read_sql.cpp:28-170, job/schema.sql) with a parser of
its own for the SQL subset the 113 JOB queries use (the reference links the Hyrise SQL parser):

    SELECT MIN(a.c) [AS x], ... FROM table [AS] alias, ... WHERE <expr>;
    <expr>: OR / AND / NOT, parentheses, a.c <op> literal (= <> != < > <= >=), a.c = b.d (a join
            condition), a.c [NOT] LIKE 'pattern', a.c [NOT] IN (literals), a.c [NOT] BETWEEN x
            AND y, a.c IS [NOT] NULL

Semantics follow the reference, not PostgreSQL:
  * the top-level AND's conjuncts are split: a = b between two tables' columns joins them (all
    columns equal through a chain of such conditions are one class, read_sql.cpp:813-852); any
    other conjunct is a filter of the one table (alias) it names, ANDed with that table's other
    filters; an AND / OR / NOT below the top must name one table only (read_sql.cpp:386-414);
  * IN (v1, ..., vn) is (= v1) OR ... OR (= vn), BETWEEN x AND y is (>= x) AND (<= y),
    IS NOT NULL is NOT (IS NULL) (read_sql.cpp:536-652);
  * a filter runs on the table before the joins (the reference filters at load time), here as
    the scan's predicate program (hj_pred): a comparison with NULL is false, NOT negates it;
  * the plan's output columns are the select list's columns (inside MIN), in order, with no
    aggregation (read_sql.cpp:785-809);
  * joins come from the plan: every "Hash Join" node with a "Hash" child is a join whose build
    side is that child's input (build_left when it is the first child); "Aggregate" and "Gather"
    nodes pass through; "Seq Scan" and "Index Only Scan" nodes are scans of the alias they name;
    the join condition is a class edge between an alias of each side (read_sql.cpp:868-1192).
Each base table is one input of the plan (shared by the aliases of the same table), so a query
over resident tables needs every table uploaded once.
"""
import json
import os
import re

from .plan import DataType, Plan

I32, STR = DataType.INT32, DataType.VARCHAR


class SQLError(ValueError):
    pass


# ------------------------------------------------------------------ schema
def parse_schema(text):
    """{table: [(column, DataType)]} from CREATE TABLE statements: integer columns are INT32,
    every other column VARCHAR (the reference's attributes_map, read_sql.cpp:28-170)"""
    schema = {}
    for m in re.finditer(r"CREATE TABLE\s+(\w+)\s*\((.*?)\);", text, re.S | re.I):
        cols = []
        for line in m.group(2).split(","):
            parts = line.split()
            if not parts:
                continue
            cols.append((parts[0], I32 if parts[1].lower() == "integer" else STR))
        schema[m.group(1)] = cols
    return schema


def load_schema(path):
    with open(path) as f:
        return parse_schema(f.read())


def nullable_columns(text):
    """{(table, column)} of the columns declared without NOT NULL"""
    out = set()
    for m in re.finditer(r"CREATE TABLE\s+(\w+)\s*\((.*?)\);", text, re.S | re.I):
        for line in m.group(2).split(","):
            parts = line.split()
            if parts and "NOT NULL" not in " ".join(parts).upper():
                out.add((m.group(1), parts[0]))
    return out


# ------------------------------------------------------------------ tokens and parse tree
_TOKEN = re.compile(r"""\s*(?:
    (?P<str>'(?:[^']|'')*')
  | (?P<num>-?\d+)
  | (?P<op><>|!=|<=|>=|=|<|>|\(|\)|,|;|\.)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  )""", re.X)


def tokenize(sql):
    out, pos = [], 0
    sql = sql.rstrip()
    while pos < len(sql):
        m = _TOKEN.match(sql, pos)
        if not m or m.end() == pos:
            if sql[pos:].strip() == "":
                break
            raise SQLError(f"cannot read SQL at: {sql[pos:pos + 30]!r}")
        pos = m.end()
        if m.group("str") is not None:
            out.append(("str", m.group("str")[1:-1].replace("''", "'")))
        elif m.group("num") is not None:
            out.append(("num", int(m.group("num"))))
        elif m.group("op") is not None:
            out.append(("op", m.group("op")))
        else:
            out.append(("id", m.group("id")))
    return out


class _Parser:
    """recursive descent over the token list; expression nodes are tuples:
    ("and", a, b) ("or", a, b) ("not", a) ("cmp", op, col, literal) ("eqcol", col, col)
    ("like", negated, col, pattern) ("in", col, [literals]) ("between", col, lo, hi)
    ("isnull", col); a column is (alias or None, name)"""

    def __init__(self, tokens):
        self.t = tokens
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else ("eof", None)

    def kw(self, word, k=0):
        tok = self.peek(k)
        return tok[0] == "id" and tok[1].upper() == word

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expect_kw(self, word):
        if not self.kw(word):
            raise SQLError(f"expected {word}, got {self.peek()}")
        self.i += 1

    def expect_op(self, op):
        tok = self.take()
        if tok != ("op", op):
            raise SQLError(f"expected {op!r}, got {tok}")

    def colref(self):
        tok = self.take()
        if tok[0] != "id":
            raise SQLError(f"expected a column, got {tok}")
        if self.peek() == ("op", "."):
            self.i += 1
            name = self.take()
            if name[0] != "id":
                raise SQLError(f"expected a column name after {tok[1]}.")
            return (tok[1], name[1])
        return (None, tok[1])

    def literal(self):
        tok = self.take()
        if tok[0] not in ("str", "num"):
            raise SQLError(f"expected a literal, got {tok}")
        return tok[1]

    def select(self):
        self.expect_kw("SELECT")
        items = []
        while True:
            if self.kw("MIN") or self.kw("MAX") or self.kw("COUNT"):
                self.i += 1
                self.expect_op("(")
                items.append(self.colref())
                self.expect_op(")")
            else:
                items.append(self.colref())
            if self.kw("AS"):
                self.i += 2
            if self.peek() != ("op", ","):
                break
            self.i += 1
        self.expect_kw("FROM")
        tables = []
        while True:
            name = self.take()
            if name[0] != "id":
                raise SQLError(f"expected a table, got {name}")
            alias = None
            if self.kw("AS"):
                self.i += 1
            if self.peek()[0] == "id" and not self.kw("WHERE"):
                alias = self.take()[1]
            tables.append((name[1], alias))
            if self.peek() != ("op", ","):
                break
            self.i += 1
        where = None
        if self.kw("WHERE"):
            self.i += 1
            where = self.expr()
        if self.peek() == ("op", ";"):
            self.i += 1
        if self.peek()[0] != "eof":
            raise SQLError(f"unexpected {self.peek()} after the query")
        return items, tables, where

    def expr(self):
        e = self.conj()
        while self.kw("OR"):
            self.i += 1
            e = ("or", e, self.conj())
        return e

    def conj(self):
        e = self.neg()
        while self.kw("AND"):
            self.i += 1
            e = ("and", e, self.neg())
        return e

    def neg(self):
        if self.kw("NOT"):
            self.i += 1
            return ("not", self.neg())
        return self.pred()

    def pred(self):
        if self.peek() == ("op", "("):
            self.i += 1
            e = self.expr()
            self.expect_op(")")
            return e
        col = self.colref()
        tok = self.peek()
        if tok[0] == "op" and tok[1] in ("=", "<>", "!=", "<", ">", "<=", ">="):
            self.i += 1
            op = {"=": "eq", "<>": "neq", "!=": "neq", "<": "lt", ">": "gt", "<=": "leq",
                  ">=": "geq"}[tok[1]]
            if self.peek()[0] == "id":
                other = self.colref()
                if op != "eq":
                    raise SQLError("Non-EuqalJoins are not supported")
                return ("eqcol", col, other)
            return ("cmp", op, col, self.literal())
        negated = False
        if self.kw("NOT"):
            negated = True
            self.i += 1
        if self.kw("LIKE"):
            self.i += 1
            pat = self.literal()
            if not isinstance(pat, str):
                raise SQLError("LIKE needs a string pattern")
            return ("like", negated, col, pat)
        if self.kw("IN"):
            self.i += 1
            self.expect_op("(")
            vals = [self.literal()]
            while self.peek() == ("op", ","):
                self.i += 1
                vals.append(self.literal())
            self.expect_op(")")
            e = ("in", col, vals)
            return ("not", e) if negated else e
        if self.kw("BETWEEN"):
            self.i += 1
            lo = self.literal()
            self.expect_kw("AND")
            hi = self.literal()
            e = ("between", col, lo, hi)
            return ("not", e) if negated else e
        if negated:
            raise SQLError("NOT must precede LIKE, IN or BETWEEN here")
        if self.kw("IS"):
            self.i += 1
            neg2 = False
            if self.kw("NOT"):
                neg2 = True
                self.i += 1
            self.expect_kw("NULL")
            return ("not", ("isnull", col)) if neg2 else ("isnull", col)
        raise SQLError(f"unsupported condition after {col}: {tok}")


# ------------------------------------------------------------------ parsed query
class ParsedSQL:
    """A parsed query over a schema: its entities (a table and its occurrence), aliases, the
    select list, the join classes and graph, and each entity's filter expression"""

    def __init__(self, sql, schema):
        items, tables, where = _Parser(tokenize(sql)).select()
        self.schema = schema
        self.entities = []   # (table, occurrence)
        self.alias = {}      # alias -> entity
        counts = {}
        for name, alias in tables:
            if name not in schema:
                raise SQLError(f"No table: {name} in schema")
            counts[name] = counts.get(name, 0) + 1
            ent = (name, counts[name] - 1)
            self.entities.append(ent)
            if alias:
                self.alias[alias] = ent
        self.counts = counts
        self.output = [self.resolve(c) for c in items]     # [(entity, column)]
        self.filters = {}    # entity -> expression
        self._parent = {}    # join classes (union-find over (entity, column))
        if where is not None:
            self._top(where)
        classes = {}
        for x in list(self._parent):
            classes.setdefault(self._find(x), []).append(x)
        self.join_graph = {}  # entity -> {entity: (my column, its column)}
        for items_ in classes.values():
            # the reference enumerates a class in column order (read_sql.cpp:826-852)
            items_.sort(key=lambda ec: (self.entities.index(ec[0]), self.column_index(*ec)))
            for i in range(len(items_)):
                for j in range(i + 1, len(items_)):
                    (le, lc), (re_, rc) = items_[i], items_[j]
                    if le == re_:
                        continue
                    if re_ in self.join_graph.get(le, {}):
                        raise SQLError("At least two conditions between a same pair of tables.")
                    self.join_graph.setdefault(le, {})[re_] = (lc, rc)
                    self.join_graph.setdefault(re_, {})[le] = (rc, lc)

    def column_index(self, ent, col):
        for i, (name, _) in enumerate(self.schema[ent[0]]):
            if name == col:
                return i
        raise SQLError(f"Cannot find column {col} in table {ent[0]}")

    def column_type(self, ent, col):
        return self.schema[ent[0]][self.column_index(ent, col)][1]

    def resolve(self, colref):
        alias, col = colref
        if alias is not None:
            if alias in self.alias:
                ent = self.alias[alias]
            elif alias in self.counts and self.counts[alias] == 1:
                ent = (alias, 0)
            else:
                raise SQLError(f"Cannot find alias: {alias}")
        else:
            owners = [e for e in self.entities if any(c == col for c, _ in self.schema[e[0]])]
            if len(owners) != 1:
                raise SQLError(f"column {col} is ambiguous or unknown")
            ent = owners[0]
        self.column_index(ent, col)
        return (ent, col)

    # union-find of join classes
    def _find(self, x):
        self._parent.setdefault(x, x)
        while self._parent[x] != x:
            self._parent[x] = self._parent[self._parent[x]]
            x = self._parent[x]
        return x

    def _top(self, e):
        """the top-level AND: join conditions into the classes, filters to their entity"""
        if e[0] == "and":
            self._top(e[1])
            self._top(e[2])
            return
        if e[0] == "eqcol":
            a, b = self.resolve(e[1]), self.resolve(e[2])
            self._parent[self._find(a)] = self._find(b)
            return
        ent, f = self._filter(e)
        if ent in self.filters:
            self.filters[ent] = ("and", self.filters[ent], f)
        else:
            self.filters[ent] = f

    def _filter(self, e):
        """(entity, expression with resolved columns) of a filter below the top level"""
        kind = e[0]
        if kind in ("and", "or"):
            la, a = self._filter(e[1])
            lb, b = self._filter(e[2])
            if la != lb:
                raise SQLError("Filter can not be pushed down")
            return la, (kind, a, b)
        if kind == "not":
            ent, a = self._filter(e[1])
            return ent, ("not", a)
        if kind == "eqcol":
            raise SQLError("Non top level contains join condition instead of filter")
        ent, col = self.resolve(e[2] if kind in ("cmp", "like") else e[1])
        if kind == "cmp":
            return ent, ("cmp", e[1], col, e[3])
        if kind == "like":
            return ent, ("like", e[1], col, e[3])
        if kind == "in":
            return ent, ("in", col, e[2])
        if kind == "between":
            return ent, ("between", col, e[2], e[3])
        if kind == "isnull":
            return ent, ("isnull", col)
        raise SQLError(f"unsupported filter {kind}")

    def program(self, ent):
        """the entity's filter as a postfix predicate program over its table's columns (the
        Plan.new_scan_node preds format), or []"""
        e = self.filters.get(ent)
        if e is None:
            return []
        out = []

        def lit(col, v):
            t = self.column_type(ent, col)
            if t == I32:
                if not isinstance(v, int):
                    raise SQLError(f"string literal compared with INT32 column {col}")
                return int(v)
            return v if isinstance(v, str) else str(v)

        def emit(x):
            k = x[0]
            if k in ("and", "or"):
                emit(x[1])
                emit(x[2])
                out.append((k,))
            elif k == "not":
                emit(x[1])
                out.append(("not",))
            elif k == "cmp":
                out.append((x[1], self.column_index(ent, x[2]), lit(x[2], x[3])))
            elif k == "like":
                out.append(("not_like" if x[1] else "like", self.column_index(ent, x[2]),
                            str(x[3])))
            elif k == "in":
                c = self.column_index(ent, x[1])
                for i, v in enumerate(x[2]):
                    out.append(("eq", c, lit(x[1], v)))
                    if i:
                        out.append(("or",))
            elif k == "between":
                c = self.column_index(ent, x[1])
                out.append(("geq", c, lit(x[1], x[2])))
                out.append(("leq", c, lit(x[1], x[3])))
                out.append(("and",))
            elif k == "isnull":
                out.append(("is_null", self.column_index(ent, x[1])))
            else:
                raise SQLError(f"unsupported filter {k}")

        emit(e)
        return out


# ------------------------------------------------------------------ PostgreSQL plan -> Plan
_PASS = {"Aggregate", "Gather"}
_JOINS = {"Nested Loop", "Hash Join", "Merge Join"}
_SCANS = {"Seq Scan", "Index Only Scan"}


def build_plan(pg, parsed, table_ids):
    """the engine Plan of a PostgreSQL plan node (plans.json's "Plan") for the parsed query;
    table_ids maps a base table name to its index in the plan's inputs (the caller fills the
    inputs). Follows load_join_pipeline (read_sql.cpp:868-1192)."""
    plan = Plan()

    def scan_entity(node):
        if "Alias" in node:
            alias = node["Alias"]
            if alias in parsed.alias:
                return parsed.alias[alias]
            if alias in parsed.counts and parsed.counts[alias] == 1:
                return (alias, 0)
            raise SQLError(f"Cannot find alias: {alias}")
        if "Relation Name" in node:
            rel = node["Relation Name"]
            if parsed.counts.get(rel) == 1:
                return (rel, 0)
            raise SQLError(f"Table {rel} is used at least twice")
        raise SQLError('Neither "Alias" nor "Relation Name" exists in scan node')

    def sides(node):
        kids = node["Plans"]
        lt, rt = kids[0]["Node Type"], kids[1]["Node Type"]
        if lt == "Hash" and rt != "Hash":
            return True, kids[0]["Plans"][0], kids[1]
        if lt != "Hash" and rt == "Hash":
            return False, kids[0], kids[1]["Plans"][0]
        raise SQLError("Hash Join should have at least one Hash child")

    def entities(node):
        t = node["Node Type"]
        if t in _PASS:
            return entities(node["Plans"][0])
        if t in _JOINS:
            if t != "Hash Join":
                raise SQLError("Not Hash Join")
            _, l, r = sides(node)
            return entities(l) | entities(r)
        if t in _SCANS:
            return {scan_entity(node)}
        raise SQLError(f"Not supported node type: {t}")

    def recurse(node, required):
        t = node["Node Type"]
        if t in _PASS:
            return recurse(node["Plans"][0], required)
        if t in _JOINS:
            if t != "Hash Join":
                raise SQLError("Not Hash Join")
            build_left, pl, pr = sides(node)
            le_set, re_set = entities(pl), entities(pr)
            cond = None
            for ent in le_set:  # the last edge found wins, as in the reference's loop
                for adj, cols in parsed.join_graph.get(ent, {}).items():
                    if adj in re_set:
                        cond = (ent, adj, cols)
            if cond is None:
                raise SQLError("Cannot find join condition")
            lent, rent, (lcol, rcol) = cond
            lreq = [x for x in required if x[0] in le_set]
            rreq = [x for x in required if x[0] in re_set]
            if len(lreq) + len(rreq) != len(required):
                raise SQLError("Required attributes cannot be found in neither left child nor "
                               "right child.")
            if (lent, lcol) not in lreq:
                lreq.append((lent, lcol))
            if (rent, rcol) not in rreq:
                rreq.append((rent, rcol))
            left, lcols = recurse(pl, lreq)
            right, rcols = recurse(pr, rreq)
            left_attr = lcols.index((lent, lcol))
            right_attr = [c for c in rcols].index((rent, rcol))
            both = lcols + rcols
            out_cols, out_attrs = [], []
            for r in required:
                i = both.index(r)
                out_cols.append(r)
                out_attrs.append((i, parsed.column_type(*r)))
            node_id = plan.new_join_node(build_left, left, right, left_attr, right_attr, out_attrs)
            return node_id, out_cols
        if t in _SCANS:
            ent = scan_entity(node)
            out_cols, out_attrs = [], []
            for r in required:
                if r[0] != ent:
                    raise SQLError(f"Cannot found the required attr: {r} in children's output")
                out_cols.append(r)
                out_attrs.append((parsed.column_index(*r), parsed.column_type(*r)))
            node_id = plan.new_scan_node(table_ids[ent[0]], out_attrs,
                                         preds=parsed.program(ent))
            return node_id, out_cols
        raise SQLError(f"Not supported node type: {t}")

    plan.root, _ = recurse(pg, list(parsed.output))
    return plan


# ------------------------------------------------------------------ the benchmark's files
def load_benchmark(plans_path, sql_dir=None):
    """[(name, sql text, PostgreSQL plan)] of a plans.json (its sql_directory relative to the
    file unless sql_dir is given)"""
    with open(plans_path) as f:
        d = json.load(f)
    base = sql_dir or os.path.join(os.path.dirname(plans_path), d["sql_directory"])
    out = []
    for name, p in zip(d["names"], d["plans"]):
        with open(os.path.join(base, f"{name}.sql")) as f:
            out.append((name, f.read(), p["Plan"]))
    return out


def load_fixture(path):
    """(schema text, [(name, sql text, PostgreSQL plan)]) of a JSON file holding "schema_sql"
    and "queries" [{"name", "sql", "plan"}] (the benchmark's files in one document)"""
    with open(path) as f:
        d = json.load(f)
    return d["schema_sql"], [(q["name"], q["sql"], q["plan"]) for q in d["queries"]]


def literals(parsed):
    """{(table, column): [literal or LIKE pattern, ...]} that the query's filters name"""
    out = {}

    def add(ent, col, v):
        out.setdefault((ent[0], col), []).append(v)

    def walk(ent, x):
        k = x[0]
        if k in ("and", "or"):
            walk(ent, x[1])
            walk(ent, x[2])
        elif k == "not":
            walk(ent, x[1])
        elif k == "cmp":
            add(ent, x[2], x[3])
        elif k == "like":
            add(ent, x[2], ("like", x[3]))
        elif k == "in":
            for v in x[2]:
                add(ent, x[1], v)
        elif k == "between":
            add(ent, x[1], x[2])
            add(ent, x[1], x[3])
        elif k == "isnull":
            out.setdefault((ent[0], x[1]), [])

    for ent, e in parsed.filters.items():
        walk(ent, e)
    return out


def translate(sql, pg_plan, schema, tables=None):
    """(Plan, table order, parsed query): the query's plan with one input per base table it
    reads, in the returned order; `tables` (name -> ColumnarTable or resident table) fills the
    plan's inputs when given"""
    parsed = ParsedSQL(sql, schema)
    order = []
    for t, _ in parsed.entities:
        if t not in order:
            order.append(t)
    ids = {t: i for i, t in enumerate(order)}
    plan = build_plan(pg_plan, parsed, ids)
    plan.inputs = [tables[t] for t in order] if tables is not None else [None] * len(order)
    return plan, order, parsed
