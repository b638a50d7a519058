"""Writes tests/golden/job_queries.json from the reference's Join Order Benchmark files:
job/schema.sql (the CREATE TABLE statements, kept as text), job/<name>.sql (the 113 queries)
and plans.json (PostgreSQL EXPLAIN plans, of which only the keys the plan loader reads are kept:
"Node Type", "Plans", "Alias", "Relation Name").

    python tests/golden/make_job_fixture.py [/root/reference]
"""
import json
import os
import sys

KEEP = ("Node Type", "Alias", "Relation Name")


def strip(node):
    out = {k: node[k] for k in KEEP if k in node}
    if "Plans" in node:
        out["Plans"] = [strip(p) for p in node["Plans"]]
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "plans.json")) as f:
        d = json.load(f)
    sql_dir = os.path.join(ref, d["sql_directory"])
    queries = []
    for name, p in zip(d["names"], d["plans"]):
        with open(os.path.join(sql_dir, f"{name}.sql")) as f:
            queries.append({"name": name, "sql": f.read().strip(), "plan": strip(p["Plan"])})
    with open(os.path.join(sql_dir, "schema.sql")) as f:
        schema = f.read()
    out = {"source": "job/schema.sql, job/*.sql and plans.json of the reference",
           "schema_sql": schema, "queries": queries}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "job_queries.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, separators=(",", ":"))
        f.write("\n")
    print(path, len(queries), "queries")


if __name__ == "__main__":
    main()
