"""Writes the query-count known-answer fixture from the reference's own
test data: tests/scripts/query-count.list (query -> hit count) over
tests/test0.csv and tests/test1.csv.  Only the queries that are
conjunctions of range terms on single columns are kept (no functions,
arithmetic, OR or IN lists); each term becomes lo OP1 x OP2 hi.  The
columns are stored with their values (int32 when every value in the file is
integral as written, else float32); a column with NULL entries is left out,
and so are the queries that use it.

    python tests/golden/make_kat.py
"""
from __future__ import annotations

import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests"

# Two entries count 98 where the rows hold 99 matches (i0 and i2 are
# permutations of 0..99).  The reference library's relic and bin indexes
# answer 99 for them (checked with oracle/_ref); the listed 98 comes from the
# index the command-line tool builds by default, which is not on this path.
EXCLUDE = {"-3 < i0 < 99", "-5 < i2 < 99"}

NUM = r"[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"
COL = r"[A-Za-z_]\w*"
OPS = {"<": 1, "<=": 2}


def load_csv(path):
    lines = [ln.strip() for ln in open(path) if ln.strip()]
    names = [x.strip().strip('"') for x in lines[0].lstrip("#").split(",")]
    raw = [[x.strip() for x in ln.split(",")] for ln in lines[1:]]
    cols = {}
    for j, nm in enumerate(names):
        vals = [r[j] for r in raw]
        if any(v == "" for v in vals):
            continue            # a column with NULLs (no null masks in this engine)
        is_float = any(("." in v or "e" in v.lower()) for v in vals)
        cols[nm] = {"type": "float32" if is_float else "int32",
                    "values": [float(v) if is_float else int(v) for v in vals]}
    return cols


def parse_term(t):
    t = t.strip()
    m = re.fullmatch(rf"({NUM})\s*(<=|<)\s*({COL})\s*(<=|<)\s*({NUM})", t)
    if m:
        return m.group(3), float(m.group(1)), OPS[m.group(2)], float(m.group(5)), OPS[m.group(4)]
    m = re.fullmatch(rf"({COL})\s*(<=|<|>=|>|==|=)\s*({NUM})", t)
    if m:
        c, op, v = m.group(1), m.group(2), float(m.group(3))
        if op in ("=", "=="):
            return c, v, 2, v, 2
        if op == "<":
            return c, 0.0, 0, v, 1
        if op == "<=":
            return c, 0.0, 0, v, 2
        if op == ">":
            return c, v, 1, 0.0, 0
        return c, v, 2, 0.0, 0
    return None


def main():
    tables = {"test0": load_csv(os.path.join(SRC, "test0.csv")),
              "test1": load_csv(os.path.join(SRC, "test1.csv"))}
    queries = []
    for ln in open(os.path.join(SRC, "scripts", "query-count.list")):
        ln = ln.strip()
        if not ln or ln.startswith("#"):
            continue
        q, cnt = ln.rsplit(";", 1)
        m = re.fullmatch(r"from\s+(\w+)\s+where\s+(.*)", q.strip())
        if not m or m.group(1) not in tables:
            continue
        expr = m.group(2)
        if expr.strip() in EXCLUDE:
            continue
        if re.search(r"\b(or|OR|in|IN)\b|\|\||[()*/]|\*\*|sqrt|log|exp", expr):
            continue
        terms = [parse_term(t) for t in re.split(r"\s+(?:and|AND)\s+|\s*&&\s*", expr)]
        if any(t is None for t in terms) or any(t[0] not in tables[m.group(1)] for t in terms):
            continue
        queries.append({"table": m.group(1), "text": expr, "terms": terms, "count": int(cnt)})
    with open(os.path.join(HERE, "kat_query_count.json"), "w") as f:
        json.dump({"tables": tables, "queries": queries}, f)
    print(len(queries), "queries")


if __name__ == "__main__":
    main()
