"""SQL text -> GPU plan -> exact result rows.

``run_sql`` plans a SELECT over TPC-H tables joined in a tree of key
equalities from its text (``gdb_plan_build_sql``:
the engine's own parser and fixed-point compiler, the counterpart of the
reference's analyzeSQL + analysis-based plan builders,
``codegen/planning/query_analyzer.cpp:640-727``), runs it on the TPC-H tables
of a ``TpchDB`` and finishes the rows the way the plan's layout says: group
keys (aggregate columns, or string columns attached from each group's
representative row), the select list evaluated exactly over the aggregate
columns (rationals; a column is shown with its fixed-point digits, an
average or quotient with six more), HAVING, ORDER BY, LIMIT. A GROUP BY with more
groups than the plan's dense capacity is rebuilt once, sized by the exact
group count the first run took.
"""

from __future__ import annotations

import functools
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from . import _native as N
from . import engine as E
from . import tpch as T


@dataclass
class SqlRun:
    sql: str
    columns: dict      # select-list name -> list of values (int, or str)
    formats: dict      # name -> (scale digits, is_date)
    timing: E.Timing
    plan: E.HipQueryPlan = None
    layout: dict = field(default_factory=dict)
    canonical: str = ""


_CMP = {"==": lambda a, b: a == b, "!=": lambda a, b: a != b, "<": lambda a, b: a < b,
        "<=": lambda a, b: a <= b, ">": lambda a, b: a > b, ">=": lambda a, b: a >= b}


def _trunc(v: Fraction, scale: int) -> int:
    x = v * (10 ** scale)
    q = abs(x.numerator) // x.denominator
    return q if x >= 0 else -q


def _text(row) -> str:
    return bytes(row).split(b"\0", 1)[0].decode()


def _value(table: dict, column: str, row: int, k: dict):
    v = table[column][row]
    if k["type"] == "str":
        return _text(v)
    if k["type"] == "char1":
        return chr(int(v)) if int(v) else ""
    if k["type"] == "date":
        return int(v)
    return Fraction(int(v), 10 ** k["scale"])


def finish(layout: dict, cols: dict, host: dict | None) -> tuple[dict, dict]:
    """Result columns of a SQL plan (the collector's aggregate columns) ->
    the select list's columns and formats, ordered and limited. host holds
    the tables ({table: {column: array}}) that string keys and keys fixed by
    another key are attached from."""
    n = len(cols["_n"]) if "_n" in cols else 0
    keys = layout["keys"]
    rep = layout.get("rep") or ""
    index: dict = {}  # (table, key column) -> {key value: row}

    def row_of(table: str, key: str, value: int) -> int:
        ix = index.get((table, key))
        if ix is None:
            ix = index[(table, key)] = {int(v): i for i, v in enumerate(host[table][key].tolist())}
        return ix[int(value)]

    def key_value(j: int, r: int):
        k = keys[j]
        if host is None and ("via" in k or "column" in k):
            raise ValueError("these GROUP BY keys need the host tables to attach")
        if "via" in k:  # fixed by key k["via"]: walk its row's foreign keys
            v = key_value(k["via"], r)
            path = k["path"]
            row = row_of(path[0]["table"], path[0]["key"], int(v))
            for prev, step in zip(path, path[1:]):
                row = row_of(step["table"], step["key"], int(host[prev["table"]][step["fk"]][row]))
            return _value(host[path[-1]["table"]], k["column"], row, k)
        if k["type"] == "str" and "column" in k:
            return _text(host[layout["table"]][k["column"]][int(cols[rep][r])])
        v = int(cols[k["agg"]][r])
        if k["type"] == "str":  # a short string, its bytes packed (byte 0 lowest)
            return v.to_bytes(8, "little")[:k["width"]].split(b"\0", 1)[0].decode()
        if k["type"] == "char1":
            return chr(v) if v else ""
        if k["type"] == "date":
            return v
        return Fraction(v, 10 ** k["scale"])

    def agg_scale(name: str) -> int:
        return int(cols.get("__scales", {}).get(name, 0))

    # An aggregate without GROUP BY over no rows is still one row (SQL):
    # COUNT 0, every other aggregate NULL (None), and NULL propagates.
    empty_global = n == 0 and not keys

    def ev(e: dict, r: int):
        if "key" in e:
            return key_value(e["key"], r)
        if "agg" in e:
            if empty_global:
                return Fraction(0) if e["agg"] == "_n" else None
            return Fraction(int(cols[e["agg"]][r]), 10 ** agg_scale(e["agg"]))
        if "avg" in e:
            if empty_global:
                return None
            num, den = e["avg"]
            return Fraction(int(cols[num][r]), 10 ** agg_scale(num)) / int(cols[den][r])
        if "const" in e:
            return Fraction(int(e["const"]), 10 ** e["scale"])
        if "str" in e:
            return e["str"]
        if "date" in e:
            return e["date"]
        op = e["op"]
        if op in ("and", "or"):  # short-circuit, as SQL may
            vals = (ev(a, r) for a in e["args"])
            return all(vals) if op == "and" else any(vals)
        args = [ev(a, r) for a in e["args"]]
        if any(a is None for a in args):
            return None
        if op == "not":
            return not args[0]
        if op in _CMP:
            return _CMP[op](args[0], args[1])
        if op == "neg":
            return -args[0]
        if op == "+":
            return args[0] + args[1]
        if op == "-":
            return args[0] - args[1]
        if op == "*":
            return args[0] * args[1]
        if args[1] == 0:
            raise ZeroDivisionError("division by zero in the select list")
        return args[0] / args[1]

    targets = layout["targets"]
    having = layout.get("having")
    rows = []
    for r in range(1 if empty_global else n):
        if having is not None and not ev(having, r):
            continue
        vals = [ev(t["expr"], r) for t in targets]
        order = []
        for o in layout["order"]:
            b = o["by"]
            order.append(vals[b["target"]] if "target" in b else ev(b["expr"], r))
        rows.append((vals, order))

    desc = [o["desc"] for o in layout["order"]]

    def cmp(a, b):
        for x, y, d in zip(a[1], b[1], desc):
            if x is None or y is None:
                continue
            if x != y:
                return (1 if x < y else -1) if d else (-1 if x < y else 1)
        for x, y in zip(a[0], b[0]):  # deterministic order for ties and no ORDER BY
            if x is not None and y is not None and x != y:
                return -1 if x < y else 1
        return 0

    rows.sort(key=functools.cmp_to_key(cmp))
    if layout["limit"] >= 0:
        rows = rows[:layout["limit"]]
    out, fmts = {}, {}
    for c, t in enumerate(targets):
        name = t["name"]
        if t["type"] in ("str", "char1"):
            out[name] = [v[c] for v, _ in rows]
            fmts[name] = (0, False)
        elif t["type"] == "date":
            out[name] = [int(v[c]) for v, _ in rows]
            fmts[name] = (0, True)
        else:
            out[name] = [None if v[c] is None else _trunc(v[c], t["scale"]) for v, _ in rows]
            fmts[name] = (t["scale"], False)
    return out, fmts


def run_sql(ctx: E.Context, db: T.TpchDB, sql: str, warmup: int = 0, measured: int = 1, options: str = "",
            executor: E.HipGenericExecutor | None = None) -> SqlRun:
    """Plan, run and finish one SELECT over the TPC-H tables of db."""
    opts = options
    first = E.build_sql_plan(sql, opts)
    lay = first.sql_layout()
    if "named" in lay:
        # a shape the front end maps to a named plan (TPC-H Q7): run it with
        # the constants taken from the text, then rename to the select list
        prm = dict(lay["params"])
        if "strings" in prm:
            prm["strings"] = tuple(x.encode() for x in prm["strings"])
        r = T.run(ctx, db, lay["named"], warmup=warmup, measured=measured, executor=executor, params=prm)
        out = {new: r.columns[old] for new, old in zip(lay["names"], T.SELECT[lay["named"]])}
        fmts = {new: r.formats[old] for new, old in zip(lay["names"], T.SELECT[lay["named"]])}
        run = SqlRun(sql, out, fmts, r.timing, r.plan, lay)
        run.canonical = T.canonical(out, fmts)
        return run
    for _ in range(8):
        plan = E.build_sql_plan(sql, opts).compile(ctx)
        layout = plan.sql_layout()
        table = layout["table"]
        if table not in db.host:
            raise ValueError(f"table {table} is not loaded")
        rows = int(next(iter(db.host[table].values())).shape[0])
        ex = executor or E.HipGenericExecutor(ctx)
        ex.register_table(db.device(table))
        ex.register_symbol("sql_slots", T._pow2(2 * max(rows, 1)))
        # Tables of the joins and subqueries, and the key domains (largest
        # value + 1) their key-indexed arrays are sized by.
        doms = [(e["table"], e["column"], e["keys"]) for e in layout.get("join", []) + layout.get("semi", [])]
        if layout.get("domain"):
            doms.append((layout["domain"]["table"], layout["domain"]["column"], "sql_key_domain"))
        for tab, col, sym in doms:
            if tab not in db.host:
                raise ValueError(f"table {tab} is not loaded")
            ex.register_table(db.device(tab))
            v = db.host[tab][col]
            if v.shape[0] and int(v.min()) < 0:
                raise ValueError(f"{col}: negative key values")
            ex.register_symbol(sym, int(v.max(initial=0)) + 1)
        try:
            t = ex.execute(plan, warmup, measured)
        except N.GdbOverflow:
            # word 3 of the group table's status counted every first arrival:
            # the exact number of groups, so one rebuild suffices
            groups = int(ex.read_buffer("d_sql_groups_status", 32, np.uint64)[3])
            if groups <= layout["groups"]:
                raise
            opts = ",".join(o for o in (options, f"groups={T._pow2(groups)}") if o)
            continue
        res = ex.collect(plan)
        break
    else:
        raise RuntimeError("SQL: group capacity did not converge")
    cols = {k: list(v) for k, v in res.columns.items()}
    cols["__scales"] = {k: f[0] for k, f in res.formats.items()}
    out, fmts = finish(layout, cols, db.host)
    run = SqlRun(sql, out, fmts, t, plan, layout)
    run.canonical = T.canonical(out, fmts)
    return run


# The reference's seven scan microbenchmarks over lineitem (sql/mb1.sql ...
# sql/mb7.sql; dispatched as chunkable single-table reductions,
# codegen/planning/metal_plan_builder.cpp:2104): column footprint, filter
# selectivity and arithmetic against bandwidth. Each is one fused
# scan-filter-reduce kernel here. MB1 sums l_returnflag's byte codes.
MICROBENCH = {
    "MB1": "select sum(l_returnflag) as s from lineitem",
    "MB2": "select sum(l_extendedprice) as total from lineitem",
    "MB3": "select sum(l_extendedprice) as revenue from lineitem "
           "where l_shipdate >= date '1994-01-01' and l_shipdate < date '1996-01-01'",
    "MB4": "select sum(l_extendedprice) as revenue from lineitem "
           "where l_shipdate >= date '1998-09-01' and l_shipdate < date '1998-10-01'",
    "MB5": "select sum(l_extendedprice) as revenue from lineitem where l_shipdate >= date '1992-01-01'",
    "MB6": "select sum(l_extendedprice * (1 - l_discount) * (1 + l_tax)) as charge from lineitem",
    "MB7": "select sum(l_quantity + l_extendedprice + l_discount + l_tax) as s from lineitem "
           "where l_shipdate >= date '1992-01-01'",
}

# lineitem columns each microbenchmark reads (its algorithmic bytes per row)
MICROBENCH_COLUMNS = {
    "MB1": ["l_returnflag"], "MB2": ["l_extendedprice"], "MB3": ["l_shipdate", "l_extendedprice"],
    "MB4": ["l_shipdate", "l_extendedprice"], "MB5": ["l_shipdate", "l_extendedprice"],
    "MB6": ["l_extendedprice", "l_discount", "l_tax"],
    "MB7": ["l_shipdate", "l_quantity", "l_extendedprice", "l_discount", "l_tax"],
}
