"""SQL front end on the CPU: parsing, exact fixed-point compilation and the
result finishing (no device needed: plans are built and their HIP source and
layout inspected; finishing runs on hand-made aggregate columns).

The reference parses SQL with libpg_query (codegen/planning/query_analyzer.cpp)
and plans the analysis-based shapes from it (metal_plan_builder.cpp:253-360);
here the engine's own parser covers single-table SELECTs.
"""

from __future__ import annotations

from fractions import Fraction
from pathlib import Path

import pytest

from gpudbmetalcodegen_amd import _native as N
from gpudbmetalcodegen_amd import engine as E
from gpudbmetalcodegen_amd import sql as S

SQL_DIR = Path(__file__).resolve().parent / "sql"
Q1 = (SQL_DIR / "tpch_q1.sql").read_text()
Q6 = (SQL_DIR / "tpch_q6.sql").read_text()


def test_q6_constants_fold_exactly():
    p = E.build_sql_plan(Q6)
    d = p.describe()
    # dates folded with calendar arithmetic; 0.06 +- 0.01 in hundredths; 24 -> 2400 hundredths
    assert "(19940101ll)" in d and "(19950101ll)" in d
    assert "(5ll)" in d and "(7ll)" in d and "(2400ll)" in d
    lay = p.sql_layout()
    assert lay["table"] == "lineitem" and lay["keys"] == []
    assert lay["targets"] == [{"name": "revenue", "expr": {"agg": "a1"}, "scale": 4, "type": "num"}]
    assert p.phases() == ["SQL_reduce"]


def test_q1_layout_and_packed_keys():
    p = E.build_sql_plan(Q1)
    assert p.phases() == ["SQL_groups", "SQL_aggregate"]
    d = p.describe()
    assert "(19980902ll)" in d                  # 1998-12-01 - 90 days
    assert "RowCheck" not in d                  # two char(1) keys pack into 16 bits
    lay = p.sql_layout()
    names = [t["name"] for t in lay["targets"]]
    assert names == ["l_returnflag", "l_linestatus", "sum_qty", "sum_base_price", "sum_disc_price", "sum_charge",
                     "avg_qty", "avg_price", "avg_disc", "count_order"]
    scales = {t["name"]: t["scale"] for t in lay["targets"]}
    assert scales["sum_disc_price"] == 4 and scales["sum_charge"] == 6 and scales["avg_qty"] == 8
    assert [o["desc"] for o in lay["order"]] == [False, False]


@pytest.mark.parametrize("expr,date", [
    ("date '1998-12-01' - interval '90' day", 19980902),
    ("date '1995-01-31' + interval '1' month", 19950228),   # clamped to the month's end
    ("date '1996-01-31' + interval '1' month", 19960229),   # leap year
    ("date '1994-03-01' - interval '1' day", 19940228),
    ("date '1993-10-01' + interval '3' month", 19940101),
    ("date '2000-02-29' + interval '1' year", 20010228),
])
def test_date_interval_folding(expr, date):
    p = E.build_sql_plan(f"select count(*) from orders where o_orderdate < {expr}")
    assert f"({date}ll)" in p.describe()


def test_string_predicates_compile():
    p = E.build_sql_plan("select count(*) from part where p_type like 'PROMO%' and p_brand <> 'Brand#45' "
                         "and p_container in ('SM BOX', 'LG CASE') and p_name not like '%gre_n%' "
                         "and p_mfgr > 'Manufacturer#2'")
    d = p.describe()
    assert 'gdb_like_prefix(p_type, 25, "PROMO", 5)' in d
    assert '!gdb_str_eq_lit(p_brand, 10, "Brand#45", 8)' in d
    assert 'gdb_str_eq_lit(p_container, 10, "SM BOX", 6)' in d
    assert '!gdb_like_match(p_name, 55, "%gre_n%", 7)' in d
    assert '(gdb_str_cmp_lit(p_mfgr, 25, "Manufacturer#2", 14) > 0)' in d


def test_quotes_and_escapes_in_literals():
    p = E.build_sql_plan("select count(*) from customer where c_name = 'O''Brien\"x'")
    assert r'"O\047Brien\042x", 9' in p.describe()


def test_join_plans_split_conjuncts():
    p = E.build_sql_plan((SQL_DIR / "tpch_q12.sql").read_text())
    d = p.describe()
    assert p.phases() == ["SQL_join0_build", "SQL_groups", "SQL_aggregate"]
    assert "GridStrideScan(orders: o_orderkey/i32)" in d          # build side: the key (and its filters)
    assert "BindColumns(orders: o_orderpriority)" in d              # read at the matched row
    assert '"MA", 2' in d and '"SH", 2' in d                          # literals cut to the 2-byte catalog width
    assert "(49ll)" in p.source()                                     # '1-URGENT' by its first byte ('1' = 49)
    lay = p.sql_layout()
    assert lay["table"] == "lineitem"
    assert lay["join"] == [{"table": "orders", "keys": "sql_join0_keys", "column": "o_orderkey"}]
    assert "DenseUniqueStore(d_sql_join0[(i64)o_orderkey] = i + 1)" in d   # row + 1 by key, repeats fail
    q = E.build_sql_plan("select count(*) from lineitem, part where p_partkey = l_partkey and p_size > 10 "
                         "and l_quantity < 5 and (p_size < 20 or l_tax > 0.02)")
    d = q.describe()
    assert "GridStrideScan(part: p_partkey/i32 p_size/i32)" in d     # p_size > 10 filters the build
    assert "(2ll)" in d                                               # cross conjunct after the join


def test_case_when_aligns_scales():
    p = E.build_sql_plan("select sum(case when l_shipmode = 'MAIL' then l_extendedprice * (1 - l_discount) "
                         "else 0 end) as r from lineitem")
    lay = p.sql_layout()
    assert lay["targets"][0]["scale"] == 4


@pytest.mark.parametrize("sql,msg", [
    ("select * from lineitem", "SELECT *"),
    ("select sum(l_quantity) from lineitem, orders, part where l_orderkey = o_orderkey",
     "without a join condition"),
    ("select count(*) from customer, orders, nation, supplier where c_custkey = o_custkey "
     "and c_nationkey = n_nationkey and s_nationkey = n_nationkey", "joined by its key twice"),
    ("select count(*) from lineitem, orders where l_quantity < 3", "no cross products"),
    ("select count(*) from lineitem, partsupp where l_partkey = ps_partkey", "neither is its table's key"),
    ("select count(*) from lineitem, orders where l_orderkey = o_orderkey and l_suppkey = o_custkey",
     "neither is its table's key"),
    ("select count(*) from lineitem, orders where l_orderkey = o_orderkey and l_partkey = o_orderkey",
     "joined by its key twice"),
    ("select p_type, count(*) from lineitem, part where l_partkey = p_partkey group by p_type",
     "a string of a joined table"),
    ("select l_quantity from lineitem", "must be an aggregate or appear in GROUP BY"),
    ("select sum(l_quantity / 2) from lineitem", "division of row values"),
    ("select sum(nope) from lineitem", "no column nope"),
    ("select count(distinct l_suppkey) from lineitem", "COUNT(DISTINCT)"),
    ("select count(*) from orders where exists (select * from orders where o_custkey = o_orderkey)",
     "outer query's own table"),
    ("select count(*) from orders where exists (select * from lineitem where l_quantity > 3)",
     "EXISTS without an equality"),
    ("select count(*) from orders where o_orderkey in (select l_orderkey, l_suppkey from lineitem)",
     "selects one column"),
    ("select count(*) from orders where exists (select * from lineitem where l_orderkey = o_orderkey) "
     "or o_orderkey < 3", "only as AND conditions"),
    ("select count(*) from orders where o_totalprice in (select l_extendedprice from lineitem)",
     "integer columns"),
    ("select count(*) from orders where o_orderkey in (select l_orderkey from lineitem group by l_orderkey)",
     "without GROUP BY"),
    ("select sum(l_quantity) from lineitem where l_shipdate < '1994-13-01'", "bad date literal"),
    ("select sum(l_quantity) from nosuch", "unknown table"),
    ("select sum(l_quantity) from lineitem where l_comment", "boolean"),
    ("select sum(l_quantity) from lineitem group by l_returnflag having l_quantity > 1",
     "must be an aggregate or appear in GROUP BY"),
    ("select sum(l_quantity) from lineitem group by l_returnflag having sum(l_quantity)", "HAVING needs a boolean"),
    ("select sum(l_quantity) from lineitem where l_shipdate = 3", "comparison of date with number"),
    ("select sum(l_quantity) from lineitem limit 2.5", "LIMIT takes an integer"),
])
def test_unsupported_sql_is_named(sql, msg):
    with pytest.raises(N.GdbError) as ei:
        E.build_sql_plan(sql)
    assert msg in str(ei.value)


def test_finish_orders_limits_and_averages():
    layout = {"keys": [{"agg": "k1", "type": "char1", "scale": 0}], "rep": "",
              "targets": [{"name": "f", "expr": {"key": 0}, "scale": 0, "type": "char1"},
                          {"name": "s", "expr": {"agg": "a0"}, "scale": 2, "type": "num"},
                          {"name": "avg", "expr": {"avg": ["a0", "_n"], "scale": 8}, "scale": 8, "type": "num"},
                          {"name": "ratio", "expr": {"op": "/", "args": [{"agg": "a0"}, {"agg": "_n"}], "scale": 8},
                           "scale": 8, "type": "num"}],
              "order": [{"by": {"target": 1}, "desc": True}], "limit": 2}
    cols = {"_n": [3, 1, 2], "a0": [100, -7, 100], "k1": [ord("A"), ord("N"), ord("R")],
            "__scales": {"a0": 2, "_n": 0}}
    out, fmts = S.finish(layout, cols, None)
    # ties on s broken by the remaining columns ascending
    assert out["f"] == ["A", "R"] and out["s"] == [100, 100]
    assert out["avg"] == [33333333, 50000000]      # 1.00 / 3 truncated to 8 digits
    assert out["ratio"] == out["avg"]
    assert fmts["s"] == (2, False) and fmts["avg"] == (8, False)
    neg = S._trunc(Fraction(-7, 300), 8)
    assert neg == -2333333                          # truncation toward zero


def test_finish_string_keys_from_representative_rows():
    import numpy as np

    host = {"p_brand": np.frombuffer(b"Brand#11\0\0Brand#22\0\0", np.uint8).reshape(2, 10)}
    layout = {"keys": [{"column": "p_brand", "type": "str", "scale": 0}], "rep": "_rep2",
              "targets": [{"name": "p_brand", "expr": {"key": 0}, "scale": 0, "type": "str"},
                          {"name": "c", "expr": {"agg": "_n"}, "scale": 0, "type": "num"}],
              "order": [], "limit": -1}
    cols = {"_n": [5, 7], "_rep2": [1, 0], "__scales": {}}
    out, _ = S.finish(dict(layout, table="part"), cols, {"part": host})
    assert out == {"p_brand": ["Brand#11", "Brand#22"], "c": [7, 5]}


def test_cli_sql_usage_errors(tmp_path):
    from gpudbmetalcodegen_amd import tpch_cli

    assert tpch_cli.main(["--sql", str(tmp_path / "missing.sql")]) == 3
    assert tpch_cli.main([]) == 3


def test_having_filters_finished_groups():
    p = E.build_sql_plan("select l_returnflag, count(*) from lineitem group by l_returnflag "
                         "having sum(l_quantity) > 2.5 and not l_returnflag = 'N' order by l_returnflag")
    lay = p.sql_layout()
    assert lay["having"]["op"] == "and"
    cols = {"_n": [4, 2, 9], "k2": [ord("A"), ord("N"), ord("R")], "a1": [300, 900, 250],
            "__scales": {"a1": 2, "_n": 0}}
    key = lay["keys"][0]["agg"]
    cols[key] = cols.pop("k2")
    sum_name = lay["having"]["args"][0]["args"][0]["agg"]
    cols[sum_name] = cols.pop("a1")
    out, _ = S.finish(lay, cols, None)
    assert out == {"l_returnflag": ["A"], "count(*)": [4]}   # N excluded by name, R by its sum (2.50)


def test_exists_and_not_in_become_semi_and_anti_joins():
    p = E.build_sql_plan((SQL_DIR / "tpch_q4.sql").read_text())
    assert p.phases() == ["SQL_semi0_build", "SQL_groups", "SQL_aggregate"]
    d = p.describe()
    assert "ByteFlagBuild(d_sql_semi0[(i64)l_orderkey] = 1)" in d
    assert "Selection(((i64)l_commitdate < (i64)l_receiptdate))" in d   # the subquery's own condition
    assert p.sql_layout()["semi"] == [{"table": "lineitem", "keys": "sql_semi0_keys", "column": "l_orderkey",
                                       "anti": False}]
    q = E.build_sql_plan("select count(*) from customer where c_custkey not in (select o_custkey from orders)")
    assert "AntiByteFlagProbe" in q.describe() and q.sql_layout()["semi"][0]["anti"] is True
    r = E.build_sql_plan("select count(*) from customer where not exists "
                         "(select * from orders where o_custkey = c_custkey)")
    assert "AntiByteFlagProbe" in r.describe()


def test_single_int_key_is_direct_addressed():
    p = E.build_sql_plan((SQL_DIR / "tpch_q10.sql").read_text())
    lay = p.sql_layout()
    assert lay["domain"] == {"table": "customer", "column": "c_custkey"}
    assert "Hash64" not in p.describe()              # joins and groups all key-indexed
    assert [k.get("via") for k in lay["keys"]] == [None, 0, 0, 0, 0, 0, 0]
    assert lay["keys"][4]["path"] == [{"table": "customer", "key": "c_custkey"},
                                      {"table": "nation", "key": "n_nationkey", "fk": "c_nationkey"}]


@pytest.mark.parametrize("mb", sorted(S.MICROBENCH))
def test_microbenchmarks_plan_as_one_fused_reduction(mb):
    """The seven scan microbenchmarks each plan as one scan-filter-reduce
    phase that reads exactly the listed lineitem columns."""
    p = E.build_sql_plan(S.MICROBENCH[mb])
    assert p.phases() == ["SQL_reduce"]
    src = p.source()
    for c in S.MICROBENCH_COLUMNS[mb]:
        assert c in src, (mb, c)
    others = {"l_shipdate", "l_quantity", "l_extendedprice", "l_discount", "l_tax", "l_returnflag"}
    for c in others - set(S.MICROBENCH_COLUMNS[mb]):
        assert f"p_{c}" not in src and f" {c}[" not in src, (mb, c)


def test_global_aggregate_over_no_rows_is_one_null_row():
    """SELECT SUM(x), COUNT(*) ... WHERE <nothing qualifies>: one row, the sum
    NULL and the count 0 (the collector holds no group for it)."""
    layout = {"keys": [], "targets": [{"name": "s", "type": "num", "scale": 2, "expr": {"agg": "a1"}},
                                      {"name": "c", "type": "num", "scale": 0, "expr": {"agg": "_n"}},
                                      {"name": "avg_s", "type": "num", "scale": 8,
                                       "expr": {"op": "/", "args": [{"agg": "a1"}, {"agg": "_n"}], "scale": 8}}],
              "order": [], "limit": -1}
    out, fmts = S.finish(layout, {"_n": [], "a1": []}, None)
    assert out == {"s": [None], "c": [0], "avg_s": [None]}
    from gpudbmetalcodegen_amd import tpch as T
    assert T.canonical(out, fmts) == "s,c,avg_s\n,0,\n"


def test_q7_text_plans_as_the_named_shape_with_its_constants():
    """TPC-H Q7's text (a derived table over nation twice) is recognised by
    shape, as the reference's buildQ7Plan detector does, and planned as the
    named Q7 plan with the nations and dates read from the text."""
    sql = (SQL_DIR / "tpch_q7.sql").read_text()
    lay = E.build_sql_plan(sql).sql_layout()
    assert lay["named"] == "Q7"
    assert lay["params"] == {"strings": ["FRANCE", "GERMANY"], "q7_date_lo": 19950101, "q7_date_hi": 19961231,
                             "q7_year0": 1995}
    assert lay["names"] == ["supp_nation", "cust_nation", "l_year", "revenue"]
    other = sql.replace("FRANCE", "JAPAN").replace("GERMANY", "CHINA").replace("1995-01-01", "1993-01-01")
    other = other.replace("1996-12-31", "1994-06-30")
    lay = E.build_sql_plan(other).sql_layout()
    assert lay["params"]["strings"] == ["JAPAN", "CHINA"] and lay["params"]["q7_date_lo"] == 19930101
    assert lay["params"]["q7_year0"] == 1993 and lay["params"]["q7_date_hi"] == 19940630
    # the plan buckets two years: a wider range is refused, not aliased
    with pytest.raises(N.GdbError, match="two consecutive years"):
        E.build_sql_plan(sql.replace("1995-01-01", "1994-01-01"))


@pytest.mark.parametrize("sql,msg", [
    ("select a, sum(b) from (select l_orderkey as a, l_quantity as b from lineitem) as t group by a",
     "TPC-H Q7's shape"),
    ("select count(*) from lineitem, (select * from orders) as o", "only FROM item"),
])
def test_other_derived_tables_are_refused_by_name(sql, msg):
    with pytest.raises(N.GdbError, match=msg):
        E.build_sql_plan(sql)
