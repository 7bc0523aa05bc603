"""TPC-H ingest (.tbl -> .colbin v2), the golden comparator, canonical CSV
and the TPC-H oracle (CPU).

The .colbin layout is checked byte by byte with struct against the format of
the reference (src/infra.h:511-771); the comparator against the reference's
own golden files (tests/golden/tpch, DuckDB results at SF1); the oracle of the
22 named queries against SQLite running the SQL text on the same rows.
"""

from __future__ import annotations

import os
import sqlite3
import struct
from pathlib import Path

import numpy as np
import pytest

from gpudbmetalcodegen_amd import tpch as T
from gpudbmetalcodegen_amd import tpch_gen as G
from oracle import tpch as OT

GOLDEN = Path(__file__).resolve().parent / "golden" / "tpch"


def q19_parts(data: dict) -> dict:
    """Give a fifth of the parts each of Q19's three brand/container/size
    classes and another fifth Q17's brand and container, so the small
    generated data selects Q17 and Q19 rows."""
    P = data["part"]
    k = P["p_partkey"] % 5
    for c, (brand, cont, size) in enumerate([("Brand#12", "SM BOX", 3), ("Brand#23", "MED BAG", 7),
                                             ("Brand#34", "LG PKG", 12), ("Brand#23", "MED BOX", None)]):
        sel = k == c
        P["p_brand"][sel] = G.fixed([brand], 10)[0]
        P["p_container"][sel] = G.fixed([cont], 10)[0]
        if size is not None:
            P["p_size"][sel] = size
    return data


def q16_rows(data: dict) -> dict:
    """Q16 cases the small generated data rarely holds: one group (brand,
    type, size) shared by every eleventh part, so many of its parts have the
    same suppliers (COUNT DISTINCT); parts of the excluded brand and type
    prefix with a size of the set; every ninth supplier complained about."""
    P, S = data["part"], data["supplier"]
    k = P["p_partkey"]
    for sel, brand, ptype, size in ((k % 11 == 0, "Brand#11", "SMALL PLATED TIN", 49),
                                    (k % 13 == 0, "Brand#45", None, 14),
                                    (k % 17 == 0, None, "MEDIUM POLISHED TIN", 23)):
        if brand:
            P["p_brand"][sel] = G.fixed([brand], 10)[0]
        if ptype:
            P["p_type"][sel] = G.fixed([ptype], 25)[0]
        P["p_size"][sel] = size
    S["s_comment"][S["s_suppkey"] % 9 == 0] = G.fixed(["slyly Customer final Complaints"], 101)[0]
    return data


@pytest.fixture(scope="module")
def gen_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("tpch")
    data = q16_rows(q19_parts(G.generate(0.005, seed=11)))
    G.write_tbl(data, d, T.table_columns)
    T.ingest(d)
    return d, data


# ---------------------------------------------------------------------------
# .tbl parser and .colbin v2
# ---------------------------------------------------------------------------

def test_catalog_matches_reference_schema():
    cols = T.table_columns("lineitem")
    assert [c[0] for c in cols][:5] == ["l_orderkey", "l_partkey", "l_suppkey", "l_linenumber", "l_quantity"]
    assert cols[4][2] == T.FLOAT and cols[10][2] == T.DATE and cols[8][2] == T.CHAR1
    assert cols[14] == ("l_shipmode", 14, T.CHAR_FIXED, 2)
    assert T.table_columns("orders")[8] == ("o_comment", 8, T.CHAR_FIXED, 79)
    assert len(T.table_columns("customer")) == 8 and len(T.table_columns("region")) == 3


def test_colbin_layout_bytes(tmp_path):
    tbl = tmp_path / "nation.tbl"
    tbl.write_text("0|ALGERIA|0|haggle furiously|\n1|ARGENTINA|1|al foxes|\n")
    assert T.tbl_to_colbin(tbl, "nation")
    raw = (tmp_path / "nation.colbin").read_bytes()
    magic, version, ncols, nrows, src_size, src_mtime, pad = struct.unpack_from("<8sIIQQqQ", raw, 0)
    st = os.stat(tbl)
    assert magic == b"TPCHCB01" and version == 2 and ncols == 4 and nrows == 2 and pad == 0
    assert src_size == st.st_size and src_mtime == st.st_mtime_ns
    descs = [struct.unpack_from("<iBBHQQQ", raw, 48 + 32 * i) for i in range(ncols)]
    assert [d[0] for d in descs] == [0, 1, 2, 3]
    assert [d[1] for d in descs] == [0, 4, 0, 4]
    assert [d[3] for d in descs] == [0, 25, 0, 152]
    assert [d[5] for d in descs] == [8, 50, 8, 304]
    for d in descs:
        assert d[4] % 16384 == 0
    assert descs[0][4] == 16384 and len(raw) == descs[-1][4] + 16384
    assert np.frombuffer(raw[descs[0][4]:descs[0][4] + 8], np.int32).tolist() == [0, 1]
    name = raw[descs[1][4]:descs[1][4] + 50]
    assert name[:7] == b"ALGERIA" and name[7:25] == b"\0" * 18 and name[25:34] == b"ARGENTINA"


def test_tbl_field_rules(tmp_path):
    tbl = tmp_path / "orders.tbl"
    long_comment = "x" * 100
    tbl.write_text(
        "7|-3|O|12.5|1996-01-02|1-URGENT|Clerk#000000951|0|" + long_comment + "|\n"
        "\n"  # a line without '|' is skipped
        "  42|+8||0.07|1992-12-31||C|5||\n"
    )
    T.tbl_to_colbin(tbl, "orders", force=True)
    cols = T.load_colbin(tmp_path / "orders.colbin", "orders", tbl_path=tbl)
    assert cols["o_orderkey"].tolist() == [7, 42]
    assert cols["o_custkey"].tolist() == [-3, 8]
    assert cols["o_orderstatus"].tolist() == [ord("O"), 0]  # empty CHAR1 -> NUL
    assert cols["o_totalprice"].tolist() == [1250, 7]
    assert cols["o_orderdate"].tolist() == [19960102, 19921231]
    assert cols["o_orderpriority"].tolist() == [ord("1"), 0]
    assert bytes(cols["o_clerk"][1]) == b"C" + b"\0" * 14
    assert bytes(cols["o_comment"][0]) == b"x" * 79  # truncated to the width
    assert bytes(cols["o_comment"][1]) == b"\0" * 79


def test_tbl_short_line_is_an_error(tmp_path):
    tbl = tmp_path / "region.tbl"
    tbl.write_text("0|AFRICA|lar deposits|\n1|AMERICA|\n")
    with pytest.raises(Exception, match="fewer fields"):
        T.tbl_to_colbin(tbl, "region")


def test_colbin_staleness(tmp_path):
    tbl = tmp_path / "region.tbl"
    tbl.write_text("0|AFRICA|c|\n")
    assert T.tbl_to_colbin(tbl, "region")
    assert not T.tbl_to_colbin(tbl, "region")  # up to date
    tbl.write_text("0|AFRICA|c|\n1|AMERICA|d|\n")
    with pytest.raises(Exception, match="stale"):
        T.load_colbin(tmp_path / "region.colbin", "region", tbl_path=tbl)
    assert T.tbl_to_colbin(tbl, "region")  # rebuilt
    assert T.load_colbin(tmp_path / "region.colbin", "region", tbl_path=tbl)["r_regionkey"].tolist() == [0, 1]
    tbl.unlink()  # without the .tbl the file's own header is trusted
    assert T.load_colbin(tmp_path / "region.colbin", "region", tbl_path=tbl)["r_regionkey"].tolist() == [0, 1]


def test_colbin_rejects_bad_files(tmp_path):
    p = tmp_path / "x.colbin"
    p.write_bytes(b"NOTMAGIC" + b"\0" * 40)
    with pytest.raises(Exception, match="magic"):
        T.Colbin(p)
    p.write_bytes(b"TPCHCB01" + struct.pack("<IIQQqQ", 1, 0, 0, 0, 0, 0))
    with pytest.raises(Exception, match="version"):
        T.Colbin(p)


def test_generated_tables_round_trip(gen_dir):
    d, data = gen_dir
    host = T.load_dir(d)
    assert set(host) == set(T.TABLES)
    for t, cols in data.items():
        for c, a in cols.items():
            b = host[t][c]
            if c == "o_totalprice":
                # float32 in the .colbin format: cents are lost above 2^17
                assert np.abs(a - b).max() <= 4
                small = np.abs(a) < 13_000_000
                assert np.array_equal(a[small], b[small])
            else:
                assert np.array_equal(a, b), (t, c)


def test_parallel_parse_matches_single_thread(gen_dir, tmp_path):
    d, _ = gen_dir
    src = (d / "lineitem.tbl").read_bytes()
    one = tmp_path / "lineitem.tbl"
    one.write_bytes(src)
    T.tbl_to_colbin(one, "lineitem", force=True, threads=1)
    a = (tmp_path / "lineitem.colbin").read_bytes()
    T.tbl_to_colbin(one, "lineitem", force=True, threads=7)
    b = (tmp_path / "lineitem.colbin").read_bytes()
    assert a == b


# ---------------------------------------------------------------------------
# Golden comparator (the reference's golden files as fixtures)
# ---------------------------------------------------------------------------

def _golden(q):
    return (GOLDEN / f"{q}_SF1.csv").read_text()


@pytest.mark.parametrize("q", ["Q1", "Q3", "Q13", "Q14", "Q18", "Q22"])
def test_compare_identical_golden(q):
    g = _golden(q)
    assert T.compare_canonical(g, g) == ""
    assert T.compare_canonical(g.replace("\n", "\r\n"), g) == ""


def test_compare_numeric_tolerance():
    g = _golden("Q1")
    # sum_qty 37734107.0 -> within abs 1e-2 + rel 1e-4 (3773 here) and beyond
    assert T.compare_canonical(g.replace("37734107.0", "37734200.0"), g) == ""
    msg = T.compare_canonical(g.replace("37734107.0", "37744107.0"), g)
    assert msg.startswith("row 1 col 'sum_qty': 37744107.0 vs 37734107.0")
    assert T.compare_canonical(g.replace("37734107.0", "37744107.0"), g, 1e-2, 1e-3) == ""


def test_compare_strings_rows_and_schema():
    g = _golden("Q1")
    assert "row 1 col 'l_returnflag': 'B' vs 'A'" in T.compare_canonical(g.replace("A,F", "B,F", 1), g)
    lines = g.strip().split("\n")
    assert T.compare_canonical("\n".join(lines[:-1]) + "\n", g) == "row count mismatch: got=3 expected=4"
    assert T.compare_canonical("x,y\n1,2\n", "a,b\n1,2\n").startswith("schema mismatch")
    assert T.compare_canonical("", "") == ""
    assert T.compare_canonical("", g).startswith("got 0 rows, expected 4")


def test_compare_shared_columns_and_quotes():
    g = _golden("Q18")
    # drop the quotes DuckDB put around c_name and an entire column: still a match
    lines = g.strip().split("\n")
    unq = [ln.replace('"', "") for ln in lines]
    assert T.compare_canonical("\n".join(unq) + "\n", g) == ""
    no_price = [",".join(f for i, f in enumerate(ln.split(",")) if i != 4) for ln in unq]
    assert T.compare_canonical("\n".join(no_price) + "\n", g) == ""
    assert T.compare_canonical('a,b\n"x,y",2\n', 'a,b\n"x,y",2\n') == ""
    # "" inside quotes is one quote character; a bare quote only toggles quoting
    assert T.compare_canonical('a\n"say ""hi"""\n', 'a\n"say ""hi"""\n') == ""
    assert T.compare_canonical('a\n"say ""hi"""\n', 'a\nsay "hi"\n') == "row 1 col 'a': 'say \"hi\"' vs 'say hi'"


def test_check_status_codes(tmp_path):
    g = _golden("Q14")
    T.save_golden(g, tmp_path, "Q14", "SF1")
    assert T.check(g, tmp_path, "Q14", "SF1") == (0, "OK")
    st, msg = T.check(g.replace("16.38", "17.38"), tmp_path, "Q14", "SF1")
    assert st == 1 and "promo_revenue" in msg
    assert T.check(g, tmp_path, "Q14", "SF10")[0] == 2


def test_canonical_rendering():
    cols = {"k": [1, -2], "v": [5, -5], "d": [19950315, 19920101], "s": ["a,b", 'q"x']}
    fm = {"k": (0, False), "v": (2, False), "d": (0, True), "s": (0, False)}
    want = 'k,v,d,s\n1,0.05,1995-03-15,"a,b"\n-2,-0.05,1992-01-01,"q""x"\n'
    assert T.canonical(cols, fm) == want
    assert OT.canonical(cols, fm) == want


# ---------------------------------------------------------------------------
# The oracle pinned against SQLite on the same rows
# ---------------------------------------------------------------------------

def _sqlite(host: dict) -> sqlite3.Connection:
    con = sqlite3.connect(":memory:")
    con.execute("PRAGMA case_sensitive_like = ON")  # LIKE as in SQL (SQLite folds ASCII case by default)
    for t, cols in host.items():
        spec = T.table_columns(t)
        con.execute(f"CREATE TABLE {t} ({', '.join(n for n, *_ in spec)})")
        conv = []
        for n, _, dt, _w in spec:
            a = cols[n]
            if dt == T.FLOAT:
                conv.append([v / 100.0 for v in a.tolist()])
            elif dt == T.DATE:
                conv.append([f"{v // 10000:04d}-{v // 100 % 100:02d}-{v % 100:02d}" for v in a.tolist()])
            elif dt == T.CHAR1:
                conv.append([chr(v) if v else "" for v in a.tolist()])
            elif dt == T.CHAR_FIXED:
                conv.append([bytes(r).split(b"\0", 1)[0].decode() for r in a])
            else:
                conv.append(a.tolist())
        con.executemany(f"INSERT INTO {t} VALUES ({', '.join('?' * len(spec))})", list(zip(*conv)))
    return con


SQL = {
    "Q1": """SELECT l_returnflag, l_linestatus, SUM(l_quantity), SUM(l_extendedprice),
        SUM(l_extendedprice * (1 - l_discount)), SUM(l_extendedprice * (1 - l_discount) * (1 + l_tax)),
        AVG(l_quantity), AVG(l_extendedprice), AVG(l_discount), COUNT(*)
        FROM lineitem WHERE l_shipdate <= '1998-09-02' GROUP BY l_returnflag, l_linestatus
        ORDER BY l_returnflag, l_linestatus""",
    "Q3": """SELECT l_orderkey, SUM(l_extendedprice * (1 - l_discount)) AS revenue, o_orderdate, o_shippriority
        FROM customer, orders, lineitem WHERE c_mktsegment = 'B' AND c_custkey = o_custkey
        AND l_orderkey = o_orderkey AND o_orderdate < '1995-03-15' AND l_shipdate > '1995-03-15'
        GROUP BY l_orderkey, o_orderdate, o_shippriority ORDER BY revenue DESC, o_orderdate, l_orderkey LIMIT 10""",
    "Q4": """SELECT o_orderpriority, COUNT(*) FROM orders WHERE o_orderdate >= '1993-07-01'
        AND o_orderdate < '1993-10-01' AND EXISTS (SELECT * FROM lineitem WHERE l_orderkey = o_orderkey
        AND l_commitdate < l_receiptdate) GROUP BY o_orderpriority ORDER BY o_orderpriority""",
    "Q6": """SELECT SUM(l_extendedprice * l_discount) FROM lineitem WHERE l_shipdate >= '1994-01-01'
        AND l_shipdate < '1995-01-01' AND l_discount BETWEEN 0.05 AND 0.07 AND l_quantity < 24""",
    "Q12": """SELECT l_shipmode, SUM(CASE WHEN o_orderpriority IN ('1', '2') THEN 1 ELSE 0 END),
        SUM(CASE WHEN o_orderpriority NOT IN ('1', '2') THEN 1 ELSE 0 END) FROM orders, lineitem
        WHERE o_orderkey = l_orderkey AND l_shipmode IN ('MA', 'SH') AND l_commitdate < l_receiptdate
        AND l_shipdate < l_commitdate AND l_receiptdate >= '1994-01-01' AND l_receiptdate < '1995-01-01'
        GROUP BY l_shipmode ORDER BY l_shipmode""",
    "Q13": """SELECT c_count, COUNT(*) AS custdist FROM (SELECT c_custkey, COUNT(o_orderkey) AS c_count
        FROM customer LEFT OUTER JOIN orders ON c_custkey = o_custkey
        AND o_comment NOT LIKE '%special%requests%' GROUP BY c_custkey) GROUP BY c_count
        ORDER BY custdist DESC, c_count DESC""",
    "Q14": """SELECT 100.00 * SUM(CASE WHEN p_type LIKE 'PROMO%' THEN l_extendedprice * (1 - l_discount) ELSE 0 END)
        / SUM(l_extendedprice * (1 - l_discount)) FROM lineitem, part WHERE l_partkey = p_partkey
        AND l_shipdate >= '1995-09-01' AND l_shipdate < '1995-10-01'""",
    "Q18": """SELECT c_name, c_custkey, o_orderkey, o_orderdate, o_totalprice, SUM(l_quantity)
        FROM customer, orders, lineitem WHERE o_orderkey IN (SELECT l_orderkey FROM lineitem
        GROUP BY l_orderkey HAVING SUM(l_quantity) > 250) AND c_custkey = o_custkey AND o_orderkey = l_orderkey
        GROUP BY c_name, c_custkey, o_orderkey, o_orderdate, o_totalprice
        ORDER BY o_totalprice DESC, o_orderdate, c_custkey, o_orderkey LIMIT 100""",
    "Q5": """SELECT n_name, SUM(l_extendedprice * (1 - l_discount)) AS revenue FROM customer, orders, lineitem,
        supplier, nation, region WHERE c_custkey = o_custkey AND l_orderkey = o_orderkey AND l_suppkey = s_suppkey
        AND c_nationkey = s_nationkey AND s_nationkey = n_nationkey AND n_regionkey = r_regionkey
        AND r_name = 'ASIA' AND o_orderdate >= '1994-01-01' AND o_orderdate < '1995-01-01'
        GROUP BY n_name ORDER BY revenue DESC, n_name""",
    "Q10": """SELECT c_custkey, c_name, SUM(l_extendedprice * (1 - l_discount)) AS revenue, c_acctbal, n_name,
        c_address, c_phone, c_comment FROM customer, orders, lineitem, nation WHERE c_custkey = o_custkey
        AND l_orderkey = o_orderkey AND o_orderdate >= '1993-10-01' AND o_orderdate < '1994-01-01'
        AND l_returnflag = 'R' AND c_nationkey = n_nationkey GROUP BY c_custkey, c_name, c_acctbal, c_phone, n_name,
        c_address, c_comment ORDER BY revenue DESC, c_custkey LIMIT 20""",
    "Q19": """SELECT SUM(l_extendedprice * (1 - l_discount)) FROM lineitem, part WHERE
        (p_partkey = l_partkey AND p_brand = 'Brand#12' AND p_container IN ('SM CASE', 'SM BOX', 'SM PACK', 'SM PKG')
         AND l_quantity >= 1 AND l_quantity <= 11 AND p_size BETWEEN 1 AND 5 AND l_shipmode IN ('AI', 'RE')
         AND l_shipinstruct = 'DELIVER IN PERSON')
        OR (p_partkey = l_partkey AND p_brand = 'Brand#23' AND p_container IN ('MED BAG', 'MED BOX', 'MED PKG', 'MED PACK')
         AND l_quantity >= 10 AND l_quantity <= 20 AND p_size BETWEEN 1 AND 10 AND l_shipmode IN ('AI', 'RE')
         AND l_shipinstruct = 'DELIVER IN PERSON')
        OR (p_partkey = l_partkey AND p_brand = 'Brand#34' AND p_container IN ('LG CASE', 'LG BOX', 'LG PACK', 'LG PKG')
         AND l_quantity >= 20 AND l_quantity <= 30 AND p_size BETWEEN 1 AND 15 AND l_shipmode IN ('AI', 'RE')
         AND l_shipinstruct = 'DELIVER IN PERSON')""",
    "Q2": """SELECT s_acctbal, s_name, n_name, p_partkey, p_mfgr, s_address, s_phone, s_comment FROM part, supplier,
        partsupp, nation, region WHERE p_partkey = ps_partkey AND s_suppkey = ps_suppkey AND p_size = 15
        AND p_type LIKE '%BRASS' AND s_nationkey = n_nationkey AND n_regionkey = r_regionkey AND r_name = 'EUROPE'
        AND ps_supplycost = (SELECT MIN(ps_supplycost) FROM partsupp, supplier, nation, region
        WHERE p_partkey = ps_partkey AND s_suppkey = ps_suppkey AND s_nationkey = n_nationkey
        AND n_regionkey = r_regionkey AND r_name = 'EUROPE') ORDER BY s_acctbal DESC, n_name, s_name, p_partkey
        LIMIT 100""",
    "Q7": """SELECT supp_nation, cust_nation, l_year, SUM(volume) AS revenue FROM (SELECT n1.n_name AS supp_nation,
        n2.n_name AS cust_nation, CAST(SUBSTR(l_shipdate, 1, 4) AS INTEGER) AS l_year,
        l_extendedprice * (1 - l_discount) AS volume FROM supplier, lineitem, orders, customer, nation n1, nation n2
        WHERE s_suppkey = l_suppkey AND o_orderkey = l_orderkey AND c_custkey = o_custkey
        AND s_nationkey = n1.n_nationkey AND c_nationkey = n2.n_nationkey AND ((n1.n_name = 'FRANCE'
        AND n2.n_name = 'GERMANY') OR (n1.n_name = 'GERMANY' AND n2.n_name = 'FRANCE'))
        AND l_shipdate BETWEEN '1995-01-01' AND '1996-12-31') GROUP BY supp_nation, cust_nation, l_year
        ORDER BY supp_nation, cust_nation, l_year""",
    "Q8": """SELECT o_year, SUM(CASE WHEN nation = 'BRAZIL' THEN volume ELSE 0 END) / SUM(volume) AS mkt_share
        FROM (SELECT CAST(SUBSTR(o_orderdate, 1, 4) AS INTEGER) AS o_year, l_extendedprice * (1 - l_discount) AS volume,
        n2.n_name AS nation FROM part, supplier, lineitem, orders, customer, nation n1, nation n2, region
        WHERE p_partkey = l_partkey AND s_suppkey = l_suppkey AND l_orderkey = o_orderkey AND o_custkey = c_custkey
        AND c_nationkey = n1.n_nationkey AND n1.n_regionkey = r_regionkey AND r_name = 'AMERICA'
        AND s_nationkey = n2.n_nationkey AND o_orderdate BETWEEN '1995-01-01' AND '1996-12-31'
        AND p_type = 'ECONOMY ANODIZED STEEL') GROUP BY o_year ORDER BY o_year""",
    "Q9": """SELECT nation, o_year, SUM(amount) AS sum_profit FROM (SELECT n_name AS nation,
        CAST(SUBSTR(o_orderdate, 1, 4) AS INTEGER) AS o_year,
        l_extendedprice * (1 - l_discount) - ps_supplycost * l_quantity AS amount FROM part, supplier, lineitem,
        partsupp, orders, nation WHERE s_suppkey = l_suppkey AND ps_suppkey = l_suppkey AND ps_partkey = l_partkey
        AND p_partkey = l_partkey AND o_orderkey = l_orderkey AND s_nationkey = n_nationkey
        AND p_name LIKE '%green%') GROUP BY nation, o_year ORDER BY nation, o_year DESC""",
    "Q20": """SELECT s_name, s_address FROM supplier, nation WHERE s_suppkey IN (SELECT ps_suppkey FROM partsupp
        WHERE ps_partkey IN (SELECT p_partkey FROM part WHERE p_name LIKE 'forest%') AND ps_availqty > (
        SELECT 0.5 * SUM(l_quantity) FROM lineitem WHERE l_partkey = ps_partkey AND l_suppkey = ps_suppkey
        AND l_shipdate >= '1994-01-01' AND l_shipdate < '1995-01-01')) AND s_nationkey = n_nationkey
        AND n_name = 'CANADA' ORDER BY s_name""",
    "Q21": """SELECT s_name, COUNT(*) AS numwait FROM supplier, lineitem l1, orders, nation
        WHERE s_suppkey = l1.l_suppkey AND o_orderkey = l1.l_orderkey AND o_orderstatus = 'F'
        AND l1.l_receiptdate > l1.l_commitdate AND EXISTS (SELECT * FROM lineitem l2 WHERE l2.l_orderkey = l1.l_orderkey
        AND l2.l_suppkey <> l1.l_suppkey) AND NOT EXISTS (SELECT * FROM lineitem l3
        WHERE l3.l_orderkey = l1.l_orderkey AND l3.l_suppkey <> l1.l_suppkey AND l3.l_receiptdate > l3.l_commitdate)
        AND s_nationkey = n_nationkey AND n_name = 'SAUDI ARABIA' GROUP BY s_name ORDER BY numwait DESC, s_name
        LIMIT 100""",
    "Q11": """SELECT ps_partkey, SUM(ps_supplycost * ps_availqty) AS value FROM partsupp, supplier, nation
        WHERE ps_suppkey = s_suppkey AND s_nationkey = n_nationkey AND n_name = 'GERMANY' GROUP BY ps_partkey
        HAVING SUM(ps_supplycost * ps_availqty) > (SELECT SUM(ps_supplycost * ps_availqty) * 0.0001
        FROM partsupp, supplier, nation WHERE ps_suppkey = s_suppkey AND s_nationkey = n_nationkey
        AND n_name = 'GERMANY') ORDER BY value DESC, ps_partkey""",
    "Q15": """WITH revenue0 (supplier_no, total_revenue) AS (SELECT l_suppkey, SUM(l_extendedprice * (1 - l_discount))
        FROM lineitem WHERE l_shipdate >= '1996-01-01' AND l_shipdate < '1996-04-01' GROUP BY l_suppkey)
        SELECT s_suppkey, s_name, s_address, s_phone, total_revenue FROM supplier, revenue0
        WHERE s_suppkey = supplier_no AND total_revenue = (SELECT MAX(total_revenue) FROM revenue0)
        ORDER BY s_suppkey""",
    "Q16": """SELECT p_brand, p_type, p_size, COUNT(DISTINCT ps_suppkey) AS supplier_cnt FROM partsupp, part
        WHERE p_partkey = ps_partkey AND p_brand <> 'Brand#45' AND p_type NOT LIKE 'MEDIUM POLISHED%'
        AND p_size IN (49, 14, 23, 45, 19, 3, 36, 9) AND ps_suppkey NOT IN (SELECT s_suppkey FROM supplier
        WHERE s_comment LIKE '%Customer%Complaints%') GROUP BY p_brand, p_type, p_size
        ORDER BY supplier_cnt DESC, p_brand, p_type, p_size""",
    "Q17": """SELECT SUM(l_extendedprice) / 7.0 FROM lineitem, part WHERE p_partkey = l_partkey
        AND p_brand = 'Brand#23' AND p_container = 'MED BOX' AND l_quantity < (SELECT 0.2 * AVG(l_quantity)
        FROM lineitem WHERE l_partkey = p_partkey)""",
    "Q22": """SELECT cntrycode, COUNT(*), SUM(c_acctbal) FROM (SELECT SUBSTR(c_phone, 1, 2) AS cntrycode, c_acctbal
        FROM customer WHERE SUBSTR(c_phone, 1, 2) IN ('13', '31', '23', '29', '30', '18', '17')
        AND c_acctbal > (SELECT AVG(c_acctbal) FROM customer WHERE c_acctbal > 0.00
        AND SUBSTR(c_phone, 1, 2) IN ('13', '31', '23', '29', '30', '18', '17'))
        AND NOT EXISTS (SELECT * FROM orders WHERE o_custkey = c_custkey)) GROUP BY cntrycode ORDER BY cntrycode""",
}


# Q18's quantity bound is lowered to 250 (scalar q18_qty) so the small
# generated data set selects orders; the plan is the same.
PARAMS = {q: dict(p) for q, p in T.PARAMS.items()}
PARAMS["Q18"]["q18_qty"] = 25000


def oracle_run(host: dict, q: str):
    return OT.run(host, q, PARAMS[q])


def _value(v, scale, is_date):
    if isinstance(v, str):
        return v
    if is_date:
        return f"{v // 10000:04d}-{v // 100 % 100:02d}-{v % 100:02d}"
    return v / 10 ** scale if scale else v


@pytest.mark.parametrize("q", T.QUERIES)
def test_oracle_matches_sqlite(gen_dir, q):
    d, _ = gen_dir
    host = T.load_dir(d)
    cols, fmts = oracle_run(host, q)
    rows = _sqlite(host).execute(SQL[q]).fetchall()
    rows = [r for r in rows if any(v is not None for v in r)]  # SUM over no rows: one NULL row
    names = list(cols)
    n = len(cols[names[0]])
    assert n == len(rows), (q, n, len(rows))
    assert n > 0, f"{q}: the generated data selects no rows"
    mode_names = {"MAIL": "MA", "SHIP": "SH"}
    for r in range(n):
        for c, name in enumerate(names):
            got = _value(cols[name][r], *fmts[name])
            exp = rows[r][c]
            if name == "l_shipmode":
                got = mode_names[got]
            if name == "o_orderpriority":
                got = got[0]
            if isinstance(got, float) or isinstance(exp, float):
                assert abs(got - exp) <= 1e-6 + 1e-9 * abs(exp), (q, name, r, got, exp)
            else:
                assert got == exp, (q, name, r, got, exp)


def test_oracle_hand_cases():
    L = {"l_shipdate": np.array([19940101, 19941231, 19950101, 19940601], np.int32),
         "l_discount": np.array([6, 7, 6, 8], np.int64), "l_quantity": np.array([100, 2399, 100, 100], np.int64),
         "l_extendedprice": np.array([1000, 2000, 4000, 8000], np.int64)}
    cols, fm = OT.q6({"lineitem": L}, 19940101, 19950101, 5, 7, 2400)
    assert cols["revenue"] == [1000 * 6 + 2000 * 7] and fm["revenue"] == (4, False)
    # Q13: NOT LIKE keeps comments without the ordered pair; customer 3 has no order
    O = {"o_custkey": np.array([1, 1, 2], np.int32),
         "o_comment": G.fixed(["special requests", "requests special", "x"], 79)}
    C = {"c_custkey": np.array([1, 2, 3], np.int32)}
    cols, _ = OT.q13({"orders": O, "customer": C}, b"special", b"requests")
    assert cols == {"c_count": [1, 0], "custdist": [2, 1]}


def test_cli_usage_errors():
    from gpudbmetalcodegen_amd import tpch_cli

    assert tpch_cli.main(["q23"]) == 3  # no such query
    assert tpch_cli.main(["bogus"]) == 3


def test_chunk_rows_parse():
    assert T.parse_rows("16K") == 16 * 1024
    assert T.parse_rows("2m") == 2 * 1024 * 1024
    assert T.parse_rows("1G") == 1 << 30
    assert T.parse_rows("12345") == 12345
    for bad in ("0", "-5K", "K", "x"):
        with pytest.raises(ValueError):
            T.parse_rows(bad)


def test_chunkable_plans_and_stream_tables():
    """The plans marked chunkable are those whose stream-table phases only
    accumulate (the GPU tests run each of them streamed); lineitem is the
    stream table wherever a query reads it."""
    from gpudbmetalcodegen_amd import engine as E
    from test_tpch_gpu import CHUNKED

    got = [q for q in T.QUERIES if E.build_hip_plan(q, T.PLAN_OPTIONS.get(q, "")).chunkable]
    assert got == CHUNKED
    assert [q for q in CHUNKED if T.stream_table(q) == "orders"] == ["Q13", "Q22"]


def test_chunk_decision_follows_the_reference_policy():
    """codegen_main.cpp:453-624: chunk when the working set exceeds 3/4 of
    the budget; rows = (budget/2 - resident) / (2 slots x bytes/row), at least
    budget/8 and 64 MiB of stream budget and 256K rows, at most the table;
    an explicit chunk for a working set that fits is downgraded unless forced."""
    GiB = 1 << 30
    # fits: 1 GiB resident + 100M rows x 20 B = 3 GiB-ish < 0.75 x 16 GiB
    assert T.chunk_decision(GiB, 100_000_000, 20, 16 * GiB)[0] == 0
    r, why = T.chunk_decision(GiB, 100_000_000, 20, 16 * GiB, requested_rows=1 << 20)
    assert r == 0 and "downgraded" in why
    assert T.chunk_decision(GiB, 100_000_000, 20, 16 * GiB, requested_rows=1 << 20, force=True)[0] == 1 << 20
    # does not fit: 1 GiB resident + 1e9 rows x 20 B against 16 GiB
    r, why = T.chunk_decision(GiB, 1_000_000_000, 20, 16 * GiB)
    assert r == (8 * GiB - GiB) // 40 and "exceeds" in why          # 187,904,819 rows
    # resident tables eat the half budget: the 1/8 floor holds
    r, _ = T.chunk_decision(10 * GiB, 1_000_000_000, 20, 16 * GiB)
    assert r == (2 * GiB) // 40
    # tiny budget: 64 MiB floor, then the 256K-row floor, capped at the table
    assert T.chunk_decision(0, 50_000_000, 64, 100 << 20)[0] == max((64 << 20) // 128, 256 * 1024)
    assert T.chunk_decision(0, 100_000, 4096, 1 << 20)[0] == 100_000


def test_projected_bytes_count_only_the_plans_columns():
    import numpy as np

    cols = {"l_a": np.zeros(4, np.int32), "l_b": np.zeros(4, np.int64), "l_c": np.zeros((4, 10), np.uint8)}
    assert T.projected_bytes_per_row("f(p_l_a, p_l_c)", cols) == 4 + 10
    assert T.projected_bytes_per_row("p_l_ab p_l_b", cols) == 8
