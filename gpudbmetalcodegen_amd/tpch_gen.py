"""Seeded TPC-H-shaped data: the eight tables at a scale factor, in the
engine's column formats, plus a writer of TPC-H ``.tbl`` text.

This is not dbgen: values follow the TPC-H specification's domains and the
properties the queries depend on (sparse order keys in groups of 8 out of 32;
customers whose key is a multiple of 3 place no orders; a lineitem ships 1-121
days after its order, commits 30-90 days after it, is received 1-30 days after
shipping; return flag R/A up to 1995-06-17 and N after; line status F/O by
ship date; retail price 900 + (key/10 mod 20001)/100 + key mod 1000; phone
country code = nation + 10), so the queries select realistic fractions, but
the rows differ from dbgen's and the reference's golden files do not apply.

Column formats (what ``DeviceTable.from_host`` uploads): INT int32, DATE int32
YYYYMMDD, money and rates int64 hundredths, CHAR1 uint8, CHAR_FIXED uint8
(rows, width), NUL-padded.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

REGIONS = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
NATIONS = [("ALGERIA", 0), ("ARGENTINA", 1), ("BRAZIL", 1), ("CANADA", 1), ("EGYPT", 4), ("ETHIOPIA", 0),
           ("FRANCE", 3), ("GERMANY", 3), ("INDIA", 2), ("INDONESIA", 2), ("IRAN", 4), ("IRAQ", 4), ("JAPAN", 2),
           ("JORDAN", 4), ("KENYA", 0), ("MOROCCO", 0), ("MOZAMBIQUE", 0), ("PERU", 1), ("CHINA", 2),
           ("ROMANIA", 3), ("SAUDI ARABIA", 4), ("VIETNAM", 2), ("RUSSIA", 3), ("UNITED KINGDOM", 3),
           ("UNITED STATES", 1)]
SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "MACHINERY", "HOUSEHOLD"]
PRIORITIES = ["1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW"]
SHIPMODES = ["REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"]
INSTRUCTIONS = ["DELIVER IN PERSON", "COLLECT COD", "NONE", "TAKE BACK RETURN"]
TYPE1 = ["STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO"]
TYPE2 = ["ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED"]
TYPE3 = ["TIN", "NICKEL", "BRASS", "STEEL", "COPPER"]
CONTAINERS1 = ["SM", "LG", "MED", "JUMBO", "WRAP"]
CONTAINERS2 = ["CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM"]
COLORS = ["almond", "antique", "aquamarine", "azure", "beige", "bisque", "black", "blanched", "blue", "blush",
          "brown", "burlywood", "burnished", "chartreuse", "chiffon", "chocolate", "coral", "cornflower",
          "cornsilk", "cream", "cyan", "dark", "deep", "dim", "dodger", "drab", "firebrick", "floral", "forest",
          "frosted", "gainsboro", "ghost", "goldenrod", "green", "grey", "honeydew", "hot", "indian", "ivory"]
WORDS = ["furiously", "carefully", "quickly", "slyly", "blithely", "final", "ironic", "regular", "express",
         "bold", "pending", "even", "silent", "unusual", "packages", "deposits", "accounts", "requests",
         "instructions", "theodolites", "pinto", "beans", "foxes", "ideas", "dependencies", "platelets",
         "special", "asymptotes", "courts", "dolphins", "sleep", "wake", "haggle", "nag", "cajole", "boost"]

START = np.datetime64("1992-01-01")
CURRENT = 19950617  # dbgen's CURRENTDATE: return flag and line status switch here


def ymd(days: np.ndarray) -> np.ndarray:
    """Day offsets from 1992-01-01 -> int32 YYYYMMDD."""
    d = START + days.astype("timedelta64[D]")
    y = d.astype("datetime64[Y]").astype(np.int64) + 1970
    m = d.astype("datetime64[M]").astype(np.int64) % 12 + 1
    dd = (d - d.astype("datetime64[M]")).astype(np.int64) + 1
    return (y * 10000 + m * 100 + dd).astype(np.int32)


def fixed(strings, width: int) -> np.ndarray:
    """Python strings -> uint8 (n, width), truncated, NUL-padded."""
    raw = np.array([s.encode()[:width] for s in strings], dtype=f"S{width}")
    return raw.view(np.uint8).reshape(len(strings), width) if len(strings) else np.zeros((0, width), np.uint8)


def char1(strings) -> np.ndarray:
    return np.array([ord(s[0]) if s else 0 for s in strings], dtype=np.uint8)


def _pick(rng, choices, n):
    return [choices[i] for i in rng.integers(0, len(choices), n)]


def _pick_fixed(rng, choices, n, width):
    """uint8 (n, width): uniformly chosen strings of `choices`."""
    return fixed(choices, width)[rng.integers(0, len(choices), n)]


def _pick_char1(rng, choices, n):
    return char1(choices)[rng.integers(0, len(choices), n)]


def _comments(rng, n, width, special_frac=0.0) -> np.ndarray:
    """uint8 (n, width) comment text: 3-8 random words; a special_frac share of
    rows gets "special <word> requests" inserted. Above 100,000 rows the
    engine's seeded text generator fills the column (words from the same
    vocabulary, so the pair occurs by chance)."""
    if n > 100_000:
        from . import workloads as W

        return W.text(int(rng.integers(1, 2**31)), 7, n, width, [w.encode() for w in WORDS])
    lo, hi = 3, 9
    k = rng.integers(lo, hi, n)
    idx = rng.integers(0, len(WORDS), (n, hi))
    out = []
    special = rng.random(n) < special_frac
    for i in range(n):
        words = [WORDS[j] for j in idx[i, :k[i]]]
        if special[i]:
            pos = int(rng.integers(0, len(words)))
            words[pos:pos] = ["special", str(rng.choice(["packages", "deposits", "pending"])), "requests"]
        out.append(" ".join(words)[:width])
    return fixed(out, width)


def _customer_remarks(seed: int, comments: np.ndarray) -> None:
    """TPC-H 4.2.3: of every 10,000 suppliers, 5 have "Customer ... Complaints"
    and 5 "Customer ... Recommends" in their comment (here at least one of
    each), written in place at a random offset. A generator of its own, so
    the other columns do not depend on it."""
    n, width = comments.shape
    rng = np.random.default_rng(seed + 0x5151)
    k = max(1, n // 2000)
    rows = rng.permutation(n)[:2 * k]
    for j, r in enumerate(rows.tolist()):
        text = b"Customer " + WORDS[int(rng.integers(0, len(WORDS)))].encode() + \
            (b" Complaints" if j < k else b" Recommends")
        if len(text) > width:
            continue
        off = int(rng.integers(0, width - len(text) + 1))
        cur = bytes(comments[r]).split(b"\0", 1)[0].ljust(off, b" ")
        line = (cur[:off] + text + cur[off + len(text):])[:width]
        comments[r] = np.frombuffer(line.ljust(width, b"\0"), np.uint8)


def retail_price(partkey: np.ndarray) -> np.ndarray:
    """Hundredths: 90000 + (key / 10) mod 20001 + 100 * (key mod 1000) (TPC-H 4.2.3)."""
    k = partkey.astype(np.int64)
    return 90000 + (k // 10) % 20001 + 100 * (k % 1000)


def generate(sf: float, seed: int = 1) -> dict:
    """All eight tables at scale factor sf: {table: {column: array}}."""
    rng = np.random.default_rng(seed)
    n_supp = max(1, int(10_000 * sf))
    n_part = max(1, int(200_000 * sf))
    n_cust = max(3, int(150_000 * sf))
    n_ord = max(1, int(1_500_000 * sf))

    region = {"r_regionkey": np.arange(5, dtype=np.int32), "r_name": fixed(REGIONS, 25),
              "r_comment": _comments(rng, 5, 152)}
    nation = {"n_nationkey": np.arange(25, dtype=np.int32), "n_name": fixed([n for n, _ in NATIONS], 25),
              "n_regionkey": np.array([r for _, r in NATIONS], dtype=np.int32),
              "n_comment": _comments(rng, 25, 152)}

    s_nat = rng.integers(0, 25, n_supp).astype(np.int32)
    supplier = {
        "s_suppkey": np.arange(1, n_supp + 1, dtype=np.int32),
        "s_name": fixed([f"Supplier#{k:09d}" for k in range(1, n_supp + 1)], 25),
        "s_address": _comments(rng, n_supp, 40),
        "s_nationkey": s_nat,
        "s_phone": fixed(_phones(rng, s_nat), 15),
        "s_acctbal": rng.integers(-99999, 1000000, n_supp).astype(np.int64),
        "s_comment": _comments(rng, n_supp, 101),
    }
    _customer_remarks(seed, supplier["s_comment"])

    pk = np.arange(1, n_part + 1, dtype=np.int32)
    types = [f"{a} {b} {c}" for a, b, c in zip(_pick(rng, TYPE1, n_part), _pick(rng, TYPE2, n_part),
                                               _pick(rng, TYPE3, n_part))]
    mfgr = rng.integers(1, 6, n_part)
    part = {
        "p_partkey": pk,
        "p_name": fixed([" ".join(_pick(rng, COLORS, 5)) for _ in range(n_part)], 55),
        "p_mfgr": fixed([f"Manufacturer#{m}" for m in mfgr], 25),
        "p_brand": fixed([f"Brand#{m}{b}" for m, b in zip(mfgr, rng.integers(1, 6, n_part))], 10),
        "p_type": fixed(types, 25),
        "p_size": rng.integers(1, 51, n_part).astype(np.int32),
        "p_container": fixed([f"{a} {b}" for a, b in zip(_pick(rng, CONTAINERS1, n_part),
                                                          _pick(rng, CONTAINERS2, n_part))], 10),
        "p_retailprice": retail_price(pk),
        "p_comment": _comments(rng, n_part, 23),
    }

    ps_part = np.repeat(pk, 4)
    ps_supp = ((ps_part.astype(np.int64) - 1 + np.tile(np.arange(4), n_part) * (n_supp // 4 + 1)) % n_supp + 1)
    partsupp = {
        "ps_partkey": ps_part, "ps_suppkey": ps_supp.astype(np.int32),
        "ps_availqty": rng.integers(1, 10000, ps_part.shape[0]).astype(np.int32),
        "ps_supplycost": rng.integers(100, 100001, ps_part.shape[0]).astype(np.int64),
        "ps_comment": _comments(rng, ps_part.shape[0], 199),
    }

    c_nat = rng.integers(0, 25, n_cust).astype(np.int32)
    customer = {
        "c_custkey": np.arange(1, n_cust + 1, dtype=np.int32),
        "c_name": fixed([f"Customer#{k:09d}" for k in range(1, n_cust + 1)], 25),
        "c_address": _comments(rng, n_cust, 40),
        "c_nationkey": c_nat,
        "c_phone": fixed(_phones(rng, c_nat), 15),
        "c_acctbal": rng.integers(-99999, 1000000, n_cust).astype(np.int64),
        "c_mktsegment": _pick_char1(rng, SEGMENTS, n_cust),
        "c_comment": _comments(rng, n_cust, 117),
    }

    # orders: sparse keys (8 used out of every 32), customers with key % 3 != 0
    i = np.arange(n_ord, dtype=np.int64)
    okey = ((i // 8) * 32 + (i % 8) + 1).astype(np.int32)
    active = np.arange(1, n_cust + 1)
    active = active[active % 3 != 0]
    ocust = active[rng.integers(0, active.shape[0], n_ord)].astype(np.int32)
    odays = rng.integers(0, 2406 - 151, n_ord)  # 1992-01-01 .. 1998-08-02
    odate = ymd(odays)

    # lineitem: 1..7 lines per order
    nl = rng.integers(1, 8, n_ord)
    n_line = int(nl.sum())
    lo = np.repeat(np.arange(n_ord), nl)
    lnum = (np.arange(n_line) - np.repeat(np.cumsum(nl) - nl, nl) + 1).astype(np.int32)
    lpart = rng.integers(1, n_part + 1, n_line).astype(np.int32)
    lsupp = ((lpart.astype(np.int64) - 1 + rng.integers(0, 4, n_line) * (n_supp // 4 + 1)) % n_supp + 1).astype(np.int32)
    qty = rng.integers(1, 51, n_line).astype(np.int64)
    price = qty * retail_price(lpart)
    disc = rng.integers(0, 11, n_line).astype(np.int64)
    tax = rng.integers(0, 9, n_line).astype(np.int64)
    ship_days = odays[lo] + rng.integers(1, 122, n_line)
    commit_days = odays[lo] + rng.integers(30, 91, n_line)
    receipt_days = ship_days + rng.integers(1, 31, n_line)
    ship, commit, receipt = ymd(ship_days), ymd(commit_days), ymd(receipt_days)
    rflag = np.where(receipt <= CURRENT, np.where(rng.random(n_line) < 0.5, ord("R"), ord("A")), ord("N"))
    lstatus = np.where(ship > CURRENT, ord("O"), ord("F"))
    lineitem = {
        "l_orderkey": okey[lo], "l_partkey": lpart, "l_suppkey": lsupp, "l_linenumber": lnum,
        "l_quantity": qty * 100, "l_extendedprice": price, "l_discount": disc, "l_tax": tax,
        "l_returnflag": rflag.astype(np.uint8), "l_linestatus": lstatus.astype(np.uint8),
        "l_shipdate": ship, "l_commitdate": commit, "l_receiptdate": receipt,
        "l_shipinstruct": _pick_fixed(rng, INSTRUCTIONS, n_line, 25),
        "l_shipmode": _pick_fixed(rng, SHIPMODES, n_line, 2),
        "l_comment": _comments(rng, n_line, 44),
    }

    # order totals and status from the lines: sum of price * (1 + tax) * (1 - disc), in cents
    line_total = (price * (100 + tax) * (100 - disc) + 5000) // 10000
    total = np.bincount(lo, weights=line_total, minlength=n_ord).astype(np.int64)
    n_f = np.bincount(lo, weights=(lstatus == ord("F")), minlength=n_ord)
    status = np.where(n_f == nl, ord("F"), np.where(n_f == 0, ord("O"), ord("P"))).astype(np.uint8)
    orders = {
        "o_orderkey": okey, "o_custkey": ocust, "o_orderstatus": status, "o_totalprice": total,
        "o_orderdate": odate, "o_orderpriority": _pick_char1(rng, PRIORITIES, n_ord),
        "o_clerk": _pick_fixed(rng, [f"Clerk#{c:09d}" for c in range(1, max(2, int(1000 * sf)) + 1)], n_ord, 15),
        "o_shippriority": np.zeros(n_ord, dtype=np.int32),
        "o_comment": _comments(rng, n_ord, 79, special_frac=0.02),
    }
    return {"region": region, "nation": nation, "supplier": supplier, "part": part, "partsupp": partsupp,
            "customer": customer, "orders": orders, "lineitem": lineitem}


def _phones(rng, nations) -> list:
    a, b, c = rng.integers(100, 1000, len(nations)), rng.integers(100, 1000, len(nations)), \
        rng.integers(1000, 10000, len(nations))
    return [f"{n + 10:02d}-{x}-{y}-{z}" for n, x, y, z in zip(nations, a, b, c)]


# ---------------------------------------------------------------------------
# .tbl text (one row per line, every field followed by '|')
# ---------------------------------------------------------------------------

# Full strings of the CHAR1 columns, so the text matches TPC-H files (the
# catalog keeps only the first byte).
FULL_CHAR1 = {
    "c_mktsegment": {ord(s[0]): s for s in SEGMENTS},
    "o_orderpriority": {ord(s[0]): s for s in PRIORITIES},
}


# Full strings of columns the catalog keeps shorter (l_shipmode is char[2]).
FULL_FIXED = {"l_shipmode": {m.encode()[:2]: m for m in SHIPMODES}}


def _field_text(table: str, col: str, dtype: int, arr: np.ndarray) -> list:
    if dtype == 2:  # DATE
        return [f"{v // 10000:04d}-{v // 100 % 100:02d}-{v % 100:02d}" for v in arr.tolist()]
    if dtype == 1:  # FLOAT (hundredths)
        return [f"{'-' if v < 0 else ''}{abs(v) // 100}.{abs(v) % 100:02d}" for v in arr.tolist()]
    if dtype == 3:  # CHAR1
        full = FULL_CHAR1.get(col)
        return [full[v] if full else chr(v) for v in arr.tolist()]
    if dtype == 4:  # CHAR_FIXED
        full = FULL_FIXED.get(col, {})
        return [full.get(bytes(r).rstrip(b"\0"), bytes(r).rstrip(b"\0").decode()) for r in arr]
    return [str(v) for v in arr.tolist()]


def write_tbl(tables: dict, out_dir, catalog) -> dict:
    """Write <out_dir>/<table>.tbl for each table; catalog(table) gives
    [(name, index, dtype, width)] in file order. Returns {table: path}."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    paths = {}
    for t, cols in tables.items():
        specs = catalog(t)
        texts = [_field_text(t, name, dtype, cols[name]) for name, _, dtype, _ in specs]
        path = out_dir / f"{t}.tbl"
        with open(path, "w") as f:
            for row in zip(*texts):
                f.write("|".join(row) + "|\n")
        paths[t] = path
    return paths
