"""CPU oracle of the named TPC-H queries (numpy, exact integers).

Test infrastructure only (tests/ and bench.py's cpu_baseline leg): a plain
restatement of each query's SQL (the reference's sql/q1.sql ... sql/q22.sql)
over the TPC-H tables in the engine's column formats (INT/DATE int32 with
dates as YYYYMMDD, money and rates int64 hundredths, CHAR1 uint8, CHAR_FIXED
uint8 (rows, width)). Results are {column: list} in the select list's order
with the same integer scaling as the engine (revenue = price * (100 -
discount), 4 decimal digits; averages trunc(sum * 10^6 / count)), plus the
display formats. Pinned in tests against SQLite running the SQL text on the
same rows (an independent engine) and against hand-computed tiny tables.
"""

from __future__ import annotations

import re

import numpy as np

PRIORITY_LABELS = {ord("1"): "1-URGENT", ord("2"): "2-HIGH", ord("3"): "3-MEDIUM", ord("4"): "4-NOT SPECIFIED",
                   ord("5"): "5-LOW"}


def _s(arr_row) -> bytes:
    return bytes(arr_row).split(b"\0", 1)[0]


def _avg(num: int, den: int) -> int:
    if den == 0:
        return 0
    q = abs(num) * 1_000_000 // abs(den)
    return q if (num >= 0) == (den > 0) else -q


def q1(T: dict, cutoff: int) -> tuple:
    """sql/q1.sql; groups ordered by (returnflag, linestatus)."""
    L = T["lineitem"]
    m = L["l_shipdate"] <= cutoff
    f, st = L["l_returnflag"][m], L["l_linestatus"][m]
    qty, price = L["l_quantity"][m].astype(object), L["l_extendedprice"][m].astype(object)
    disc, tax = L["l_discount"][m].astype(object), L["l_tax"][m].astype(object)
    cols = {k: [] for k in ("l_returnflag", "l_linestatus", "sum_qty", "sum_base_price", "sum_disc_price",
                            "sum_charge", "avg_qty", "avg_price", "avg_disc", "count_order")}
    for fv in sorted(set(f.tolist())):
        for sv in sorted(set(st[f == fv].tolist())):
            g = (f == fv) & (st == sv)
            n = int(g.sum())
            sq, sp, sd = int(qty[g].sum()), int(price[g].sum()), int(disc[g].sum())
            cols["l_returnflag"].append(chr(fv))
            cols["l_linestatus"].append(chr(sv))
            cols["sum_qty"].append(sq)
            cols["sum_base_price"].append(sp)
            cols["sum_disc_price"].append(int((price[g] * (100 - disc[g])).sum()))
            cols["sum_charge"].append(int((price[g] * (100 - disc[g]) * (100 + tax[g])).sum()))
            cols["avg_qty"].append(_avg(sq, n))
            cols["avg_price"].append(_avg(sp, n))
            cols["avg_disc"].append(_avg(sd, n))
            cols["count_order"].append(n)
    fmt = {"sum_qty": 2, "sum_base_price": 2, "sum_disc_price": 4, "sum_charge": 6, "avg_qty": 8, "avg_price": 8,
           "avg_disc": 8}
    return cols, {k: (fmt.get(k, 0), False) for k in cols}


def q3(T: dict, seg: int, date: int, k: int = 10) -> tuple:
    """sql/q3.sql; ties of (revenue, o_orderdate) broken by l_orderkey asc."""
    C, O, L = T["customer"], T["orders"], T["lineitem"]
    cust = set(C["c_custkey"][C["c_mktsegment"] == seg].tolist())
    om = O["o_orderdate"] < date
    odate, oship = {}, {}
    for key, ck, d, sp in zip(O["o_orderkey"][om].tolist(), O["o_custkey"][om].tolist(), O["o_orderdate"][om].tolist(),
                              O["o_shippriority"][om].tolist()):
        if ck in cust:
            odate[key] = d
            oship[key] = sp
    lm = L["l_shipdate"] > date
    rev: dict = {}
    for key, p, d in zip(L["l_orderkey"][lm].tolist(), L["l_extendedprice"][lm].tolist(), L["l_discount"][lm].tolist()):
        if key in odate:
            rev[key] = rev.get(key, 0) + p * (100 - d)
    top = sorted(rev.items(), key=lambda kv: (-kv[1], odate[kv[0]], kv[0]))[:k]
    cols = {"l_orderkey": [t[0] for t in top], "revenue": [t[1] for t in top],
            "o_orderdate": [odate[t[0]] for t in top], "o_shippriority": [oship[t[0]] for t in top]}
    return cols, {"l_orderkey": (0, False), "revenue": (4, False), "o_orderdate": (0, True),
                  "o_shippriority": (0, False)}


def q4(T: dict, lo: int, hi: int) -> tuple:
    """sql/q4.sql."""
    O, L = T["orders"], T["lineitem"]
    late = set(L["l_orderkey"][L["l_commitdate"] < L["l_receiptdate"]].tolist())
    m = (O["o_orderdate"] >= lo) & (O["o_orderdate"] < hi)
    counts: dict = {}
    for key, pr in zip(O["o_orderkey"][m].tolist(), O["o_orderpriority"][m].tolist()):
        if key in late:
            counts[pr] = counts.get(pr, 0) + 1
    prs = sorted(p for p in counts if p in PRIORITY_LABELS)
    cols = {"o_orderpriority": [PRIORITY_LABELS[p] for p in prs], "order_count": [counts[p] for p in prs]}
    return cols, {k: (0, False) for k in cols}


def q6(T: dict, lo: int, hi: int, disc_lo: int, disc_hi: int, qty_hi: int) -> tuple:
    """sql/q6.sql."""
    L = T["lineitem"]
    m = ((L["l_shipdate"] >= lo) & (L["l_shipdate"] < hi) & (L["l_discount"] >= disc_lo) & (L["l_discount"] <= disc_hi)
         & (L["l_quantity"] < qty_hi))
    rev = int((L["l_extendedprice"][m].astype(object) * L["l_discount"][m].astype(object)).sum())
    return {"revenue": [rev] if m.any() else []}, {"revenue": (4, False)}


def q12(T: dict, modes: tuple, lo: int, hi: int) -> tuple:
    """sql/q12.sql; modes are the full names (the catalog keeps 2 bytes)."""
    O, L = T["orders"], T["lineitem"]
    pri = dict(zip(O["o_orderkey"].tolist(), O["o_orderpriority"].tolist()))
    code = {m.encode()[:2]: m for m in modes}
    out: dict = {}
    m = ((L["l_commitdate"] < L["l_receiptdate"]) & (L["l_shipdate"] < L["l_commitdate"]) &
         (L["l_receiptdate"] >= lo) & (L["l_receiptdate"] < hi))
    for i in np.flatnonzero(m):
        mode = code.get(_s(L["l_shipmode"][i]))
        key = int(L["l_orderkey"][i])
        if mode is None or key not in pri:
            continue
        hi_lo = out.setdefault(mode, [0, 0])
        hi_lo[0 if pri[key] in (ord("1"), ord("2")) else 1] += 1
    ms = sorted(out)
    cols = {"l_shipmode": ms, "high_line_count": [out[x][0] for x in ms], "low_line_count": [out[x][1] for x in ms]}
    return cols, {k: (0, False) for k in cols}


def q13(T: dict, w1: bytes, w2: bytes) -> tuple:
    """sql/q13.sql: per customer the orders whose comment is NOT LIKE
    '%w1%w2%' (the comment up to its first NUL), then the histogram."""
    O, C = T["orders"], T["customer"]
    pat = re.compile(re.escape(w1) + b".*" + re.escape(w2), re.S)
    per: dict = {}
    for ck, com in zip(O["o_custkey"].tolist(), O["o_comment"]):
        if not pat.search(_s(com)):
            per[ck] = per.get(ck, 0) + 1
    hist: dict = {}
    for ck in C["c_custkey"].tolist():
        c = per.get(ck, 0)
        hist[c] = hist.get(c, 0) + 1
    rows = sorted(hist.items(), key=lambda kv: (-kv[1], -kv[0]))
    cols = {"c_count": [r[0] for r in rows], "custdist": [r[1] for r in rows]}
    return cols, {k: (0, False) for k in cols}


def q14(T: dict, prefix: bytes, lo: int, hi: int) -> tuple:
    """sql/q14.sql: 100 * promo revenue / revenue, truncated to 6 digits."""
    P, L = T["part"], T["lineitem"]
    parts = set(P["p_partkey"].tolist())
    promo = {int(k) for k, t in zip(P["p_partkey"], P["p_type"]) if bytes(t[:len(prefix)]) == prefix}
    m = (L["l_shipdate"] >= lo) & (L["l_shipdate"] < hi)
    tot = pro = n = 0
    for pk, p, d in zip(L["l_partkey"][m].tolist(), L["l_extendedprice"][m].tolist(), L["l_discount"][m].tolist()):
        if pk in parts:
            r = p * (100 - d)
            tot += r
            n += 1
            if pk in promo:
                pro += r
    return {"promo_revenue": [_avg(100 * pro, tot)] if n else []}, {"promo_revenue": (6, False)}


def q18(T: dict, qty: int, k: int = 100) -> tuple:
    """sql/q18.sql; ordered o_totalprice desc, o_orderdate asc, then by
    (c_custkey, o_orderkey)."""
    C, O, L = T["customer"], T["orders"], T["lineitem"]
    keys, inv = np.unique(L["l_orderkey"], return_inverse=True)
    sums = np.bincount(inv, weights=L["l_quantity"].astype(np.float64)).astype(np.int64)
    big = dict(zip(keys[sums > qty].tolist(), sums[sums > qty].tolist()))
    names = {int(ck): _s(n).decode() for ck, n in zip(C["c_custkey"], C["c_name"])}
    rows = []
    for ok, ck, d, tp in zip(O["o_orderkey"].tolist(), O["o_custkey"].tolist(), O["o_orderdate"].tolist(),
                             O["o_totalprice"].tolist()):
        if ok in big:
            rows.append((ck, ok, d, tp, big[ok]))
    rows.sort(key=lambda r: (-r[3], r[2], r[0], r[1]))
    rows = rows[:k]
    cols = {"c_name": [names[r[0]] for r in rows], "c_custkey": [r[0] for r in rows], "o_orderkey": [r[1] for r in rows],
            "o_orderdate": [r[2] for r in rows], "o_totalprice": [r[3] for r in rows],
            "sum(l_quantity)": [r[4] for r in rows]}
    return cols, {"c_name": (0, False), "c_custkey": (0, False), "o_orderkey": (0, False), "o_orderdate": (0, True),
                  "o_totalprice": (2, False), "sum(l_quantity)": (2, False)}


def q22(T: dict, codes) -> tuple:
    """sql/q22.sql."""
    C, O = T["customer"], T["orders"]
    ph = C["c_phone"]
    cc = (ph[:, 0].astype(np.int64) - 48) * 10 + (ph[:, 1].astype(np.int64) - 48)
    ins = np.isin(cc, list(codes))
    bal = C["c_acctbal"]
    pos = ins & (bal > 0)
    s, n = int(bal[pos].sum()), int(pos.sum())
    has = set(O["o_custkey"].tolist())
    out: dict = {}
    for ck, c, b, i in zip(C["c_custkey"].tolist(), cc.tolist(), bal.tolist(), ins.tolist()):
        if i and b * n > s and ck not in has:
            e = out.setdefault(c, [0, 0])
            e[0] += 1
            e[1] += b
    ks = sorted(out)
    cols = {"cntrycode": [f"{c:02d}" for c in ks], "numcust": [out[c][0] for c in ks],
            "totacctbal": [out[c][1] for c in ks]}
    return cols, {"cntrycode": (0, False), "numcust": (0, False), "totacctbal": (2, False)}


def q5(T: dict, region: bytes, lo: int, hi: int) -> tuple:
    """sql/q5.sql; ordered revenue desc, then n_name."""
    R, N, S, C, O, L = (T[t] for t in ("region", "nation", "supplier", "customer", "orders", "lineitem"))
    rk = {int(k) for k, nm in zip(R["r_regionkey"], R["r_name"]) if _s(nm) == region}
    nations = {int(k): _s(nm).decode() for k, nm, r in zip(N["n_nationkey"], N["n_name"], N["n_regionkey"]) if int(r) in rk}
    snat = {int(k): int(n) for k, n in zip(S["s_suppkey"], S["s_nationkey"]) if int(n) in nations}
    cnat = {int(k): int(n) for k, n in zip(C["c_custkey"], C["c_nationkey"]) if int(n) in nations}
    m = (O["o_orderdate"] >= lo) & (O["o_orderdate"] < hi)
    onat = {int(k): cnat[int(c)] for k, c in zip(O["o_orderkey"][m], O["o_custkey"][m]) if int(c) in cnat}
    rev: dict = {}
    for ok, sk, p, d in zip(L["l_orderkey"].tolist(), L["l_suppkey"].tolist(), L["l_extendedprice"].tolist(),
                            L["l_discount"].tolist()):
        n = onat.get(ok)
        if n is not None and snat.get(sk) == n:
            rev[n] = rev.get(n, 0) + p * (100 - d)
    rows = sorted(((nations[n], r) for n, r in rev.items()), key=lambda x: (-x[1], x[0]))
    return {"n_name": [r[0] for r in rows], "revenue": [r[1] for r in rows]}, {"n_name": (0, False),
                                                                                "revenue": (4, False)}


def q10(T: dict, lo: int, hi: int, flag: int, k: int = 20) -> tuple:
    """sql/q10.sql; ties of revenue broken by c_custkey asc."""
    C, O, L, N = T["customer"], T["orders"], T["lineitem"], T["nation"]
    m = (O["o_orderdate"] >= lo) & (O["o_orderdate"] < hi)
    ocust = dict(zip(O["o_orderkey"][m].tolist(), O["o_custkey"][m].tolist()))
    rev: dict = {}
    lm = L["l_returnflag"] == flag
    for ok, p, d in zip(L["l_orderkey"][lm].tolist(), L["l_extendedprice"][lm].tolist(), L["l_discount"][lm].tolist()):
        c = ocust.get(ok)
        if c is not None:
            rev[c] = rev.get(c, 0) + p * (100 - d)
    top = sorted(rev.items(), key=lambda kv: (-kv[1], kv[0]))[:k]
    pos = {int(ck): i for i, ck in enumerate(C["c_custkey"].tolist())}
    nat = {int(n): _s(nm).decode() for n, nm in zip(N["n_nationkey"], N["n_name"])}
    cols = {"c_custkey": [], "c_name": [], "revenue": [], "c_acctbal": [], "n_name": [], "c_address": [], "c_phone": [],
            "c_comment": []}
    for ck, r in top:
        i = pos[ck]
        cols["c_custkey"].append(ck)
        cols["c_name"].append(_s(C["c_name"][i]).decode())
        cols["revenue"].append(r)
        cols["c_acctbal"].append(int(C["c_acctbal"][i]))
        cols["n_name"].append(nat[int(C["c_nationkey"][i])])
        cols["c_address"].append(_s(C["c_address"][i]).decode())
        cols["c_phone"].append(_s(C["c_phone"][i]).decode())
        cols["c_comment"].append(_s(C["c_comment"][i]).decode())
    fm = {c: (0, False) for c in cols}
    fm["revenue"] = (4, False)
    fm["c_acctbal"] = (2, False)
    return cols, fm


def q19(T: dict, strings, modes) -> tuple:
    """sql/q19.sql; strings = 3 brands, 3 x 4 containers, ship instruction."""
    P, L = T["part"], T["lineitem"]
    brands, conts, instr = strings[:3], [set(strings[3 + 4 * c: 7 + 4 * c]) for c in range(3)], strings[15]
    sizes, qty = (5, 10, 15), ((100, 1100), (1000, 2000), (2000, 3000))
    cls = {}
    for k, b, c, sz in zip(P["p_partkey"].tolist(), P["p_brand"], P["p_container"], P["p_size"].tolist()):
        for j in range(3):
            if _s(b) == brands[j] and _s(c) in conts[j] and 1 <= sz <= sizes[j]:
                cls[k] = j
    mode_codes = {m.encode()[:2] for m in modes}
    rev = n = 0
    for i in range(L["l_partkey"].shape[0]):
        j = cls.get(int(L["l_partkey"][i]))
        if j is None or _s(L["l_shipmode"][i]) not in mode_codes or _s(L["l_shipinstruct"][i]) != instr:
            continue
        q = int(L["l_quantity"][i])
        if qty[j][0] <= q <= qty[j][1]:
            rev += int(L["l_extendedprice"][i]) * (100 - int(L["l_discount"][i]))
            n += 1
    # no qualifying row: no result row (the engine skips an empty group)
    return {"revenue": [rev] if n else []}, {"revenue": (4, False)}


def q11(T: dict, nation: bytes, frac_den: int) -> tuple:
    """sql/q11.sql with HAVING value > total / frac_den; value desc, then key."""
    N, S, PS = T["nation"], T["supplier"], T["partsupp"]
    nk = {int(k) for k, nm in zip(N["n_nationkey"], N["n_name"]) if _s(nm) == nation}
    sup = {int(k) for k, n in zip(S["s_suppkey"], S["s_nationkey"]) if int(n) in nk}
    val: dict = {}
    for pk, sk, c, a in zip(PS["ps_partkey"].tolist(), PS["ps_suppkey"].tolist(), PS["ps_supplycost"].tolist(),
                            PS["ps_availqty"].tolist()):
        if sk in sup:
            val[pk] = val.get(pk, 0) + c * a
    total = sum(val.values())
    rows = sorted(((k, v) for k, v in val.items() if v * frac_den > total), key=lambda r: (-r[1], r[0]))
    return {"ps_partkey": [r[0] for r in rows], "value": [r[1] for r in rows]}, {"ps_partkey": (0, False),
                                                                                "value": (2, False)}


def q15(T: dict, lo: int, hi: int) -> tuple:
    """sql/q15.sql; suppliers with the maximum quarter revenue, by key."""
    S, L = T["supplier"], T["lineitem"]
    m = (L["l_shipdate"] >= lo) & (L["l_shipdate"] < hi)
    rev: dict = {}
    for sk, p, d in zip(L["l_suppkey"][m].tolist(), L["l_extendedprice"][m].tolist(), L["l_discount"][m].tolist()):
        rev[sk] = rev.get(sk, 0) + p * (100 - d)
    suppliers = set(S["s_suppkey"].tolist())
    rev = {k: v for k, v in rev.items() if k in suppliers}
    cols = {"s_suppkey": [], "s_name": [], "s_address": [], "s_phone": [], "total_revenue": []}
    if rev:
        mx = max(rev.values())
        pos = {int(k): i for i, k in enumerate(S["s_suppkey"].tolist())}
        for k in sorted(k for k, v in rev.items() if v == mx):
            i = pos[k]
            cols["s_suppkey"].append(k)
            cols["s_name"].append(_s(S["s_name"][i]).decode())
            cols["s_address"].append(_s(S["s_address"][i]).decode())
            cols["s_phone"].append(_s(S["s_phone"][i]).decode())
            cols["total_revenue"].append(mx)
    fm = {c: (0, False) for c in cols}
    fm["total_revenue"] = (4, False)
    return cols, fm


def q16(T: dict, brand: bytes, type_prefix: bytes, sizes, w1: bytes, w2: bytes) -> tuple:
    """sql/q16.sql: per (p_brand, p_type, p_size) of the parts passing the
    brand / type-prefix / size-set filters, the number of distinct suppliers
    of their partsupp rows whose comment is not LIKE '%w1%w2%'; ordered by
    supplier_cnt desc, brand, type, size."""
    P, PS, S = T["part"], T["partsupp"], T["supplier"]
    pat = re.compile(re.escape(w1) + b".*" + re.escape(w2), re.S)
    complained = {int(k) for k, c in zip(S["s_suppkey"], S["s_comment"]) if pat.search(_s(c))}
    sizes = set(sizes)
    group = {}
    for k, b, t, z in zip(P["p_partkey"].tolist(), P["p_brand"], P["p_type"], P["p_size"].tolist()):
        b, t = _s(b), _s(t)
        if b != brand and not t.startswith(type_prefix) and z in sizes:
            group[k] = (b.decode(), t.decode(), z)
    supp: dict = {}
    for pk, sk in zip(PS["ps_partkey"].tolist(), PS["ps_suppkey"].tolist()):
        g = group.get(pk)
        if g is not None and sk not in complained:
            supp.setdefault(g, set()).add(sk)
    rows = sorted(((g[0], g[1], g[2], len(v)) for g, v in supp.items()), key=lambda r: (-r[3], r[0], r[1], r[2]))
    names = ["p_brand", "p_type", "p_size", "supplier_cnt"]
    return {n: [r[c] for r in rows] for c, n in enumerate(names)}, {n: (0, False) for n in names}


def q17(T: dict, brand: bytes, container: bytes) -> tuple:
    """sql/q17.sql: sum(l_extendedprice) / 7 as trunc(sum * 10^6 / 7), 8 digits."""
    P, L = T["part"], T["lineitem"]
    parts = {int(k) for k, b, c in zip(P["p_partkey"], P["p_brand"], P["p_container"])
             if _s(b) == brand and _s(c) == container}
    qs: dict = {}
    qc: dict = {}
    for pk, q in zip(L["l_partkey"].tolist(), L["l_quantity"].tolist()):
        if pk in parts:
            qs[pk] = qs.get(pk, 0) + q
            qc[pk] = qc.get(pk, 0) + 1
    tot = n = 0
    for pk, q, p in zip(L["l_partkey"].tolist(), L["l_quantity"].tolist(), L["l_extendedprice"].tolist()):
        if pk in parts and 5 * q * qc[pk] < qs[pk]:
            tot += p
            n += 1
    return {"avg_yearly": [tot * 1_000_000 // 7] if n else []}, {"avg_yearly": (8, False)}


def q2(T: dict, region: bytes, suffix: bytes, size: int, k: int = 100) -> tuple:
    """sql/q2.sql; ordered s_acctbal desc, n_name, s_name, p_partkey."""
    R, N, S, P, PS = (T[t] for t in ("region", "nation", "supplier", "part", "partsupp"))
    rk = {int(x) for x, nm in zip(R["r_regionkey"], R["r_name"]) if _s(nm) == region}
    nat = {int(x): _s(nm).decode() for x, nm, r in zip(N["n_nationkey"], N["n_name"], N["n_regionkey"]) if int(r) in rk}
    spos = {int(x): i for i, (x, n) in enumerate(zip(S["s_suppkey"], S["s_nationkey"])) if int(n) in nat}
    parts = {int(x): i for i, (x, sz, t) in enumerate(zip(P["p_partkey"], P["p_size"], P["p_type"]))
             if int(sz) == size and _s(t).endswith(suffix)}
    mn: dict = {}
    cand = []
    for pk, sk, c in zip(PS["ps_partkey"].tolist(), PS["ps_suppkey"].tolist(), PS["ps_supplycost"].tolist()):
        if pk in parts and sk in spos:
            mn[pk] = min(mn.get(pk, c), c)
            cand.append((pk, sk, c))
    rows = []
    for pk, sk, c in cand:
        if c == mn[pk]:
            i, j = spos[sk], parts[pk]
            rows.append((int(S["s_acctbal"][i]), _s(S["s_name"][i]).decode(), nat[int(S["s_nationkey"][i])], pk,
                         _s(P["p_mfgr"][j]).decode(), _s(S["s_address"][i]).decode(), _s(S["s_phone"][i]).decode(),
                         _s(S["s_comment"][i]).decode()))
    rows.sort(key=lambda r: (-r[0], r[2], r[1], r[3]))
    rows = rows[:k]
    names = ["s_acctbal", "s_name", "n_name", "p_partkey", "p_mfgr", "s_address", "s_phone", "s_comment"]
    fm = {c: (0, False) for c in names}
    fm["s_acctbal"] = (2, False)
    return {c: [r[i] for r in rows] for i, c in enumerate(names)}, fm


def q7(T: dict, n1: bytes, n2: bytes, lo: int, hi: int) -> tuple:
    """sql/q7.sql; ordered supp_nation, cust_nation, l_year."""
    N, S, C, O, L = (T[t] for t in ("nation", "supplier", "customer", "orders", "lineitem"))
    code = {}
    for x, nm in zip(N["n_nationkey"], N["n_name"]):
        if _s(nm) in (n1, n2):
            code[int(x)] = _s(nm).decode()
    snat = {int(x): code[int(n)] for x, n in zip(S["s_suppkey"], S["s_nationkey"]) if int(n) in code}
    cnat = {int(x): code[int(n)] for x, n in zip(C["c_custkey"], C["c_nationkey"]) if int(n) in code}
    onat = {int(x): cnat[int(c)] for x, c in zip(O["o_orderkey"], O["o_custkey"]) if int(c) in cnat}
    m = (L["l_shipdate"] >= lo) & (L["l_shipdate"] <= hi)
    rev: dict = {}
    for ok, sk, d8, p, d in zip(L["l_orderkey"][m].tolist(), L["l_suppkey"][m].tolist(), L["l_shipdate"][m].tolist(),
                                L["l_extendedprice"][m].tolist(), L["l_discount"][m].tolist()):
        a, b = snat.get(sk), onat.get(ok)
        if a is not None and b is not None and a != b:
            key = (a, b, d8 // 10000)
            rev[key] = rev.get(key, 0) + p * (100 - d)
    rows = sorted(rev.items())
    cols = {"supp_nation": [r[0][0] for r in rows], "cust_nation": [r[0][1] for r in rows],
            "l_year": [r[0][2] for r in rows], "revenue": [r[1] for r in rows]}
    fm = {c: (0, False) for c in cols}
    fm["revenue"] = (4, False)
    return cols, fm


def q8(T: dict, region: bytes, nation: bytes, ptype: bytes, lo: int, hi: int) -> tuple:
    """sql/q8.sql; mkt_share = trunc(marked * 10^6 / volume), 6 digits."""
    R, N, S, P, C, O, L = (T[t] for t in ("region", "nation", "supplier", "part", "customer", "orders", "lineitem"))
    rk = {int(x) for x, nm in zip(R["r_regionkey"], R["r_name"]) if _s(nm) == region}
    amer = {int(x) for x, r in zip(N["n_nationkey"], N["n_regionkey"]) if int(r) in rk}
    mark = {int(x) for x, nm in zip(N["n_nationkey"], N["n_name"]) if _s(nm) == nation}
    smark = {int(x): int(n) in mark for x, n in zip(S["s_suppkey"], S["s_nationkey"])}
    parts = {int(x) for x, t in zip(P["p_partkey"], P["p_type"]) if _s(t) == ptype}
    cust = {int(x) for x, n in zip(C["c_custkey"], C["c_nationkey"]) if int(n) in amer}
    m = (O["o_orderdate"] >= lo) & (O["o_orderdate"] <= hi)
    oyear = {int(x): int(d) // 10000 for x, c, d in zip(O["o_orderkey"][m], O["o_custkey"][m], O["o_orderdate"][m])
             if int(c) in cust}
    vol: dict = {}
    for pk, sk, ok, p, d in zip(L["l_partkey"].tolist(), L["l_suppkey"].tolist(), L["l_orderkey"].tolist(),
                                L["l_extendedprice"].tolist(), L["l_discount"].tolist()):
        y = oyear.get(ok)
        if y is None or pk not in parts or sk not in smark:
            continue
        e = vol.setdefault(y, [0, 0])
        e[0] += p * (100 - d)
        if smark[sk]:
            e[1] += p * (100 - d)
    ys = sorted(vol)
    return {"o_year": ys, "mkt_share": [_avg(vol[y][1], vol[y][0]) for y in ys]}, {"o_year": (0, False),
                                                                                  "mkt_share": (6, False)}


def q21(T: dict, nation: bytes, status: int, k: int = 100) -> tuple:
    """sql/q21.sql; ordered numwait desc, s_name."""
    N, S, O, L = T["nation"], T["supplier"], T["orders"], T["lineitem"]
    nk = {int(x) for x, nm in zip(N["n_nationkey"], N["n_name"]) if _s(nm) == nation}
    sname = {int(x): _s(nm).decode() for x, nm, n in zip(S["s_suppkey"], S["s_name"], S["s_nationkey"]) if int(n) in nk}
    fo = set(O["o_orderkey"][O["o_orderstatus"] == status].tolist())
    supp: dict = {}
    late: dict = {}
    lines = list(zip(L["l_orderkey"].tolist(), L["l_suppkey"].tolist(),
                     (L["l_receiptdate"] > L["l_commitdate"]).tolist()))
    for ok, sk, lt in lines:
        supp.setdefault(ok, set()).add(sk)
        if lt:
            late.setdefault(ok, set()).add(sk)
    wait: dict = {}
    for ok, sk, lt in lines:
        if not lt or sk not in sname or ok not in fo:
            continue
        if supp[ok] - {sk} and not (late[ok] - {sk}):
            wait[sk] = wait.get(sk, 0) + 1
    rows = sorted(((sname[x], n) for x, n in wait.items()), key=lambda r: (-r[1], r[0]))[:k]
    return {"s_name": [r[0] for r in rows], "numwait": [r[1] for r in rows]}, {"s_name": (0, False),
                                                                              "numwait": (0, False)}


def q9(T: dict, word: bytes) -> tuple:
    """sql/q9.sql; ordered nation asc, o_year desc."""
    P, S, PS, O, L, N = (T[t] for t in ("part", "supplier", "partsupp", "orders", "lineitem", "nation"))
    parts = {int(k) for k, nm in zip(P["p_partkey"], P["p_name"]) if word in _s(nm)}
    cost = {(int(p), int(sk)): int(c) for p, sk, c in zip(PS["ps_partkey"], PS["ps_suppkey"], PS["ps_supplycost"])
            if int(p) in parts}
    nat = {int(k): _s(nm).decode() for k, nm in zip(N["n_nationkey"], N["n_name"])}
    snat = {int(k): nat[int(n)] for k, n in zip(S["s_suppkey"], S["s_nationkey"])}
    oyear = dict(zip(O["o_orderkey"].tolist(), (O["o_orderdate"] // 10000).tolist()))
    prof: dict = {}
    for pk, sk, ok, q, p, d in zip(L["l_partkey"].tolist(), L["l_suppkey"].tolist(), L["l_orderkey"].tolist(),
                                   L["l_quantity"].tolist(), L["l_extendedprice"].tolist(), L["l_discount"].tolist()):
        c = cost.get((pk, sk))
        if c is None or sk not in snat or ok not in oyear:
            continue
        key = (snat[sk], oyear[ok])
        prof[key] = prof.get(key, 0) + p * (100 - d) - c * q
    rows = sorted(prof.items(), key=lambda kv: (kv[0][0], -kv[0][1]))
    return ({"nation": [r[0][0] for r in rows], "o_year": [r[0][1] for r in rows], "sum_profit": [r[1] for r in rows]},
            {"nation": (0, False), "o_year": (0, False), "sum_profit": (4, False)})


def q20(T: dict, prefix: bytes, nation: bytes, lo: int, hi: int) -> tuple:
    """sql/q20.sql; ordered by s_name."""
    P, PS, L, N, S = (T[t] for t in ("part", "partsupp", "lineitem", "nation", "supplier"))
    parts = {int(k) for k, nm in zip(P["p_partkey"], P["p_name"]) if _s(nm).startswith(prefix)}
    m = (L["l_shipdate"] >= lo) & (L["l_shipdate"] < hi)
    qty: dict = {}
    for pk, sk, q in zip(L["l_partkey"][m].tolist(), L["l_suppkey"][m].tolist(), L["l_quantity"][m].tolist()):
        if pk in parts:
            qty[(pk, sk)] = qty.get((pk, sk), 0) + q
    ok = set()
    for pk, sk, a in zip(PS["ps_partkey"].tolist(), PS["ps_suppkey"].tolist(), PS["ps_availqty"].tolist()):
        s_ = qty.get((pk, sk))
        if pk in parts and s_ is not None and 200 * a > s_:
            ok.add(sk)
    nk = {int(k) for k, nm in zip(N["n_nationkey"], N["n_name"]) if _s(nm) == nation}
    rows = sorted((_s(nm).decode(), _s(ad).decode()) for k, nm, ad, n in zip(S["s_suppkey"], S["s_name"], S["s_address"],
                                                                           S["s_nationkey"])
                  if int(k) in ok and int(n) in nk)
    return {"s_name": [r[0] for r in rows], "s_address": [r[1] for r in rows]}, {"s_name": (0, False),
                                                                                "s_address": (0, False)}


def run(T: dict, q: str, P: dict) -> tuple:
    """Query q with the constants P (the names of the engine's scalars:
    q1_cutoff, q3_seg, ... plus words / prefix for Q13 / Q14)."""
    if q == "Q1":
        return q1(T, P["q1_cutoff"])
    if q == "Q3":
        return q3(T, P["q3_seg"], P["q3_date"])
    if q == "Q4":
        return q4(T, P["q4_date_lo"], P["q4_date_hi"])
    if q == "Q6":
        return q6(T, P["q6_date_lo"], P["q6_date_hi"], P["q6_disc_lo"], P["q6_disc_hi"], P["q6_qty_hi"])
    if q == "Q12":
        modes = tuple(bytes([c & 255, c >> 8]).rstrip(b"\0").decode() for c in (P["q12_mode1"], P["q12_mode2"]))
        full = {"MA": "MAIL", "SH": "SHIP", "RE": "REG AIR", "AI": "AIR", "RA": "RAIL", "TR": "TRUCK", "FO": "FOB"}
        return q12(T, tuple(full[m] for m in modes), P["q12_date_lo"], P["q12_date_hi"])
    if q == "Q13":
        return q13(T, *P["words"])
    if q == "Q14":
        return q14(T, P["prefix"], P["q14_date_lo"], P["q14_date_hi"])
    if q == "Q18":
        return q18(T, P["q18_qty"])
    if q == "Q5":
        return q5(T, P["strings"][0], P["q5_date_lo"], P["q5_date_hi"])
    if q == "Q10":
        return q10(T, P["q10_date_lo"], P["q10_date_hi"], P["q10_flag"])
    if q == "Q9":
        return q9(T, P["words"][0])
    if q == "Q20":
        return q20(T, P["prefix"], P["strings"][0], P["q20_date_lo"], P["q20_date_hi"])
    if q == "Q21":
        return q21(T, P["strings"][0], P["q21_status"])
    if q == "Q2":
        return q2(T, P["strings"][0], P["strings"][1], P["q2_size"])
    if q == "Q7":
        return q7(T, P["strings"][0], P["strings"][1], P["q7_date_lo"], P["q7_date_hi"])
    if q == "Q8":
        return q8(T, P["strings"][0], P["strings"][1], P["strings"][2], P["q8_date_lo"], P["q8_date_hi"])
    if q == "Q11":
        return q11(T, P["strings"][0], P["q11_frac_den"])
    if q == "Q15":
        return q15(T, P["q15_date_lo"], P["q15_date_hi"])
    if q == "Q16":
        sizes = [v for v in range(64) if (P["q16_sizes"] >> v) & 1]
        return q16(T, P["strings"][0], P["strings"][1], sizes, *P["words"])
    if q == "Q17":
        return q17(T, P["strings"][0], P["strings"][1])
    if q == "Q19":
        full = {"AI": "AIR", "RE": "REG AIR", "MA": "MAIL", "SH": "SHIP", "RA": "RAIL", "TR": "TRUCK", "FO": "FOB"}
        modes = tuple(full[bytes([c & 255, c >> 8]).decode()] for c in (P["q19_mode1"], P["q19_mode2"]))
        return q19(T, P["strings"], modes)
    if q == "Q22":
        codes = [c for c in range(64) if (P["q22_mask_lo"] >> c) & 1] + \
                [c + 64 for c in range(64) if (P["q22_mask_hi"] >> c) & 1]
        return q22(T, codes)
    raise ValueError(q)


def canonical(cols: dict, fmts: dict) -> str:
    """CSV with exact decimals for scaled integers and YYYY-MM-DD dates."""

    def field(v, scale, is_date):
        if isinstance(v, str):
            return '"' + v.replace('"', '""') + '"' if any(ch in v for ch in ',"\n\r') else v
        if is_date:
            return f"{v // 10000:04d}-{v // 100 % 100:02d}-{v % 100:02d}"
        if scale <= 0:
            return str(v)
        sgn = "-" if v < 0 else ""
        digits = str(abs(v)).rjust(scale + 1, "0")
        return f"{sgn}{digits[:-scale]}.{digits[-scale:]}"

    names = list(cols)
    n = len(cols[names[0]]) if names else 0
    lines = [",".join(names)] + [",".join(field(cols[c][r], *fmts[c]) for c in names) for r in range(n)]
    return "\n".join(lines) + "\n"
