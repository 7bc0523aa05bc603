"""TPC-H front end of the engine: ``.tbl`` -> ``.colbin`` ingest, loading
``.colbin`` tables into the engine's column formats, the named TPC-H plans
(Q1, Q3, Q4, Q6, Q12, Q13, Q14, Q18, Q22) with the constants of their SQL text,
canonical CSV and the golden check.

Counterparts in the reference: ``tools/tbl_to_colbin.cpp:106-172`` (ingest),
``loadColumnsFromBinary`` (``src/infra.h:579-664``), the per-query driver
``runCodegenQuery`` (``codegen/codegen_main.cpp:271-1776``) with its CPU
post-processing (top-k ties, names), and ``--check`` / ``--save-golden``
(``codegen/codegen_main.cpp:1722-1760``, golden files named
``<query>_<scale>.csv``).

Column formats after loading: INT int32, DATE int32 YYYYMMDD, FLOAT int64
hundredths (round(float32 * 100): exact for |v| < 131072, i.e. every TPC-H
money column except o_totalprice, whose float32 in the .colbin format already
lost cents above that), CHAR1 uint8, CHAR_FIXED uint8 (rows, width).
"""

from __future__ import annotations

import ctypes as C
import datetime as _dt
import os
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native as N
from . import engine as E

INT, FLOAT, DATE, CHAR1, CHAR_FIXED = 0, 1, 2, 3, 4
TABLES = ["lineitem", "orders", "customer", "supplier", "part", "partsupp", "nation", "region"]
QUERIES = ["Q1", "Q2", "Q3", "Q4", "Q5", "Q6", "Q7", "Q8", "Q9", "Q10", "Q11", "Q12", "Q13", "Q14", "Q15", "Q16",
           "Q17", "Q18", "Q19", "Q20", "Q21", "Q22"]
CHECK_ABS_TOL = 1e-2  # the reference's --check defaults (codegen_main.cpp:47-48)
CHECK_REL_TOL = 1e-4


def _eng():
    return N.load()[1]


# ---------------------------------------------------------------------------
# Catalog, ingest, .colbin
# ---------------------------------------------------------------------------

def table_columns(table: str) -> list:
    """[(name, index, dtype, width)] of a TPC-H table in file order."""
    names = N.two_call_string(_eng().gdb_tpch_table_columns, table.encode()).split(",")
    out = []
    for n in names:
        idx, dt, w = C.c_int(), C.c_int(), C.c_int()
        N.eng_check(_eng().gdb_tpch_column(table.encode(), n.encode(), C.byref(idx), C.byref(dt), C.byref(w)))
        out.append((n, idx.value, dt.value, w.value))
    return out


def tbl_to_colbin(tbl_path, table: str, force: bool = False, threads: int = 0) -> bool:
    """Build <base>.colbin next to the .tbl unless it is up to date; True when written."""
    built = C.c_int()
    N.eng_check(_eng().gdb_tbl_to_colbin(str(tbl_path).encode(), table.encode(), int(force), threads, C.byref(built)))
    return bool(built.value)


def ingest(data_dir, tables=None, force: bool = False, threads: int = 0) -> dict:
    """The tbl_to_colbin tool over a directory: {table: built?} for every
    <table>.tbl present."""
    d = Path(data_dir)
    out = {}
    for t in tables or TABLES:
        p = d / f"{t}.tbl"
        if p.exists():
            out[t] = tbl_to_colbin(p, t, force, threads)
    return out


class Colbin:
    """Memory-mapped .colbin file; tbl_path (optional) enables the staleness check."""

    def __init__(self, path, tbl_path=None):
        self._h = C.c_void_p()
        N.eng_check(_eng().gdb_colbin_open(str(path).encode(), str(tbl_path).encode() if tbl_path else None,
                                           C.byref(self._h)))
        rows, ncols = C.c_uint64(), C.c_int()
        N.eng_check(_eng().gdb_colbin_shape(self._h, C.byref(rows), C.byref(ncols)))
        self.rows, self.ncols = rows.value, ncols.value

    def raw(self, index: int):
        """(bytes copy as uint8 array, dtype, width) of the column with this index."""
        p, nb, dt, w = C.c_void_p(), C.c_uint64(), C.c_int(), C.c_int()
        N.eng_check(_eng().gdb_colbin_column(self._h, index, C.byref(p), C.byref(nb), C.byref(dt), C.byref(w)))
        buf = (C.c_uint8 * nb.value).from_address(p.value) if nb.value else b""
        return np.frombuffer(buf, dtype=np.uint8).copy(), dt.value, w.value

    def close(self):
        if self._h:
            _eng().gdb_colbin_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_engine_format(raw: np.ndarray, dtype: int, width: int, rows: int) -> np.ndarray:
    if dtype in (INT, DATE):
        return raw.view(np.int32)
    if dtype == FLOAT:
        return np.rint(raw.view(np.float32).astype(np.float64) * 100.0).astype(np.int64)
    if dtype == CHAR1:
        return raw
    return raw.reshape(rows, width)


def load_colbin(path, table: str, columns=None, tbl_path=None) -> dict:
    """Columns of a .colbin file in the engine's formats, {name: array}."""
    want = set(columns) if columns else None
    out = {}
    with Colbin(path, tbl_path) as f:
        for name, idx, dt, w in table_columns(table):
            if want is not None and name not in want:
                continue
            raw, fdt, fw = f.raw(idx)
            if fdt != dt or (dt == CHAR_FIXED and fw != w):
                raise ValueError(f"{path}: column {name} has type {fdt}/{fw}, expected {dt}/{w}")
            out[name] = to_engine_format(raw, dt, w, f.rows)
    return out


def load_dir(data_dir, tables=None) -> dict:
    d = Path(data_dir)
    return {t: load_colbin(d / f"{t}.colbin", t, tbl_path=d / f"{t}.tbl") for t in (tables or TABLES)
            if (d / f"{t}.colbin").exists()}


# ---------------------------------------------------------------------------
# The queries' constants (from the SQL text, sql/q*.sql)
# ---------------------------------------------------------------------------

def ymd(y: int, m: int, d: int) -> int:
    return y * 10000 + m * 100 + d


def shift(date: int, days: int = 0, months: int = 0, years: int = 0) -> int:
    """YYYYMMDD +/- an interval (month arithmetic keeps the day, as SQL's
    DATE + INTERVAL on the first of a month)."""
    y, m, d = date // 10000, date // 100 % 100, date % 100
    m0 = (y + years) * 12 + (m - 1) + months
    t = _dt.date(m0 // 12, m0 % 12 + 1, d) + _dt.timedelta(days=days)
    return ymd(t.year, t.month, t.day)


def mode_code(mode: str) -> int:
    b = mode.encode()[:2].ljust(2, b"\0")
    return b[0] | (b[1] << 8)


Q22_CODES = (13, 31, 23, 29, 30, 18, 17)
Q16_SIZES = (49, 14, 23, 45, 19, 3, 36, 9)

PARAMS = {
    "Q1": {"q1_cutoff": shift(ymd(1998, 12, 1), days=-90)},
    "Q3": {"q3_seg": ord("B"), "q3_date": ymd(1995, 3, 15)},  # BUILDING (CHAR1 keeps the first byte)
    "Q4": {"q4_date_lo": ymd(1993, 7, 1), "q4_date_hi": shift(ymd(1993, 7, 1), months=3)},
    "Q6": {"q6_date_lo": ymd(1994, 1, 1), "q6_date_hi": shift(ymd(1994, 1, 1), years=1), "q6_disc_lo": 5,
           "q6_disc_hi": 7, "q6_qty_hi": 2400},
    "Q12": {"q12_mode1": mode_code("MAIL"), "q12_mode2": mode_code("SHIP"), "q12_date_lo": ymd(1994, 1, 1),
            "q12_date_hi": shift(ymd(1994, 1, 1), years=1)},
    "Q13": {"words": (b"special", b"requests")},
    "Q14": {"prefix": b"PROMO", "q14_date_lo": ymd(1995, 9, 1), "q14_date_hi": shift(ymd(1995, 9, 1), months=1)},
    "Q18": {"q18_qty": 30000},
    "Q5": {"strings": (b"ASIA",), "q5_date_lo": ymd(1994, 1, 1), "q5_date_hi": shift(ymd(1994, 1, 1), years=1)},
    "Q11": {"strings": (b"GERMANY",), "q11_frac_den": 10000},  # value > total * 0.0001
    "Q15": {"q15_date_lo": ymd(1996, 1, 1), "q15_date_hi": shift(ymd(1996, 1, 1), months=3)},
    "Q16": {"words": (b"Customer", b"Complaints"), "strings": (b"Brand#45", b"MEDIUM POLISHED"),
            "q16_sizes": sum(1 << v for v in Q16_SIZES)},
    "Q17": {"strings": (b"Brand#23", b"MED BOX")},
    "Q2": {"strings": (b"EUROPE", b"BRASS"), "q2_size": 15},
    "Q21": {"strings": (b"SAUDI ARABIA",), "q21_status": ord("F")},
    "Q9": {"words": (b"green", b""), "q9_year0": 1992},
    "Q20": {"prefix": b"forest", "strings": (b"CANADA",), "q20_date_lo": ymd(1994, 1, 1),
            "q20_date_hi": shift(ymd(1994, 1, 1), years=1)},
    "Q7": {"strings": (b"FRANCE", b"GERMANY"), "q7_date_lo": ymd(1995, 1, 1), "q7_date_hi": ymd(1996, 12, 31),
           "q7_year0": 1995},
    "Q8": {"strings": (b"AMERICA", b"BRAZIL", b"ECONOMY ANODIZED STEEL"), "q8_date_lo": ymd(1995, 1, 1),
           "q8_date_hi": ymd(1996, 12, 31), "q8_year0": 1995},
    "Q10": {"q10_date_lo": ymd(1993, 10, 1), "q10_date_hi": shift(ymd(1993, 10, 1), months=3), "q10_flag": ord("R")},
    "Q19": {"q19_mode1": mode_code("AIR"), "q19_mode2": mode_code("REG AIR"),
            "strings": (b"Brand#12", b"Brand#23", b"Brand#34", b"SM CASE", b"SM BOX", b"SM PACK", b"SM PKG",
                        b"MED BAG", b"MED BOX", b"MED PKG", b"MED PACK", b"LG CASE", b"LG BOX", b"LG PACK", b"LG PKG",
                        b"DELIVER IN PERSON")},
    "Q22": {"q22_mask_lo": sum(1 << c for c in Q22_CODES if c < 64),
            "q22_mask_hi": sum(1 << (c - 64) for c in Q22_CODES if c >= 64)},
}

PLAN_OPTIONS = {"Q1": "schema=tpch", "Q3": "schema=tpch", "Q6": "schema=tpch"}
SCANNED = {"Q1": ["lineitem"], "Q3": ["customer", "orders", "lineitem"], "Q4": ["lineitem", "orders"],
           "Q6": ["lineitem"], "Q12": ["orders", "lineitem"], "Q13": ["orders", "customer"],
           "Q14": ["part", "lineitem"], "Q18": ["lineitem", "orders"], "Q22": ["customer", "orders"],
           "Q5": ["region", "nation", "supplier", "customer", "orders", "lineitem"],
           "Q10": ["orders", "lineitem", "customer"], "Q19": ["part", "lineitem"],
           "Q11": ["nation", "supplier", "partsupp", "part"], "Q15": ["lineitem", "supplier"],
           "Q16": ["supplier", "part", "partsupp"], "Q17": ["part", "lineitem"], "Q2": ["region", "nation", "supplier", "part", "partsupp"],
           "Q7": ["nation", "supplier", "customer", "orders", "lineitem"],
           "Q8": ["region", "nation", "supplier", "part", "customer", "orders", "lineitem"],
           "Q21": ["nation", "supplier", "orders", "lineitem"],
           "Q9": ["part", "partsupp", "supplier", "orders", "lineitem"],
           "Q20": ["part", "partsupp", "lineitem", "nation", "supplier"]}
# Output columns in the SQL select list's order.
SELECT = {
    "Q1": ["l_returnflag", "l_linestatus", "sum_qty", "sum_base_price", "sum_disc_price", "sum_charge", "avg_qty",
           "avg_price", "avg_disc", "count_order"],
    "Q3": ["l_orderkey", "revenue", "o_orderdate", "o_shippriority"],
    "Q4": ["o_orderpriority", "order_count"],
    "Q6": ["revenue"],
    "Q12": ["l_shipmode", "high_line_count", "low_line_count"],
    "Q13": ["c_count", "custdist"],
    "Q14": ["promo_revenue"],
    "Q18": ["c_name", "c_custkey", "o_orderkey", "o_orderdate", "o_totalprice", "sum(l_quantity)"],
    "Q22": ["cntrycode", "numcust", "totacctbal"],
    "Q5": ["n_name", "revenue"],
    "Q10": ["c_custkey", "c_name", "revenue", "c_acctbal", "n_name", "c_address", "c_phone", "c_comment"],
    "Q19": ["revenue"],
    "Q11": ["ps_partkey", "value"],
    "Q15": ["s_suppkey", "s_name", "s_address", "s_phone", "total_revenue"],
    "Q16": ["p_brand", "p_type", "p_size", "supplier_cnt"],
    "Q17": ["avg_yearly"],
    "Q2": ["s_acctbal", "s_name", "n_name", "p_partkey", "p_mfgr", "s_address", "s_phone", "s_comment"],
    "Q7": ["supp_nation", "cust_nation", "l_year", "revenue"],
    "Q8": ["o_year", "mkt_share"],
    "Q21": ["s_name", "numwait"],
    "Q9": ["nation", "o_year", "sum_profit"],
    "Q20": ["s_name", "s_address"],
}


def _pow2(n: int) -> int:
    return 1 << max(6, int(n - 1).bit_length())


# ---------------------------------------------------------------------------
# Database and runs
# ---------------------------------------------------------------------------

class TpchDB:
    """TPC-H tables in host memory (engine formats) and, on first use, in HBM;
    key-range symbols computed once (registerMaxKeySymbols,
    codegen/execution/max_key_symbols.cpp:188-235)."""

    def __init__(self, ctx: E.Context, host: dict):
        self.ctx = ctx
        self.host = host
        self.dev: dict = {}
        h = host
        self.symbols = {}
        if "customer" in h:
            self.symbols["c_bits"] = int(h["customer"]["c_custkey"].max(initial=0)) + 1
        if "orders" in h:
            self.symbols["o_bits"] = int(h["orders"]["o_orderkey"].max(initial=0)) + 1
        if "lineitem" in h:
            self.symbols["o_bits"] = max(self.symbols.get("o_bits", 1), int(h["lineitem"]["l_orderkey"].max(initial=0)) + 1)
        if "part" in h:
            self.symbols["p_bits"] = int(h["part"]["p_partkey"].max(initial=0)) + 1
        if "lineitem" in h:
            self.symbols["p_bits"] = max(self.symbols.get("p_bits", 1), int(h["lineitem"]["l_partkey"].max(initial=0)) + 1)
        if "orders" in h:
            self.symbols["c_bits"] = max(self.symbols.get("c_bits", 1), int(h["orders"]["o_custkey"].max(initial=0)) + 1)
        for t, key, sym in (("region", "r_regionkey", "r_bits"), ("nation", "n_nationkey", "n_bits"),
                            ("supplier", "s_suppkey", "s_bits")):
            if t in h:
                self.symbols[sym] = int(h[t][key].max(initial=0)) + 1
        if "nation" in h and "region" in h:
            self.symbols["r_bits"] = max(self.symbols["r_bits"], int(h["nation"]["n_regionkey"].max(initial=0)) + 1)
        if "lineitem" in h and "supplier" in h:
            self.symbols["s_bits"] = max(self.symbols["s_bits"], int(h["lineitem"]["l_suppkey"].max(initial=0)) + 1)
        if "partsupp" in h:
            ps = h["partsupp"]
            self.symbols["p_bits"] = max(self.symbols.get("p_bits", 1), int(ps["ps_partkey"].max(initial=0)) + 1)
            self.symbols["s_bits"] = max(self.symbols.get("s_bits", 1), int(ps["ps_suppkey"].max(initial=0)) + 1)

    def device(self, table: str) -> E.DeviceTable:
        t = self.dev.get(table)
        if t is None:
            t = E.DeviceTable.from_host(self.ctx, table, self.host[table])
            self.dev[table] = t
        return t

    def free(self) -> None:
        for t in self.dev.values():
            t.free(self.ctx)
        self.dev.clear()


@dataclass
class TpchRun:
    query: str
    columns: dict            # name -> list of Python values (int, or str)
    formats: dict            # name -> (scale digits, is_date)
    timing: E.Timing
    plan: E.HipQueryPlan = None
    executor: E.HipGenericExecutor = None
    raw: E.GenericResult = None
    canonical: str = field(default="")


STREAM_TABLE = {"Q11": "partsupp"}


def stream_table(query: str) -> str:
    """The table a chunked run streams: lineitem where the query reads it,
    partsupp for Q11 (the reference's Tier C, metal_plan_builder.cpp:2089-2102),
    otherwise orders (the reference's chunk loop streams one table per query,
    codegen_main.cpp:686-808)."""
    if query in STREAM_TABLE:
        return STREAM_TABLE[query]
    return "lineitem" if "lineitem" in SCANNED[query] else "orders"


def parse_rows(text: str) -> int:
    """N[K|M|G] row counts of --chunk (K = 1024, as codegen_main.cpp parses them)."""
    t = text.strip().upper()
    mult = {"K": 1 << 10, "M": 1 << 20, "G": 1 << 30}.get(t[-1:], 1)
    n = int(t[:-1] if t[-1:] in "KMG" else t) * mult
    if n <= 0:
        raise ValueError(f"chunk rows must be > 0: {text}")
    return n


# Automatic chunking (codegen_main.cpp:453-624): a chunkable plan streams its
# stream table when the projected working set (the columns its kernels read,
# resident tables plus the whole stream table) exceeds 3/4 of the memory
# budget; the chunk then takes (budget/2 - resident bytes) over two slots,
# at least budget/8 and 64 MiB, at least 256K rows and at most the table.
# An explicit chunk size for a working set that fits is downgraded to a
# direct run unless forced (the reference's GPUDB_FORCE_CHUNK). The budget
# here is free HBM (the reference takes min(physical memory, the Metal
# working-set limit) of its unified memory).
AUTO_CHUNK_THRESHOLD = 0.75
AUTO_CHUNK_BUDGET_FRACTION = 0.50
AUTO_CHUNK_MIN_ROWS = 256 * 1024
AUTO_CHUNK_MIN_STREAM_BYTES = 64 << 20


def chunk_decision(resident_bytes: int, stream_rows: int, stream_bytes_per_row: int, budget_bytes: int,
                   requested_rows: int = 0, force: bool = False) -> tuple[int, str]:
    """Chunk rows for one chunkable query (0 = run directly) and the reason."""
    ws = resident_bytes + stream_rows * stream_bytes_per_row
    fits = budget_bytes > 0 and ws <= int(budget_bytes * AUTO_CHUNK_THRESHOLD)
    if requested_rows:
        if fits and not force:
            return 0, (f"requested chunk of {requested_rows} rows downgraded to a direct run: working set "
                       f"{ws / 1e9:.2f} GB fits in {budget_bytes * AUTO_CHUNK_THRESHOLD / 1e9:.2f} GB")
        return requested_rows, "requested"
    if fits:
        return 0, f"working set {ws / 1e9:.2f} GB fits"
    bpr = max(1, stream_bytes_per_row)
    stream_budget = int(budget_bytes * AUTO_CHUNK_BUDGET_FRACTION) - resident_bytes
    stream_budget = max(stream_budget, budget_bytes // 8, AUTO_CHUNK_MIN_STREAM_BYTES)
    rows = max(stream_budget // (bpr * 2), AUTO_CHUNK_MIN_ROWS)
    rows = min(rows, max(1, stream_rows))
    return int(rows), (f"working set {ws / 1e9:.2f} GB exceeds {budget_bytes * AUTO_CHUNK_THRESHOLD / 1e9:.2f} GB: "
                       f"chunks of {rows} rows ({rows * bpr / 2**20:.0f} MiB a slot)")


def projected_bytes_per_row(plan_source: str, columns: dict) -> int:
    """Bytes a row of one table contributes to the working set: the columns
    the plan's kernels take (p_<column> parameters)."""
    n = 0
    for c, a in columns.items():
        if re.search(r"\bp_" + c + r"\b", plan_source):
            n += a.itemsize * (a.shape[1] if a.ndim == 2 else 1)
    return n


def auto_chunk_rows(db: "TpchDB", query: str, plan: E.HipQueryPlan, requested_rows: int = 0,
                    budget_bytes: int | None = None, force: bool | None = None) -> tuple[int, str]:
    """chunk_decision for a named query on db's tables; the budget defaults
    to the device's free HBM, force to GDB_FORCE_CHUNK=1 in the environment."""
    if not plan.chunkable:
        return 0, "plan not chunkable"
    if force is None:
        force = os.environ.get("GDB_FORCE_CHUNK", "0") not in ("", "0")
    src = plan.source()
    stream = stream_table(query)
    resident = 0
    for t in SCANNED[query]:
        if t != stream:
            cols = db.host[t]
            resident += int(next(iter(cols.values())).shape[0]) * projected_bytes_per_row(src, cols)
    scols = db.host[stream]
    rows = int(next(iter(scols.values())).shape[0])
    if budget_bytes is None:
        budget_bytes = db.ctx.mem_info()[0]
    return chunk_decision(resident, rows, projected_bytes_per_row(src, scols), budget_bytes, requested_rows, force)


def run(ctx: E.Context, db: TpchDB, query: str, warmup: int = 0, measured: int = 1, options: str = "",
        executor: E.HipGenericExecutor | None = None, params: dict | None = None, chunk_rows: int = 0,
        colbin_dir=None) -> TpchRun:
    """Run a named TPC-H query; the result columns follow the SQL select
    list. params overrides constants of the query (e.g. q18_qty).
    chunk_rows > 0: out-of-core form; the stream table stays in host memory
    and its phases run once per chunk of that many rows, copied into HBM
    with double-buffered copies (plans not marked chunkable are refused, as
    codegen_main.cpp:320-335 refuses an explicit --chunk). With colbin_dir,
    the stream table's chunks are read from <colbin_dir>/<table>.colbin
    instead of host memory."""
    if query not in QUERIES:
        raise ValueError(f"unknown TPC-H query {query}; known: {QUERIES}")
    opts = ",".join(o for o in (PLAN_OPTIONS.get(query, ""), options) if o)
    from .queries import get_plan, like_words_buffer

    plan = get_plan(ctx, query, opts)
    streamed = None
    if chunk_rows:
        if not plan.chunkable:
            raise ValueError(f"{query} does not support chunked execution (its plan is not marked chunkable)")
        streamed = stream_table(query)
    ex = executor or E.HipGenericExecutor(ctx)
    for t in SCANNED[query]:
        if t != streamed:
            ex.register_table(db.device(t))
    prm = dict(PARAMS[query], **(params or {}))
    words_ptr = None
    strings_ptr = None
    for k, v in prm.items():
        if k.startswith("q"):
            ex.register_scalar(k, int(v))
    for k, v in db.symbols.items():
        ex.register_symbol(k, v)
    if query == "Q3":
        od = db.host["orders"]["o_orderdate"]
        ex.register_symbol("o_slots", _pow2(2 * int((od < prm["q3_date"]).sum())))
    if query in ("Q9", "Q20"):  # partsupp pairs of the selected parts: at most every row
        ex.register_symbol(f"{query.lower()}_slots", _pow2(2 * int(db.host["partsupp"]["ps_partkey"].shape[0])))
    if query == "Q16":  # groups: at most one per part; (group, supplier) pairs: at most one per partsupp row
        ex.register_symbol("q16_slots", _pow2(2 * int(db.host["part"]["p_partkey"].shape[0])))
        ex.register_symbol("q16_pair_slots", _pow2(2 * int(db.host["partsupp"]["ps_partkey"].shape[0])))
    try:
        if "words" in prm:  # LIKE '%w1%w2%'
            w1, w2 = prm["words"]
            words_ptr = like_words_buffer(ctx, w1, w2)
            ex.register_buffer("d_like_words", words_ptr, 64)
            ex.register_scalar("like_l1", len(w1))
            ex.register_scalar("like_l2", len(w2))
        elif "prefix" in prm:  # LIKE 'p%'
            words_ptr = like_words_buffer(ctx, prm["prefix"], b"")
            ex.register_buffer("d_like_words", words_ptr, 64)
            ex.register_scalar("like_l1", len(prm["prefix"]))
        if "strings" in prm:
            buf = string_slots(prm["strings"])
            strings_ptr = ctx.upload(buf)
            ex.register_buffer("d_str_consts", strings_ptr, buf.nbytes)
        if streamed is None:
            t = ex.execute(plan, warmup, measured)
        elif colbin_dir is not None:
            st = ex.execute_streamed_colbin(plan, streamed, Path(colbin_dir) / f"{streamed}.colbin", chunk_rows)
            t = E.Timing(st.total_ms, 0.0, [], [])
            t.chunks, t.stream_ms, t.bytes_copied = int(st.chunks), float(st.stream_ms), int(st.bytes_copied)
        else:
            # only the columns the plan's kernels take cross the link
            src = plan.source()
            cols = {c: a for c, a in db.host[streamed].items() if re.search(r"\bp_" + c + r"\b", src)}
            host = E.HostTable(streamed, cols)
            try:
                st = ex.execute_streamed(plan, host, chunk_rows)
            finally:
                host.close()
            t = E.Timing(st.total_ms, 0.0, [], [])
            t.chunks, t.stream_ms, t.bytes_copied = int(st.chunks), float(st.stream_ms), int(st.bytes_copied)
        res = ex.collect(plan)
    finally:
        if words_ptr is not None:
            ctx.free(words_ptr)
        if strings_ptr is not None:
            ctx.free(strings_ptr)
    cols = {k: [int(x) if not isinstance(x, str) else x for x in v] for k, v in res.columns.items()}
    fmts = dict(res.formats)
    _finish(query, db, cols, fmts, prm)
    out = {name: cols[name] for name in SELECT[query]}
    run_ = TpchRun(query, out, {k: fmts.get(k, (0, False)) for k in out}, t, plan, ex, res)
    run_.canonical = canonical(run_.columns, run_.formats)
    return run_


def string_slots(strings) -> np.ndarray:
    """32-byte slots: a length byte, then the string's bytes (at most 31)."""
    buf = np.zeros(32 * max(1, len(strings)), np.uint8)
    for k, w in enumerate(strings):
        if len(w) > 31:
            raise ValueError(f"string constant {w!r} is longer than 31 bytes")
        buf[32 * k] = len(w)
        buf[32 * k + 1:32 * k + 1 + len(w)] = np.frombuffer(w, np.uint8)
    return buf


def _text(arr_row) -> str:
    return bytes(arr_row).split(b"\0", 1)[0].decode()


# Queries whose result the host finishes (attached columns, reordering, a
# final division); the others are the engine's result as collected.
HOST_FINISH = {"Q2", "Q3", "Q5", "Q7", "Q8", "Q9", "Q10", "Q15", "Q16", "Q17", "Q18", "Q20", "Q21"}


def _finish(query: str, db: TpchDB, cols: dict, fmts: dict, prm: dict) -> None:
    """Host post-processing on the few result rows (the reference's CPU
    post-processing in runCodegenQuery): attach o_shippriority (Q3) and c_name
    (Q18) by key."""
    if query == "Q3":
        o = db.host["orders"]
        sub = np.flatnonzero(np.isin(o["o_orderkey"], cols["l_orderkey"]))
        sp = {int(o["o_orderkey"][i]): int(o["o_shippriority"][i]) for i in sub}
        cols["o_shippriority"] = [sp[k] for k in cols["l_orderkey"]]
        fmts["o_shippriority"] = (0, False)
    elif query == "Q5":
        n = db.host["nation"]
        names = {int(k): _text(nm) for k, nm in zip(n["n_nationkey"], n["n_name"])}
        rows = sorted(zip(cols["bucket"], cols["revenue"]), key=lambda r: (-r[1], names[r[0]]))
        cols["n_name"] = [names[r[0]] for r in rows]
        cols["revenue"] = [r[1] for r in rows]
    elif query == "Q10":
        c, n = db.host["customer"], db.host["nation"]
        nat = {int(k): _text(nm) for k, nm in zip(n["n_nationkey"], n["n_name"])}
        sub = np.flatnonzero(np.isin(c["c_custkey"], cols["c_custkey"]))
        pos = {int(c["c_custkey"][i]): i for i in sub}
        idx = [pos[k] for k in cols["c_custkey"]]
        cols["c_name"] = [_text(c["c_name"][i]) for i in idx]
        cols["c_acctbal"] = [int(c["c_acctbal"][i]) for i in idx]
        fmts["c_acctbal"] = (2, False)
        cols["n_name"] = [nat[int(c["c_nationkey"][i])] for i in idx]
        cols["c_address"] = [_text(c["c_address"][i]) for i in idx]
        cols["c_phone"] = [_text(c["c_phone"][i]) for i in idx]
        cols["c_comment"] = [_text(c["c_comment"][i]) for i in idx]
    elif query == "Q2":
        sp, n, pt = db.host["supplier"], db.host["nation"], db.host["part"]
        nat = {int(k): _text(nm) for k, nm in zip(n["n_nationkey"], n["n_name"])}
        spos = {int(k): i for i, k in enumerate(sp["s_suppkey"].tolist())}
        ppos = {int(k): i for i, k in enumerate(pt["p_partkey"].tolist())}
        rows = []
        for sk, pk in zip(cols["s_suppkey"], cols["p_partkey"]):
            i, j = spos[sk], ppos[pk]
            rows.append((int(sp["s_acctbal"][i]), _text(sp["s_name"][i]), nat[int(sp["s_nationkey"][i])], pk,
                         _text(pt["p_mfgr"][j]), _text(sp["s_address"][i]), _text(sp["s_phone"][i]),
                         _text(sp["s_comment"][i])))
        rows.sort(key=lambda r: (-r[0], r[2], r[1], r[3]))
        rows = rows[:100]
        for c, name in enumerate(SELECT["Q2"]):
            cols[name] = [r[c] for r in rows]
        fmts["s_acctbal"] = (2, False)
    elif query == "Q9":
        n = db.host["nation"]
        names = {int(k): _text(nm) for k, nm in zip(n["n_nationkey"], n["n_name"])}
        rows = sorted(((names[b // 8], prm["q9_year0"] + b % 8, v) for b, v in zip(cols["bucket"], cols["sum_profit"])),
                      key=lambda r: (r[0], -r[1]))
        for c, name in enumerate(SELECT["Q9"]):
            cols[name] = [r[c] for r in rows]
        fmts["o_year"] = (0, False)
    elif query == "Q20":
        sp = db.host["supplier"]
        pos = {int(k): i for i, k in enumerate(sp["s_suppkey"].tolist())}
        rows = sorted((_text(sp["s_name"][pos[k]]), _text(sp["s_address"][pos[k]])) for k in cols["s_suppkey"])
        cols["s_name"] = [r[0] for r in rows]
        cols["s_address"] = [r[1] for r in rows]
    elif query == "Q21":
        sp = db.host["supplier"]
        names = {int(k): _text(nm) for k, nm in zip(sp["s_suppkey"], sp["s_name"])}
        rows = sorted(((names[k], n) for k, n in zip(cols["s_suppkey"], cols["numwait"])), key=lambda r: (-r[1], r[0]))
        rows = rows[:100]
        cols["s_name"] = [r[0] for r in rows]
        cols["numwait"] = [r[1] for r in rows]
    elif query == "Q7":
        names = [w.decode() for w in prm["strings"]]
        y0 = prm["q7_year0"]
        rows = sorted((names[b // 4], names[(b // 2) % 2], y0 + b % 2, r) for b, r in zip(cols["bucket"], cols["revenue"]))
        for c, name in enumerate(SELECT["Q7"]):
            cols[name] = [r[c] for r in rows]
        fmts["l_year"] = (0, False)
    elif query == "Q8":
        cols["o_year"] = [prm["q8_year0"] + b for b in cols["bucket"]]
        fmts["o_year"] = (0, False)
    elif query == "Q15":
        sp = db.host["supplier"]
        sub = np.flatnonzero(np.isin(sp["s_suppkey"], cols["s_suppkey"]))
        pos = {int(sp["s_suppkey"][i]): i for i in sub}
        idx = [pos[k] for k in cols["s_suppkey"]]
        for c in ("s_name", "s_address", "s_phone"):
            cols[c] = [_text(sp[c][i]) for i in idx]
    elif query == "Q16":
        pt = db.host["part"]
        sub = np.flatnonzero(np.isin(pt["p_partkey"], cols["p_partkey"]))
        pos = {int(pt["p_partkey"][i]): i for i in sub}
        rows = sorted(((_text(pt["p_brand"][pos[k]]), _text(pt["p_type"][pos[k]]), int(pt["p_size"][pos[k]]), n)
                       for k, n in zip(cols["p_partkey"], cols["supplier_cnt"])),
                      key=lambda r: (-r[3], r[0], r[1], r[2]))
        for c, name in enumerate(SELECT["Q16"]):
            cols[name] = [r[c] for r in rows]
        fmts["p_size"] = (0, False)
    elif query == "Q17":
        # SUM(l_extendedprice) / 7.0, exact to 6 more digits: trunc(sum * 10^6 / 7)
        cols["avg_yearly"] = [v * 1_000_000 // 7 for v in cols["sum_price"]]
        fmts["avg_yearly"] = (fmts["sum_price"][0] + 6, False)
    elif query == "Q18":
        c = db.host["customer"]
        sub = np.flatnonzero(np.isin(c["c_custkey"], cols["c_custkey"]))
        names = {int(c["c_custkey"][i]): bytes(c["c_name"][i]).rstrip(b"\0").decode() for i in sub}
        cols["c_name"] = [names[k] for k in cols["c_custkey"]]


# ---------------------------------------------------------------------------
# Canonical CSV and the golden check
# ---------------------------------------------------------------------------

def fixed_str(v: int, scale: int) -> str:
    if scale <= 0:
        return str(v)
    s = str(abs(v)).rjust(scale + 1, "0")
    return ("-" if v < 0 else "") + s[:-scale] + "." + s[-scale:]


def _field(v, scale: int, is_date: bool) -> str:
    if v is None:  # SQL NULL (an aggregate over no rows): an empty field
        return ""
    if isinstance(v, str):
        if any(ch in v for ch in ',"\n\r'):
            return '"' + v.replace('"', '""') + '"'
        return v
    if is_date:
        return f"{v // 10000:04d}-{v // 100 % 100:02d}-{v % 100:02d}"
    return fixed_str(int(v), scale)


def canonical(columns: dict, formats: dict) -> str:
    """Header line, then one line per row: exact decimals for scaled integers,
    dates as YYYY-MM-DD, strings quoted when they hold ',', '"' or a line
    break (the rules of GenericResult::toCanonical in the engine)."""
    names = list(columns)
    lines = [",".join(names)]
    n = len(columns[names[0]]) if names else 0
    for r in range(n):
        lines.append(",".join(_field(columns[c][r], *formats.get(c, (0, False))) for c in names))
    return "\n".join(lines) + "\n"


def compare_canonical(got: str, expected: str, abs_tol: float = CHECK_ABS_TOL, rel_tol: float = CHECK_REL_TOL) -> str:
    """'' on a match, else the first difference (compareCanonical,
    codegen/codegen_main.cpp:63-186)."""
    return N.two_call_string(_eng().gdb_compare_canonical, got.encode(), expected.encode(), C.c_double(abs_tol),
                             C.c_double(rel_tol))


def golden_name(query: str, scale_label: str) -> str:
    return f"{query}_{scale_label}.csv"


def save_golden(canonical_csv: str, golden_dir, query: str, scale_label: str) -> Path:
    d = Path(golden_dir)
    d.mkdir(parents=True, exist_ok=True)
    p = d / golden_name(query, scale_label)
    p.write_text(canonical_csv)
    return p


def check(canonical_csv: str, golden_dir, query: str, scale_label: str, abs_tol: float = CHECK_ABS_TOL,
          rel_tol: float = CHECK_REL_TOL) -> tuple:
    """(status, message): status 0 OK, 1 mismatch, 2 golden file missing (the
    reference's --check exit codes)."""
    p = Path(golden_dir) / golden_name(query, scale_label)
    if not p.exists():
        return 2, f"golden file missing: {p}"
    diff = compare_canonical(canonical_csv, p.read_text(), abs_tol, rel_tol)
    return (0, "OK") if not diff else (1, diff)


def data_dir_from_env() -> str | None:
    return os.environ.get("GDB_TPCH_DIR")
