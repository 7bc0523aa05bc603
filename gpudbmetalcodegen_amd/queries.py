"""Query drivers: plan + runner + sizing for each workload, over the C-ABI.

Each driver mirrors one step of the reference's ``runCodegenQuery``
(codegen/codegen_main.cpp:271-1776): build the plan, compile it, register
tables / scalars / size symbols, execute, collect. Sizing follows the
reference's exact pre-pass idea (codegen/execution/max_key_symbols.cpp:188-235)
generalised to hash tables.
"""

from __future__ import annotations

import threading
from dataclasses import dataclass

import numpy as np

from . import engine as E
from . import workloads as W

Q1_LABELS = [("A", "F"), ("A", "O"), ("N", "F"), ("N", "O"), ("R", "F"), ("R", "O")]


def next_pow2(x: int) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


@dataclass
class QueryRun:
    result: E.GenericResult
    timing: E.Timing
    plan: E.HipQueryPlan
    executor: E.HipGenericExecutor
    stream_stats: object = None


@dataclass
class Stream:
    """Out-of-core form: this host table is streamed into HBM in chunks of
    chunk_rows rows instead of being read from a resident copy."""
    table: E.HostTable
    chunk_rows: int


def _register(ex: E.HipGenericExecutor, dev: dict, names, stream: Stream | None) -> None:
    for t in names:
        if stream is not None and t == stream.table.name:
            continue
        ex.register_table(dev[t])


def _run(ex: E.HipGenericExecutor, plan: E.HipQueryPlan, warmup: int, measured: int, stream: Stream | None):
    if stream is None:
        return ex.execute(plan, warmup, measured), None
    return None, ex.execute_streamed(plan, stream.table, stream.chunk_rows)


_plan_cache: dict = {}
_plan_lock = threading.Lock()


def get_plan(ctx: E.Context, name: str, options: str = "") -> E.HipQueryPlan:
    """Built and compiled plans, one per (context, name, options); shared by
    every runner and thread (plans are read-only once compiled)."""
    key = (id(ctx), name, options)
    with _plan_lock:
        p = _plan_cache.get(key)
        if p is None:
            p = E.build_hip_plan(name, options).compile(ctx)
            _plan_cache[key] = p
    return p


def upload_tables(ctx: E.Context, tables: dict) -> dict:
    return {name: E.DeviceTable.from_host(ctx, name, cols) for name, cols in tables.items() if isinstance(cols, dict)}


def max_key(ctx: E.Context, dev: dict, table: str, column: str) -> int:
    """Largest value of an int32 key column, from the MAXKEY plan's scan on the
    device (cached on the table): the key-range symbol of a bitmap or
    key-indexed array is this + 1, as the reference registers max + 1 from an
    exact scan (codegen/execution/max_key_symbols.cpp:188-235). 0 for an
    empty table."""
    t = dev.get(table)
    if t is None:
        raise ValueError(f"max_key: table {table} is not resident; pass the key bound explicitly")
    if column in t.key_max:
        return t.key_max[column]
    mk = 0
    if t.rows > 0:
        plan = get_plan(ctx, "MAXKEY", f"table={table},column={column}")
        ex = E.HipGenericExecutor(ctx)
        try:
            ex.register_table(t)
            ex.execute(plan)
            res = ex.collect(plan)
        finally:
            ex.close()
        if int(res.columns["rows"][0]) != t.rows:
            raise RuntimeError(f"max_key: scanned {int(res.columns['rows'][0])} rows of {table}, expected {t.rows}")
        mk = max(0, int(res.columns["max_key"][0]))
    t.key_max[column] = mk
    return mk


def _key_bits(ctx: E.Context, dev: dict, table: str, column: str, given: int | None) -> int:
    return (given if given is not None else max_key(ctx, dev, table, column)) + 1


# ---------------------------------------------------------------------------
# Workload 0: Q1
# ---------------------------------------------------------------------------
def run_q1(ctx: E.Context, dev: dict, cutoff: int, warmup: int = 0, measured: int = 1,
           options: str = "", executor: E.HipGenericExecutor | None = None, stream: Stream | None = None) -> QueryRun:
    plan = get_plan(ctx, "Q1", options)
    ex = executor or E.HipGenericExecutor(ctx)
    _register(ex, dev, ["F"], stream)
    ex.register_scalar("q1_cutoff", cutoff)
    t, st = _run(ex, plan, warmup, measured, stream)
    return QueryRun(ex.collect(plan), t, plan, ex, st)


def _executor(ctx, executor):
    return executor or E.HipGenericExecutor(ctx)


# ---------------------------------------------------------------------------
# Workload 1: Q3 analog (semi-join bitmap -> hash build -> probe -> top-10)
# ---------------------------------------------------------------------------
@dataclass
class ColumnStats:
    """Load-time statistics of an integer column (computed once, like the
    reference's cached max-key scan, codegen/execution/max_key_symbols.cpp:28-41):
    exact value frequencies when the column has at most 64 distinct values,
    otherwise an equi-width histogram of `bins` buckets over [lo, hi]."""
    lo: int
    hi: int
    n: int
    counts: np.ndarray
    values: np.ndarray | None = None   # categorical: the distinct values

    def frac_eq(self, v: int) -> float:
        if self.n == 0:
            return 0.0
        if self.values is not None:
            hit = np.nonzero(self.values == v)[0]
            return float(self.counts[hit[0]]) / self.n if hit.size else 0.0
        return self.frac_lt(v + 1) - self.frac_lt(v)

    def frac_lt(self, v: int) -> float:
        """Estimated fraction of rows < v (uniform within a bucket)."""
        if self.n == 0 or v <= self.lo:
            return 0.0
        if v > self.hi:
            return 1.0
        if self.values is not None:
            return float(self.counts[self.values < v].sum()) / self.n
        nb = self.counts.size
        width = (self.hi - self.lo + 1) / nb
        pos = (v - self.lo) / width
        k = int(pos)
        part = float(self.counts[:k].sum()) + (pos - k) * float(self.counts[k] if k < nb else 0)
        return part / self.n


def column_stats(col: np.ndarray, bins: int = 256) -> ColumnStats:
    col = np.asarray(col)
    n = int(col.shape[0])
    if n == 0:
        return ColumnStats(0, 0, 0, np.zeros(1, np.int64))
    lo, hi = int(col.min()), int(col.max())
    if hi - lo < 64:
        vals, cnt = np.unique(col, return_counts=True)
        return ColumnStats(lo, hi, n, cnt.astype(np.int64), vals.astype(np.int64))
    cnt, _ = np.histogram(col, bins=bins, range=(lo, hi + 1))
    return ColumnStats(lo, hi, n, cnt.astype(np.int64))


def q3_table_stats(T: dict) -> dict:
    """The statistics the join's table sizing uses, gathered at load time."""
    return {"b_date": column_stats(T["B"]["b_date"]), "a_seg": column_stats(T["A"]["a_seg"]),
            "n_b": int(T["B"]["b_key"].shape[0])}


def q3_slots_from_stats(stats: dict, seg: int, d: int, margin: float = 2.0) -> int:
    """B hash-table slots from statistics: estimated qualifying B rows (date
    selectivity x segment frequency of the referenced A rows, independence
    assumed) x margin, as a power of two."""
    est = stats["n_b"] * stats["b_date"].frac_lt(d) * stats["a_seg"].frac_eq(seg)
    return next_pow2(max(64, int(margin * est) + 1))


def run_q3(ctx: E.Context, dev: dict, seg: int = W.W1_SEG, d: int = W.W1_DATE, options: str = "",
           max_a: int | None = None, max_b: int | None = None, b_slots: int | None = None, warmup: int = 0,
           measured: int = 1, executor: E.HipGenericExecutor | None = None, stream: Stream | None = None,
           max_resizes: int = 8) -> QueryRun:
    """Plain form: the plan's own exact-count pre-pass sizes the B hash table
    (options ``sizing=given`` takes ``b_slots`` instead, e.g. from
    q3_slots_from_stats; a reported overflow doubles it and reruns, so results
    stay exact). Key-range symbols come from table statistics (max key), as
    the reference's cached max-key mode does
    (codegen/execution/max_key_symbols.cpp:28-41): given, or else the
    device max-key scan of max_key()."""
    plan = get_plan(ctx, "Q3", options)
    ex = _executor(ctx, executor)
    _register(ex, dev, ("A", "B", "F"), stream)
    ex.register_scalar("q3_seg", seg)
    ex.register_scalar("q3_date", d)
    ex.register_symbol("a_bits", _key_bits(ctx, dev, "A", "a_key", max_a))
    ex.register_symbol("b_bits", _key_bits(ctx, dev, "B", "b_key", max_b))
    if "sizing=given" not in options:
        t, st = _run(ex, plan, warmup, measured, stream)
        return QueryRun(ex.collect(plan), t, plan, ex, st)
    if b_slots is None:
        raise ValueError("sizing=given needs b_slots")
    resizes = 0
    while True:
        ex.register_symbol("b_slots", b_slots)
        try:
            t, st = _run(ex, plan, warmup, measured, stream)
            break
        except E.GdbOverflow:
            resizes += 1
            if resizes > max_resizes:
                raise
            b_slots *= 2                      # the estimate was too small: double and rerun
    run = QueryRun(ex.collect(plan), t, plan, ex, st)
    run.resizes = resizes
    return run


def q3_hash_groups(run: QueryRun) -> dict:
    """Every group held in the B hash table after a Q3 run (key asc):
    key, rev, date — used to check the join beyond the top-10."""
    slots = run.executor.get_symbol("b_slots")
    tab = run.executor.read_buffer("d_q3_b_ht", slots * 16, np.uint64).reshape(slots, 2)
    w0 = tab[:, 0]
    keys = (w0 & np.uint64(0xFFFFFFFF)).astype(np.int64)
    occ = (keys != 0xFFFFFFFF) & (tab[:, 1].view(np.int64) > 0)
    k = keys[occ]
    order = np.argsort(k)
    return {"key": k[order], "rev": tab[occ, 1].view(np.int64)[order],
            "date": (w0[occ] >> np.uint64(32)).astype(np.int64)[order]}


# ---------------------------------------------------------------------------
# Workload 2: hash group-by with sizing by pre-pass, by estimate, or resize
# ---------------------------------------------------------------------------
@dataclass
class GroupByRun(QueryRun):
    slots: int = 0
    resizes: int = 0
    distinct: int = 0
    times_ms: list = None
    narrow: bool = False      # the partitioned form's cold rows as 8-byte records


def groupby_options(variant: str = "lds", lds_slots: int = 2048, filtered: bool = False, prepass: bool = False,
                    block: int | None = None, window: int | None = None, narrow: bool = False) -> str:
    o = [f"variant={variant}", f"lds_slots={lds_slots}"]
    if narrow and variant == "part":
        o.append("rec8=1")
    if window:
        o.append(f"window={window}")
    if filtered:
        o.append("filter=1")
    if prepass:
        o.append("sizing=prepass")
    if block:
        o.append(f"block={block}")
    return ",".join(o)


def part_max_parts(lds_slots: int) -> int:
    """Largest partition count of the partitioned group-by whose per-partition
    LDS counters fit beside a pass-1 LDS table of lds_slots slots (mirrors
    PartitionConfig::maxParts)."""
    room = (150 * 1024 - 32 * lds_slots - 256) // 4
    p = 1
    while 2 * p <= room and p < 16384:
        p *= 2
    return p


def part_lds_for(slots: int, lds_slots: int) -> int:
    """The largest pass-1 LDS table <= lds_slots that fits slots / 4096 partitions."""
    parts = max(1, slots // 4096)
    if parts > 16384:
        raise ValueError(f"partitioned group-by supports at most 16384 * 4096 slots, got {slots}")
    while lds_slots > 64 and part_max_parts(lds_slots) < parts:
        lds_slots //= 2
    return lds_slots


def run_groupby(ctx: E.Context, dev: dict, *, slots: int | None = None, prepass: bool = False,
                variant: str = "lds", lds_slots: int = 2048, val_cut: int | None = None, key_min: int = 1,
                key_max: int | None = None, max_resizes: int = 8, block: int | None = None,
                warmup: int = 0, measured: int = 1, executor: E.HipGenericExecutor | None = None,
                collect: bool = True, stream: Stream | None = None, window: int | None = None,
                narrow: bool = True) -> GroupByRun:
    """prepass=True: the plan counts distinct keys exactly first (key-range
    bitmap + popcount) and sizes the table at 2x (plain form). Otherwise
    `slots` (e.g. from the learned estimator) sizes it, and a reported
    overflow doubles the table and reruns (resize-on-overflow).
    variant="part" is the partitioned form without global atomics: slots >=
    4096, and its pass-1 LDS table is shrunk when slots / 4096 partitions
    would not leave room for the per-partition LDS counters. With `narrow`
    its cold rows travel as 8-byte records while every value fits in 32 bits;
    a value that does not is reported and the run repeated with 16-byte
    records."""
    part = variant == "part"
    narrow = narrow and part
    ex = _executor(ctx, executor)
    _register(ex, dev, ["G"], stream)
    if val_cut is not None:
        ex.register_scalar("gb_val_cut", val_cut)
    if part:
        ex.register_symbol("gbp_rows", stream.table.rows if stream is not None else dev["G"].rows)
    times = []
    resizes = 0
    sst = None
    if prepass:
        if key_max is None:
            raise ValueError("prepass sizing needs the key range (key_min, key_max)")
        ex.register_scalar("gb_key_min", key_min)
        ex.register_symbol("gb_key_bits", key_max - key_min + 1)
    else:
        if slots is None:
            raise ValueError("give slots or prepass=True")
        slots = max(4096 if part else 64, next_pow2(int(slots)))
        if part:
            lds_slots = part_lds_for(slots, lds_slots)
    while True:
        plan = get_plan(ctx, "GROUPBY", groupby_options(variant, lds_slots, val_cut is not None, prepass, block, window,
                                                        narrow))
        if not prepass:
            ex.register_symbol("gb_slots", slots)
        try:
            t, sst = _run(ex, plan, warmup, measured, stream)   # a streamed pre-pass plan is refused by the engine
            times.append(t.total_ms if t else sst.total_ms)
            break
        except E.GdbOverflow:
            status = ex.read_buffer("d_gb_ht_status", 32, np.uint64)
            times.append(float("nan"))
            resizes += 1
            if resizes > max_resizes:
                raise
            if narrow and ex.read_buffer("d_gbp_rec_status", 32, np.uint64)[0]:
                narrow = False                  # a value needs 64 bits: 16-byte records
            elif part and status[1]:
                if lds_slots <= 64:
                    raise
                lds_slots //= 2                 # more partitions than LDS counters beside the table
            elif prepass:
                raise                           # exact sizing cannot overflow the table
            else:
                slots = next_pow2(max(2 * slots, 2 * int(status[2])))
                if part:
                    lds_slots = part_lds_for(slots, lds_slots)
    if prepass:
        slots = ex.get_symbol("gb_slots")
    res = ex.collect(plan) if collect else None
    return GroupByRun(res, t, plan, ex, sst, slots=slots, resizes=resizes, distinct=res.rows() if res else 0,
                      times_ms=times, narrow=narrow)


# ---------------------------------------------------------------------------
# Workload 4 plan types: LIKE + per-key count + histogram (Q13), anti-join (Q22)
# ---------------------------------------------------------------------------
def like_words_buffer(ctx: E.Context, w1: bytes, w2: bytes) -> int:
    if len(w1) > 31 or len(w2) > 31:
        raise ValueError("LIKE words are limited to 31 bytes")
    buf = np.zeros(64, np.uint8)
    buf[:len(w1)] = np.frombuffer(w1, np.uint8)
    buf[32:32 + len(w2)] = np.frombuffer(w2, np.uint8)
    return ctx.upload(buf)


def run_q13(ctx: E.Context, dev: dict, w1: bytes, w2: bytes, negate: bool = False, max_a: int | None = None,
            words_ptr: int | None = None, warmup: int = 0, measured: int = 1,
            executor: E.HipGenericExecutor | None = None, stream: Stream | None = None) -> QueryRun:
    plan = get_plan(ctx, "Q13L", "negate=1" if negate else "")
    ex = _executor(ctx, executor)
    _register(ex, dev, ("S", "A"), stream)
    own = words_ptr is None
    wp = like_words_buffer(ctx, w1, w2) if own else words_ptr
    try:
        ex.register_buffer("d_like_words", wp, 64)
        ex.register_scalar("like_l1", len(w1))
        ex.register_scalar("like_l2", len(w2))
        ex.register_symbol("a_bits", _key_bits(ctx, dev, "A", "a_key", max_a))
        t, st = _run(ex, plan, warmup, measured, stream)
        res = ex.collect(plan)
    finally:
        if own:
            ctx.free(wp)
    return QueryRun(res, t, plan, ex, st)


def run_q22(ctx: E.Context, dev: dict, date_lo: int, date_hi: int, max_a: int | None = None, warmup: int = 0,
            measured: int = 1, executor: E.HipGenericExecutor | None = None, stream: Stream | None = None) -> QueryRun:
    plan = get_plan(ctx, "Q22A")
    ex = _executor(ctx, executor)
    _register(ex, dev, ("B", "A"), stream)
    ex.register_scalar("q22_date_lo", date_lo)
    ex.register_scalar("q22_date_hi", date_hi)
    ex.register_symbol("a_bits", _key_bits(ctx, dev, "A", "a_key", max_a))
    t, st = _run(ex, plan, warmup, measured, stream)
    return QueryRun(ex.collect(plan), t, plan, ex, st)


def result_as_lists(res: E.GenericResult) -> dict:
    return {k: [x if isinstance(x, str) else int(x) for x in v] for k, v in res.columns.items()}
