"""Workload 4: a served query stream over HBM-resident tables.

Four plan types arrive as a Poisson stream (25% each): scan-aggregate (Q1 on
F), three-way join (Q3 over A/B/F), bitmap anti-join (Q22A over B/A) and
byte-exact LIKE '%w1%w2%' + per-key count + histogram (Q13L over S/A).
Parameters are drawn from a per-kind menu (prepared statements with a fixed
set of bindings), so each distinct query can be checked once against an
independent evaluation and every served answer compared with it.

Plain form: one query at a time in arrival order.
Served form (continuous batching): a dispatcher admits the oldest waiting
query as soon as a worker is idle and its HBM footprint fits the scratch
budget next to the queries already running, with at most one heavy
(bandwidth-bound) query on the device at a time, light ones overlapping it; each worker is a persistent
thread that owns a stream and its own runners (disjoint buffers), so
admitted queries overlap on the device and a query that arrives while
others run starts the moment a slot and memory free up. The dispatcher
sleeps on a condition variable until the next arrival or the next
completion; nothing polls.

Latency = completion (result on the host) - arrival; reported as p50/p99
(nearest rank) with throughput.
"""

from __future__ import annotations

import ctypes as C
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import engine as E
from . import queries as Q
from . import stats as ST
from . import workloads as W

KINDS = ("scan", "join", "anti", "like")


@dataclass
class Query:
    qid: int
    kind: str
    params: dict
    arrival: float = 0.0          # seconds from stream start
    start: float = 0.0
    done: float = 0.0
    result: str = ""              # canonical CSV of the answer
    columns: dict = field(default_factory=dict)   # the answer's columns (ints / strings)
    groups: dict | None = None    # join only, with Worker(verify=True): every group of the hash table
    rows: int = 0
    footprint: int = 0
    error: str = ""

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3

    @property
    def key(self) -> tuple:
        """Identity of the query's answer: kind and parameters."""
        return (self.kind,) + tuple(sorted(self.params.items()))


def percentile(xs, p):
    """Nearest-rank percentile as the reference's trial statistics take it
    (codegen_main.cpp:1012-1020; stats.percentile)."""
    return ST.percentile(list(xs), p)


class ResidentDB:
    """Workload-4 tables uploaded once; key statistics for sizing."""

    def __init__(self, ctx: E.Context, tables: dict):
        self.ctx = ctx
        self.vocab = tables.get("vocab", [])
        self.dev = Q.upload_tables(ctx, {k: v for k, v in tables.items() if k in ("F", "A", "B", "S")})
        self.n = {k: t.rows for k, t in self.dev.items()}
        # key-range symbols from an exact max-key scan on the device (the
        # reference's registerMaxKeySymbols, max_key_symbols.cpp:188-235)
        self.max_a = Q.max_key(ctx, self.dev, "A", "a_key")
        self.max_b = Q.max_key(ctx, self.dev, "B", "b_key")
        self.resident_bytes = sum(t.nbytes for t in self.dev.values())

    def rows_for(self, kind: str) -> int:
        return {"scan": self.n["F"], "join": self.n["A"] + self.n["B"] + self.n["F"],
                "anti": self.n["A"] + self.n["B"], "like": self.n["S"] + self.n["A"]}[kind]

    def footprint(self, q: Query) -> int:
        """Upper estimate of the query's scratch HBM (bytes) from catalog statistics."""
        if q.kind == "scan":
            return 4096
        if q.kind == "anti":
            return self.max_a // 8 + 4096
        if q.kind == "like":
            return 4 * (self.max_a + 1) + 8192
        # join: A/B bitmaps + a hash table for the qualifying B rows, with
        # selectivity from the uniform date/segment statistics, 2x safety
        frac = (q.params["d"] / float(W.DATE_DAYS_W1)) * 0.2
        slots = Q.next_pow2(int(2 * 2 * frac * self.n["B"]) + 64)
        return self.max_a // 8 + self.max_b // 8 + 16 * slots

    def close(self):
        for t in self.dev.values():
            t.free(self.ctx)


def param_menu(db: ResidentDB, seed: int = 4, per_kind: int = 8) -> dict:
    """A fixed menu of `per_kind` parameter bindings for each plan type."""
    r = np.random.default_rng(seed)
    menu = {k: [] for k in KINDS}
    for _ in range(per_kind):
        menu["scan"].append({"cutoff": int(r.integers(2000, W.DATE_DAYS_W0))})
        menu["join"].append({"seg": int(r.integers(0, 5)), "d": int(r.integers(600, 1800))})
        lo = int(r.integers(0, 2000))
        menu["anti"].append({"lo": lo, "hi": lo + int(r.integers(50, 400))})
        a = db.vocab[int(r.integers(0, len(db.vocab)))]
        b = db.vocab[int(r.integers(0, len(db.vocab)))]
        if r.random() < 0.5:
            a, b = a[:2], b[:2]
        menu["like"].append({"w1": a, "w2": b})
    return menu


def make_stream(db: ResidentDB, rate: float, duration: float, seed: int, backlog: bool = False,
                menu: dict | None = None) -> list[Query]:
    """Poisson arrivals at `rate` q/s over `duration` s (backlog: all at t=0);
    plan types uniform, parameters uniform over the menu."""
    menu = menu or param_menu(db)
    r = np.random.default_rng(seed)
    qs, t, i = [], 0.0, 0
    while True:
        t += r.exponential(1.0 / rate)
        if t > duration:
            break
        kind = KINDS[int(r.integers(0, 4))]
        p = dict(menu[kind][int(r.integers(0, len(menu[kind])))])
        q = Query(i, kind, p, 0.0 if backlog else t, rows=db.rows_for(kind))
        q.footprint = db.footprint(q)
        qs.append(q)
        i += 1
    return qs


def _plain_columns(res: E.GenericResult) -> dict:
    return {k: [x if isinstance(x, str) else int(x) for x in v] for k, v in res.columns.items()}


class Worker:
    """One device stream with its own runners (one per plan type).
    verify=True also reads every group of a join's hash table after the run
    (a check of the whole join, not only its top-10; costs a table read)."""

    def __init__(self, db: ResidentDB, verify: bool = False):
        hip, _ = N.load()
        self.db = db
        self.ctx = db.ctx
        self.verify = verify
        s = C.c_void_p()
        N.hip_check(hip.gdb_stream_create(self.ctx.handle, C.byref(s)))
        self.stream = s
        self.ex = {k: E.HipGenericExecutor(self.ctx, s) for k in KINDS}
        self.words = self.ctx.alloc(64)
        self.hwords = np.zeros(64, np.uint8)

    def run(self, q: Query) -> None:
        db, ex = self.db, self.ex[q.kind]
        p = q.params
        if q.kind == "scan":
            run = Q.run_q1(self.ctx, db.dev, p["cutoff"], executor=ex)
        elif q.kind == "join":
            run = Q.run_q3(self.ctx, db.dev, p["seg"], p["d"], max_a=db.max_a, max_b=db.max_b, executor=ex)
            if self.verify:
                q.groups = Q.q3_hash_groups(run)
        elif q.kind == "anti":
            run = Q.run_q22(self.ctx, db.dev, p["lo"], p["hi"], max_a=db.max_a, executor=ex)
        else:
            hip, _ = N.load()
            self.hwords[:] = 0
            self.hwords[:len(p["w1"])] = np.frombuffer(p["w1"], np.uint8)
            self.hwords[32:32 + len(p["w2"])] = np.frombuffer(p["w2"], np.uint8)
            N.hip_check(hip.gdb_memcpy_h2d_async(C.c_void_p(self.words), self.hwords.ctypes.data_as(C.c_void_p), 64,
                                                 self.stream))
            run = Q.run_q13(self.ctx, db.dev, p["w1"], p["w2"], max_a=db.max_a, words_ptr=self.words, executor=ex)
        q.result = run.result.canonical
        q.columns = _plain_columns(run.result)

    def warm(self, menu: dict) -> None:
        """Run every query of the menu once on this worker, so that its plans
        are loaded and its buffers have their largest sizes before it serves:
        a worker that first meets a query while serving would allocate device
        memory inside that query's latency."""
        for kind, bindings in menu.items():
            for i, params in enumerate(bindings):
                self.run(Query(-1 - i, kind, dict(params)))
        hip, _ = N.load()
        N.hip_check(hip.gdb_stream_sync(self.stream))

    def close(self):
        hip, _ = N.load()
        for e in self.ex.values():
            e.close()
        self.ctx.free(self.words)
        hip.gdb_stream_destroy(self.stream)


class _Runner(threading.Thread):
    """Persistent thread of one worker: runs the queries the dispatcher hands it."""

    def __init__(self, worker: Worker, t0: float, on_done):
        super().__init__(daemon=True)
        self.worker, self.t0, self.on_done = worker, t0, on_done
        self.jobs: deque = deque()
        self.cv = threading.Condition()

    def submit(self, q: Query | None) -> None:
        with self.cv:
            self.jobs.append(q)
            self.cv.notify()

    def run(self) -> None:
        while True:
            with self.cv:
                while not self.jobs:
                    self.cv.wait()
                q = self.jobs.popleft()
            if q is None:
                return
            try:
                self.worker.run(q)
            except Exception as e:  # surfaced in the report
                q.error = str(e)[:300]
            q.done = time.perf_counter() - self.t0
            self.on_done(self, q)


def serve(db: ResidentDB, stream: list[Query], workers: int = 4, budget_bytes: int = 2 << 30,
          pool: list[Worker] | None = None, max_heavy: int = 1, heavy_rows: int | None = None,
          min_overlap_ms: float = 0.0, service_ms: dict | None = None) -> dict:
    """Run a query stream. workers=1 with an unlimited budget is the plain
    one-at-a-time form; workers>1 is the served (continuous batching) form.
    Served admission: the oldest waiting query that fits starts as soon as a
    worker is idle, except that at most max_heavy heavy queries (at least
    heavy_rows rows scanned; default half the largest plan type's) run at
    once: heavy scans are bandwidth-bound, so running two at once only
    stretches both, while a light query (anti-join, LIKE) overlaps a heavy
    one almost for free and may pass a heavy query that waits.
    min_overlap_ms > 0: a query does not start beside running ones when one
    of them is expected to finish within that time; min_overlap_ms < 0: within
    the waiting query's own expected service time (waiting is then shorter
    than running it stretched beside the other). Expected service time per
    plan type: service_ms, then a running mean over completed queries. Sharing
    the device stretches both queries, which costs more than a short wait.
    The window holds only while at most `workers` queries wait: in a backlog
    every query admits as soon as it fits (measured on the saturated four-type
    stream: 91 G rows/s with the window held, 107-113 G without)."""
    own = pool is None
    pool = (pool or [Worker(db) for _ in range(workers)])[:workers]
    arrivals = sorted(stream, key=lambda q: (q.arrival, q.qid))
    if heavy_rows is None:
        heavy_rows = max(db.rows_for(k) for k in KINDS) // 2
    cv = threading.Condition()
    state = {"used": 0, "done": 0, "inflight": 0, "max_inflight": 0, "heavy": 0}
    idle: deque = deque()
    running: dict = {}                          # runner -> its query
    est = dict(service_ms or {})               # plan type -> expected service ms

    def on_done(runner: _Runner, q: Query) -> None:
        with cv:
            state["used"] -= q.footprint
            state["done"] += 1
            state["inflight"] -= 1
            state["heavy"] -= q.rows >= heavy_rows
            running.pop(runner, None)
            ms = (q.done - q.start) * 1e3
            est[q.kind] = ms if q.kind not in est else 0.8 * est[q.kind] + 0.2 * ms
            idle.append(runner)
            cv.notify()

    def frees_soon(now: float, kind: str) -> float | None:
        """When a running query is expected to finish within the overlap
        window of a waiting query of this kind (or is overdue by less than
        that), the seconds until admission should be looked at again (the
        completion usually comes first); else None — a query overdue by more
        no longer holds others back."""
        if min_overlap_ms == 0 or not running:
            return None
        w = min_overlap_ms / 1e3 if min_overlap_ms > 0 else est.get(kind, 0.0) / 1e3
        if w <= 0:
            return None
        soon = None
        for r in running.values():
            left = est.get(r.kind, 0.0) / 1e3 - (now - r.start)
            if -w <= left < w:
                recheck = left + w
                soon = recheck if soon is None else min(soon, recheck)
        return soon

    t0 = time.perf_counter()
    runners = [_Runner(w, t0, on_done) for w in pool]
    for r in runners:
        r.start()
        idle.append(r)
    pending: deque = deque()          # arrived, not yet admitted (FIFO)
    ai = 0
    with cv:
        while state["done"] < len(stream):
            now = time.perf_counter() - t0
            while ai < len(arrivals) and arrivals[ai].arrival <= now:
                pending.append(arrivals[ai])
                ai += 1
            # admit, oldest first, every waiting query that fits while a
            # worker is idle (an empty device always admits one); a heavy
            # query waits while max_heavy heavy ones run, and light queries
            # behind it may pass
            wake = None
            while pending and idle:
                pick = None
                now = time.perf_counter() - t0
                for j, c in enumerate(pending):
                    fits = state["used"] + c.footprint <= budget_bytes or state["inflight"] == 0
                    heavy_ok = c.rows < heavy_rows or state["heavy"] < max_heavy or state["inflight"] == 0
                    # a running query about to finish: this one waits for it,
                    # unless more queries wait than there are workers (a
                    # backlog: then throughput decides, and every wait adds up)
                    backlog = len(pending) > len(pool)
                    soon = frees_soon(now, c.kind) if state["inflight"] > 0 and not backlog else None
                    if soon is not None:
                        wake = soon if wake is None else min(wake, soon)
                    if fits and heavy_ok and soon is None:
                        pick = j
                        break
                    if not fits:
                        break  # memory order is kept: nothing passes a query that does not fit
                if pick is None:
                    break
                q = pending[pick]
                del pending[pick]
                state["used"] += q.footprint
                state["inflight"] += 1
                state["heavy"] += q.rows >= heavy_rows
                state["max_inflight"] = max(state["max_inflight"], state["inflight"])
                q.start = time.perf_counter() - t0
                # the most recently idle worker takes it (LIFO): at low load
                # one worker, its thread and its stream serve every query, as
                # in the plain form, instead of rotating through cold ones
                runner = idle.pop()
                running[runner] = q
                runner.submit(q)
            if state["done"] >= len(stream):
                break
            # sleep until the next arrival or a completion (a waiting query
            # that cannot start now needs a completion; a new arrival may be
            # a light query that can)
            timeout = None
            if ai < len(arrivals):
                timeout = max(0.0, arrivals[ai].arrival - (time.perf_counter() - t0))
            if wake is not None:            # re-check when the expected completion is due
                timeout = wake if timeout is None else min(timeout, wake)
                timeout += 20e-6
            cv.wait(timeout)
    for r in runners:
        r.submit(None)
    for r in runners:
        r.join()
    makespan = max(q.done for q in stream) if stream else 0.0
    lat = [q.latency_ms for q in stream]
    rows = sum(q.rows for q in stream)
    out = {
        "queries": len(stream), "workers": len(pool), "budget_bytes": budget_bytes,
        "p50_ms": percentile(lat, 0.50), "p99_ms": percentile(lat, 0.99), "mean_ms": float(np.mean(lat)) if lat else 0,
        "makespan_s": makespan, "queries_per_s": len(stream) / makespan if makespan else 0.0,
        "rows_per_s": rows / makespan if makespan else 0.0, "max_inflight": state["max_inflight"],
        "errors": sum(1 for q in stream if q.error),
        "per_kind_p99_ms": {k: percentile([q.latency_ms for q in stream if q.kind == k], 0.99) for k in KINDS},
        "latency_ms": ST.trial_stats(lat),
        "service_ms": {k: round(v, 4) for k, v in est.items()},
    }
    if own:
        for w in pool:
            w.close()
    return out
