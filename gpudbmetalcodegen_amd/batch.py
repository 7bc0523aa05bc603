"""Workload 3: 64 independent pipelines of uneven size, run either one query
at a time (plain form) or all together in one persistent fused kernel
captured in a graph (batched form), with launch shapes from the autotuner.

Pipelines 0..31 filter + hash group-by (PIPE_AGG: p_date range, p_flag
equality, group by p_key, sum/count/min/max of p_val); 32..63 alternate
semi-join and anti-join against a 2,000,000-bit key bitmap with 30% of bits
set (PIPE_SEMI / PIPE_ANTI: count and sum of p_val). Sizes are
round(10^(4 + 3u)), u uniform, from the workload seed.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import engine as E
from . import workloads as W

BITMAP_BITS = 2_000_000
SIZE_CLASSES = [(0, 100_000, "small"), (100_000, 1_000_000, "medium"), (1_000_000, 1 << 62, "large")]


def _e():
    return N.load()[1]


@dataclass
class PipeSpec:
    index: int
    kind: str                 # "agg" | "semi" | "anti"
    n: int
    cols: dict
    params: dict = field(default_factory=dict)
    bitmap: np.ndarray | None = None

    @property
    def size_class(self) -> str:
        for lo, hi, name in SIZE_CLASSES:
            if lo <= self.n < hi:
                return name
        return "large"

    @property
    def shape(self) -> str:
        return f"{self.kind}/{self.size_class}"


def workload3(n_pipes: int = 64, scale: float = 1.0, seed: int | None = None) -> list[PipeSpec]:
    sd = W.seed_for(3) if seed is None else seed
    r = np.random.default_rng(sd)
    sizes = np.round(10 ** (4 + 3 * r.random(n_pipes)) * scale).astype(np.int64)
    specs = []
    half = n_pipes // 2
    for i in range(n_pipes):
        n = int(max(1, sizes[i]))
        s = sd * 1000 + i
        if i < half:
            K = int(round(2 ** r.uniform(4, 12)))           # 16 .. 4096 groups
            k = W.uniform_i64(s, 1, n, 0, K - 1)
            keys = ((k * 2654435761) % (1 << 31)).astype(np.int32)   # sparse int32 keys
            cols = {"p_key": keys, "p_date": W.uniform_i32(s, 2, n, 0, 2405), "p_val": W.uniform_i64(s, 3, n, 0, 10**7),
                    "p_flag": W.uniform_i32(s, 4, n, 0, 3).astype(np.uint8)}
            lo = int(r.integers(0, 1200))
            params = {"groups_bound": K, "pipe_date_lo": lo, "pipe_date_hi": lo + 1200, "pipe_flag": i % 4}
            specs.append(PipeSpec(i, "agg", n, cols, params))
        else:
            cols = {"p_key": W.uniform_i32(s, 1, n, 0, BITMAP_BITS - 1), "p_val": W.uniform_i64(s, 3, n, 0, 10**7)}
            bits = W.uniform_i32(s, 9, BITMAP_BITS, 0, 99) < 30
            words = np.packbits(bits, bitorder="little")
            words = np.concatenate([words, np.zeros((-words.size) % 4 + 4, np.uint8)]).view(np.uint32)
            specs.append(PipeSpec(i, "semi" if (i - half) % 2 == 0 else "anti", n, cols, {}, words))
    return specs


def plan_name(kind: str) -> str:
    return {"agg": "PIPE_AGG", "semi": "PIPE_SEMI", "anti": "PIPE_ANTI"}[kind]


def phase_name(kind: str) -> str:
    return {"agg": "PIPE_agg", "semi": "PIPE_semi", "anti": "PIPE_anti"}[kind]


def next_pow2(x: int) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


class Workload3:
    """Device-resident pipelines, each with its own runner."""

    def __init__(self, ctx: E.Context, specs: list[PipeSpec], lds_cap: int = 4096, window: int = 4,
                 partials: int = 256, lds_bitmap_words: int = 32768):
        """partials: in the batched form the aggregate pipelines store each
        work item's LDS table as a partial table, merged into the group table
        by a second phase (a second persistent kernel), up to this many per
        pipeline; 0: every work item merges its LDS table into the group
        table itself. The plain form always merges per block: its one launch
        per pipeline has few blocks to merge, and a merge phase would add a
        launch per pipeline."""
        self.ctx = ctx
        self.specs = specs
        self.lds_cap = lds_cap
        agg_opts = f"lds_cap={lds_cap},window={window}"
        kinds = {s.kind for s in specs}
        # batched join bodies keep the first 32768 bitmap words (1M keys) in
        # the LDS the group-by bodies size the batched kernel for
        join_opts = f"lds_bitmap_words={lds_bitmap_words}" if lds_bitmap_words else ""
        self.plans = {k: E.build_hip_plan(plan_name(k), agg_opts if k == "agg" else join_opts).compile(ctx)
                      for k in kinds}
        self.batch_plans = dict(self.plans)
        if partials > 0 and "agg" in kinds:
            self.batch_plans["agg"] = E.build_hip_plan("PIPE_AGG", f"{agg_opts},partials={partials}").compile(ctx)
        self.tables, self.execs, self.bitmaps = [], [], []
        self.launch = {}   # pipe index -> (block, items, lds_bits)
        self.batch_launch = {}   # the batched form's own choices, when tune_batch has run
        for s in specs:
            t = E.DeviceTable.from_host(ctx, "P", s.cols)
            ex = E.HipGenericExecutor(ctx)
            ex.register_table(t)
            if s.kind == "agg":
                for k in ("pipe_date_lo", "pipe_date_hi", "pipe_flag"):
                    ex.register_scalar(k, s.params[k])
                ex.register_scalar("pipe_lds_bits", 10)
                ex.register_symbol("pipe_slots", max(64, next_pow2(2 * s.params["groups_bound"])))
                bm = None
            else:
                bm = ctx.upload(s.bitmap)
                ex.register_buffer("d_pipe_bitmap", bm, s.bitmap.nbytes)
                ex.register_symbol("pipe_bits", BITMAP_BITS)
            self.tables.append(t)
            self.execs.append(ex)
            self.bitmaps.append(bm)
            self.launch[s.index] = (256, 4, 10)

    @property
    def total_rows(self) -> int:
        return sum(s.n for s in self.specs)

    def launch_for(self, i: int, batched: bool = False) -> tuple[int, int, int]:
        s = self.specs[i]
        return self.batch_launch.get(s.index, self.launch[s.index]) if batched else self.launch[s.index]

    def _apply(self, i: int, batched: bool = False) -> E.HipQueryPlan:
        s = self.specs[i]
        plan = (self.batch_plans if batched else self.plans)[s.kind]
        b, it, lb = self.launch_for(i, batched)
        plan.set_launch(phase_name(s.kind), b, it)
        if s.kind == "agg":
            self.execs[i].register_scalar("pipe_lds_bits", lb)
        return plan

    # --- autotuner (per pipeline shape) ----------------------------------
    def autotune(self, blocks=(64, 128, 256, 512, 1024), items=(1, 2, 4, 8, 16), lds_bits=(8, 10, 12),
                 trials=5) -> dict:
        # Each shape is tuned on its largest pipeline: that is where the
        # class's time goes, and a configuration that is as fast on a small
        # table (a tie within noise) can be much slower on a large one.
        reps = {}
        for i, s in enumerate(self.specs):
            if s.shape not in reps or s.n > self.specs[reps[s.shape]].n:
                reps[s.shape] = i
        chosen = {}
        for shape, i in sorted(reps.items()):
            s = self.specs[i]
            plan = self.plans[s.kind]
            bl = (C.c_int * len(blocks))(*blocks)
            it = (C.c_int * len(items))(*items)
            scal = "pipe_lds_bits" if s.kind == "agg" else None
            cand = [b for b in lds_bits if (1 << b) <= self.lds_cap]
            sv = (C.c_int64 * len(cand))(*cand) if scal else None
            nsv = len(cand) if scal else 0
            ncap = len(blocks) * len(items) * max(1, nsv)
            table = (C.c_double * (4 * ncap))()
            bb, bi, bs, ms, nc = C.c_int(), C.c_int(), C.c_int64(), C.c_float(), C.c_int()
            N.eng_check(_e().gdb_autotune_phase(self.execs[i].handle, plan.handle, phase_name(s.kind).encode(), bl,
                                                len(blocks), it, len(items), scal.encode() if scal else None, sv, nsv,
                                                trials, C.byref(bb), C.byref(bi), C.byref(bs), C.byref(ms), table,
                                                C.byref(nc)))
            chosen[shape] = {"block": bb.value, "items": bi.value, "lds_slots": (1 << bs.value) if scal else None,
                             "lds_bits": int(bs.value) if scal else None, "p50_ms": ms.value,
                             "candidates": nc.value, "tuned_on_rows": s.n}
        for s in self.specs:
            c = chosen[s.shape]
            self.launch[s.index] = (c["block"], c["items"], c["lds_bits"] if c["lds_bits"] is not None else 10)
        return chosen

    def tune_batch(self, block: int = 1024, chunk_rows: int = 32768, items=(1, 2, 4, 8), lds_bits=(10, 12),
                   runs: int = 7) -> dict:
        """Per-kind body settings for the batched form, chosen on the batched
        replay itself (device time, median of `runs` graph replays): first
        ITEMS x LDS table bits shared by every aggregate body, then ITEMS
        shared by every join body. Settings tuned on one pipeline launched
        alone do not carry over: in the shared persistent kernel the LDS
        table size also sets how many partial tables the merge level folds
        and how many blocks share a CU. Starts from the per-pipeline choices
        of autotune() (or the defaults); the plain form keeps those."""
        self.batch_launch = {}
        kinds = {s.kind for s in self.specs}
        cand_bits = [b for b in lds_bits if (1 << b) <= self.lds_cap]
        trace = []

        def measure(tag: str) -> float:
            b = self.build_batch(block, chunk_rows)
            try:
                b.run(graph=True)
                ms = sorted(b.run(graph=True) for _ in range(runs))[runs // 2]
            finally:
                b.close()
            trace.append({"setting": tag, "device_ms": round(ms, 4)})
            return ms

        def assign(kind_test, it: int, lb: int | None) -> None:
            for i, s in enumerate(self.specs):
                if kind_test(s.kind):
                    bb, _, old_lb = self.launch_for(i, True)
                    self.batch_launch[s.index] = (bb, it, old_lb if lb is None else lb)

        chosen = {}
        groups = [("agg", lambda k: k == "agg", cand_bits or [None]),
                  ("join", lambda k: k != "agg", [None])]
        for name, test, bits in groups:
            if not any(test(k) for k in kinds):
                continue
            best = None
            for lb in bits:
                for it in items:
                    assign(test, it, lb)
                    ms = measure(f"{name} items={it}" + (f" lds_bits={lb}" if lb is not None else ""))
                    if best is None or ms < best[0]:
                        best = (ms, it, lb)
            assign(test, best[1], best[2])
            chosen[name] = {"items": best[1], "lds_bits": best[2], "device_ms": round(best[0], 4)}
        return {"block": block, "chunk_rows": chunk_rows, "chosen": chosen, "trace": trace}

    # --- plain form: one query at a time ----------------------------------
    def run_plain(self) -> float:
        self.ctx.sync()
        t0 = time.perf_counter()
        for i in range(len(self.specs)):
            self.execs[i].execute(self._apply(i), 0, 1)
        return (time.perf_counter() - t0) * 1e3

    def results(self) -> list[dict]:
        out = []
        for i, s in enumerate(self.specs):
            res = self.execs[i].collect(self.plans[s.kind])
            out.append({k: np.asarray(v) for k, v in res.columns.items()})
        return out

    # --- batched form ----------------------------------------------------
    def build_batch(self, block: int, chunk_rows: int = 32768, grid: int = 0, fold_rows: int = 0) -> "Batch":
        """fold_rows > 0: rows per work item of the partial-table fold level
        (phase level 1) instead of chunk_rows."""
        return Batch(self, block, chunk_rows, grid, fold_rows)

    def close(self):
        for t in self.tables:
            t.free(self.ctx)
        for bm in self.bitmaps:
            if bm:
                self.ctx.free(bm)
        for ex in self.execs:
            ex.close()


class Batch:
    def __init__(self, w: Workload3, block: int, chunk_rows: int, grid: int, fold_rows: int = 0):
        self.w = w
        self._h = C.c_void_p()
        N.eng_check(_e().gdb_batch_create(w.ctx.handle, C.byref(self._h)))
        for i, s in enumerate(w.specs):
            plan = w._apply(i, batched=True)
            idx = C.c_int()
            N.eng_check(_e().gdb_batch_add(self._h, plan.handle, w.execs[i].handle, w.launch_for(i, True)[1],
                                           C.byref(idx)))
        if fold_rows:
            N.eng_check(_e().gdb_batch_set_level_chunk(self._h, 1, fold_rows))
        g, nitems = C.c_uint(), C.c_uint64()
        N.eng_check(_e().gdb_batch_build(self._h, block, chunk_rows, grid, C.byref(g), C.byref(nitems)))
        self.grid, self.work_items, self.block = g.value, nitems.value, block

    def source(self) -> str:
        return N.two_call_string(_e().gdb_batch_source, self._h)

    def capture(self) -> None:
        N.eng_check(_e().gdb_batch_capture(self._h))

    def run(self, graph: bool = True) -> float:
        ms = C.c_float()
        N.eng_check(_e().gdb_batch_run(self._h, 1 if graph else 0, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            _e().gdb_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main(argv=None) -> int:
    """Run the 64-pipeline workload's batched form on device 0: per-shape
    autotune, the batched bodies' own tuning, one batch build, then
    `--replays` graph replays; prints one JSON line (settings and device
    milliseconds per replay). A fixed command for profilers."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m gpudbmetalcodegen_amd.batch")
    ap.add_argument("--replays", type=int, default=20)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--chunk", type=int, default=131072)
    ap.add_argument("--fold", type=int, default=32768, help="rows per work item of the fold level (0: --chunk)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default body settings")
    ap.add_argument("--set", default="", metavar="AGG_ITEMS,LDS_BITS,JOIN_ITEMS",
                    help="batched body settings given instead of tuned (e.g. 4,12,2)")
    args = ap.parse_args(argv)
    from pathlib import Path
    ctx = E.Context(0)
    ctx.set_cache_dir(str(Path(__file__).resolve().parent / "_cocache"))
    w = Workload3(ctx, workload3(), lds_cap=4096)
    try:
        tuned = None
        if args.set:
            ai, lb, ji = (int(x) for x in args.set.split(","))
            for i, s in enumerate(w.specs):
                w.batch_launch[s.index] = (256, ai, lb) if s.kind == "agg" else (256, ji, 10)
            tuned = {"agg": {"items": ai, "lds_bits": lb}, "join": {"items": ji}, "given": True}
        elif not args.no_tune:
            w.autotune(trials=3)
            tuned = w.tune_batch(args.block, 32768)["chosen"]
        b = w.build_batch(args.block, args.chunk, fold_rows=args.fold)
        try:
            b.run(graph=True)
            ms = [b.run(graph=True) for _ in range(args.replays)]
        finally:
            b.close()
        print(json.dumps({"rows": w.total_rows, "block": args.block, "chunk_rows": args.chunk, "fold_rows": args.fold,
                          "grid": b.grid,
                          "work_items": b.work_items, "batch_tuned": tuned,
                          "device_ms_median": sorted(ms)[len(ms) // 2], "device_ms": [round(x, 4) for x in ms]}),
              flush=True)
    finally:
        w.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
