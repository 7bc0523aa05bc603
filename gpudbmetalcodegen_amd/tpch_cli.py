"""Command-line driver of the named TPC-H plans, the counterpart of the
reference's ``GPUDBCodegen [flags] [sf1|sf10|...] q<N>|all``
(``codegen/codegen_main.cpp:1782-1960``) for the 22 TPC-H queries.

    python -m gpudbmetalcodegen_amd.tpch_cli [flags] [sf1|sf10|sf20|sf50|sf100] q<N>|all
    python -m gpudbmetalcodegen_amd.tpch_cli [flags] [sf1|...] --sql FILE [--sql FILE ...]

Data comes from ``data/SF-<n>/<table>.colbin`` (built from ``<table>.tbl`` on
first use, as tools/tbl_to_colbin does), from ``--data-dir DIR``, or from
``--generate SF`` (seeded TPC-H-shaped tables in memory). Flags kept from the
reference: ``--warmup N`` (3), ``--repeat N`` (1, the median is reported),
``--csv``, ``--print-plan``, ``--dump-src DIR`` (generated HIP source, the
counterpart of ``--dump-msl``), ``--check DIR``, ``--save-golden DIR``,
``--check-abs-tol``, ``--check-rel-tol``, ``--threadgroup-size N``,
``--chunk N[K|M|G]`` (stream lineitem, or orders for queries without it, in
chunks of N rows from host memory; queries whose plan is not chunkable are
refused with status 3, as codegen_main.cpp:320-335 refuses them). Exit
status: 0 on success, 1 when a check fails, 2 when a golden file is missing
(the reference's g_checkExitCode), 3 on a usage error. ``--sql FILE`` runs
a single-table SELECT through the engine's SQL front end (the file's stem
names it for --check / --save-golden), the counterpart of the reference's
analysis-based path from sql/q<N>.sql text.
"""

from __future__ import annotations

import argparse
import os
import statistics
import sys
import time
from pathlib import Path

from . import engine as E
from . import tpch as T

SF_DIRS = {"sf1": ("data/SF-1/", "SF1"), "sf10": ("data/SF-10/", "SF10"), "sf20": ("data/SF-20/", "SF20"),
           "sf50": ("data/SF-50/", "SF50"), "sf100": ("data/SF-100/", "SF100")}
# Plans that take a block-size option (the reference's --threadgroup-size).
BLOCK_OPTION = {"Q1", "Q6"}


def parse(argv):
    ap = argparse.ArgumentParser(prog="tpch_cli", description=__doc__.split("\n\n")[0])
    ap.add_argument("args", nargs="*", help="[sf1|sf10|sf20|sf50|sf100] q<N>|all")
    ap.add_argument("--sql", action="append", default=[], help="a file holding one single-table SELECT")
    ap.add_argument("--data-dir", default=None)
    ap.add_argument("--generate", type=float, default=None, help="generate TPC-H-shaped tables at this scale factor")
    ap.add_argument("--sf-label", default=None, help="scale label of golden file names (default from the data set)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--csv", action="store_true")
    ap.add_argument("--print-plan", action="store_true")
    ap.add_argument("--dump-src", default=None)
    ap.add_argument("--check", default=None)
    ap.add_argument("--save-golden", default=None)
    ap.add_argument("--check-abs-tol", type=float, default=T.CHECK_ABS_TOL)
    ap.add_argument("--check-rel-tol", type=float, default=T.CHECK_REL_TOL)
    ap.add_argument("--threadgroup-size", type=int, default=None)
    ap.add_argument("--scalar-atomic", action="store_true",
                    help="reduction ablation: one global atomic per row instead of the block reduction "
                         "(sets GDB_SCALAR_ATOMIC=1, as the reference's --scalar-atomic)")
    ap.add_argument("--chunk", default=None, help="N[K|M|G]: stream the query's largest table in chunks of N rows")
    return ap.parse_args(argv)


def _golden(a, canonical: str, q: str, label: str) -> int:
    """--save-golden / --check for one result; the check's exit status."""
    if a.save_golden:
        p = T.save_golden(canonical, a.save_golden, q, label)
        print(f"[GOLDEN] saved {p}")
    if not a.check:
        return 0
    st, msg = T.check(canonical, a.check, q, label, a.check_abs_tol, a.check_rel_tol)
    if st == 0:
        print(f"[CHECK] {q} @ {label}: OK")
        return 0
    print(f"[CHECK] {q} @ {label}: {'FAIL — ' if st == 1 else ''}{msg}", file=sys.stderr)
    return st


def main(argv=None) -> int:
    a = parse(argv if argv is not None else sys.argv[1:])
    data_dir, label = None, a.sf_label
    queries = []
    for x in a.args:
        lx = x.lower()
        if lx in SF_DIRS:
            data_dir, lab = SF_DIRS[lx]
            label = label or lab
        elif lx == "all":
            queries = list(T.QUERIES)
        elif lx.startswith("q") and lx[1:].isdigit():
            q = "Q" + lx[1:]
            if q not in T.QUERIES:
                print(f"{q}: not planned by this engine (supported: {', '.join(T.QUERIES)})", file=sys.stderr)
                return 3
            queries.append(q)
        else:
            print(f"unexpected argument {x}", file=sys.stderr)
            return 3
    sql_files = []
    for f in a.sql:
        if not Path(f).is_file():
            print(f"tpch_cli: no SQL file {f}", file=sys.stderr)
            return 3
        sql_files.append(Path(f))
    if not queries and not sql_files:
        print("usage: tpch_cli [flags] [sf1|sf10|sf20|sf50|sf100] q<N>|all | --sql FILE", file=sys.stderr)
        return 3
    if a.scalar_atomic:
        os.environ["GDB_SCALAR_ATOMIC"] = "1"
    data_dir = a.data_dir or data_dir
    try:
        chunk = T.parse_rows(a.chunk) if a.chunk else 0
    except ValueError as e:
        print(f"tpch_cli: --chunk: {e}", file=sys.stderr)
        return 3
    if a.generate is None and data_dir is None:
        data_dir, label = SF_DIRS["sf1"][0], label or "SF1"
    if E.device_count() < 1:
        print("tpch_cli: no GPU visible", file=sys.stderr)
        return 3

    t0 = time.perf_counter()
    if a.generate is not None:
        from . import tpch_gen as G

        host = G.generate(a.generate, seed=1)
        label = label or f"SF{a.generate:g}"
        src = f"generated SF {a.generate:g}"
    else:
        built = T.ingest(data_dir)
        host = T.load_dir(data_dir)
        missing = [t for t in T.TABLES if t not in host]
        if missing:
            print(f"tpch_cli: {data_dir}: no .tbl/.colbin for {', '.join(missing)}", file=sys.stderr)
            return 3
        src = f"{data_dir} (.colbin built for: {', '.join(t for t, b in built.items() if b) or 'none'})"
    load_ms = (time.perf_counter() - t0) * 1e3
    if not a.csv:
        print(f"data: {src}, load {load_ms:.0f} ms")

    ctx = E.Context(0)
    cache = Path(__file__).resolve().parent / "_cocache"
    if cache.exists():
        ctx.set_cache_dir(str(cache))
    db = T.TpchDB(ctx, host)
    status = 0
    try:
        for f in sql_files:
            from . import sql as S

            q = f.stem
            run = S.run_sql(ctx, db, f.read_text(), warmup=a.warmup, measured=1)
            times = [run.timing.total_ms]
            for _ in range(max(0, a.repeat - 1)):
                times.append(S.run_sql(ctx, db, f.read_text()).timing.total_ms)
            med = statistics.median(times)
            if a.print_plan:
                print(run.plan.describe())
            if a.dump_src:
                Path(a.dump_src).mkdir(parents=True, exist_ok=True)
                (Path(a.dump_src) / f"{q}.hip").write_text(run.plan.source())
            if a.csv:
                print(f"TIMING_CSV,{q},{label},{med:.4f},{run.plan.compile_ms:.2f},{len(times)}")
            else:
                print(f"{q} (SQL) @ {label}: GPU {med:.3f} ms (median of {len(times)}), "
                      f"{len(next(iter(run.columns.values())))} rows")
                print(run.canonical.rstrip("\n"))
            status = max(status, _golden(a, run.canonical, q, label))
        for q in queries:
            opts = f"block={a.threadgroup_size}" if a.threadgroup_size and q in BLOCK_OPTION else ""
            from .queries import get_plan

            plan = get_plan(ctx, q, ",".join(o for o in (T.PLAN_OPTIONS.get(q, ""), opts) if o))
            if a.print_plan:
                print(plan.describe())
            if a.dump_src:
                Path(a.dump_src).mkdir(parents=True, exist_ok=True)
                (Path(a.dump_src) / f"{q}.hip").write_text(plan.source())
            ex = E.HipGenericExecutor(ctx)
            if chunk and not plan.chunkable:
                print(f"tpch_cli: {q} does not support chunked execution yet (--chunk is unsafe for this query)",
                      file=sys.stderr)
                status = max(status, 3)
                continue
            # the reference's chunk decision: automatic when the working set
            # does not fit, an explicit --chunk downgraded when it does
            # (GDB_FORCE_CHUNK=1 keeps it)
            q_chunk, why = T.auto_chunk_rows(db, q, plan, requested_rows=chunk)
            if plan.chunkable and (q_chunk != chunk or q_chunk) and not a.csv:
                print(f"[auto-chunk] {q}: {why}")
            # chunks of tables read from .colbin files come straight from the file
            cdir = data_dir if (q_chunk and a.generate is None) else None
            run = T.run(ctx, db, q, warmup=a.warmup, measured=1, options=opts, executor=ex, chunk_rows=q_chunk,
                        colbin_dir=cdir)
            times = [run.timing.total_ms]
            for _ in range(max(0, a.repeat - 1)):
                times.append(T.run(ctx, db, q, warmup=0, measured=1, options=opts, executor=ex,
                                   chunk_rows=q_chunk, colbin_dir=cdir).timing.total_ms)
            if q_chunk and not a.csv:
                print(f"{q}: streamed {T.stream_table(q)} in {run.timing.chunks} chunks of {q_chunk} rows, "
                      f"{run.timing.bytes_copied / 1e6:.1f} MB copied")
            med = statistics.median(times)
            if a.csv:
                print(f"TIMING_CSV,{q},{label},{med:.4f},{plan.compile_ms:.2f},{len(times)}")
            else:
                print(f"{q} @ {label}: GPU {med:.3f} ms (median of {len(times)}), compile {plan.compile_ms:.1f} ms"
                      f"{' (cache hit)' if plan.cache_hit else ''}, {len(next(iter(run.columns.values())))} rows")
                print(run.canonical.rstrip("\n"))
            status = max(status, _golden(a, run.canonical, q, label))
            ex.release()
    finally:
        db.free()
        ctx.close()
    return status


if __name__ == "__main__":
    sys.exit(main())
