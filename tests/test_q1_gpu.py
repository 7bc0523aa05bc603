"""Workload 0 (fused filter + group-by, reference Q1) on the GPU vs the oracle.

Bit-exact: every aggregate is an exact integer (int128 for the charge sum),
averages are trunc(sum * 10^6 / count) on both sides.
"""

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import engine as E
from gpudbmetalcodegen_amd import queries as Q
from gpudbmetalcodegen_amd import workloads as W
from tests.helpers import as_lists, q1_expected

pytestmark = pytest.mark.gpu


def _run(ctx, F, cutoff, options=""):
    dev = Q.upload_tables(ctx, {"F": F})
    try:
        run = Q.run_q1(ctx, dev, cutoff, options=options)
        return as_lists(run.result.columns), run
    finally:
        dev["F"].free(ctx)


@pytest.mark.parametrize("n", [0, 1, 63, 1000, 4097, 100_003])
def test_q1_small_sizes(ctx, n):
    F = W.fact_f(n, W.seed_for(0) + n)
    cutoff = W.q1_cutoff()
    got, _ = _run(ctx, F, cutoff)
    assert got == q1_expected(oracle.q1(F, cutoff))


@pytest.mark.parametrize("cutoff", [-1, 0, 1200, 10_000])
def test_q1_cutoffs(ctx, cutoff):
    F = W.fact_f(50_000, W.seed_for(0) + 7)
    got, _ = _run(ctx, F, cutoff)
    assert got == q1_expected(oracle.q1(F, cutoff))


@pytest.mark.parametrize("options", ["items=2", "items=8,block=128", "items=4,block=64", "prefetch=1", "prefetch=1,items=2"])
def test_q1_launch_variants(ctx, options):
    F = W.fact_f(333_333, W.seed_for(0) + 3)
    got, _ = _run(ctx, F, W.q1_cutoff(), options)
    assert got == q1_expected(oracle.q1(F, W.q1_cutoff()))


def test_q1_int128_charge_and_negative_values(ctx):
    # 100 rows at price 8e14 (disc 0, tax 8): each charge term is 8.64e18, the
    # sum 8.64e20 needs the 128-bit accumulator; the negative copy checks the
    # sign handling of the high limb.
    n = 100
    for sign in (1, -1):
        F = {
            "f_qty": np.full(n, 50, np.int32),
            "f_price": np.full(n, sign * 800_000_000_000_000, np.int64),
            "f_disc": np.zeros(n, np.int32),
            "f_tax": np.full(n, 8, np.int32),
            "f_flag": np.full(n, ord("N"), np.uint8),
            "f_status": np.full(n, ord("O"), np.uint8),
            "f_date": np.zeros(n, np.int32),
        }
        got, _ = _run(ctx, F, 0)
        exp = q1_expected(oracle.q1(F, 0))
        assert got == exp
        assert got["sum_charge"] == [sign * 864 * 10**18]


def test_q1_full_6m(ctx):
    F = W.workload0()["F"]
    got, run = _run(ctx, F, W.q1_cutoff())
    assert got == q1_expected(oracle.q1(F, W.q1_cutoff()))
    assert sum(got["count_order"]) == int((F["f_date"] <= W.q1_cutoff()).sum())
    print("Q1 6M timing ms", run.timing.total_ms, run.timing.phases)


def test_scalar_atomic_ablation_is_bit_exact(ctx, monkeypatch):
    """The reduction ablation (one global atomic per row and slot,
    GDB_SCALAR_ATOMIC=1) gives the same exact Q1 as the block reduction."""
    from gpudbmetalcodegen_amd import engine as E

    F = W.fact_f(200_000, W.seed_for(0) + 7)
    dev = Q.upload_tables(ctx, {"F": F})
    try:
        monkeypatch.setenv("GDB_SCALAR_ATOMIC", "1")
        plan = E.build_hip_plan("Q1", "").compile(ctx)
        assert "reduction ablation" in plan.source()
        ex = E.HipGenericExecutor(ctx)
        ex.register_table(dev["F"])
        ex.register_scalar("q1_cutoff", W.q1_cutoff())
        ex.execute(plan, 0, 1)
        got = ex.collect(plan).columns
        ex.close()
        exp = q1_expected(oracle.q1(F, W.q1_cutoff()))
        assert as_lists(got) == exp
    finally:
        dev["F"].free(ctx)


def test_phase_timing_off_keeps_results_and_total(ctx):
    """Without per-phase event records a run still reports its whole device
    time and the same exact result; per-phase times come back with them on."""
    F = W.fact_f(200_000, W.seed_for(0) + 11)
    dev = Q.upload_tables(ctx, {"F": F})
    try:
        run = Q.run_q1(ctx, dev, W.q1_cutoff())
        ex, plan = run.executor, run.plan
        ex.set_phase_timing(False)
        t = ex.execute(plan, 0, 1)
        assert t.total_ms > 0 and all(ms == 0.0 for _, ms, *_ in t.phases)
        assert as_lists(ex.collect(plan).columns) == q1_expected(oracle.q1(F, W.q1_cutoff()))
        ex.set_phase_timing(True)
        t = ex.execute(plan, 0, 1)
        assert t.total_ms > 0 and all(ms > 0.0 for _, ms, *_ in t.phases)
    finally:
        dev["F"].free(ctx)
