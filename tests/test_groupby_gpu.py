"""Workload 2 (hash group-by of sum/count/min/max per key) on the GPU vs the
oracle, for both table variants and all three sizing methods (exact pre-pass,
given size, resize-on-overflow). Results are compared group by group."""

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import queries as Q
from gpudbmetalcodegen_amd import workloads as W

pytestmark = pytest.mark.gpu


def _cmp(got, exp):
    cols = got.columns
    assert np.array_equal(cols["key"], exp["key"])
    for c in ("count", "sum", "min", "max"):
        assert np.array_equal(cols[c], exp[c]), c


def _gen(n, s, nkeys, seed):
    return W.workload2(n=n, s=s, nkeys=nkeys, seed=seed)["G"]


@pytest.mark.parametrize("variant", ["lds", "global", "part"])
@pytest.mark.parametrize("n,s,nkeys", [(1, 1.1, 10), (5000, 1.1, 1000), (300_000, 1.1, 200_000),
                                       (300_000, 0.0, 50_000), (200_000, 1.5, 20)])
def test_groupby_prepass(ctx, variant, n, s, nkeys):
    G = _gen(n, s, nkeys, 100 + n)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        run = Q.run_groupby(ctx, dev, prepass=True, variant=variant, key_min=1, key_max=nkeys)
        exp = oracle.groupby(G)
        _cmp(run.result, exp)
        assert run.resizes == 0
        assert run.slots >= 2 * len(exp["key"])
    finally:
        dev["G"].free(ctx)


@pytest.mark.parametrize("variant", ["lds", "part"])
def test_groupby_filter_selectivity(ctx, variant):
    G = _gen(200_000, 1.1, 100_000, 7)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        for cut in (0, 10_000, 500_000, 1_000_001):
            run = Q.run_groupby(ctx, dev, prepass=True, val_cut=cut, key_min=1, key_max=100_000, variant=variant)
            _cmp(run.result, oracle.groupby(G, cut))
    finally:
        dev["G"].free(ctx)


def test_groupby_resize_on_overflow_stays_exact(ctx):
    G = _gen(400_000, 0.5, 300_000, 9)
    exp = oracle.groupby(G)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        run = Q.run_groupby(ctx, dev, slots=1024, variant="lds")   # far too small on purpose
        assert run.resizes >= 1
        _cmp(run.result, exp)
        run2 = Q.run_groupby(ctx, dev, slots=4 * len(exp["key"]), variant="global")
        assert run2.resizes == 0
        _cmp(run2.result, exp)
    finally:
        dev["G"].free(ctx)


@pytest.mark.parametrize("s,lds,window", [(1.1, 4096, 2), (0.0, 1024, 1), (1.5, 64, 16)])
def test_groupby_partitioned_given_sizes(ctx, s, lds, window):
    """Two-pass partitioned form at several partition counts (slots / 4096)
    and pass-1 LDS sizes, skewed and uniform keys; every row is counted once
    (hot partials + cold rows)."""
    G = _gen(1_000_003, s, 300_000, 21)
    exp = oracle.groupby(G)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        for slots in (1 << 20, 1 << 21):
            run = Q.run_groupby(ctx, dev, slots=slots, variant="part", lds_slots=lds, window=window)
            assert run.resizes == 0
            _cmp(run.result, exp)
            assert int(run.result.columns["count"].sum()) == 1_000_003
            assert run.timing.status[0][3] > 0 or s > 1.0      # cold rows took the partition path
            assert run.timing.status[-1][2] == len(exp["key"])  # groups written by the last pass
    finally:
        dev["G"].free(ctx)


def test_groupby_partitioned_table_overflow_recovers_exactly(ctx):
    """A partition slice holding more than 4096 groups (table far too small)
    is reported and retried with a doubled table; results stay exact."""
    G = _gen(1_000_000, 0.0, 200_000, 23)
    exp = oracle.groupby(G)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        run = Q.run_groupby(ctx, dev, slots=4096, variant="part", lds_slots=64)
        assert run.resizes >= 1 and run.slots > 4096
        _cmp(run.result, exp)
    finally:
        dev["G"].free(ctx)


def test_groupby_partitioned_lds_shrinks_for_many_partitions(ctx):
    """slots / 4096 partitions beyond the LDS counters that fit beside a
    4096-slot pass-1 table: the driver shrinks the pass-1 table; exact."""
    G = _gen(600_000, 1.1, 400_000, 29)
    exp = oracle.groupby(G)
    assert Q.part_lds_for(1 << 25, 4096) < 4096
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        run = Q.run_groupby(ctx, dev, slots=1 << 25, variant="part", lds_slots=4096)
        assert run.resizes == 0
        _cmp(run.result, exp)
    finally:
        dev["G"].free(ctx)


def test_groupby_lds_slot_sizes(ctx):
    G = _gen(250_000, 1.2, 30_000, 13)
    exp = oracle.groupby(G)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        for ls in (256, 1024, 4096):
            run = Q.run_groupby(ctx, dev, prepass=True, lds_slots=ls, key_min=1, key_max=30_000)
            _cmp(run.result, exp)
    finally:
        dev["G"].free(ctx)


@pytest.mark.parametrize("big", [False, True])
def test_groupby_partitioned_narrow_records_fall_back_exactly(ctx, big):
    """Cold rows travel as 8-byte records (key | value << 32) while every
    value fits in 32 bits. With 50 values beyond that range among the rows
    (and a 64-slot pass-1 table, so nearly every row is cold) the scatter
    reports them and the run repeats with 16-byte records. Exact either way,
    and equal to a run that asks for 16-byte records from the start."""
    G = dict(_gen(500_000, 0.0, 400_000, 31))
    if big:
        v = G["g_val"].copy()
        idx = np.arange(50) * 9_973 + 11
        v[idx] = np.where(idx % 2 == 0, -(1 << 40) - idx, (1 << 40) + idx)
        v[idx[:3]] = [-(1 << 31) - 1, 1 << 31, (1 << 62) + 5]
        G["g_val"] = v
    exp = oracle.groupby(G)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        run = Q.run_groupby(ctx, dev, slots=1 << 20, variant="part", lds_slots=64)
        assert run.narrow == (not big) and run.resizes == (1 if big else 0)
        _cmp(run.result, exp)
        wide = Q.run_groupby(ctx, dev, slots=1 << 20, variant="part", lds_slots=64, narrow=False)
        assert not wide.narrow and wide.resizes == 0
        _cmp(wide.result, exp)
    finally:
        dev["G"].free(ctx)


def test_groupby_fused_prepass_beyond_lds_counters_recovers_exactly(ctx):
    """Plain form, partitioned: 20M uniform rows over 20M keys give ~12.6M
    groups, a 2^25-slot table and 8192 partitions — more than the LDS
    counters beside a 4096-slot pass-1 table. The offsets pass reports it,
    the run is rebuilt with a smaller LDS table (more counters) and the
    result stays exact."""
    G = _gen(20_000_000, 0.0, 20_000_000, 21)
    exp = oracle.groupby(G, threads=16)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        run = Q.run_groupby(ctx, dev, prepass=True, key_min=1, key_max=20_000_000, variant="part", lds_slots=4096)
        assert run.resizes >= 1 and run.slots == 1 << 25
        _cmp(run.result, exp)
    finally:
        dev["G"].free(ctx)


@pytest.mark.slow
def test_groupby_full_size(ctx):
    G = W.workload2()["G"]
    exp = oracle.groupby(G, threads=16)
    dev = Q.upload_tables(ctx, {"G": G})
    try:
        for variant, lds in (("lds", 2048), ("part", 4096)):
            run = Q.run_groupby(ctx, dev, prepass=True, key_min=1, key_max=20_000_000, variant=variant, lds_slots=lds)
            _cmp(run.result, exp)
            assert int(run.result.columns["count"].sum()) == G["g_key"].shape[0]
        print("GROUPBY full", run.timing.total_ms, run.timing.phases, run.timing.status)
    finally:
        dev["G"].free(ctx)
