"""Workload 3: many uneven pipelines — plain per-pipeline runs vs one
persistent fused launch (eager and graph-replayed), after autotuning.
Every pipeline's result must equal the oracle and be identical across
forms and across graph replays."""

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import batch as B
from gpudbmetalcodegen_amd import engine as E

pytestmark = pytest.mark.gpu


def expected(s: B.PipeSpec) -> dict:
    c = s.cols
    if s.kind == "agg":
        p = s.params
        m = (c["p_date"] >= p["pipe_date_lo"]) & (c["p_date"] < p["pipe_date_hi"]) & (c["p_flag"] == p["pipe_flag"])
        g = oracle.groupby({"g_key": c["p_key"][m], "g_val": c["p_val"][m]}) if m.any() else \
            {k: np.zeros(0, np.int64) for k in ("key", "count", "sum", "min", "max")}
        return g
    words = s.bitmap
    k = c["p_key"].astype(np.int64)
    hit = ((words[k >> 5] >> (k & 31).astype(np.uint32)) & 1).astype(bool)
    if s.kind == "anti":
        hit = ~hit
    cnt = int(hit.sum())
    return {"rows": [cnt] if cnt else [], "sum_val": [int(c["p_val"][hit].sum())] if cnt else []}


def same(got: dict, exp: dict) -> bool:
    for k, v in exp.items():
        if not np.array_equal(np.asarray(got[k], dtype=np.int64), np.asarray(v, dtype=np.int64)):
            return False
    return True


@pytest.fixture(scope="module")
def w3(ctx):
    specs = B.workload3(n_pipes=12, scale=0.01, seed=4242)
    w = B.Workload3(ctx, specs, lds_cap=4096)
    yield w
    w.close()


def test_plain_form_matches_oracle(w3):
    w3.run_plain()
    for s, got in zip(w3.specs, w3.results()):
        assert same(got, expected(s)), s.index


@pytest.mark.parametrize("block,chunk", [(256, 4096), (64, 1000), (1024, 1 << 20)])
def test_batched_eager_and_graph_match_plain(w3, block, chunk):
    w3.run_plain()
    plain = w3.results()
    b = w3.build_batch(block, chunk)
    try:
        b.run(graph=False)
        for got, exp in zip(w3.results(), plain):
            assert same(got, exp)
        for _ in range(3):   # graph replays, with eager host work in between
            b.run(graph=True)
            for got, exp in zip(w3.results(), plain):
                assert same(got, exp)
    finally:
        b.close()


def test_autotune_then_batch(w3):
    chosen = w3.autotune(blocks=(64, 256, 1024), items=(1, 4, 16), lds_bits=(8, 12), trials=3)
    assert chosen and all(c["block"] in (64, 256, 1024) and c["items"] in (1, 4, 16) for c in chosen.values())
    w3.run_plain()
    plain = w3.results()
    for s, got in zip(w3.specs, plain):
        assert same(got, expected(s))
    b = w3.build_batch(256, 8192)
    try:
        b.run(graph=True)
        for got, exp in zip(w3.results(), plain):
            assert same(got, exp)
    finally:
        b.close()


@pytest.mark.parametrize("fold_rows", [4096, 12288])
def test_fold_level_work_item_size(w3, fold_rows):
    """Rows per work item of the partial-table fold level set apart from the
    scan level's: more (smaller) fold items, the same results."""
    w3.run_plain()
    plain = w3.results()
    base = w3.build_batch(256, 8192)
    n0 = base.work_items
    base.close()
    b = w3.build_batch(256, 8192, fold_rows=fold_rows)
    try:
        assert b.work_items >= n0 if fold_rows < 8192 else b.work_items <= n0
        for _ in range(2):
            b.run(graph=True)
            for got, exp in zip(w3.results(), plain):
                assert same(got, exp)
    finally:
        b.close()


def test_batch_tuning_keeps_plain_settings_and_results(w3):
    """tune_batch picks the batched bodies' settings on the batched replay;
    the plain form keeps its own, and both forms still equal the oracle."""
    plain_launch = dict(w3.launch)
    tb = w3.tune_batch(block=256, chunk_rows=8192, items=(2, 4), lds_bits=(8, 12), runs=1)
    assert w3.launch == plain_launch
    assert tb["chosen"]["agg"]["items"] in (2, 4) and tb["chosen"]["agg"]["lds_bits"] in (8, 12)
    assert tb["chosen"]["join"]["items"] in (2, 4)
    assert len(tb["trace"]) == 4 + 2
    for i, s in enumerate(w3.specs):
        kind = "agg" if s.kind == "agg" else "join"
        assert w3.launch_for(i, True)[1] == tb["chosen"][kind]["items"]
    w3.run_plain()
    plain = w3.results()
    for s, got in zip(w3.specs, plain):
        assert same(got, expected(s))
    b = w3.build_batch(256, 8192)
    try:
        for _ in range(2):
            b.run(graph=True)
            for got, exp in zip(w3.results(), plain):
                assert same(got, exp)
    finally:
        b.close()
        w3.batch_launch = {}


@pytest.mark.parametrize("partials,bits", [(0, None), (256, None), (2, 6)])
def test_partial_tables_and_direct_merges_agree(ctx, partials, bits):
    """The aggregate pipelines' LDS tables reach the group table either as
    partial tables folded by the merge phase (at most `partials` per
    pipeline; the blocks past them merge directly) or merged by each block;
    with 64-slot LDS tables most rows also spill to the group table. Every
    combination equals the oracle, in the plain and the batched form."""
    specs = [s for s in B.workload3(n_pipes=12, scale=0.01, seed=77) if s.kind == "agg"]
    w = B.Workload3(ctx, specs, lds_cap=4096, partials=partials)
    try:
        if bits is not None:
            for s in specs:
                w.launch[s.index] = (256, 4, bits)
        w.run_plain()
        plain = w.results()
        for s, got in zip(specs, plain):
            assert same(got, expected(s)), s.index
        b = w.build_batch(256, 4096)
        try:
            for _ in range(2):
                b.run(graph=True)
                for got, exp in zip(w.results(), plain):
                    assert same(got, exp)
        finally:
            b.close()
    finally:
        w.close()
