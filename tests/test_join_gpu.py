"""Workload 1 (semi-join bitmap -> hash build -> probe -> per-key revenue ->
top-10; reference Q3 with a real hash join) on the GPU vs the oracle.
Integer results are compared exactly: the full set of groups held in the hash
table and the top-10 under (revenue desc, b_date asc, b_key asc)."""

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import queries as Q
from gpudbmetalcodegen_amd import workloads as W
from tests.test_oracle import small_join

pytestmark = pytest.mark.gpu


def _check(ctx, A, B, F, seg, d, options=""):
    dev = Q.upload_tables(ctx, {"A": A, "B": B, "F": F})
    try:
        max_a, max_b = int(A["a_key"].max()), int(B["b_key"].max())
        if "sizing=given" in options:
            run = Q.run_q3(ctx, dev, seg, d, options, max_a, max_b, b_slots=Q.next_pow2(2 * B["b_key"].shape[0]))
        else:
            run = Q.run_q3(ctx, dev, seg, d, options, max_a, max_b)
        top = {k: [int(x) for x in v] for k, v in run.result.columns.items()}
        exp_top = oracle.q3_topk(A, B, F, seg, d, 10)
        assert top == {k: v.tolist() for k, v in exp_top.items()}
        groups = Q.q3_hash_groups(run)
        exp = oracle.q3(A, B, F, seg, d)
        for c in ("key", "rev", "date"):
            assert np.array_equal(groups[c], exp[c]), c
        st = run.timing.status
        assert all(s[0] == 0 and s[1] == 0 for s in st)  # no overflow, no duplicate key
        return run
    finally:
        for t in dev.values():
            t.free(ctx)


@pytest.mark.parametrize("seed,seg,d", [(1, 2, 50), (2, 0, 30), (3, 4, 99), (4, 1, 0)])
def test_q3_small(ctx, seed, seg, d):
    A, B, F = small_join(seed)
    _check(ctx, A, B, F, seg, d)


@pytest.mark.parametrize("options", ["probe_bitmap=0", "sizing=given", "items=8", "staged=0", "staged=1", "prefetch=1",
                                     "late=0", "late=0,items=8", "late=0,staged=1", "late=0,staged=1,probe_bitmap=0",
                                     "late=0,prefetch=1", "late=0,prefetch=1,items=8", "late=0,compact=1",
                                     "late=0,compact=1,items=8", "late=0,compact=1,probe_bitmap=0", "late=1,items=8",
                                     "late=1,probe_bitmap=0", "late=1,sizing=given,block=128", "late=1,compact=0",
                                     "late=1,compact=0,items=8", "bflags=1", "bflags=1,sizing=given",
                                     "a_lds_words=32", "a_lds_words=32768", "a_lds_words=64,block_b=256,block_count=512",
                                     "a_lds_words=32,sizing=given"])
def test_q3_plan_variants(ctx, options):
    A, B, F = small_join(11, n_a=3000, n_b=40_000)
    _check(ctx, A, B, F, 2, 60, options)


def test_q3_empty_build_side(ctx):
    A, B, F = small_join(12)
    run = _check(ctx, A, B, F, 9, 50)   # no A row has segment 9
    assert run.result.rows() == 0


def test_q3_medium_generated(ctx):
    T = W.workload1(n_a=150_000, n_b=1_500_000)
    _check(ctx, T["A"], T["B"], T["F"], W.W1_SEG, W.W1_DATE)


@pytest.mark.slow
def test_q3_full_size(ctx):
    T = W.workload1()
    run = _check(ctx, T["A"], T["B"], T["F"], W.W1_SEG, W.W1_DATE)
    print("Q3 full timing", run.timing.total_ms, run.timing.phases)
