"""Out-of-core (streamed) execution: the stream table stays in host memory
and is copied into HBM chunk by chunk (double-buffered, copy stream) while
the previous chunk's kernels run. Results must equal the resident run and
the oracle bit for bit, for ragged chunk sizes, pinned and pageable host
memory, every chunkable plan shape (scan-aggregate, join probe, LIKE
counter, anti-join bitmap, hash group-by); non-chunkable plans are refused.
"""

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import _native as N
from gpudbmetalcodegen_amd import engine as E
from gpudbmetalcodegen_amd import queries as Q
from gpudbmetalcodegen_amd import workloads as W
from tests.helpers import as_lists, q1_expected

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,chunk,pinned", [(100_003, 1000, True), (100_003, 4096, False), (100_003, 1 << 20, True),
                                            (1, 64, True), (0, 64, True), (2_000_000, 333_333, True)])
def test_q1_streamed_matches_oracle(ctx, n, chunk, pinned):
    F = W.fact_f(n, W.seed_for(0) + 3 * n + 1)
    cutoff = W.q1_cutoff()
    host = E.HostTable("F", F, pinned=pinned)
    try:
        run = Q.run_q1(ctx, {}, cutoff, stream=Q.Stream(host, chunk))
        assert as_lists(run.result.columns) == q1_expected(oracle.q1(F, cutoff))
        st = run.stream_stats
        assert st.chunks == max(1, -(-n // chunk)) and st.bytes_copied == host.nbytes()
        run2 = Q.run_q1(ctx, {}, cutoff, stream=Q.Stream(host, chunk), executor=run.executor)   # runner reuse
        assert as_lists(run2.result.columns) == as_lists(run.result.columns)
    finally:
        host.close()


def test_q3_streamed_fact_matches_resident(ctx):
    T = W.workload1(n_a=20_000, n_b=200_000)
    dev = Q.upload_tables(ctx, {"A": T["A"], "B": T["B"]})
    devF = Q.upload_tables(ctx, {"F": T["F"]})
    host = E.HostTable("F", T["F"])
    try:
        ref = Q.run_q3(ctx, {**dev, **devF})
        got = Q.run_q3(ctx, dev, stream=Q.Stream(host, 77_777))
        assert got.result.canonical == ref.result.canonical
        assert got.stream_stats.chunks == -(-T["F"]["f_fkey"].shape[0] // 77_777)
        exp = oracle.q3_topk(T["A"], T["B"], T["F"], W.W1_SEG, W.W1_DATE, 10)
        assert {k: [int(x) for x in v] for k, v in got.result.columns.items()} == {k: v.tolist() for k, v in exp.items()}
        groups, full = Q.q3_hash_groups(got), oracle.q3(T["A"], T["B"], T["F"], W.W1_SEG, W.W1_DATE)
        for c in ("key", "rev", "date"):
            assert np.array_equal(groups[c], full[c]), c
    finally:
        host.close()
        for t in list(dev.values()) + list(devF.values()):
            t.free(ctx)


def test_q22_streamed_build_side_matches_resident(ctx):
    rng = np.random.default_rng(5)
    n_a, n_b = 50_000, 400_000
    A = {"a_key": W.permutation_i32(9, 11, n_a, 1), "a_seg": W.uniform_i32(9, 12, n_a, 0, 4).astype(np.uint8)}
    B = {"b_fkey": W.uniform_i32(9, 22, n_b, 1, n_a), "b_date": W.uniform_i32(9, 23, n_b, 0, 2405)}
    dev = Q.upload_tables(ctx, {"A": A})
    devB = Q.upload_tables(ctx, {"B": B})
    host = E.HostTable("B", B)
    try:
        ref = Q.run_q22(ctx, {**dev, **devB}, 100, 400)
        got = Q.run_q22(ctx, dev, 100, 400, stream=Q.Stream(host, int(rng.integers(5_000, 90_000))))
        assert got.result.canonical == ref.result.canonical
    finally:
        host.close()
        for t in list(dev.values()) + list(devB.values()):
            t.free(ctx)


def test_q13_streamed_strings_match_oracle(ctx):
    vocab = W.vocabulary(41, 500)
    n_s, n_a = 150_001, 30_000
    S = {"s_fkey": W.uniform_i32(41, 51, n_s, 1, n_a), "s_text": W.text(41, 52, n_s, 25, vocab)}
    A = {"a_key": W.permutation_i32(41, 11, n_a, 1), "a_seg": W.uniform_i32(41, 12, n_a, 0, 4).astype(np.uint8)}
    dev = Q.upload_tables(ctx, {"A": A})
    host = E.HostTable("S", S)
    try:
        w1, w2 = vocab[3][:2], vocab[9][:2]
        got = Q.run_q13(ctx, dev, w1, w2, stream=Q.Stream(host, 10_007))
        exp = oracle.q13(S, A, n_a, w1, w2)
        assert {k: [int(x) for x in v] for k, v in got.result.columns.items()} == {k: v.tolist() for k, v in exp.items()}
    finally:
        host.close()
        dev["A"].free(ctx)


def test_groupby_streamed_matches_oracle(ctx):
    G = W.workload2(n=1_000_003, nkeys=200_000, seed=13)["G"]
    host = E.HostTable("G", G)
    try:
        got = Q.run_groupby(ctx, {}, slots=1 << 19, stream=Q.Stream(host, 123_457))
        exp = oracle.groupby(G)
        for c in exp:
            assert np.array_equal(got.result.columns[c], exp[c]), c
    finally:
        host.close()


def test_non_chunkable_plan_is_refused(ctx):
    G = W.workload2(n=10_000, nkeys=1000, seed=3)["G"]
    host = E.HostTable("G", G)
    try:
        assert not E.build_hip_plan("GROUPBY", "sizing=prepass").chunkable
        assert E.build_hip_plan("Q1").chunkable
        with pytest.raises(N.GdbError, match="not chunkable"):
            Q.run_groupby(ctx, {}, prepass=True, key_min=1, key_max=1000, stream=Q.Stream(host, 1000))
    finally:
        host.close()
