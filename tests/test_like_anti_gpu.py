"""Workload 4 plan types on the GPU vs the oracle: byte-exact LIKE '%w1%w2%'
per-key count + histogram (reference Q13 shape) and the bitmap anti-join
(reference Q22 shape)."""

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import queries as Q
from gpudbmetalcodegen_amd import workloads as W

pytestmark = pytest.mark.gpu


def _tables(n_s, n_a, seed):
    vocab = W.vocabulary(seed, 500)
    S = {"s_fkey": W.uniform_i32(seed, 51, n_s, 1, n_a), "s_text": W.text(seed, 52, n_s, 25, vocab)}
    A = {"a_key": W.permutation_i32(seed, 11, n_a, 1), "a_seg": W.uniform_i32(seed, 12, n_a, 0, 4).astype(np.uint8)}
    return S, A, vocab


@pytest.mark.parametrize("n_s,n_a", [(1, 10), (1000, 300), (100_001, 20_000)])
@pytest.mark.parametrize("negate", [False, True])
def test_q13_like(ctx, n_s, n_a, negate):
    S, A, vocab = _tables(n_s, n_a, 31 + n_s)
    dev = Q.upload_tables(ctx, {"S": S, "A": A})
    try:
        for w1, w2 in [(vocab[0], vocab[1]), (vocab[5][:2], vocab[5][1:3]), (b"a", b"e"), (vocab[7], vocab[7])]:
            run = Q.run_q13(ctx, dev, w1, w2, negate=negate)
            exp = oracle.q13(S, A, n_a, w1, w2, negate=negate)
            got = {k: [int(x) for x in v] for k, v in run.result.columns.items()}
            assert got == {k: v.tolist() for k, v in exp.items()}, (w1, w2)
            assert sum(got["custdist"]) == n_a
    finally:
        for t in dev.values():
            t.free(ctx)


def test_q13_like_edge_strings(ctx):
    rows = [b"special requests", b"requests special", b"specialrequests", b"specialrequest", b"aaa",
            b"xx\0ab cd", b"ab" + b" " * 21 + b"cd", b"", b"abab", b"ab cd ab cd ab cd ab cd a",
            b"abcdefghijklmnopqrstuvwxy", b"cd ab", b"a" * 25]
    texts = np.zeros((len(rows), 25), np.uint8)
    for i, r in enumerate(rows):
        texts[i, :min(25, len(r))] = np.frombuffer(r[:25], np.uint8)
    S = {"s_fkey": np.arange(1, len(rows) + 1, dtype=np.int32), "s_text": texts}
    A = {"a_key": np.arange(1, len(rows) + 1, dtype=np.int32), "a_seg": np.zeros(len(rows), np.uint8)}
    dev = Q.upload_tables(ctx, {"S": S, "A": A})
    try:
        for w1, w2 in [(b"special", b"requests"), (b"aa", b"a"), (b"ab", b"cd"), (b"ab", b"ab"), (b"", b""),
                       (b"", b"cd"), (b"ab", b""), (b"abcdefghijklmnopqrstuvwxy", b""), (b"a" * 24, b"a"),
                       (b"a" * 25, b"a"), (b"y", b"z")]:
            run = Q.run_q13(ctx, dev, w1, w2)
            exp = oracle.q13(S, A, len(rows), w1, w2)
            got = {k: [int(x) for x in v] for k, v in run.result.columns.items()}
            assert got == {k: v.tolist() for k, v in exp.items()}, (w1, w2)
    finally:
        for t in dev.values():
            t.free(ctx)


@pytest.mark.parametrize("n_b,n_a", [(0, 100), (5000, 1000), (1_000_000, 200_000)])
def test_q22_anti_join(ctx, n_b, n_a):
    s = 77 + n_b
    A = {"a_key": W.permutation_i32(s, 11, n_a, 1), "a_seg": W.uniform_i32(s, 12, n_a, 0, 4).astype(np.uint8)}
    B = {"b_key": W.permutation_i32(s, 21, max(n_b, 1), 1)[:n_b],
         "b_fkey": W.uniform_i32(s, 22, n_b, 1, n_a), "b_date": W.uniform_i32(s, 23, n_b, 0, 2405)}
    dev = Q.upload_tables(ctx, {"A": A, "B": B})
    try:
        for lo, hi in [(0, 2406), (100, 900), (5, 5)]:
            run = Q.run_q22(ctx, dev, lo, hi)
            cnt, sm = oracle.q22(B, A, n_a, lo, hi)
            cols = run.result.columns
            segs = [int(x) for x in cols["a_seg"]]
            exp_segs = [i for i in range(5) if cnt[i] > 0]
            assert segs == exp_segs
            assert [int(x) for x in cols["num_keys"]] == [int(cnt[i]) for i in exp_segs]
            assert [int(x) for x in cols["sum_keys"]] == [int(sm[i]) for i in exp_segs]
    finally:
        for t in dev.values():
            t.free(ctx)


def test_q22_bitmap_cache_collisions(ctx):
    """Referenced keys concentrated on bitmap words that share a few entries of
    the build kernel's per-block LDS word cache (word index mod 1024 in
    {0..3}), so lanes of one block replace each other's entries while others
    add bits to them: every referenced key must still be set in the bitmap."""
    rng = np.random.default_rng(2024)
    n_a, n_b = 4_000_000, 2_000_000
    A = {"a_key": np.arange(1, n_a + 1, dtype=np.int32), "a_seg": (np.arange(n_a) % 5).astype(np.uint8)}
    words = rng.integers(0, n_a // (32 * 1024), n_b) * 1024 + rng.integers(0, 4, n_b)
    fkey = words * 32 + rng.integers(0, 32, n_b)
    fkey = np.clip(fkey, 1, n_a).astype(np.int32)
    B = {"b_key": np.arange(1, n_b + 1, dtype=np.int32), "b_fkey": fkey,
         "b_date": rng.integers(0, 2406, n_b).astype(np.int32)}
    dev = Q.upload_tables(ctx, {"A": A, "B": B})
    try:
        for lo, hi in [(0, 2406), (300, 1700)]:
            run = Q.run_q22(ctx, dev, lo, hi)
            cnt, sm = oracle.q22(B, A, n_a, lo, hi)
            cols = run.result.columns
            exp_segs = [i for i in range(5) if cnt[i] > 0]
            assert [int(x) for x in cols["a_seg"]] == exp_segs
            assert [int(x) for x in cols["num_keys"]] == [int(cnt[i]) for i in exp_segs]
            assert [int(x) for x in cols["sum_keys"]] == [int(sm[i]) for i in exp_segs]
    finally:
        for t in dev.values():
            t.free(ctx)
