"""Pinning the CPU oracle (runs without a GPU).

The reference holds no fixtures for these synthetic workloads (its golden
CSVs need TPC-H data), so each oracle function is checked against what the
mathematics fixes: closed forms on constructed tables, invariants, library
routines (numpy reductions, Python sets/dicts/sorted, re) on small random
inputs, and hand cases for the reference's encodings. Each check is chosen so
that a plausible mistake (wrong bucket, dropped term, sign, off-by-one on a
date bound, wrong tie-break) fails it.
"""

import re

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import workloads as W

rng = np.random.default_rng(12345)


def rand_f(n, seed):
    r = np.random.default_rng(seed)
    return {
        "f_qty": r.integers(1, 51, n).astype(np.int32),
        "f_price": r.integers(90_000, 10_500_001, n).astype(np.int64),
        "f_disc": r.integers(0, 11, n).astype(np.int32),
        "f_tax": r.integers(0, 9, n).astype(np.int32),
        "f_flag": r.choice(np.frombuffer(b"ARN", np.uint8), n),
        "f_status": r.choice(np.frombuffer(b"FO", np.uint8), n),
        "f_date": r.integers(0, 2526, n).astype(np.int32),
    }


# ------------------------------------------------------------------ Q1 ----

def test_q1_bucket_encoding_hand_cases():
    # metal_plan_builder.cpp:335: flag A->0, N->2, other->4; status F->0, other->1
    cases = {(b"A", b"F"): 0, (b"A", b"O"): 1, (b"N", b"F"): 2, (b"N", b"O"): 3, (b"R", b"F"): 4,
             (b"R", b"O"): 5, (b"X", b"Z"): 5}
    for (fl, st), bucket in cases.items():
        F = {"f_qty": np.array([7], np.int32), "f_price": np.array([100], np.int64),
             "f_disc": np.array([0], np.int32), "f_tax": np.array([0], np.int32),
             "f_flag": np.frombuffer(fl, np.uint8).copy(), "f_status": np.frombuffer(st, np.uint8).copy(),
             "f_date": np.array([0], np.int32)}
        out = oracle.q1(F, 0)
        assert out[bucket][6] == 1 and out[bucket][0] == 7
        assert out[:, 6].sum() == 1


def test_q1_closed_form_single_group_all_pass():
    n = 10_000
    F = rand_f(n, 1)
    F["f_flag"][:] = ord("N")
    F["f_status"][:] = ord("O")
    out = oracle.q1(F, 10_000)
    w = out[3]
    p = F["f_price"].astype(object)
    d = F["f_disc"].astype(object)
    t = F["f_tax"].astype(object)
    assert w[0] == int(F["f_qty"].astype(np.int64).sum())
    assert w[1] == int(F["f_price"].sum())
    assert w[2] == int(np.sum(p * (100 - d)))
    assert oracle.i128(w[3], w[4]) == int(np.sum(p * (100 - d) * (100 + t)))
    assert w[5] == int(F["f_disc"].astype(np.int64).sum())
    assert w[6] == n
    assert out[[0, 1, 2, 4, 5]].sum() == 0


def test_q1_constant_columns_closed_form():
    n = 12_345
    F = {"f_qty": np.full(n, 3, np.int32), "f_price": np.full(n, 1_000_000, np.int64),
         "f_disc": np.full(n, 4, np.int32), "f_tax": np.full(n, 2, np.int32),
         "f_flag": np.full(n, ord("A"), np.uint8), "f_status": np.full(n, ord("F"), np.uint8),
         "f_date": np.full(n, 5, np.int32)}
    w = oracle.q1(F, 5)[0]
    assert list(w[:3]) == [3 * n, 1_000_000 * n, 1_000_000 * 96 * n]
    assert oracle.i128(w[3], w[4]) == 1_000_000 * 96 * 102 * n
    assert w[5] == 4 * n and w[6] == n
    assert oracle.q1(F, 4)[0][6] == 0  # date 5 > cutoff 4: the bound is inclusive only at equality


@pytest.mark.parametrize("seed", [2, 3, 4])
def test_q1_vs_python_bigint_bruteforce(seed):
    F = rand_f(700, seed)
    cutoff = 1800
    exp = [[0] * 6 for _ in range(6)]
    for i in range(700):
        if int(F["f_date"][i]) > cutoff:
            continue
        fl, st = chr(F["f_flag"][i]), chr(F["f_status"][i])
        b = {"A": 0, "N": 2}.get(fl, 4) + (0 if st == "F" else 1)
        p, d, t = int(F["f_price"][i]), int(F["f_disc"][i]), int(F["f_tax"][i])
        e = exp[b]
        e[0] += int(F["f_qty"][i]); e[1] += p; e[2] += p * (100 - d); e[3] += p * (100 - d) * (100 + t)
        e[4] += d; e[5] += 1
    out = oracle.q1(F, cutoff)
    for b in range(6):
        w = out[b]
        assert [int(w[0]), int(w[1]), int(w[2]), oracle.i128(w[3], w[4]), int(w[5]), int(w[6])] == exp[b]


def test_q1_int128_adversarial():
    n = 1000
    F = {"f_qty": np.full(n, 50, np.int32), "f_price": np.full(n, 9_000_000_000_000, np.int64),
         "f_disc": np.zeros(n, np.int32), "f_tax": np.full(n, 8, np.int32),
         "f_flag": np.full(n, ord("R"), np.uint8), "f_status": np.full(n, ord("O"), np.uint8),
         "f_date": np.zeros(n, np.int32)}
    w = oracle.q1(F, 0)[5]
    assert oracle.i128(w[3], w[4]) == 9_000_000_000_000 * 100 * 108 * n  # 9.72e22 > 2^63
    F["f_price"] = -F["f_price"]
    w = oracle.q1(F, 0)[5]
    assert oracle.i128(w[3], w[4]) == -9_000_000_000_000 * 100 * 108 * n


def test_q1_multithreaded_equals_single():
    F = rand_f(200_000, 9)
    assert np.array_equal(oracle.q1(F, 2000), oracle.q1(F, 2000, threads=7))


def test_q6_vs_numpy():
    F = rand_f(50_000, 11)
    m = (F["f_date"] >= 100) & (F["f_date"] < 465) & (F["f_disc"] >= 5) & (F["f_disc"] <= 7) & (F["f_qty"] < 24)
    rev, cnt = oracle.q6(F, 100, 465, 5, 7, 24)
    assert cnt == int(m.sum())
    assert rev == int((F["f_price"][m] * F["f_disc"][m].astype(np.int64)).sum())


# ------------------------------------------------------------------ Q3 ----

def small_join(seed, n_a=200, n_b=2000, cmax=7):
    r = np.random.default_rng(seed)
    A = {"a_key": (r.permutation(n_a) + 1).astype(np.int32), "a_seg": r.integers(0, 5, n_a).astype(np.uint8)}
    B = {"b_key": (r.permutation(n_b) + 1).astype(np.int32), "b_fkey": r.integers(1, n_a + 1, n_b).astype(np.int32),
         "b_date": r.integers(0, 100, n_b).astype(np.int32)}
    counts = r.integers(1, cmax + 1, n_b)
    fk = np.repeat(np.arange(1, n_b + 1), counts).astype(np.int32)
    r.shuffle(fk)
    n_f = fk.shape[0]
    F = {"f_fkey": fk, "f_price": r.integers(90_000, 10_500_001, n_f).astype(np.int64),
         "f_disc": r.integers(0, 11, n_f).astype(np.int32), "f_date": r.integers(0, 100, n_f).astype(np.int32)}
    return A, B, F


def q3_python(A, B, F, seg, d):
    aset = {int(k) for k, s in zip(A["a_key"], A["a_seg"]) if int(s) == seg}
    bq = {int(k): int(dt) for k, fk, dt in zip(B["b_key"], B["b_fkey"], B["b_date"]) if int(dt) < d and int(fk) in aset}
    rev = {}
    for fk, p, dc, dt in zip(F["f_fkey"], F["f_price"], F["f_disc"], F["f_date"]):
        if int(dt) > d and int(fk) in bq:
            rev[int(fk)] = rev.get(int(fk), 0) + int(p) * (100 - int(dc))
    groups = sorted((k, v, bq[k]) for k, v in rev.items() if v > 0)
    return groups


@pytest.mark.parametrize("seed,seg,d", [(1, 2, 50), (2, 0, 30), (3, 4, 99), (4, 1, 0)])
def test_q3_groups_vs_python_join(seed, seg, d):
    A, B, F = small_join(seed)
    exp = q3_python(A, B, F, seg, d)
    got = oracle.q3(A, B, F, seg, d)
    assert list(zip(got["key"].tolist(), got["rev"].tolist(), got["date"].tolist())) == exp
    got_mt = oracle.q3(A, B, F, seg, d, threads=5)
    assert np.array_equal(got["rev"], got_mt["rev"])


@pytest.mark.parametrize("seed", [5, 6])
def test_q3_topk_is_sorted_prefix(seed):
    A, B, F = small_join(seed)
    groups = q3_python(A, B, F, 2, 60)
    exp = sorted(groups, key=lambda g: (-g[1], g[2], g[0]))[:10]
    got = oracle.q3_topk(A, B, F, 2, 60, k=10)
    assert list(zip(got["b_key"].tolist(), got["revenue"].tolist(), got["b_date"].tolist())) == exp


def test_q3_topk_ties_break_by_date_then_key():
    # three B keys with identical revenue: order is date asc, then key asc
    A = {"a_key": np.array([1], np.int32), "a_seg": np.array([0], np.uint8)}
    B = {"b_key": np.array([10, 11, 12, 13], np.int32), "b_fkey": np.ones(4, np.int32),
         "b_date": np.array([5, 3, 3, 1], np.int32)}
    F = {"f_fkey": np.array([10, 11, 12, 13], np.int32), "f_price": np.array([100, 100, 100, 50], np.int64),
         "f_disc": np.zeros(4, np.int32), "f_date": np.full(4, 9, np.int32)}
    got = oracle.q3_topk(A, B, F, 0, 8, k=3)
    assert got["b_key"].tolist() == [11, 12, 10]


# ------------------------------------------------------------- group-by ----

@pytest.mark.parametrize("seed,cut", [(1, 1 << 62), (2, 500_000), (3, 10)])
def test_groupby_vs_numpy(seed, cut):
    r = np.random.default_rng(seed)
    n = 20_000
    G = {"g_key": r.integers(1, 3000, n).astype(np.int32), "g_val": r.integers(-1000, 1_000_001, n).astype(np.int64)}
    got = oracle.groupby(G, cut)
    m = G["g_val"] < cut
    k, v = G["g_key"][m], G["g_val"][m]
    order = np.argsort(k, kind="stable")
    k, v = k[order], v[order]
    uk, start = np.unique(k, return_index=True)
    assert np.array_equal(got["key"], uk)
    assert np.array_equal(got["count"], np.diff(np.append(start, k.shape[0])))
    assert np.array_equal(got["sum"], np.add.reduceat(v, start)) if uk.size else got["sum"].size == 0
    if uk.size:
        assert np.array_equal(got["min"], np.minimum.reduceat(v, start))
        assert np.array_equal(got["max"], np.maximum.reduceat(v, start))
    assert oracle.distinct(G, cut) == uk.size
    mt = oracle.groupby(G, cut, threads=6)
    assert all(np.array_equal(got[c], mt[c]) for c in got)


def test_groupby_negative_values_min_max_signed():
    G = {"g_key": np.array([5, 5, 5, 7], np.int32), "g_val": np.array([-3, 2, -9, 4], np.int64)}
    got = oracle.groupby(G)
    assert got["min"].tolist() == [-9, 4] and got["max"].tolist() == [2, 4] and got["sum"].tolist() == [-10, 4]


# ----------------------------------------------------------------- LIKE ----

LIKE_CASES = [
    (b"special requests", b"special", b"requests", True),
    (b"requests special", b"special", b"requests", False),   # w2 before w1
    (b"specialrequests", b"special", b"requests", True),     # adjacency
    (b"specialrequest", b"special", b"requests", False),     # w2 cut at the end
    (b"aaa", b"aa", b"a", True),                              # overlap resolved by leftmost w1
    (b"aa", b"aa", b"a", False),
    (b"abab", b"ab", b"ab", True),                            # w1 == w2
    (b"ab", b"ab", b"ab", False),
    (b"xx\0ab cd", b"ab", b"cd", False),                      # NUL ends the value
    (b"ab" + b" " * 21 + b"cd", b"ab", b"cd", True),          # w2 at the last bytes
    (b"", b"a", b"b", False),
    (b"anything", b"", b"", True),
]


@pytest.mark.parametrize("s,w1,w2,exp", LIKE_CASES)
def test_like_hand_cases(s, w1, w2, exp):
    assert oracle.like2(s, 25, w1, w2) == exp


def test_like_vs_regex_random():
    r = np.random.default_rng(7)
    alphabet = np.frombuffer(b"ab c", np.uint8)
    for _ in range(3000):
        ln = int(r.integers(0, 26))
        s = bytes(r.choice(alphabet, ln))
        w1 = bytes(r.choice(alphabet[:2], int(r.integers(1, 4))))
        w2 = bytes(r.choice(alphabet[:2], int(r.integers(1, 4))))
        val = s.split(b"\0")[0][:25]
        exp = re.fullmatch(b".*" + re.escape(w1) + b".*" + re.escape(w2) + b".*", val, re.S) is not None
        assert oracle.like2(s, 25, w1, w2) == exp, (s, w1, w2)


def test_q13_vs_python_counter():
    r = np.random.default_rng(8)
    vocab = [b"ab", b"cd", b"abc", b"x", b"cab"]
    n_s, n_a = 3000, 400
    texts = np.zeros((n_s, 25), np.uint8)
    for i in range(n_s):
        words = [vocab[j] for j in r.integers(0, len(vocab), int(r.integers(1, 6)))]
        t = b" ".join(words)[:25]
        texts[i, :len(t)] = np.frombuffer(t, np.uint8)
    S = {"s_fkey": r.integers(1, n_a + 1, n_s).astype(np.int32), "s_text": texts}
    A = {"a_key": (r.permutation(n_a) + 1).astype(np.int32)}
    for neg in (False, True):
        got = oracle.q13(S, A, n_a, b"ab", b"cd", negate=neg)
        counts = {}
        for fk, row in zip(S["s_fkey"], texts):
            val = bytes(row).split(b"\0")[0]
            m = re.fullmatch(b".*ab.*cd.*", val, re.S) is not None
            if m != neg:
                counts[int(fk)] = counts.get(int(fk), 0) + 1
        hist = {}
        for k in A["a_key"]:
            c = counts.get(int(k), 0)
            hist[c] = hist.get(c, 0) + 1
        exp = sorted(hist.items(), key=lambda kv: (-kv[1], -kv[0]))
        assert list(zip(got["c_count"].tolist(), got["custdist"].tolist())) == exp


def test_q22_vs_python_sets():
    r = np.random.default_rng(9)
    n_a, n_b = 500, 3000
    A = {"a_key": (r.permutation(n_a) + 1).astype(np.int32), "a_seg": r.integers(0, 5, n_a).astype(np.uint8)}
    B = {"b_fkey": r.integers(1, n_a + 1, n_b).astype(np.int32), "b_date": r.integers(0, 100, n_b).astype(np.int32)}
    cnt, sm = oracle.q22(B, A, n_a, 20, 40)
    bset = {int(k) for k, d in zip(B["b_fkey"], B["b_date"]) if 20 <= int(d) < 40}
    ec, es = [0] * 5, [0] * 5
    for k, s in zip(A["a_key"], A["a_seg"]):
        if int(k) not in bset:
            ec[int(s)] += 1
            es[int(s)] += int(k)
    assert cnt.tolist() == ec and sm.tolist() == es


# ----------------------------------------------------------- generators ----

def test_generators_counter_based_and_deterministic():
    s = W.seed_for(0)
    a = W.uniform_i32(s, 1, 100_000, 1, 50)
    b = W.uniform_i32(s, 1, 1000, 1, 50)
    assert np.array_equal(a[:1000], b)
    assert a.min() >= 1 and a.max() <= 50
    assert np.array_equal(W.uniform_i64(s, 2, 5000, -7, 7), W.uniform_i64(s, 2, 5000, -7, 7))


def test_permutation_is_bijection():
    for n in [1, 2, 3, 1000, 70_001]:
        p = W.permutation_i32(42, 3, n, 1)
        assert np.array_equal(np.sort(p), np.arange(1, n + 1))


def test_fanout_counts_in_range():
    fk = W.fanout_fkey_i32(5, 6, 5000, 1, 7)
    c = np.bincount(fk, minlength=5001)[1:]
    assert c.min() >= 1 and c.max() <= 7 and c.sum() == fk.shape[0]


def test_zipf_head_frequency():
    n, N, s = 400_000, 1000, 1.1
    z = W.zipf_i32(77, 1, n, s, N)
    assert z.min() >= 1 and z.max() <= N
    h = sum(k ** -s for k in range(1, N + 1))
    for k in (1, 2, 10):
        p = k ** -s / h
        assert abs((z == k).mean() - p) < 5 * np.sqrt(p * (1 - p) / n)


def test_text_is_words_from_vocabulary():
    vocab = W.vocabulary(3, 50)
    t = W.text(3, 1, 200, 25, vocab)
    vs = set(vocab)
    for row in t:
        val = bytes(row).split(b"\0")[0]
        assert 0 < len(val) <= 25
        assert all(w in vs for w in val.split(b" "))
