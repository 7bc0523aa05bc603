"""Learned group-count estimator.

CPU: the oracle's analytic MLP gradient is pinned by central finite
differences (double precision); features are pinned on constructed samples.
GPU: the HIP forward/backward kernels match the oracle within FP32 tolerance
(stated below), training decreases the loss, and — with the committed
weights and held-out run log — the network's estimate beats the estimators
available without it (upper bound, GEE-style sample scaling) and sizes tables
that rarely need a resize, at a fraction of the upper-bound memory.
"""

import ctypes as C

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import _native as N
from gpudbmetalcodegen_amd import engine as E
from gpudbmetalcodegen_amd import estimator as EST

# FP32 device vs double-precision oracle: |gpu - cpu| <= RTOL * max|cpu| + ATOL
GRAD_RTOL, GRAD_ATOL = 2e-4, 1e-6
FWD_RTOL = 1e-4


def rand_params(seed):
    r = np.random.default_rng(seed)
    p = np.zeros(oracle.MLP_NPARAMS, np.float32)
    p[0:1024] = r.normal(0, np.sqrt(2 / 16), 1024)
    p[1024:1088] = r.normal(0, 0.1, 64)
    p[1088:5184] = r.normal(0, np.sqrt(2 / 64), 4096)
    p[5184:5248] = r.normal(0, 0.1, 64)
    p[5248:5312] = r.normal(0, np.sqrt(1 / 64), 64)
    p[5312] = 0.3
    return p


def test_oracle_gradient_matches_finite_differences():
    r = np.random.default_rng(3)
    P = rand_params(4).astype(np.float64)
    X = r.normal(0, 1, (7, 16)).astype(np.float32)
    T = r.normal(0, 1, 7).astype(np.float32)
    g, loss = oracle.mlp_grad(P.astype(np.float32), X, T, 1 / 7)
    # float32 parameters: perturb in float32 steps large enough to be exact
    idx = list(range(0, 1024, 97)) + list(range(1024, 1088, 13)) + list(range(1088, 5184, 401)) + \
        list(range(5184, 5248, 11)) + list(range(5248, 5312, 9)) + [5312]
    def fd(i, h):
        Pp = P.astype(np.float32).copy()
        Pm = Pp.copy()
        Pp[i] += h
        Pm[i] -= h
        hp = float(Pp[i]) - float(P[i])
        hm = float(P[i]) - float(Pm[i])
        lp = ((oracle.mlp_forward(Pp, X) - T) ** 2).sum() / 7
        lm = ((oracle.mlp_forward(Pm, X) - T) ** 2).sum() / 7
        return (lp - lm) / (hp + hm)

    for i in idx:
        # two step sizes: a step that straddles a ReLU kink spoils one of them
        ok = [abs(fd(i, h) - g[i]) <= 1e-6 * max(1.0, abs(g[i])) for h in (1e-4, 2e-5)]
        assert any(ok), (i, g[i], fd(i, 1e-4), fd(i, 2e-5))


def test_oracle_forward_closed_form_linear_region():
    # all weights of layers 1/2 identity-like and positive inputs: the net is linear
    P = np.zeros(oracle.MLP_NPARAMS, np.float32)
    for j in range(16):
        P[j * 16 + j] = 1.0
    for j in range(16):
        P[1088 + j * 64 + j] = 1.0
    P[5248:5264] = 2.0
    P[5312] = 0.5
    X = np.abs(np.random.default_rng(1).normal(0, 1, (5, 16))).astype(np.float32)
    assert np.allclose(oracle.mlp_forward(P, X), 2.0 * X.astype(np.float64).sum(1) + 0.5)


def test_features_constructed_sample():
    keys = np.array([1] * 50 + list(range(2, 52)), np.int32)   # one hot key, 50 singletons
    x, ub = EST.features(1000, 1, 1000, keys)
    assert ub == 1000.0
    assert abs(x[4] - 51 / 100) < 1e-6       # distinct / sample
    assert abs(x[5] - 50 / 100) < 1e-6       # singletons
    assert abs(x[7] - 50 / 100) < 1e-6       # top-1 frequency
    vals = np.arange(100, dtype=np.int64)
    x2, ub2 = EST.features(1000, 1, 1000, keys, vals, val_cut=50)
    assert abs(x2[2] - 0.5) < 1e-6 and ub2 == 500.0   # sample selectivity scales the bound


def test_committed_estimator_beats_sample_scaling_and_upper_bound():
    if not EST.WEIGHTS.exists():
        pytest.skip("estimator weights not trained yet")
    runs = np.load(EST.HERE / "estimator_runs.npz") if hasattr(EST, "HERE") else np.load(
        EST.WEIGHTS.parent / "estimator_runs.npz")
    X, ub, g = runs["heldout_X"], runs["heldout_ub"], runs["heldout_groups"]
    blob = np.fromfile(EST.WEIGHTS, np.uint8)
    mean = blob[16:80].view(np.float32)
    std = blob[80:144].view(np.float32)
    P = blob[144:].view(np.float32)
    t = oracle.mlp_forward(P, (X - mean) / std)
    pred = np.maximum(1.0, np.exp2(t) * np.maximum(1.0, ub))
    q_mlp = np.maximum(pred / g, g / pred)
    q_ub = np.maximum(ub / g, g / ub)
    gee = np.exp2(X[:, 9] * 32.0)            # the GEE-style sample estimate among the features
    gee = np.minimum(gee, ub)
    q_gee = np.maximum(gee / g, g / np.maximum(1, gee))
    assert np.median(q_mlp) < np.median(q_gee) and np.median(q_mlp) < np.median(q_ub)
    assert np.percentile(q_mlp, 90) < np.percentile(q_gee, 90)
    slots = np.array([EST.next_pow2(2 * p * 1.25) for p in pred])
    no_resize = (slots >= g).mean()          # the table holds every group without overflow
    assert no_resize >= 0.95
    ub_slots = np.array([EST.next_pow2(2 * u) for u in ub])
    exact_slots = np.array([EST.next_pow2(2 * x) for x in g])
    loose = ub >= 4 * g                      # runs where the catalog bound is far from the truth
    assert loose.sum() >= 20
    assert np.median(slots[loose] / ub_slots[loose]) <= 0.25    # memory saved against upper-bound sizing
    assert slots.sum() <= 1.3 * exact_slots.sum()              # close to exact pre-pass sizing


# ------------------------------------------------------------------ GPU ----

def _dev(ctx, a):
    return ctx.upload(np.ascontiguousarray(a))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 5, 64, 257])
def test_gpu_mlp_forward_and_gradient_match_oracle(ctx, n):
    hip, _ = N.load()
    r = np.random.default_rng(n)
    P = rand_params(n)
    X = r.normal(0, 1, (n, 16)).astype(np.float32)
    T = r.normal(0, 1, n).astype(np.float32)
    dP, dX, dT = _dev(ctx, P), _dev(ctx, X), _dev(ctx, T)
    dY = ctx.alloc(4 * n)
    dG = ctx.alloc(4 * n * oracle.MLP_NPARAMS)
    dL = ctx.alloc(4 * n)
    dg = ctx.alloc(4 * oracle.MLP_NPARAMS)
    try:
        N.hip_check(hip.gdb_mlp_forward(C.c_void_p(dP), C.c_void_p(dX), C.c_void_p(dY), n, ctx.stream))
        y = ctx.download(dY, 4 * n, np.float32)
        ref = oracle.mlp_forward(P, X)
        assert np.all(np.abs(y - ref) <= FWD_RTOL * np.maximum(1, np.abs(ref)))
        N.hip_check(hip.gdb_mlp_backward(C.c_void_p(dP), C.c_void_p(dX), C.c_void_p(dT), C.c_void_p(dG),
                                         C.c_void_p(dL), n, 1.0 / n, ctx.stream))
        N.hip_check(hip.gdb_mlp_reduce_adam(C.c_void_p(dP), None, None, C.c_void_p(dG), C.c_void_p(dg), n, 1,
                                            0.0, 0.9, 0.999, 1e-8, ctx.stream))
        g = ctx.download(dg, 4 * oracle.MLP_NPARAMS, np.float32).astype(np.float64)
        gref, _ = oracle.mlp_grad(P, X, T, 1.0 / n)
        assert np.max(np.abs(g - gref)) <= GRAD_RTOL * np.max(np.abs(gref)) + GRAD_ATOL
        # deterministic: a second run gives the same bits
        N.hip_check(hip.gdb_mlp_backward(C.c_void_p(dP), C.c_void_p(dX), C.c_void_p(dT), C.c_void_p(dG),
                                         C.c_void_p(dL), n, 1.0 / n, ctx.stream))
        N.hip_check(hip.gdb_mlp_reduce_adam(C.c_void_p(dP), None, None, C.c_void_p(dG), C.c_void_p(dg), n, 1,
                                            0.0, 0.9, 0.999, 1e-8, ctx.stream))
        assert np.array_equal(ctx.download(dg, 4 * oracle.MLP_NPARAMS, np.float32).astype(np.float64), g)
    finally:
        for p in (dP, dX, dT, dY, dG, dL, dg):
            ctx.free(p)


@pytest.mark.gpu
def test_gpu_training_reduces_loss_on_logged_runs(ctx):
    runs = EST.collect_runs(ctx, 300, 99, nmin=1 << 12, nmax=1 << 16, kmin=1 << 4, kmax=1 << 16)
    assert np.all(runs["groups"] >= 1)
    assert np.all(runs["groups"] <= runs["spec"][:, 2])  # never more groups than keys in the range
    y = EST.target(runs["groups"], runs["ub"])
    est = EST.CardinalityEstimator(ctx)
    est.init(3)
    losses = est.train(runs["X"], y, epochs=60, batch=64, lr=2e-3, seed=5)
    assert losses[-1] < 0.5 * losses[0]
    # GPU inference agrees with the double-precision oracle on the trained weights
    p, m, s = est.state()
    Xs = (runs["X"] - m) / s
    assert np.allclose(est.predict(runs["X"]), oracle.mlp_forward(p, Xs), rtol=1e-4, atol=1e-4)


def test_features_equal_a_direct_restatement():
    """The feature vector (counts of counts from an open-addressing table,
    entropy summed per distinct count) equals a direct numpy restatement —
    np.unique frequencies, per-key entropy — on random uniform and Zipf
    samples with and without a value filter."""
    import math

    def direct(n, kmax, keys, vals=None, cut=(1 << 63) - 1):
        S = max(1, len(keys))
        if vals is not None:
            keys = keys[vals < cut]
        s = len(keys)
        sel = s / S
        n_est = max(1.0, n * sel)
        ub = max(1.0, min(n_est, kmax))
        u, c = np.unique(keys, return_counts=True)
        d, f1, f2 = len(u), float((c == 1).sum()), float((c == 2).sum())
        cs = np.sort(c)[::-1]
        top1, top8 = (float(cs[0]) if d else 0.0), float(cs[:8].sum())
        ss = max(1.0, s)
        ent = float(-(c / ss * np.log2(c / ss)).sum()) if d else 0.0
        lg2 = lambda x: math.log2(max(1.0, x))
        gee = math.sqrt(n_est / ss) * f1 + (d - f1)
        chao = d + (f1 * f1 / (2 * f2) if f2 > 0 else f1 * (f1 - 1) / 2)
        x = [lg2(n) / 32, lg2(kmax) / 32, sel, lg2(n_est) / 32, d / ss, f1 / ss, f2 / ss, top1 / ss, top8 / ss,
             lg2(gee) / 32, lg2(chao) / 32, lg2(ub) / 32, lg2(d) / 32, ent / math.log2(ss) if ss > 1 else 0.0,
             min(1, max(-1, (lg2(n_est) - lg2(kmax)) / 10)), min(0, max(-2, (lg2(min(gee, ub)) - lg2(ub)) / 10))]
        return np.array(x, np.float32), ub

    r = np.random.default_rng(5)
    for t in range(60):
        m, K = int(r.integers(1, 9000)), int(r.integers(1, 100000))
        keys = ((r.zipf(1.1 + r.random(), m) % K + 1) if t % 2 else r.integers(1, K + 1, m)).astype(np.int32)
        vals = r.integers(0, 100, m).astype(np.int64) if t % 3 == 0 else None
        cut = int(r.integers(1, 100)) if vals is not None else (1 << 63) - 1
        x, ub = EST.features(10**7, 1, K, keys, vals, cut)
        xr, ubr = direct(10**7, K, keys, vals, cut)
        assert ub == ubr and np.array_equal(x, xr), t
