"""The allocation-free training step (gdb_mlp_train_step_noalloc) against the
existing entry points it stands for (gdb_mlp_backward with a caller
workspace, then gdb_mlp_reduce_adam): the same parameters, Adam moments,
summed gradient and per-sample losses, bit for bit, on two calls in a row
(the second with a smaller batch, so stale workspace rows must not leak in);
batches above the static workspace are rejected."""

import ctypes as C

import numpy as np
import pytest

import oracle
from gpudbmetalcodegen_amd import _native as N

pytestmark = pytest.mark.gpu

NP = oracle.MLP_NPARAMS
MAX_SAMPLES = 256


def _params(seed):
    r = np.random.default_rng(seed)
    return (r.normal(0, 0.2, NP)).astype(np.float32)


def _ptr(x):
    return C.c_void_p(x)


def test_noalloc_step_matches_workspace_step(ctx):
    hip, _ = N.load()
    P0 = _params(7)
    M0 = np.abs(_params(8)) * 1e-3
    V0 = np.abs(_params(9)) * 1e-4
    bufs = []

    def up(a):
        p = ctx.upload(np.ascontiguousarray(a))
        bufs.append(p)
        return p

    def alloc(nbytes):
        p = ctx.alloc(nbytes)
        bufs.append(p)
        return p

    try:
        # state A: existing entry points with a caller workspace; state B: the allocation-free step
        dPa, dMa, dVa = up(P0), up(M0), up(V0)
        dPb, dMb, dVb = up(P0), up(M0), up(V0)
        dG = alloc(4 * MAX_SAMPLES * NP)
        dga, dgb = alloc(4 * NP), alloc(4 * NP)
        for step, n in ((1, MAX_SAMPLES), (2, 37)):
            print(f"no-alloc training step {step}: {n} samples", flush=True)
            r = np.random.default_rng(100 + step)
            X = r.normal(0, 1, (n, 16)).astype(np.float32)
            T = r.normal(0, 1, n).astype(np.float32)
            dX, dT = up(X), up(T)
            dLa, dLb = alloc(4 * n), alloc(4 * n)
            lr, b1, b2, eps, scale = 1e-3, 0.9, 0.999, 1e-8, 1.0 / n
            N.hip_check(hip.gdb_mlp_backward(_ptr(dPa), _ptr(dX), _ptr(dT), _ptr(dG), _ptr(dLa), n, scale,
                                             ctx.stream))
            N.hip_check(hip.gdb_mlp_reduce_adam(_ptr(dPa), _ptr(dMa), _ptr(dVa), _ptr(dG), _ptr(dga), n, step, lr,
                                                b1, b2, eps, ctx.stream))
            st = hip.gdb_mlp_train_step_noalloc(_ptr(dPb), _ptr(dMb), _ptr(dVb), _ptr(dX), _ptr(dT), _ptr(dLb),
                                                _ptr(dgb), n, scale, step, lr, b1, b2, eps, ctx.stream)
            assert st == N.GDB_OK, hip.gdb_kernels_last_error()
            ctx.sync()
            for a, b, cnt in ((dPa, dPb, NP), (dMa, dMb, NP), (dVa, dVb, NP), (dga, dgb, NP), (dLa, dLb, n)):
                ga = ctx.download(a, 4 * cnt, np.float32)
                gb = ctx.download(b, 4 * cnt, np.float32)
                assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), (step, n)
            if step == 1:  # the first step's summed gradient against the double-precision oracle
                g = ctx.download(dgb, 4 * NP, np.float32).astype(np.float64)
                gref, _ = oracle.mlp_grad(P0, X, T, scale)
                assert np.max(np.abs(g - gref)) <= 2e-4 * np.max(np.abs(gref)) + 1e-6  # FP32 vs double
        # more samples than the workspace holds: rejected before any launch
        st = hip.gdb_mlp_train_step_noalloc(_ptr(dPb), _ptr(dMb), _ptr(dVb), _ptr(dX), _ptr(dT), _ptr(dLb),
                                            _ptr(dgb), MAX_SAMPLES + 1, 1.0, 3, 1e-3, 0.9, 0.999, 1e-8, ctx.stream)
        assert st != N.GDB_OK
        assert b"exceeds the static workspace" in hip.gdb_kernels_last_error()
    finally:
        ctx.sync()
        for p in bufs:
            ctx.free(p)
