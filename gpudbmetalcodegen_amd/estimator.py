"""Learned group-count estimator for hash group-by sizing (Python front end
of the C++ CardinalityEstimator, csrc/runtime/estimator.*).

Policy (``size_for``): slots = next_pow2(2 * predicted_groups * margin); the
LDS variant when the prediction fits a block's LDS table or the sample shows
heavy hitters, the global variant otherwise. The exact pre-pass with
resize-on-overflow remains the fallback and the judge of the estimate.

The committed weights (``estimator_weights.bin``) were produced by
``python -m gpudbmetalcodegen_amd.train_estimator`` on an MI355X: training
runs logged by the engine (device-generated Zipf keys, exact counts from the
GB_COUNT plan), training by the project's HIP forward/backward/Adam kernels.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native as N
from . import engine as E

NFEAT = 16
NPARAMS = 5313
WEIGHTS = Path(__file__).resolve().parent / "estimator_weights.bin"
SAMPLE = 8192


def _e():
    return N.load()[1]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def features(n_rows: int, key_min: int, key_max: int, sample_keys: np.ndarray, sample_vals: np.ndarray | None = None,
             val_cut: int = (1 << 63) - 1) -> tuple[np.ndarray, float]:
    k = np.ascontiguousarray(sample_keys, np.int32)
    v = None if sample_vals is None else np.ascontiguousarray(sample_vals, np.int64)
    out = np.zeros(NFEAT, np.float32)
    ub = C.c_double()
    N.eng_check(_e().gdb_est_features(n_rows, key_min, key_max, val_cut, _p(k), None if v is None else _p(v),
                                      k.shape[0], _p(out), C.byref(ub)))
    return out, ub.value


def device_sample(ctx: E.Context, dkeys: int, dvals: int | None, n: int, m: int = SAMPLE):
    m = min(m, n)
    hk = np.zeros(m, np.int32)
    hv = np.zeros(m, np.int64)
    N.eng_check(_e().gdb_est_sample(ctx.handle, C.c_void_p(dkeys), C.c_void_p(dvals) if dvals else None, n, m,
                                    _p(hk), _p(hv)))
    return hk, (hv if dvals else None)


def collect_runs(ctx: E.Context, nruns: int, seed: int, nmin=1 << 14, nmax=1 << 22, kmin=1 << 6, kmax=1 << 25,
                 smax=1.5, selmin=0.01) -> dict:
    X = np.zeros((nruns, NFEAT), np.float32)
    ub = np.zeros(nruns, np.float64)
    g = np.zeros(nruns, np.float64)
    spec = np.zeros((nruns, 4), np.float64)
    N.eng_check(_e().gdb_est_collect_runs(ctx.handle, nruns, seed, nmin, nmax, kmin, kmax, smax, selmin, _p(X),
                                          _p(ub), _p(g), _p(spec)))
    return {"X": X, "ub": ub, "groups": g, "spec": spec}


def target(groups: np.ndarray, ub: np.ndarray) -> np.ndarray:
    return (np.log2(np.maximum(1.0, groups)) - np.log2(np.maximum(1.0, ub))).astype(np.float32)


class CardinalityEstimator:
    def __init__(self, ctx: E.Context):
        self.ctx = ctx
        self._h = C.c_void_p()
        N.eng_check(_e().gdb_est_create(ctx.handle, C.byref(self._h)))

    def init(self, seed: int) -> None:
        N.eng_check(_e().gdb_est_init(self._h, seed))

    def train(self, X: np.ndarray, y: np.ndarray, epochs: int = 300, batch: int = 256, lr: float = 1e-3,
              seed: int = 1) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.zeros(epochs, np.float32)
        N.eng_check(_e().gdb_est_train(self._h, _p(X), _p(y), X.shape[0], epochs, batch, lr, seed, _p(losses)))
        return losses

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(np.atleast_2d(X), np.float32)
        out = np.zeros(X.shape[0], np.float32)
        N.eng_check(_e().gdb_est_predict(self._h, _p(X), X.shape[0], _p(out)))
        return out

    def predict_groups(self, X: np.ndarray, ub) -> np.ndarray:
        t = self.predict(X).astype(np.float64)
        return np.maximum(1.0, np.exp2(t) * np.maximum(1.0, np.asarray(ub, np.float64)))

    def save(self, path=WEIGHTS) -> None:
        N.eng_check(_e().gdb_est_save(self._h, str(path).encode()))

    def load(self, path=WEIGHTS) -> "CardinalityEstimator":
        N.eng_check(_e().gdb_est_load(self._h, str(path).encode()))
        return self

    def state(self):
        p = np.zeros(NPARAMS, np.float32)
        m = np.zeros(NFEAT, np.float32)
        s = np.zeros(NFEAT, np.float32)
        N.eng_check(_e().gdb_est_get_state(self._h, _p(p), _p(m), _p(s)))
        return p, m, s

    def set_state(self, p, m, s) -> None:
        p = np.ascontiguousarray(p, np.float32)
        m = np.ascontiguousarray(m, np.float32)
        s = np.ascontiguousarray(s, np.float32)
        N.eng_check(_e().gdb_est_set_state(self._h, _p(p), _p(m), _p(s)))

    def close(self):
        if self._h:
            _e().gdb_est_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Sizing:
    groups_pred: float
    slots: int
    variant: str
    lds_slots: int
    features: np.ndarray
    upper_bound: float


def next_pow2(x: float) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


# Tables above this many slots (2 MiB of 32-byte slots) use the partitioned
# group-by: below it the table stays cache-resident and per-row atomics on it
# are cheap.
PART_MIN_SLOTS = 65536


def size_for(est: CardinalityEstimator, X: np.ndarray, ub: float, margin: float = 1.25,
             lds_slots: int = 2048) -> Sizing:
    g = float(est.predict_groups(X, [ub])[0])
    slots = max(64, next_pow2(2.0 * g * margin))
    top8 = float(X[8])  # sample mass of the 8 most frequent keys
    if slots > PART_MIN_SLOTS:
        # table far beyond the caches: two-pass partitioned form (no per-row
        # atomics on the table); a large LDS pass-1 table when the sample
        # shows heavy hitters
        return Sizing(g, slots, "part", 4096 if top8 >= 0.02 else 1024, X, ub)
    variant = "lds" if (g <= lds_slots // 2 or top8 >= 0.02) else "global"
    return Sizing(g, slots, variant, lds_slots, X, ub)


def run_groupby_learned(ctx: E.Context, dev: dict, est: CardinalityEstimator, key_min: int, key_max: int,
                        val_cut: int | None = None, executor: E.HipGenericExecutor | None = None,
                        collect: bool = True, slots_scale: float = 1.0):
    """The learned-sizing form of the hash group-by over table G: a device
    row sample -> 16 features -> MLP -> table slots and variant -> one
    group-by run (a reported overflow doubles the table and reruns, so the
    result stays exact whatever the estimate). slots_scale scales the
    predicted table (below 1: a deliberately low estimate, to exercise the
    resize path). Returns (GroupByRun, Sizing)."""
    from . import queries as Q

    t = dev["G"]
    kptr = t.columns["g_key"]
    vptr = t.columns["g_val"] if val_cut is not None else None
    hk, hv = device_sample(ctx, kptr, vptr, t.rows)
    cut = (1 << 63) - 1 if val_cut is None else val_cut
    x, ub = features(t.rows, key_min, key_max, hk, hv, cut)
    sz = size_for(est, x, ub)
    slots = max(64, int(sz.slots * slots_scale))
    run = Q.run_groupby(ctx, dev, slots=slots, variant=sz.variant, lds_slots=sz.lds_slots, val_cut=val_cut,
                        executor=executor, collect=collect)
    return run, sz
