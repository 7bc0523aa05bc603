"""ctypes loader of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and the timed CPU baseline.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
    return _lib


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _cols(cols: dict, names):
    return [np.ascontiguousarray(cols[n]) for n in names]


def q1(F: dict, cutoff: int, threads: int = 0) -> np.ndarray:
    """Per-bucket [sum_qty, sum_base, sum_disc_price, charge_lo, charge_hi, sum_disc, count, 0] (6 x 8 int64)."""
    lib = load()
    qty, price, disc, tax, flag, status, date = _cols(F, ["f_qty", "f_price", "f_disc", "f_tax", "f_flag", "f_status", "f_date"])
    out = np.zeros(48, dtype=np.int64)
    n = qty.shape[0]
    args = [C.c_uint64(n), ptr(qty), ptr(price), ptr(disc), ptr(tax), ptr(flag), ptr(status), ptr(date),
            C.c_int64(cutoff), ptr(out)]
    if threads > 0:
        lib.oracle_q1_mt.restype = None
        lib.oracle_q1_mt(*args, C.c_int(threads))
    else:
        lib.oracle_q1.restype = None
        lib.oracle_q1(*args)
    return out.reshape(6, 8)


def q6(F: dict, date_lo, date_hi, disc_lo, disc_hi, qty_hi) -> tuple[int, int]:
    lib = load()
    qty, price, disc, date = _cols(F, ["f_qty", "f_price", "f_disc", "f_date"])
    out = np.zeros(2, dtype=np.int64)
    lib.oracle_q6.restype = None
    lib.oracle_q6(C.c_uint64(qty.shape[0]), ptr(qty), ptr(price), ptr(disc), ptr(date), C.c_int64(date_lo),
                  C.c_int64(date_hi), C.c_int64(disc_lo), C.c_int64(disc_hi), C.c_int64(qty_hi), ptr(out))
    return int(out[0]), int(out[1])


def i128(lo, hi) -> int:
    return int(np.uint64(np.int64(lo))) + (int(np.int64(hi)) << 64)


def _u64(x):
    return C.c_uint64(int(x))


def _i64(x):
    return C.c_int64(int(x))


def q3(A: dict, B: dict, F: dict, seg: int, d: int, threads: int = 0):
    """All groups (key asc): dict of numpy arrays key, rev, date."""
    lib = load()
    a_key, a_seg = _cols(A, ["a_key", "a_seg"])
    b_key, b_fkey, b_date = _cols(B, ["b_key", "b_fkey", "b_date"])
    f_fkey, f_price, f_disc, f_date = _cols(F, ["f_fkey", "f_price", "f_disc", "f_date"])
    cap = b_key.shape[0]
    k = np.zeros(cap, np.int64)
    r = np.zeros(cap, np.int64)
    dt = np.zeros(cap, np.int64)
    lib.oracle_q3.restype = C.c_int64
    g = lib.oracle_q3(_u64(a_key.shape[0]), ptr(a_key), ptr(a_seg), _u64(b_key.shape[0]), ptr(b_key), ptr(b_fkey),
                      ptr(b_date), _u64(f_fkey.shape[0]), ptr(f_fkey), ptr(f_price), ptr(f_disc), ptr(f_date),
                      _i64(seg), _i64(d), ptr(k), ptr(r), ptr(dt), _i64(cap), C.c_int(threads))
    return {"key": k[:g], "rev": r[:g], "date": dt[:g]}


def q3_topk(A: dict, B: dict, F: dict, seg: int, d: int, k: int = 10, threads: int = 0):
    lib = load()
    a_key, a_seg = _cols(A, ["a_key", "a_seg"])
    b_key, b_fkey, b_date = _cols(B, ["b_key", "b_fkey", "b_date"])
    f_fkey, f_price, f_disc, f_date = _cols(F, ["f_fkey", "f_price", "f_disc", "f_date"])
    ok = np.zeros(k, np.int64)
    orv = np.zeros(k, np.int64)
    od = np.zeros(k, np.int64)
    lib.oracle_q3_topk.restype = C.c_int64
    m = lib.oracle_q3_topk(_u64(a_key.shape[0]), ptr(a_key), ptr(a_seg), _u64(b_key.shape[0]), ptr(b_key),
                           ptr(b_fkey), ptr(b_date), _u64(f_fkey.shape[0]), ptr(f_fkey), ptr(f_price), ptr(f_disc),
                           ptr(f_date), _i64(seg), _i64(d), C.c_int(k), ptr(ok), ptr(orv), ptr(od), C.c_int(threads))
    return {"b_key": ok[:m], "revenue": orv[:m], "b_date": od[:m]}


INT64_MAX = (1 << 63) - 1


def groupby(G: dict, val_cut: int = INT64_MAX, threads: int = 0):
    lib = load()
    key, val = _cols(G, ["g_key", "g_val"])
    n = key.shape[0]
    cap = int(min(n, (int(key.max()) - int(key.min()) + 1) if n else 0))
    outs = [np.zeros(max(cap, 1), np.int64) for _ in range(5)]
    lib.oracle_groupby.restype = C.c_int64
    g = lib.oracle_groupby(_u64(n), ptr(key), ptr(val), _i64(val_cut), *[ptr(o) for o in outs], _i64(cap),
                           C.c_int(threads))
    names = ["key", "count", "sum", "min", "max"]
    return {nm: o[:g] for nm, o in zip(names, outs)}


def distinct(G: dict, val_cut: int = INT64_MAX) -> int:
    lib = load()
    key, val = _cols(G, ["g_key", "g_val"])
    lib.oracle_distinct.restype = C.c_int64
    return int(lib.oracle_distinct(_u64(key.shape[0]), ptr(key), ptr(val), _i64(val_cut)))


def like2(s: bytes, width: int, w1: bytes, w2: bytes) -> bool:
    lib = load()
    buf = C.create_string_buffer(s.ljust(width, b"\0")[:width], width)
    lib.oracle_like2.restype = C.c_int
    return bool(lib.oracle_like2(buf, C.c_int(width), C.c_char_p(w1), C.c_int(len(w1)), C.c_char_p(w2),
                                 C.c_int(len(w2))))


def q13(S: dict, A: dict, max_key: int, w1: bytes, w2: bytes, negate: bool = False, bins: int = 1024,
        threads: int = 0):
    lib = load()
    s_fkey, s_text = _cols(S, ["s_fkey", "s_text"])
    (a_key,) = _cols(A, ["a_key"])
    width = s_text.shape[1]
    oc = np.zeros(bins, np.int64)
    od = np.zeros(bins, np.int64)
    lib.oracle_q13.restype = C.c_int64
    m = lib.oracle_q13(_u64(s_fkey.shape[0]), ptr(s_fkey), ptr(s_text), C.c_int(width), _u64(a_key.shape[0]),
                       ptr(a_key), _i64(max_key), C.c_char_p(w1), C.c_int(len(w1)), C.c_char_p(w2),
                       C.c_int(len(w2)), C.c_int(1 if negate else 0), C.c_int(bins), ptr(oc), ptr(od),
                       C.c_int(threads))
    return {"c_count": oc[:m], "custdist": od[:m]}


def q22(B: dict, A: dict, max_key: int, date_lo: int, date_hi: int):
    lib = load()
    b_fkey, b_date = _cols(B, ["b_fkey", "b_date"])
    a_key, a_seg = _cols(A, ["a_key", "a_seg"])
    out = np.zeros(10, np.int64)
    lib.oracle_q22.restype = None
    lib.oracle_q22(_u64(b_fkey.shape[0]), ptr(b_fkey), ptr(b_date), _u64(a_key.shape[0]), ptr(a_key), ptr(a_seg),
                   _i64(max_key), _i64(date_lo), _i64(date_hi), ptr(out))
    return out[:5], out[5:]


MLP_NPARAMS = 5313


def mlp_forward(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    lib = load()
    P = np.ascontiguousarray(params, np.float32)
    Xc = np.ascontiguousarray(X, np.float32)
    y = np.zeros(Xc.shape[0], np.float64)
    lib.oracle_mlp_forward.restype = None
    lib.oracle_mlp_forward(ptr(P), ptr(Xc), C.c_int(Xc.shape[0]), ptr(y))
    return y


def mlp_grad(params: np.ndarray, X: np.ndarray, T: np.ndarray, scale: float):
    lib = load()
    P = np.ascontiguousarray(params, np.float32)
    Xc = np.ascontiguousarray(X, np.float32)
    Tc = np.ascontiguousarray(T, np.float32)
    g = np.zeros(MLP_NPARAMS, np.float64)
    lib.oracle_mlp_grad.restype = C.c_double
    loss = lib.oracle_mlp_grad(ptr(P), ptr(Xc), ptr(Tc), C.c_int(Xc.shape[0]), C.c_double(scale), ptr(g))
    return g, loss
