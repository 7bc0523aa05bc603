"""Seeded synthetic tables for the five workloads.

Every column is produced by the engine's counter-based generators
(``gdb_gen_*`` in include/gdb/gdb_engine.h), so the same (seed, size) always
yields the same bytes, on any thread count, for the GPU path and the oracle.
Seeds: ``0x6D693335 + workload`` with one stream id per column.

Workloads (sizes are parameters; the defaults are the full shapes):
  0  fact F for the fused filter + group-by aggregate (6,000,000 rows; also
     run at 120,000,000 rows)
  1  A (1,500,000) / B (15,000,000) / F (~60,000,000, 1-7 rows per B key)
  2  G: 120,000,000 Zipf(1.1) keys over 1..20,000,000 with values 0..10^6
  3  64 uneven pipelines (see gpudbmetalcodegen_amd.batch)
  4  resident F (120M) / B (30M) / A (3M) / S (4M x char[25])
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

SEED_BASE = 0x6D693335

# date domains (days)
DATE_DAYS_W0 = 2526          # workload 0: shipdate uniform over 2,526 days
DATE_DAYS_W1 = 2406          # workload 1: dates uniform over 2,406 days
PRICE_LO, PRICE_HI = 90_000, 10_500_000


def seed_for(workload: int) -> int:
    return SEED_BASE + workload


def _e():
    return N.load()[1]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def uniform_i32(seed, stream, n, lo, hi) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    N.eng_check(_e().gdb_gen_uniform_i32(seed, stream, n, lo, hi, _ptr(out)))
    return out


def uniform_i64(seed, stream, n, lo, hi) -> np.ndarray:
    out = np.empty(n, dtype=np.int64)
    N.eng_check(_e().gdb_gen_uniform_i64(seed, stream, n, lo, hi, _ptr(out)))
    return out


def categorical_u8(seed, stream, n, values: bytes, weights) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    v = np.frombuffer(values, dtype=np.uint8).copy()
    w = np.asarray(weights, dtype=np.uint32)
    N.eng_check(_e().gdb_gen_categorical_u8(seed, stream, n, _ptr(v), _ptr(w), len(v), _ptr(out)))
    return out


def permutation_i32(seed, stream, n, base=1) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    N.eng_check(_e().gdb_gen_permutation_i32(seed, stream, n, base, _ptr(out)))
    return out


def fanout_fkey_i32(seed, stream, nparent, cmin, cmax) -> np.ndarray:
    total = C.c_uint64()
    N.eng_check(_e().gdb_gen_fanout_total(seed, stream, nparent, cmin, cmax, C.byref(total)))
    out = np.empty(total.value, dtype=np.int32)
    N.eng_check(_e().gdb_gen_fanout_fkey_i32(seed, stream, nparent, cmin, cmax, total.value, _ptr(out)))
    return out


def zipf_i32(seed, stream, n, s, nkeys) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    N.eng_check(_e().gdb_gen_zipf_i32(seed, stream, n, float(s), nkeys, _ptr(out)))
    return out


def vocabulary(seed, nwords=500, minlen=3, maxlen=10) -> list[bytes]:
    out = np.zeros(nwords * 16, dtype=np.uint8)
    N.eng_check(_e().gdb_gen_vocabulary(seed, nwords, minlen, maxlen, _ptr(out)))
    raw = out.tobytes()
    return [raw[i * 16:(i + 1) * 16].rstrip(b"\0") for i in range(nwords)]


def text(seed, stream, n, width, vocab: list[bytes]) -> np.ndarray:
    v = np.zeros(len(vocab) * 16, dtype=np.uint8)
    for i, w in enumerate(vocab):
        v[i * 16:i * 16 + len(w)] = np.frombuffer(w, dtype=np.uint8)
    out = np.empty((n, width), dtype=np.uint8)
    N.eng_check(_e().gdb_gen_text(seed, stream, n, width, _ptr(v), len(vocab), _ptr(out)))
    return out


# ---------------------------------------------------------------------------
# Workload 0 — fact table F (cfg columns of the fused scan-aggregate)
# ---------------------------------------------------------------------------
def fact_f(n: int, seed: int, with_fkey_into: int = 0) -> dict:
    cols = {
        "f_qty": uniform_i32(seed, 1, n, 1, 50),
        "f_price": uniform_i64(seed, 2, n, PRICE_LO, PRICE_HI),
        "f_disc": uniform_i32(seed, 3, n, 0, 10),
        "f_tax": uniform_i32(seed, 4, n, 0, 8),
        "f_flag": categorical_u8(seed, 5, n, b"ARN", [1, 1, 2]),
        "f_status": categorical_u8(seed, 6, n, b"FO", [1, 1]),
        "f_date": uniform_i32(seed, 7, n, 0, DATE_DAYS_W0 - 1),
    }
    if with_fkey_into:
        cols["f_fkey"] = uniform_i32(seed, 8, n, 1, with_fkey_into)
    return cols


def q1_cutoff() -> int:
    """shipdate <= max - 90 over the 2,526-day range."""
    return (DATE_DAYS_W0 - 1) - 90


def workload0(n: int = 6_000_000) -> dict:
    return {"F": fact_f(n, seed_for(0))}


# ---------------------------------------------------------------------------
# Workload 1 — three-way join
# ---------------------------------------------------------------------------
W1_SEG = 2
W1_DATE = 1203


def workload1(n_a: int = 1_500_000, n_b: int = 15_000_000, cmin: int = 1, cmax: int = 7) -> dict:
    s = seed_for(1)
    a = {"a_key": permutation_i32(s, 11, n_a, 1), "a_seg": uniform_i32(s, 12, n_a, 0, 4).astype(np.uint8)}
    b = {
        "b_key": permutation_i32(s, 21, n_b, 1),
        "b_fkey": uniform_i32(s, 22, n_b, 1, n_a),
        "b_date": uniform_i32(s, 23, n_b, 0, DATE_DAYS_W1 - 1),
    }
    fk = fanout_fkey_i32(s, 31, n_b, cmin, cmax)
    n_f = fk.shape[0]
    f = {
        "f_fkey": fk,
        "f_price": uniform_i64(s, 32, n_f, PRICE_LO, PRICE_HI),
        "f_disc": uniform_i32(s, 33, n_f, 0, 10),
        "f_date": uniform_i32(s, 34, n_f, 0, DATE_DAYS_W1 - 1),
    }
    return {"A": a, "B": b, "F": f}


# ---------------------------------------------------------------------------
# Workload 2 — skewed group-by
# ---------------------------------------------------------------------------
def workload2(n: int = 120_000_000, s: float = 1.1, nkeys: int = 20_000_000, seed: int | None = None) -> dict:
    sd = seed_for(2) if seed is None else seed
    return {"G": {"g_key": zipf_i32(sd, 41, n, s, nkeys), "g_val": uniform_i64(sd, 42, n, 0, 1_000_000)}}


# ---------------------------------------------------------------------------
# Workload 4 — served tables
# ---------------------------------------------------------------------------
S_WIDTH = 25


def workload4(n_f: int = 120_000_000, n_b: int = 30_000_000, n_a: int = 3_000_000, n_s: int = 4_000_000) -> dict:
    s = seed_for(4)
    f = fact_f(n_f, s, with_fkey_into=n_b)
    a = {"a_key": permutation_i32(s, 11, n_a, 1), "a_seg": uniform_i32(s, 12, n_a, 0, 4).astype(np.uint8)}
    b = {
        "b_key": permutation_i32(s, 21, n_b, 1),
        "b_fkey": uniform_i32(s, 22, n_b, 1, n_a),
        "b_date": uniform_i32(s, 23, n_b, 0, DATE_DAYS_W1 - 1),
    }
    vocab = vocabulary(s)
    st = {"s_fkey": uniform_i32(s, 51, n_s, 1, n_a), "s_text": text(s, 52, n_s, S_WIDTH, vocab)}
    return {"F": f, "A": a, "B": b, "S": st, "vocab": vocab}
