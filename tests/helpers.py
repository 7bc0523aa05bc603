"""Shared test helpers: oracle outputs reshaped to the collector's layout."""

from __future__ import annotations

import numpy as np

import oracle

Q1_LABELS = [("A", "F"), ("A", "O"), ("N", "F"), ("N", "O"), ("R", "F"), ("R", "O")]


def q1_expected(words: np.ndarray) -> dict:
    """Oracle 6x8 words -> the collector's Q1 columns (empty buckets skipped)."""
    cols = {k: [] for k in ["f_flag", "f_status", "sum_qty", "sum_base_price", "sum_disc_price", "sum_charge",
                            "sum_disc", "count_order", "avg_qty", "avg_price", "avg_disc"]}
    for b in range(6):
        w = words[b]
        cnt = int(w[6])
        if cnt == 0:
            continue
        cols["f_flag"].append(Q1_LABELS[b][0])
        cols["f_status"].append(Q1_LABELS[b][1])
        cols["sum_qty"].append(int(w[0]))
        cols["sum_base_price"].append(int(w[1]))
        cols["sum_disc_price"].append(int(w[2]))
        cols["sum_charge"].append(oracle.i128(w[3], w[4]))
        cols["sum_disc"].append(int(w[5]))
        cols["count_order"].append(cnt)
        cols["avg_qty"].append(int(w[0]) * 10**6 // cnt)
        cols["avg_price"].append(int(w[1]) * 10**6 // cnt)
        cols["avg_disc"].append(int(w[5]) * 10**6 // cnt)
    return cols


def as_lists(columns: dict) -> dict:
    return {k: [x if isinstance(x, str) else int(x) for x in v] for k, v in columns.items()}


def served_expected(T: dict, kind: str, params: dict, max_a: int, threads: int = 0) -> dict:
    """The oracle's answer to one served query of workload 4, in the
    collector's columns: scan = Q1 over F, join = top-10 of the three-way
    join, anti = the anti-join's per-segment count and key sum, like = the
    LIKE count histogram."""
    if kind == "scan":
        return q1_expected(oracle.q1(T["F"], params["cutoff"], threads=threads))
    if kind == "join":
        top = oracle.q3_topk(T["A"], T["B"], T["F"], params["seg"], params["d"], 10, threads=threads)
        return {k: [int(x) for x in v] for k, v in top.items()}
    if kind == "anti":
        cnt, sm = oracle.q22(T["B"], T["A"], max_a, params["lo"], params["hi"])
        segs = [i for i in range(5) if cnt[i] > 0]
        # the plan labels its five segment buckets "0".."4" (string labels)
        return {"a_seg": [str(i) for i in segs], "num_keys": [int(cnt[i]) for i in segs], "sum_keys": [int(sm[i]) for i in segs]}
    if kind == "like":
        h = oracle.q13(T["S"], T["A"], max_a, params["w1"], params["w2"], threads=threads)
        return {k: [int(x) for x in v] for k, v in h.items()}
    raise ValueError(kind)
