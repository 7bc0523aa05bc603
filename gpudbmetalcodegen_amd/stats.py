"""Trial statistics over per-run times.

The summary follows the reference runner's per-query statistics
(codegen/codegen_main.cpp:1005-1037): the median as element n/2 of the
sorted sample (the reference names it the lower median; for an even n it is
the upper of the two middle elements), nearest-rank percentiles taken as element
clamp(p*n - 1, 0, n-1) truncated to an integer index, and the median absolute
deviation about the lower median. p50/p99 use the same percentile rule.

`mann_whitney_u` is the two-sided Mann-Whitney U test with mid-ranks for
ties, the tie-corrected variance, a continuity correction and
p = erfc(z / sqrt(2)) (the reference's significance harness,
scripts/significance.py:30-75); the p-value is clipped to 1.
"""

from __future__ import annotations

import math
from typing import Sequence


def lower_median(xs: Sequence[float]) -> float:
    if not xs:
        return 0.0
    v = sorted(xs)
    return float(v[len(v) // 2])


def percentile(xs: Sequence[float], p: float) -> float:
    """Nearest-rank percentile as the reference computes it: index
    floor(clamp(p*n - 1, 0, n - 1)) of the sorted sample."""
    if not xs:
        return 0.0
    v = sorted(xs)
    idx = min(max(p * len(v) - 1.0, 0.0), float(len(v) - 1))
    return float(v[int(idx)])


def mad(xs: Sequence[float]) -> float:
    """Median absolute deviation about the lower median (0 for n < 2)."""
    if len(xs) < 2:
        return 0.0
    m = lower_median(xs)
    return lower_median([abs(x - m) for x in xs])


def trial_stats(xs: Sequence[float], digits: int = 4) -> dict:
    """Summary of a sample of per-trial times (ms): n, mean, lower median,
    p10/p90 (nearest rank), MAD, p50/p99, min and max."""
    n = len(xs)
    if n == 0:
        return {"n": 0}
    r = lambda x: round(float(x), digits)  # noqa: E731
    return {"n": n, "mean": r(sum(xs) / n), "median": r(lower_median(xs)), "p10": r(percentile(xs, 0.10)),
            "p90": r(percentile(xs, 0.90)), "mad": r(mad(xs)), "p50": r(percentile(xs, 0.50)),
            "p99": r(percentile(xs, 0.99)), "min": r(min(xs)), "max": r(max(xs))}


def mann_whitney_u(a: Sequence[float], b: Sequence[float]) -> tuple[float, float, float]:
    """Two-sided Mann-Whitney U of sample a against sample b.

    Returns (U_a, p, rank_biserial) with rank_biserial = 1 - 2 U_a / (n_a n_b),
    positive when a tends to be smaller (faster) than b."""
    na, nb = len(a), len(b)
    if na == 0 or nb == 0:
        return float("nan"), float("nan"), float("nan")
    pooled = sorted([(float(x), 0) for x in a] + [(float(x), 1) for x in b], key=lambda t: t[0])
    n = na + nb
    rank_sum_a = 0.0
    ties = 0.0
    i = 0
    while i < n:
        j = i
        while j + 1 < n and pooled[j + 1][0] == pooled[i][0]:
            j += 1
        mid = (i + j) / 2.0 + 1.0          # 1-based mid-rank of the tie run
        t = j - i + 1
        rank_sum_a += mid * sum(1 for k in range(i, j + 1) if pooled[k][1] == 0)
        if t > 1:
            ties += t ** 3 - t
        i = j + 1
    u_a = rank_sum_a - na * (na + 1) / 2.0
    mu = na * nb / 2.0
    var = na * nb * (n + 1) / 12.0
    if ties:
        var -= na * nb * ties / (12.0 * n * (n - 1))
    if var <= 0:
        return u_a, 1.0, 0.0
    z = (abs(u_a - mu) - 0.5) / math.sqrt(var)
    p = min(1.0, math.erfc(z / math.sqrt(2.0)))
    return u_a, p, 1.0 - 2.0 * u_a / (na * nb)


def compare(a: Sequence[float], base: Sequence[float]) -> dict:
    """a against a baseline sample: speedup of the lower medians, the U test
    and the p99 ratio."""
    u, p, eff = mann_whitney_u(a, base)
    ma, mb = lower_median(a), lower_median(base)
    return {"speedup_median": round(mb / ma, 4) if ma > 0 else None,
            "speedup_p99": round(percentile(base, 0.99) / percentile(a, 0.99), 4) if percentile(a, 0.99) > 0 else None,
            "mann_whitney_u": round(u, 2), "p_value": p, "rank_biserial": round(eff, 4)}
