"""Trial statistics: hand-computed values for the reference's lower median,
nearest-rank percentiles and MAD; the U test against scipy's asymptotic
Mann-Whitney U (continuity-corrected, tie-corrected)."""

import numpy as np
import pytest

from gpudbmetalcodegen_amd import stats as S


def test_lower_median_and_percentiles_by_hand():
    xs = [5.0, 1.0, 4.0, 2.0, 3.0, 6.0]          # sorted 1..6
    assert S.lower_median(xs) == 4.0              # element n/2 = 3 of the sorted sample (0-based)
    assert S.percentile(xs, 0.10) == 1.0          # 0.6 - 1 -> 0
    assert S.percentile(xs, 0.90) == 5.0          # 5.4 - 1 = 4.4 -> 4
    assert S.percentile(xs, 0.99) == 5.0          # 5.94 - 1 = 4.94 -> 4
    assert S.percentile(xs, 1.00) == 6.0
    assert S.percentile(list(range(1, 101)), 0.99) == 99.0
    assert S.percentile(list(range(1, 1001)), 0.99) == 990.0
    assert S.lower_median([7.0]) == 7.0 and S.mad([7.0]) == 0.0


def test_mad_by_hand():
    xs = [1.0, 2.0, 3.0, 4.0, 100.0]              # lower median 3; deviations 2,1,0,1,97 -> sorted 0,1,1,2,97
    assert S.mad(xs) == 1.0
    st = S.trial_stats(xs)
    assert st["n"] == 5 and st["median"] == 3.0 and st["p10"] == 1.0 and st["p90"] == 4.0 and st["max"] == 100.0
    assert st["mean"] == pytest.approx(22.0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mann_whitney_matches_scipy(seed):
    from scipy.stats import mannwhitneyu

    rng = np.random.default_rng(seed)
    a = np.round(rng.normal(1.0, 0.1, 30), 2)     # rounding makes ties
    b = np.round(rng.normal(1.05, 0.1, 25), 2)
    u, p, eff = S.mann_whitney_u(a.tolist(), b.tolist())
    ref = mannwhitneyu(a, b, alternative="two-sided", method="asymptotic", use_continuity=True)
    assert u == pytest.approx(ref.statistic)
    assert p == pytest.approx(ref.pvalue, rel=1e-9)
    assert eff == pytest.approx(1 - 2 * ref.statistic / (len(a) * len(b)))


def test_mann_whitney_separated_samples():
    u, p, eff = S.mann_whitney_u([1.0, 2.0, 3.0, 4.0, 5.0], [10.0, 11.0, 12.0, 13.0, 14.0])
    assert u == 0.0 and eff == 1.0 and p < 0.02
    u, p, eff = S.mann_whitney_u([2.0] * 4, [2.0] * 4)   # all tied: zero variance
    assert p == 1.0 and eff == 0.0
