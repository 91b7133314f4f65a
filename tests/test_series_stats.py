"""The lag-1 corrected standard error (tests/series_stats.py, the method of
pot.R:172-178) against what fixes it: R's acf definition on a hand-worked
series, and the spread of the mean over many independent AR(1) series of
known autocorrelation, for which (1 + rho) / (1 - rho) is the asymptotic
variance inflation of the mean."""
import numpy as np

from series_stats import across_runs, lag1_stats


def test_lag1_autocorrelation_is_r_acf():
    # x = 1..5: deviations -2..2, lag products 2 + 0 + 0 + 2 = 4, squares 10
    s = lag1_stats([1, 2, 3, 4, 5])
    assert abs(s["r1"] - 0.4) < 1e-15
    assert abs(s["mean"] - 3.0) < 1e-15
    plain = np.std([1, 2, 3, 4, 5], ddof=1) / np.sqrt(5)
    assert abs(s["se"] - plain * np.sqrt(1.4 / 0.6)) < 1e-14


def _ar1(rng, n, rho, reps):
    e = rng.normal(size=(reps, n))
    x = np.empty_like(e)
    x[:, 0] = e[:, 0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[:, t] = rho * x[:, t - 1] + e[:, t]
    return x


def test_corrected_se_matches_spread_of_ar1_means():
    rng = np.random.default_rng(3)
    for rho in (0.0, 0.5, 0.9):
        x = _ar1(rng, 2000, rho, 400)
        true_sd = x.mean(axis=1).std(ddof=1)  # spread of the mean over replicates
        ses = np.array([lag1_stats(row)["se"] for row in x])
        r1 = np.median([lag1_stats(row)["r1"] for row in x])
        assert abs(r1 - rho) < 0.02, (rho, r1)
        # the corrected se tracks the true spread (400 replicates: ~4% noise)
        assert abs(np.median(ses) / true_sd - 1.0) < 0.15, (rho, np.median(ses), true_sd)
        if rho > 0:
            plain = np.median([row.std(ddof=1) / np.sqrt(len(row)) for row in x])
            assert plain < 0.8 * true_sd  # uncorrected it would understate


def test_across_runs():
    s = across_runs([1.0, 3.0])
    assert s["mean"] == 2.0 and abs(s["se"] - 1.0) < 1e-15 and s["n"] == 2
