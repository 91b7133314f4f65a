"""Mean and standard error of a time series with the lag-1 autocorrelation
correction the reference's analysis script applies (pot.R:172-178, after
Foster & Rahmstorf 2011): the plain standard error of the mean times
sqrt((1 + r1) / (1 - r1)), r1 the sample lag-1 autocorrelation.  Used by the
ensemble test and by bench.py's charge-series mode."""
import math

import numpy as np


def lag1_stats(x):
    x = np.asarray(x, float)
    n = len(x)
    m = float(np.mean(x)) if n else float("nan")
    if n < 3 or float(np.var(x)) == 0.0:
        return {"mean": m, "se": 0.0, "n": n, "r1": 0.0}
    d = x - m
    r1 = float(np.sum(d[1:] * d[:-1]) / np.sum(d * d))  # acf(x)$acf[2] in R
    r1 = min(r1, 0.999)
    se = float(np.std(x, ddof=1) / math.sqrt(n) * math.sqrt((1.0 + r1) / (1.0 - r1)))
    return {"mean": m, "se": se, "n": n, "r1": r1}


def across_runs(values):
    """Mean over independent runs and its standard error."""
    v = np.asarray(values, float)
    return {"mean": float(v.mean()), "se": float(v.std(ddof=1) / math.sqrt(len(v))),
            "n": int(len(v))}
