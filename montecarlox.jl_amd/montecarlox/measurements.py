"""Host-side analysis of the device-recorded series.

integrated_autocorrelation_time / integrated_autocorrelation_times / tau_int
    restate src/measurements/autocorrelations.jl:28-110 (self-consistent
    window, positive-sequence truncation, 0.5 for zero variance).
binning_error
    standard error of the mean from non-overlapping bins (new: the reference
    leaves error bars to the user).
observables_from_moments
    <E>, <|m|>, specific heat and Binder cumulant from the FP64 moments the
    GPU accumulates per chain.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from ._lib import MCXArgumentError


def integrated_autocorrelation_time(samples, max_lag: Optional[int] = None, c: float = 5.0) -> float:
    x = np.asarray(samples, np.float64)
    n = len(x)
    if n < 2:
        raise MCXArgumentError("integrated_autocorrelation_time requires at least 2 samples")
    if not c > 0:
        raise MCXArgumentError("c must be positive")
    cap_max = n // 2
    cap = cap_max if max_lag is None else int(max_lag)
    if not 1 <= cap <= cap_max:
        raise MCXArgumentError("max_lag must satisfy 1 <= max_lag <= floor(length(samples)/2)")
    d = x - x.sum() / n
    C0 = float(np.dot(d, d)) / n
    if not C0 > 0:
        return 0.5
    tau = 0.5
    for lag in range(1, cap + 1):
        C = float(np.dot(d[: n - lag], d[lag:])) / (n - lag) / C0
        if C <= 0:
            break
        nxt = tau + C
        if lag > c * nxt:
            break
        tau = nxt
    return max(0.5, tau)


def tau_int(samples, **kw) -> float:
    return integrated_autocorrelation_time(samples, **kw)


def integrated_autocorrelation_times(traces: Sequence, min_points: int = 2,
                                     max_lag: Optional[int] = None, c: float = 5.0):
    if min_points < 2:
        raise MCXArgumentError("min_points must be >= 2")
    out = np.full(len(traces), np.nan)
    for i, t in enumerate(traces):
        n = len(t)
        if n < min_points:
            continue
        ml = None if max_lag is None else min(int(max_lag), n // 2)
        if ml is not None and ml < 1:
            continue
        out[i] = integrated_autocorrelation_time(t, max_lag=ml, c=c)
    return out


def binning_error(x, n_bins: int = 32) -> float:
    """Standard error of the mean of x from n_bins equal consecutive bins."""
    x = np.asarray(x, np.float64)
    nb = min(n_bins, len(x))
    if nb < 2:
        raise MCXArgumentError("need at least 2 samples")
    m = len(x) // nb
    b = x[: nb * m].reshape(nb, m).mean(1)
    return float(b.std(ddof=1) / math.sqrt(nb))


def jackknife(f, columns: Sequence[np.ndarray], n_bins: int = 32):
    """Jackknife estimate and error of f(*means) over binned columns."""
    cols = [np.asarray(c, np.float64) for c in columns]
    n = len(cols[0])
    m = n // n_bins
    binned = [c[: n_bins * m].reshape(n_bins, m).mean(1) for c in cols]
    tot = [b.sum() for b in binned]
    est = f(*[t / n_bins for t in tot])
    jk = np.array([f(*[(t - b[i]) / (n_bins - 1) for t, b in zip(tot, binned)])
                   for i in range(n_bins)])
    err = math.sqrt((n_bins - 1) / n_bins * ((jk - jk.mean()) ** 2).sum())
    return est, err


def observables_from_moments(mom: np.ndarray, N: int, beta: float):
    """Per chain: <E>, <|m|>, C = beta^2 (<E^2> - <E>^2), U4 = 1 - <m^4>/(3<m^2>^2)."""
    mom = np.atleast_2d(np.asarray(mom, np.float64))
    cnt = mom[:, 0]
    E = mom[:, 1] / cnt
    E2 = mom[:, 2] / cnt
    absm = mom[:, 3] / cnt / N
    m2 = mom[:, 4] / cnt / N ** 2
    m4 = mom[:, 5] / cnt / N ** 4
    beta = np.asarray(beta, np.float64)
    return {"E": E, "absm": absm, "C": beta ** 2 * (E2 - E * E), "U4": 1 - m4 / (3 * m2 * m2)}
