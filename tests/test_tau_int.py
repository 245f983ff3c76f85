"""tau_int and the exchange-interval retune on the CPU.

The oracle's exact-sum tau_int (oracle/mcx_oracle.c: orc_tau_int) is pinned
against exact rational arithmetic on small series (Fractions: the centred
products of autocorrelations.jl:47-73 with no rounding at all), against the
reference's own FP64 evaluation (orc_tau_int_centred, and the host mirror) on
long series, against the AR(1) closed form, and against the known answers of
test/test_measurements.jl:151-187.  _retune_exchange_sweeps! and
optimize_exchange_interval! (parallel_tempering.jl:53-135) are checked on
hand-computed cases."""
from fractions import Fraction
from types import SimpleNamespace

import numpy as np
import pytest


def _tau_fraction(x, cap, c, D=None, y=None):
    """tau_int in exact rationals (window decisions exact too)."""
    if y is not None:
        x = [Fraction(int(a)) + Fraction(D) * Fraction(int(b)) for a, b in zip(x, y)]
    else:
        x = [Fraction(int(a)) for a in x]
    n = len(x)
    mu = sum(x) / n
    d = [a - mu for a in x]
    C0 = sum(v * v for v in d) / n
    if C0 <= 0:
        return 0.5
    tau = Fraction(1, 2)
    cc = Fraction(c)
    for t in range(1, cap + 1):
        C = (sum(d[i] * d[i + t] for i in range(n - t)) / (n - t)) / C0
        if C <= 0:
            break
        nxt = tau + C
        if t > cc * nxt:
            break
        tau = nxt
    return float(max(Fraction(1, 2), tau))


def _ar1_ints(rng, n, phi, scale=1000.0):
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0]
    for i in range(1, n):
        x[i] = phi * x[i - 1] + np.sqrt(1 - phi * phi) * e[i]
    return np.rint(x * scale).astype(np.int32)


def test_oracle_tau_exact_vs_fractions(orc):
    rng = np.random.default_rng(3)
    for trial in range(60):
        n = int(rng.integers(4, 41))
        x = rng.integers(-20, 21, n).astype(np.int32)
        if trial % 3 == 0:
            x = np.cumsum(rng.integers(-3, 4, n)).astype(np.int32)  # correlated walk
        c = [5.0, 2.0, 3.5][trial % 3]
        cap = n // 2 if trial % 2 else max(1, n // 3)
        want = _tau_fraction(x, cap, c)
        got = orc.tau_int(x, max_lag=cap, c=c)
        assert got == pytest.approx(want, rel=1e-14, abs=0), (trial, x.tolist(), cap, c)
        # the energy of a Blume-Capel series, E = E_pair + D Q, with a dyadic D
        y = rng.integers(0, n + 1, n).astype(np.int32)
        want = _tau_fraction(x, cap, c, D=0.75, y=y)
        got = orc.tau_int(x, y=y, D=0.75, max_lag=cap, c=c)
        assert got == pytest.approx(want, rel=1e-14, abs=0), (trial, "energy")
        # |x|
        want = _tau_fraction(np.abs(x), cap, c)
        assert orc.tau_int(x, absx=True, max_lag=cap, c=c) == pytest.approx(want, rel=1e-14, abs=0)


def test_oracle_tau_exact_vs_reference_evaluation(orc):
    from montecarlox.measurements import integrated_autocorrelation_time
    rng = np.random.default_rng(11)
    for phi, n in [(0.0, 4000), (0.5, 3000), (0.8, 6000), (0.95, 5000), (-0.3, 2000)]:
        x = _ar1_ints(rng, n, phi)
        for max_lag, c in [(None, 5.0), (50, 5.0), (200, 3.0)]:
            exact = orc.tau_int(x, max_lag=max_lag, c=c)
            assert exact == pytest.approx(orc.tau_int_centred(x.astype(np.float64), max_lag, c),
                                          rel=1e-12)
            assert exact == pytest.approx(integrated_autocorrelation_time(x, max_lag=max_lag, c=c),
                                          rel=1e-12)


def test_oracle_tau_known_answers(orc):
    """test_measurements.jl:151-187 on integer series, plus closed forms."""
    rng = np.random.default_rng(123)
    assert orc.tau_int(np.ones(64, np.int32)) == 0.5
    assert orc.tau_int(np.array([3, -3] * 40, np.int32)) == 0.5  # C(1) = -1 stops at once
    white = np.rint(rng.standard_normal(4000) * 1000).astype(np.int32)
    assert 0.5 <= orc.tau_int(white) <= 2.0
    ar = _ar1_ints(rng, 6000, 0.8)
    assert 2.5 <= orc.tau_int(ar) <= 7.0
    capped = orc.tau_int(ar, max_lag=50)
    assert np.isfinite(capped) and capped >= 0.5
    t = orc.tau_int_times([rng.integers(-5, 5, 20)], min_points=2, max_lag=0)
    assert t.shape == (1,) and np.isnan(t[0])
    t = orc.tau_int_times([rng.integers(-5, 5, 20), rng.integers(-5, 5, 500)], min_points=400)
    assert np.isnan(t[0]) and np.isfinite(t[1])
    # AR(1): tau_int = (1 + phi) / (2 (1 - phi)) = 4.5 at phi = 0.8
    long_ar = _ar1_ints(np.random.default_rng(5), 200_000, 0.8)
    assert orc.tau_int(long_ar) == pytest.approx(4.5, abs=0.4)


def test_retune_exchange_sweeps_known_answers(mcx):
    from montecarlox.pt import _retune_exchange_sweeps_
    s = np.zeros(4, np.int64)
    _retune_exchange_sweeps_(s, [1.0, 2.0, 3.0, np.nan], 100, 1, 10 ** 9)
    assert s.tolist() == [50, 100, 150, 100]  # median of finite taus = 2; NaN keeps scale 1
    _retune_exchange_sweeps_(s, [1.0, 2.0, 3.0, np.nan], 100, 60, 120)
    assert s.tolist() == [60, 100, 120, 100]
    s2 = np.zeros(2, np.int64)
    _retune_exchange_sweeps_(s2, [1.0, 3.0], 5, 1, 100)  # 2.5 -> 2, 7.5 -> 8 (half to even)
    assert s2.tolist() == [2, 8]
    _retune_exchange_sweeps_(s2, [np.nan, np.nan], 7, 1, 100)  # no finite tau: tau_ref = 1
    assert s2.tolist() == [7, 7]


def test_optimize_exchange_interval_host_samples(mcx):
    from montecarlox.measurements import integrated_autocorrelation_times
    rng = np.random.default_rng(8)
    R = 4
    pt = SimpleNamespace(R=R, n_sets=1, indices=np.array([[2, 1, 4, 3]], np.int32))
    traces = [rng.standard_normal(500).cumsum() * 0.05 + rng.standard_normal(500) for _ in range(R)]
    flat = [(i + 1, v) for t in range(500) for i in range(R) for v in [traces[i][t]]]
    sw = np.full(R, 100, np.int64)
    got = mcx.optimize_exchange_interval_(pt, flat, sw, base_sweeps=100, min_sweeps=10,
                                          max_sweeps=400, min_points=400, max_lag=200)
    taus = integrated_autocorrelation_times(traces, min_points=400, max_lag=200)
    ref = np.clip(np.rint(100 * taus / np.median(taus)), 10, 400).astype(np.int64)
    assert sw.tolist() == ref.tolist()
    assert got.tolist() == ref[pt.indices[0] - 1].tolist()  # sweeps_after_exchange[index(pt)]
    sw2 = np.full(R, 100, np.int64)  # pre-grouped traces give the same answer
    mcx.optimize_exchange_interval_(pt, traces, sw2, base_sweeps=100, min_sweeps=10,
                                    max_sweeps=400, min_points=400, max_lag=200)
    assert sw2.tolist() == sw.tolist()
    with pytest.raises(mcx.MCXArgumentError):
        mcx.optimize_exchange_interval_(pt, [(5, 1.0)], sw, base_sweeps=100)
    with pytest.raises(mcx.MCXArgumentError):
        mcx.optimize_exchange_interval_(pt, flat, np.zeros(3, np.int64), base_sweeps=100)
    few = [(1, 0.5), (2, 0.1)]  # below min_points: every tau NaN, every scale 1
    sw3 = np.zeros(R, np.int64)
    mcx.optimize_exchange_interval_(pt, few, sw3, base_sweeps=33)
    assert sw3.tolist() == [33] * R
