"""Statistical agreement of the GPU chains with exact results and with the
reference CPU path.

* 2D Ising L=16: <E> and C = beta^2 var(E) from 64 independent GPU chains
  against Kaufman's exact finite-lattice partition function (FP64 from a
  40-digit evaluation), Metropolis, heat bath and Glauber, three betas.
* 2D Ising L=16 at beta_c: <E>, <|m|>, C and the Binder cumulant U4 of the GPU
  chains against the reference CPU path (random-site Metropolis with
  Xoshiro256++, restated in C), in units of the combined error.

* 2D Ising L=64 (the 4096-chain workload's lattice) at T = 1.5, 2.5002 and 3.5
  (chains 0, 2048 and 4095 of its temperature grid): <E> and C of 256 GPU
  chains against Kaufman's exact partition function at L=64, and <|m|> and U4
  against the reference CPU path, from an ordered start.

Errors come from the scatter of independent chains (each chain's mean is one
sample), so the correlation time does not enter them.  Seeds are fixed.  The
L=16 tests make 22 z-tests; a false-alarm rate of 1 % for the whole family
gives each one the two-sided level 0.01 / 22, i.e. |z| < 3.5 (Bonferroni).  The
L=64 test is a family of 12 (|z| < 3.35).  A per-test 2-sigma cut would fail
somewhere in such a family about two times in three with correct sampling, so
no result here is a "within 2 sigma" claim: every z is printed instead.
"""

Z_MAX = 3.5
Z_MAX_L64 = 3.35
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BETA_C = 0.4406868
THERM, MEAS, CHAINS, L = 2000, 20000, 64, 16


def _gpu_chain_moments(mcx, alg, beta, seed, L=L, chains=CHAINS, therm=THERM, meas=MEAS,
                       start="random"):
    b = mcx.LatticeBatch("ising", 2, L, n_chains=chains, seed=seed, series_capacity=meas)
    b.set_betas(np.full(chains, beta))
    b.init(start)
    b.run(alg, therm)
    b.run(alg, meas, record=True)
    mom = b.moments()
    b.close()
    assert np.all(mom[:, 0] == meas)
    return mom[:, 1:] / meas  # per chain <E>, <E^2>, <|M|>, <M^2>, <M^4>


def _mean_err(x):
    x = np.asarray(x, np.float64)
    return float(x.mean()), float(x.std(ddof=1) / math.sqrt(len(x)))


def _jackknife(f, cols):
    """Jackknife over chains of f(column means)."""
    cols = [np.asarray(c, np.float64) for c in cols]
    n = len(cols[0])
    tot = [c.sum() for c in cols]
    est = f(*[t / n for t in tot])
    jk = np.array([f(*[(t - c[i]) / (n - 1) for t, c in zip(tot, cols)]) for i in range(n)])
    return float(est), float(math.sqrt((n - 1) / n * ((jk - jk.mean()) ** 2).sum()))


def _observables(m, beta, L=L):
    """<E>, <|m|>, C, U4 with errors from per-chain moments m[:, 0..4]."""
    N = L * L
    E, E2, AM, M2, M4 = (m[:, k] for k in range(5))
    return {
        "E": _mean_err(E),
        "absm": _mean_err(AM / N),
        "C": _jackknife(lambda e, e2: beta ** 2 * (e2 - e * e), [E, E2]),
        "U4": _jackknife(lambda m2, m4: 1 - m4 / (3 * m2 * m2), [M2, M4]),
    }


@pytest.mark.parametrize("alg", ["metropolis", "heatbath", "glauber"])
@pytest.mark.parametrize("beta", [0.3, BETA_C, 0.55])
def test_l16_matches_kaufman(mcx, alg, beta):
    from oracle import exact
    Ek, _, Ck = exact.kaufman_observables(L, beta)
    seed = 1000 + int(beta * 1e4) + {"metropolis": 0, "heatbath": 7, "glauber": 13}[alg]
    print(f"statistics: L=16 {alg} beta={beta} vs Kaufman", flush=True)
    m = _gpu_chain_moments(mcx, alg, beta, seed)
    obs = _observables(m, beta)
    (e, de), (c, dc) = obs["E"], obs["C"]
    ze, zc = (e - Ek) / de, (c - Ck) / dc
    print(f"  <E> {e:.4f} +- {de:.4f} (exact {Ek:.4f}, z={ze:+.2f});"
          f" C {c:.3f} +- {dc:.3f} (exact {Ck:.3f}, z={zc:+.2f})", flush=True)
    assert abs(ze) < Z_MAX and abs(zc) < Z_MAX


def test_l16_matches_reference_cpu_path(mcx, orc):
    beta = BETA_C
    print("statistics: L=16 GPU checkerboard vs reference CPU path", flush=True)
    gpu = _observables(_gpu_chain_moments(mcx, "metropolis", beta, 4242), beta)
    n_cpu = 16
    rows = []
    for k in range(n_cpu):
        s = np.ones(L * L, np.int8)
        orc.refpath_series_ising2d(L, beta, 500 + k, THERM, s)
        E, M, _ = orc.refpath_series_ising2d(L, beta, 900 + k, MEAS, s)
        e, mm = E.astype(np.float64), M.astype(np.float64)
        rows.append([e.mean(), (e * e).mean(), np.abs(mm).mean(), (mm * mm).mean(),
                     ((mm * mm) ** 2).mean()])
    cpu = _observables(np.array(rows), beta)
    for k in ("E", "absm", "C", "U4"):
        (a, da), (b, db) = gpu[k], cpu[k]
        z = (a - b) / math.hypot(da, db)
        print(f"  {k}: gpu {a:.5f} +- {da:.5f}  cpu {b:.5f} +- {db:.5f}  z={z:+.2f}", flush=True)
        assert abs(z) < Z_MAX, k


def test_l64_workload_temperatures_match_kaufman_and_reference_path(mcx, orc):
    """The 4096-chain workload's lattice (L=64) at three of its temperatures."""
    from oracle import exact
    L64, n, chains, therm, meas = 64, 4096, 256, 2000, 20000
    cpu_chains, cpu_meas = 16, 5000
    temps = [1.5 + 2.0 * k / (n - 1) for k in (0, 2048, n - 1)]
    zs = []
    for T in temps:
        beta = 1.0 / T
        print(f"statistics: L=64 T={T:.4f} ({chains} GPU chains x {meas}, "
              f"{cpu_chains} CPU chains x {cpu_meas})", flush=True)
        gpu = _observables(_gpu_chain_moments(mcx, "metropolis", beta, 64000 + int(T * 100),
                                              L=L64, chains=chains, therm=therm, meas=meas,
                                              start="up"), beta, L=L64)
        Ek, _, Ck = exact.kaufman_observables(L64, beta)
        for k, ref in (("E", Ek), ("C", Ck)):
            v, dv = gpu[k]
            z = (v - ref) / dv
            zs.append(z)
            print(f"  {k}: gpu {v:.4f} +- {dv:.4f}  exact {ref:.4f}  z={z:+.2f}", flush=True)

        def cpu_chain(k):
            s = np.ones(L64 * L64, np.int8)
            orc.refpath_series_ising2d(L64, beta, 7000 + 31 * k + int(T * 100), therm, s)
            E, M, _ = orc.refpath_series_ising2d(L64, beta, 9000 + 31 * k + int(T * 100),
                                                 cpu_meas, s)
            e, mm = E.astype(np.float64), M.astype(np.float64)
            return [e.mean(), (e * e).mean(), np.abs(mm).mean(), (mm * mm).mean(),
                    ((mm * mm) ** 2).mean()]

        with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
            rows = list(ex.map(cpu_chain, range(cpu_chains)))
        cpu = _observables(np.array(rows), beta, L=L64)
        for k in ("absm", "U4"):
            (a, da), (b, db) = gpu[k], cpu[k]
            z = (a - b) / math.hypot(da, db)
            zs.append(z)
            print(f"  {k}: gpu {a:.5f} +- {da:.5f}  cpu {b:.5f} +- {db:.5f}  z={z:+.2f}",
                  flush=True)
    print("  all z:", " ".join(f"{z:+.2f}" for z in zs), flush=True)
    assert len(zs) == 12
    assert max(abs(z) for z in zs) < Z_MAX_L64
