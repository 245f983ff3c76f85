"""Exact finite-lattice results -- TEST INFRASTRUCTURE ONLY (oracle side).

kaufman_logZ(L, beta)
    2D Ising on an L x L torus, J = 1: Kaufman's (1949) closed form in the
    arrangement of Beale (PRL 76, 78 (1996)), eq. (2)-(4):
      Z = 1/2 (2 sinh 2K)^(N/2) (Z1 + Z2 + Z3 + Z4)
      Z1 = prod_{r=0}^{L-1} 2 cosh(L g_{2r+1} / 2),  Z2 = prod 2 sinh(L g_{2r+1} / 2)
      Z3 = prod_{r=0}^{L-1} 2 cosh(L g_{2r} / 2),    Z4 = prod 2 sinh(L g_{2r} / 2)
      cosh g_l = cosh 2K coth 2K - cos(l pi / L)  (l >= 1),  g_0 = 2K + ln tanh K
    g_0 is negative below T_c, which flips the sign of Z4.  Evaluated with
    mpmath at 40 digits.  Pinned by brute force at L = 4 and by the
    reference's Beale exact DOS at L = 8 (SpinSystems/data/exact_solutions/
    ising2D_8x8.csv, copied as a fixture to tests/golden/).

kaufman_observables(L, beta)
    <E>, <E^2> - <E>^2 and C = beta^2 var(E) from derivatives of log Z.

dos_ising_bruteforce(dim, L) / dos_blume_capel_bruteforce(L)
    Exact joint densities of states by enumeration of every configuration:
    g[(E_pair, M)] for Ising, g[(pair_sum, M, Q)] for 2D Blume-Capel.

moments_from_dos(...)
    Exact canonical averages from those densities.
"""
from __future__ import annotations

import itertools

import numpy as np

try:
    import mpmath as mp
except ImportError:  # pragma: no cover
    mp = None


def kaufman_logZ(L: int, beta, dps: int = 40):
    if mp is None:
        raise RuntimeError("mpmath is required for the Kaufman closed form")
    with mp.workdps(dps):
        K = mp.mpf(beta)
        n = L
        c = mp.cosh(2 * K) / mp.tanh(2 * K)

        def gam(l):
            if l == 0:
                return 2 * K + mp.log(mp.tanh(K))
            return mp.acosh(c - mp.cos(mp.pi * l / n))

        # work with logs of |Z_i| and signs
        terms = []
        for odd, fn in ((1, mp.cosh), (1, mp.sinh), (0, mp.cosh), (0, mp.sinh)):
            logabs = mp.mpf(0)
            sign = 1
            for r in range(n):
                g = gam(2 * r + odd)
                v = 2 * fn(L * g / 2)
                if v == 0:
                    sign = 0
                    break
                if v < 0:
                    sign = -sign
                logabs += mp.log(abs(v))
            terms.append((sign, logabs))
        m = max(t[1] for t in terms if t[0] != 0)
        s = mp.fsum(t[0] * mp.e ** (t[1] - m) for t in terms if t[0] != 0)
        N = L * L
        return mp.log(mp.mpf(1) / 2) + (N / mp.mpf(2)) * mp.log(2 * mp.sinh(2 * K)) + m + mp.log(s)


def kaufman_observables(L: int, beta: float, dps: int = 40):
    """Returns (<E>, var(E), C) with E = -sum_<ij> s_i s_j, from central
    differences of log Z at 2*dps digits (step 1e-12: truncation ~1e-24)."""
    hi = 2 * dps
    with mp.workdps(hi):
        b = mp.mpf(beta)
        h = mp.mpf("1e-12")
        fp, f0, fm = (kaufman_logZ(L, b + h, hi), kaufman_logZ(L, b, hi),
                      kaufman_logZ(L, b - h, hi))
        d1 = (fp - fm) / (2 * h)
        d2 = (fp - 2 * f0 + fm) / (h * h)
        return float(-d1), float(d2), float(b * b * d2)


def logZ_from_dos(energies, logdos, beta: float) -> float:
    e = np.asarray(energies, np.float64)
    g = np.asarray(logdos, np.float64)
    ok = np.isfinite(g)
    x = g[ok] - beta * e[ok]
    m = x.max()
    return float(m + np.log(np.exp(x - m).sum()))


def _torus_pairs(dim: int, L: int):
    N = L ** dim
    idx = np.arange(N).reshape((L,) * dim)
    pairs = []
    for ax in range(dim):
        pairs.append(np.stack([idx.ravel(), np.roll(idx, -1, axis=dim - 1 - ax).ravel()], 1))
    return np.concatenate(pairs, 0)


def dos_ising_bruteforce(dim: int, L: int):
    """dict (E_pair, M) -> count over all 2^N states; E_pair = -sum_<ij> s_i s_j."""
    N = L ** dim
    if N > 24:
        raise ValueError("brute force limited to N <= 24")
    pairs = _torus_pairs(dim, L)
    codes = np.arange(2 ** N, dtype=np.int64)
    s = 1 - 2 * ((codes[:, None] >> np.arange(N)) & 1).astype(np.int64)
    pair = (s[:, pairs[:, 0]] * s[:, pairs[:, 1]]).sum(1)
    M = s.sum(1)
    keys, counts = np.unique(np.stack([-pair, M], 1), axis=0, return_counts=True)
    return {(int(a), int(b)): int(c) for (a, b), c in zip(keys, counts)}


def dos_blume_capel_bruteforce(L: int = 4):
    """dict (pair_sum, M, Q) -> count over all 3^(L*L) states of the 2D torus."""
    N = L * L
    pairs = _torus_pairs(2, L)
    out = {}
    # enumerate row-by-row in chunks of the first row
    states = np.array(list(itertools.product((-1, 0, 1), repeat=L)), np.int64)  # 3^L rows
    nr = len(states)
    rest = np.array(list(itertools.product(range(nr), repeat=L - 1)), np.int64)
    tail = states[rest].reshape(len(rest), -1)  # rows 1..L-1 flattened
    for r0 in range(nr):
        s = np.concatenate([np.broadcast_to(states[r0], (len(tail), L)), tail], 1)
        pr = (s[:, pairs[:, 0]] * s[:, pairs[:, 1]]).sum(1)
        M = s.sum(1)
        Q = (s * s).sum(1)
        keys, counts = np.unique(np.stack([pr, M, Q], 1), axis=0, return_counts=True)
        for (a, b, c), k in zip(keys, counts):
            key = (int(a), int(b), int(c))
            out[key] = out.get(key, 0) + int(k)
    return out


def ising_moments(dos: dict, beta: float, N: int):
    """Exact <E>, <E^2>, <|m|>, <m^2>, <m^4>, C, U4 (m = M/N) from dict (E, M) -> g."""
    E = np.array([k[0] for k in dos], np.float64)
    M = np.array([k[1] for k in dos], np.float64)
    g = np.array(list(dos.values()), np.float64)
    x = np.log(g) - beta * E
    w = np.exp(x - x.max())
    w /= w.sum()
    m = M / N
    out = {"E": (w * E).sum(), "E2": (w * E * E).sum(), "absm": (w * np.abs(m)).sum(),
           "m2": (w * m * m).sum(), "m4": (w * m ** 4).sum()}
    out["C"] = beta ** 2 * (out["E2"] - out["E"] ** 2)
    out["U4"] = 1 - out["m4"] / (3 * out["m2"] ** 2)
    return out


def blume_capel_moments(dos: dict, beta: float, D: float, N: int):
    """Exact averages for E = -pair + D Q from dict (pair, M, Q) -> g."""
    P = np.array([k[0] for k in dos], np.float64)
    M = np.array([k[1] for k in dos], np.float64)
    Q = np.array([k[2] for k in dos], np.float64)
    g = np.array(list(dos.values()), np.float64)
    E = -P + D * Q
    x = np.log(g) - beta * E
    w = np.exp(x - x.max())
    w /= w.sum()
    m = M / N
    out = {"E": (w * E).sum(), "E2": (w * E * E).sum(), "absm": (w * np.abs(m)).sum(),
           "m2": (w * m * m).sum(), "q": (w * Q / N).sum()}
    out["C"] = beta ** 2 * (out["E2"] - out["E"] ** 2)
    return out
