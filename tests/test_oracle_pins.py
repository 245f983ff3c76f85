"""Pins of the CPU oracle: every oracle function is checked against something
the reference or the mathematics fixes -- published known-answer vectors, the
reference's own known-answer tests and data files, exact enumerations and
closed forms -- so a dropped term, wrong sign, index or operand fails here.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ISING, BC = 0, 1
METRO, HB, GLAUBER = 0, 1, 2


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# Philox4x32-10
# --------------------------------------------------------------------------

def test_philox_known_answers(orc):
    vecs = _load("philox_kat.json")
    assert sum(v["source"] == "Random123 kat_vectors" for v in vecs) == 3
    for v in vecs:
        assert list(orc.philox4x32_10(v["ctr"], v["key"])) == v["out"], v


# --------------------------------------------------------------------------
# geometry and random words
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dim,L", [(2, 4), (2, 6), (2, 66), (3, 4), (3, 6)])
def test_site_group_is_a_bijection(orc, dim, L):
    N = L ** dim
    nw = (L // 2 + 31) // 32
    rows = L ** (dim - 1)
    seen = set()
    for i in range(N):
        g, t, c = orc.site_group(dim, L, i)
        x, y, z = i % L, (i // L) % L, i // (L * L)
        assert c == (x + y + (z if dim == 3 else 0)) % 2
        assert 0 <= g < rows * 2 * nw and 0 <= t < 32
        assert ((g // nw) % 2) == c
        seen.add((g, t))
    assert len(seen) == N


@pytest.mark.parametrize("dim,L", [(2, 8), (2, 24), (2, 56), (3, 8), (3, 16), (3, 24)])
def test_ising_group_fills_whole_words(orc, dim, L):
    """L < 64, L % 8 == 0: the Ising sweep stream lays each colour's sites out
    flat -- every group's 32 bits are used (N/32 groups in all), a colour's
    groups have its parity, and along x the next site of the colour takes the
    next bit (or bit 0 of the colour's next group) except where its row wraps."""
    N = L ** dim
    at = {}
    for i in range(N):
        g, t, c = orc.ising_group(dim, L, i)
        x, y, z = i % L, (i // L) % L, i // (L * L)
        assert c == (x + y + (z if dim == 3 else 0)) % 2 and g % 2 == c and 0 <= t < 32
        at[(g, t)] = (x, y, z)
    assert len(at) == N and len({g for g, _ in at}) == N // 32
    for (g, t), (x, y, z) in at.items():
        nxt = (g, t + 1) if t < 31 else (g + 2, 0)
        if x + 2 < L:
            assert at[nxt] == (x + 2, y, z)
    # the row words of other sizes are the site words' groups
    for i in range(6 ** dim):
        assert orc.ising_group(dim, 6, i) == orc.site_group(dim, 6, i)


def test_bitsliced_word_is_built_from_bit_planes(orc):
    seed, key, g, sweep = (5 << 32) | 9, 77, 123, 4321
    for t in (0, 5, 31):
        w = 0
        for chunk in range(8):
            ctr = [g, sweep & 0xFFFFFFFF, ((sweep >> 32) & 0xFFFF) | (chunk << 16) | (orc.P_METRO << 24),
                   seed >> 32]
            out = orc.philox4x32_10(ctr, [seed & 0xFFFFFFFF, key])
            for b in range(4):
                w |= ((int(out[b]) >> t) & 1) << (31 - (4 * chunk + b))
        assert orc.bitsliced_word(seed, key, g, t, sweep, orc.P_METRO) == w


def test_words_are_uniform(orc):
    rng = np.random.default_rng(0)
    for fn, P in ((orc.bitsliced_word, orc.P_METRO), (orc.site_word, orc.P_INIT)):
        ws = np.array([fn(7, int(k), int(g), int(t), 3, P) for k, g, t in
                       rng.integers(0, [100, 1000, 32], size=(4000, 3))], np.uint64)
        bits = ((ws[:, None] >> np.arange(32, dtype=np.uint64)) & 1).mean(0)
        assert np.all(np.abs(bits - 0.5) < 4 * 0.5 / math.sqrt(len(ws)))


def test_bc_site_fields_use_each_prefix_bit_once(orc):
    # The 32 sites of a word take the 16 words of prefix blocks 0..3 half by
    # half: the multiset of their fields is exactly those 32 halves, so no
    # random bit is shared or dropped; the tail comes from block 5, chunk 1 + bit.
    seed, key, g, sweep = (3 << 32) | 21, 1234, 77, (5 << 32) | 99
    for P in (orc.P_BC_METRO, orc.P_BC_HB):
        halves = []
        for i in range(4):
            ctr = [8 * g + i, sweep & 0xFFFFFFFF, ((sweep >> 32) & 0xFFFF) | (P << 24), seed >> 32]
            out = orc.philox4x32_10(ctr, [seed & 0xFFFFFFFF, key])
            halves += [int(w) & 0xFFFF for w in out] + [int(w) >> 16 for w in out]
        fields, tails = [], []
        for t in range(32):
            U, alt = orc.bc_site_uniform(seed, key, g, t, sweep, P)
            f = (U >> 16) ^ 0x8000
            assert alt == f & 1
            fields.append(f)
            ctr = [8 * g + 5, sweep & 0xFFFFFFFF, ((sweep >> 32) & 0xFFFF) | ((1 + t) << 16) | (P << 24),
                   seed >> 32]
            assert U & 0xFFFF == int(orc.philox4x32_10(ctr, [seed & 0xFFFFFFFF, key])[0]) & 0xFFFF
            tails.append(U & 0xFFFF)
        assert sorted(fields) == sorted(halves)
        assert len(set(tails)) > 28  # 32 distinct draws (a collision is possible, not 4)


def test_bc_site_uniform_is_uniform(orc):
    rng = np.random.default_rng(3)
    for P in (orc.P_BC_METRO, orc.P_BC_HB):
        draws = [orc.bc_site_uniform(11, int(k), int(g), int(t), 5, P) for k, g, t in
                 rng.integers(0, [100, 1000, 32], size=(4000, 3))]
        ws = np.array([u for u, _ in draws], np.uint64)
        bits = ((ws[:, None] >> np.arange(32, dtype=np.uint64)) & 1).mean(0)
        assert np.all(np.abs(bits - 0.5) < 4 * 0.5 / math.sqrt(len(ws)))
        alts = np.array([a for _, a in draws])
        assert abs(alts.mean() - 0.5) < 4 * 0.5 / math.sqrt(len(alts))


def test_threshold_is_exact_ceiling(orc):
    rng = np.random.default_rng(1)
    ps = list(rng.random(200)) + [math.exp(-8 * 0.4406868), 0.5, 1 - 2 ** -40, 2 ** -40, 0.0, 1.0]
    for p in ps:
        assert orc.threshold32(p) == math.ceil(Fraction(p) * 2 ** 32)


def test_logistic_known_answers(orc):
    # test/test_utils.jl:56-67
    assert orc.logistic(0.0) == 0.5
    assert abs(orc.logistic(20.0) - 1.0) < 1e-8
    assert abs(orc.logistic(-20.0)) < 1e-8
    assert abs(orc.logistic(3.7) + orc.logistic(-3.7) - 1.0) < 1e-12


# --------------------------------------------------------------------------
# bookkeeping (SpinSystems/test/test_ising.jl:8-35, test_blume_capel.jl:8-31)
# --------------------------------------------------------------------------

def test_ising_bookkeeping_known_answers(orc):
    s = np.ones(16, np.int8)
    assert -orc.pair_sum(2, 4, s) == -32
    assert orc.magnetization(2, 4, s) == 16
    s[0] = -1
    assert -orc.pair_sum(2, 4, s) == -24
    assert orc.magnetization(2, 4, s) == 14


def test_blume_capel_bookkeeping_known_answers(orc):
    D = 0.5
    s = np.ones(16, np.int8)
    E0 = -orc.pair_sum(2, 4, s) + D * orc.sum_sq(2, 4, s)
    assert E0 == -24.0
    s[0] = 0
    E1 = -orc.pair_sum(2, 4, s) + D * orc.sum_sq(2, 4, s)
    assert E1 - E0 == 3.5
    assert orc.magnetization(2, 4, s) == 15


def test_energy_series_equal_full_recompute(orc):
    # cached == full after random updates (test_ising.jl:149-172, test_blume_capel.jl:87-107)
    for model, alg, dim, L, D in ((ISING, METRO, 2, 6, 0), (ISING, HB, 3, 4, 0),
                                  (BC, METRO, 2, 6, 0.7), (BC, HB, 2, 4, 1.3),
                                  (ISING, GLAUBER, 3, 4, 0), (BC, GLAUBER, 2, 6, 0.9)):
        s = orc.init_random(model, dim, L, 11, 2)
        for k in range(4):
            E, M, Q, _ = orc.canonical_sweeps(model, alg, dim, L, 0.45, s, 11, 2, 10 * k, 1, D=D)
            assert E[0] == -orc.pair_sum(dim, L, s)
            assert M[0] == orc.magnetization(dim, L, s)
            assert Q[0] == orc.sum_sq(dim, L, s)


# --------------------------------------------------------------------------
# exact references
# --------------------------------------------------------------------------

def test_kaufman_matches_bruteforce_and_beale():
    from oracle import exact
    d4 = exact.dos_ising_bruteforce(2, 4)
    agg = {}
    for (e, m), c in d4.items():
        agg[e] = agg.get(e, 0) + c
    es = np.array(sorted(agg))
    raw = np.loadtxt(os.path.join(GOLD, "ising2D_8x8.csv"), delimiter=",", skiprows=1)
    # reference pins: logdos(-128) = log 2, logdos(0) = 42.41274640460084 (test_ising.jl:174-191)
    beale = dict(zip(raw[:, 0].astype(int), raw[:, 1]))
    assert beale[-128] == math.log(2) and abs(beale[0] - 42.41274640460084) < 1e-12
    assert -124 not in beale
    for beta in (0.1, 0.3, 0.4406868, 0.6, 1.0):
        z4 = exact.logZ_from_dos(es, np.log([agg[e] for e in es]), beta)
        assert abs(float(exact.kaufman_logZ(4, beta)) - z4) < 1e-12 * abs(z4)
        z8 = exact.logZ_from_dos(raw[:, 0], raw[:, 1], beta)
        assert abs(float(exact.kaufman_logZ(8, beta)) - z8) < 1e-12 * abs(z8)
    E, var, C = exact.kaufman_observables(4, 0.4406868)
    m = exact.ising_moments(d4, 0.4406868, 16)
    assert abs(E - m["E"]) < 1e-9 and abs(C - m["C"]) < 1e-8


def test_bc_dos_fixture_is_complete():
    d = _load("bc_4x4_dos.json")
    rows = np.array(d["rows"])
    assert rows[:, 3].sum() == 3 ** 16
    # all-up state: pair_sum 32, M 16, Q 16, exactly one configuration
    assert [32, 16, 16, 1] in d["rows"]


def _bc_dos():
    d = _load("bc_4x4_dos.json")
    return {(r[0], r[1], r[2]): r[3] for r in d["rows"]}


def _chain_stats(E, M, N, beta, nb=40):
    from montecarlox.measurements import binning_error, jackknife
    e = E.astype(np.float64)
    m = np.abs(M.astype(np.float64)) / N
    out = {"E": (e.mean(), binning_error(e, nb)), "absm": (m.mean(), binning_error(m, nb))}
    out["C"] = jackknife(lambda a, b: beta ** 2 * (b - a * a), [e, e * e], nb)
    mm = M.astype(np.float64) / N
    out["U4"] = jackknife(lambda a, b: 1 - b / (3 * a * a), [mm ** 2, mm ** 4], nb)
    return out


@pytest.mark.parametrize("alg", [METRO, HB, GLAUBER])
@pytest.mark.parametrize("beta", [0.3, 0.4406868, 0.6])
def test_ising_4x4_sampling_matches_exact(orc, alg, beta):
    from oracle import exact
    ex = exact.ising_moments(exact.dos_ising_bruteforce(2, 4), beta, 16)
    s = orc.init_random(ISING, 2, 4, 1, 0)
    orc.canonical_sweeps(ISING, alg, 2, 4, beta, s, 1, 0, 0, 500)
    E, M, _, _ = orc.canonical_sweeps(ISING, alg, 2, 4, beta, s, 1, 0, 500, 60000)
    st = _chain_stats(E, M, 16, beta)
    for k in ("E", "absm", "C", "U4"):
        v, err = st[k]
        assert abs(v - ex[k]) < 4 * err + 1e-12, (k, v, ex[k], err)


@pytest.mark.parametrize("alg", [METRO, HB, GLAUBER])
@pytest.mark.parametrize("beta,D", [(0.5, 0.5), (1.1, 1.6), (0.8, 0.0)])
def test_blume_capel_4x4_sampling_matches_exact(orc, alg, beta, D):
    from montecarlox.measurements import binning_error
    from oracle import exact
    ex = exact.blume_capel_moments(_bc_dos(), beta, D, 16)
    s = orc.init_random(BC, 2, 4, 3, 1)
    orc.canonical_sweeps(BC, alg, 2, 4, beta, s, 3, 1, 0, 500, D=D)
    E, M, Q, _ = orc.canonical_sweeps(BC, alg, 2, 4, beta, s, 3, 1, 500, 60000, D=D)
    e = E.astype(np.float64) + D * Q
    for name, x in (("E", e), ("absm", np.abs(M) / 16.0), ("q", Q / 16.0)):
        err = binning_error(x, 40)
        assert abs(x.mean() - ex[name]) < 4 * err, (name, x.mean(), ex[name], err)


def test_ising_3d_expansions(orc):
    # low temperature: E/N = -3 + 12 p / (1 + p), p = exp(-12 beta) (single flips)
    from montecarlox.measurements import binning_error
    L, N = 4, 64
    s = np.ones(N, np.int8)
    beta = 0.6
    E, _, _, _ = orc.canonical_sweeps(ISING, METRO, 3, L, beta, s, 5, 0, 0, 30000)
    p = math.exp(-12 * beta)
    e = E / N
    assert abs(e.mean() - (-3 + 12 * p / (1 + p))) < 4 * binning_error(e) + 2e-4
    # high temperature: E/N = -3t - 12t^3 + O(t^5), t = tanh(beta)
    beta = 0.05
    s = orc.init_random(ISING, 3, L, 5, 1)
    for alg in (METRO, HB):
        E, _, _, _ = orc.canonical_sweeps(ISING, alg, 3, L, beta, s, 5, 1, 0, 30000)
        t = math.tanh(beta)
        e = E / N
        assert abs(e.mean() - (-3 * t - 12 * t ** 3)) < 4 * binning_error(e) + 1e-4


def test_ising_16_matches_kaufman(orc):
    from montecarlox.measurements import binning_error, jackknife
    from oracle import exact
    beta = 0.4406868
    Ek, _, Ck = exact.kaufman_observables(16, beta)
    s = orc.init_random(ISING, 2, 16, 9, 0)
    orc.canonical_sweeps(ISING, METRO, 2, 16, beta, s, 9, 0, 0, 2000)
    E, _, _, _ = orc.canonical_sweeps(ISING, METRO, 2, 16, beta, s, 9, 0, 2000, 60000)
    e = E.astype(np.float64)
    assert abs(e.mean() - Ek) < 4 * binning_error(e, 40)
    C, dC = jackknife(lambda a, b: beta ** 2 * (b - a * a), [e, e * e], 40)
    assert abs(C - Ck) < 4 * dC


def test_refpath_matches_exact(orc):
    from oracle import exact
    beta = 0.4406868
    ex = exact.ising_moments(exact.dos_ising_bruteforce(2, 4), beta, 16)
    s = np.ones(16, np.int8)
    orc.refpath_series_ising2d(4, beta, 17, 500, s)
    E, M, _ = orc.refpath_series_ising2d(4, beta, 18, 60000, s)
    st = _chain_stats(E, M, 16, beta)
    for k in ("E", "absm", "C"):
        v, err = st[k]
        assert abs(v - ex[k]) < 4 * err, (k, v, ex[k], err)


# --------------------------------------------------------------------------
# generalised ensembles (test/test_multicanonical.jl, test_wang_landau.jl)
# --------------------------------------------------------------------------

def test_muca_flat_weights_accept_everything(orc):
    N = 36
    nb = N + 1  # bins -2N:4:2N
    lw = np.zeros(nb)
    hist = np.zeros(nb, np.int64)
    s = orc.init_random(ISING, 2, 6, 1, 0)
    E, M, Q, steps, acc = orc.generalized_sweeps(ISING, 0, 2, 6, lw, hist, s, 1, 0, 0, 20,
                                                 bin_start=-2 * N, bin_step=4)
    assert steps == 20 * N and acc == steps  # acceptance_rate == 1.0
    assert hist.sum() == steps  # histogram total == steps
    assert E[-1] == -orc.pair_sum(2, 6, s)


def test_muca_bounds_error_leaves_counters(orc):
    lw = np.zeros(3)
    hist = np.zeros(3, np.int64)
    s = np.ones(16, np.int8)  # E = -32 is far below the domain start
    with pytest.raises(orc.BoundsError):
        orc.generalized_sweeps(ISING, 0, 2, 4, lw, hist, s, 1, 0, 0, 1, bin_start=0, bin_step=4)
    assert hist.sum() == 0


def test_muca_update_known_answer(orc):
    # test_multicanonical.jl:53-73
    lw = np.zeros(4)
    h = np.array([0.2, 0.8, 1.1, 2.5])
    orc.muca_update(lw, h)
    assert np.allclose(lw, -np.log(h))
    lw2 = np.array([1.0, 2.0, 3.0])
    orc.muca_update(lw2, np.array([0.0, 1.0, np.e]))
    assert np.allclose(lw2, [1.0, 2.0, 2.0])


def test_wang_landau_decrements_once_per_attempt(orc):
    N = 16
    lw = np.zeros(N + 1)
    hist = np.zeros(N + 1, np.int64)
    s = orc.init_random(ISING, 2, 4, 2, 0)
    _, _, _, steps, _ = orc.generalized_sweeps(ISING, 1, 2, 4, lw, hist, s, 2, 0, 0, 50,
                                               bin_start=-2 * N, bin_step=4, logf=0.5)
    assert -lw.sum() == 0.5 * steps and hist.sum() == steps


def test_wang_landau_converges_to_beale(orc):
    raw = np.loadtxt(os.path.join(GOLD, "ising2D_8x8.csv"), delimiter=",", skiprows=1)
    L, N = 8, 64
    energies = np.arange(-2 * N, 2 * N + 1, 4)
    exact = np.full(len(energies), np.nan)
    exact[((raw[:, 0] + 2 * N) // 4).astype(int)] = raw[:, 1]
    lw = np.zeros(len(energies))
    hist = np.zeros(len(energies), np.int64)
    s = orc.init_random(ISING, 2, L, 4, 0)
    logf, sweep = 1.0, 0
    while logf > 2e-4:
        orc.generalized_sweeps(ISING, 1, 2, L, lw, hist, s, 4, 0, sweep, 3000,
                               bin_start=-2 * N, bin_step=4, logf=logf)
        sweep += 3000
        logf *= 0.5
    ok = np.isfinite(exact)
    est = -(lw - lw[N // 2])  # log g relative to E = 0
    ref = exact - exact[N // 2]
    assert hist[~ok].sum() == 0  # forbidden energies never visited
    rmse = math.sqrt(np.mean((est[ok] - ref[ok]) ** 2))
    assert rmse < 0.15, rmse


def test_blume_capel_muca_histogram_and_bins(orc):
    L, N, D = 4, 16, 1.0
    nb = 4 * N + 1  # E in [-N, 3N]
    lw = np.zeros(nb)
    hist = np.zeros(nb, np.int64)
    s = orc.init_random(BC, 2, L, 6, 0)
    E, M, Q, steps, acc = orc.generalized_sweeps(BC, 0, 2, L, lw, hist, s, 6, 0, 0, 40,
                                                 bin_start=-N, bin_step=1, D=D)
    assert acc == steps == 40 * N and hist.sum() == steps
    assert E[-1] + D * Q[-1] == -orc.pair_sum(2, L, s) + D * orc.sum_sq(2, L, s)


# --------------------------------------------------------------------------
# replica exchange (test_parallel_ensembles.jl:183-206; replica_exchange.jl:158-178)
# --------------------------------------------------------------------------

def test_exchange_known_answers(orc):
    ok, lr = orc.exchange_pair(1.0, 0.5, 0.0, -5.0, 0.0)
    assert ok and abs(lr - 2.5) < 1e-12
    ok, _ = orc.exchange_pair(1.0, 0.5, -5.0, 0.0, 1.0)
    assert not ok


def test_pt_round_schedule(orc):
    n = 4
    idx = np.arange(1, n + 1, dtype=np.int32)
    steps = np.zeros(n - 1, np.int32)
    acc = np.zeros(n - 1, np.int32)
    keys = np.arange(n, dtype=np.uint32)
    same = np.full(n, 0.5)  # equal betas: log ratio 0, every attempt accepted
    st = orc.pt_round(same, np.zeros(n), idx, steps, acc, 0, 1, keys, 0)
    assert st == 1 and list(idx) == [2, 1, 4, 3] and list(steps) == [1, 0, 1]
    st = orc.pt_round(same, np.zeros(n), idx, steps, acc, st, 1, keys, 1)
    assert st == 0 and list(idx) == [3, 1, 4, 2] and list(steps) == [1, 1, 1] and list(acc) == [1, 1, 1]
    # energies strongly favouring the current order are never exchanged
    idx = np.arange(1, n + 1, dtype=np.int32)
    steps[:] = 0
    acc[:] = 0
    betas = np.array([1.0, 0.8, 0.6, 0.4])
    x = np.array([-100.0, -50.0, 0.0, 50.0])
    for r in range(6):
        st = orc.pt_round(betas, x, idx, steps, acc, r % 2, 1, keys, r)
    assert list(idx) == [1, 2, 3, 4] and acc.sum() == 0 and steps.sum() == 9


def test_glauber_acceptance_is_the_logistic_rule(orc):
    """accept!(::Glauber) (metropolis.jl:121-127): a flip is taken iff
    u < logistic(-beta dE).  On one sweep of a 2D lattice, replay every site
    of colour 0 with the site's own uniform and compare; logistic satisfies
    detailed balance p(dE) / p(-dE) = exp(-beta dE)."""
    for beta in (0.2, 0.7, 3.0):
        for dE in (-8, -4, 0, 4, 8):
            pa, pb = orc.logistic(-beta * dE), orc.logistic(beta * dE)
            assert pa / pb == pytest.approx(math.exp(-beta * dE), rel=1e-13)
    L, beta, seed, key = 8, 0.45, 77, 3
    s = orc.init_random(ISING, 2, L, seed, key)
    before = s.copy()
    orc.canonical_sweeps(ISING, GLAUBER, 2, L, beta, s, seed, key, 0, 1)
    # colour 0 sites see only colour-1 neighbours, which are unchanged when colour 0 runs
    for i in range(L * L):
        x, y = i % L, i // L
        g, t, colour = orc.ising_group(2, L, i)
        if colour != 0:
            continue
        nb = (int(before[y * L + (x + 1) % L]) + int(before[y * L + (x - 1) % L])
              + int(before[((y + 1) % L) * L + x]) + int(before[((y - 1) % L) * L + x]))
        dE = 2 * int(before[i]) * nb
        u = orc.bitsliced_word(seed, key, g, t, 0, 1) / 2.0 ** 32  # Metropolis stream
        flipped = s[i] != before[i]
        assert flipped == (u < orc.logistic(-beta * dE)), i


@pytest.mark.parametrize("beta,lw_kind", [(1 / 0.9, "flat"), (0.5, "tilted"), (1 / 0.9, "tilted")])
def test_pair_spin2_sampling_matches_exact(orc, beta, lw_kind):
    """The CustomEnsemble walk of muca_BlumeCapel.jl (weight exp(-beta * H[1] +
    lw(H[2])) with H = (sum_pair_interactions, sum_spins2)) samples, on the
    4x4 lattice, the distribution fixed by the exact joint density of states:
    <P>, <Q>, <|m|> and P(Q) within 4 sigma."""
    from montecarlox.measurements import binning_error
    N, L = 16, 4
    dos = _bc_dos()
    P = np.array([k[0] for k in dos], np.float64)
    Mv = np.array([k[1] for k in dos], np.float64)
    Qv = np.array([k[2] for k in dos], np.int64)
    g = np.array(list(dos.values()), np.float64)
    lw = np.zeros(N + 1)
    if lw_kind == "tilted":
        q = np.arange(N + 1, dtype=np.float64)
        lw = 0.3 * q - 0.02 * q * q
    x = np.log(g) - beta * P + lw[Qv]
    w = np.exp(x - x.max())
    w /= w.sum()
    ex = {"P": (w * P).sum(), "Q": (w * Qv).sum(), "absm": (w * np.abs(Mv) / N).sum()}
    pq = np.bincount(Qv, weights=w, minlength=N + 1)
    s = orc.init_random(BC, 2, L, 21, 4)
    hist = np.zeros(N + 1, np.int64)
    orc.pair_spin2_sweeps(2, L, beta, lw, hist, s, 21, 4, 0, 500, 0, 1)
    hist[:] = 0
    E, M, Q, steps, acc = orc.pair_spin2_sweeps(2, L, beta, lw, hist, s, 21, 4, 500, 60000, 0, 1)
    assert steps == 60000 * N and hist.sum() == steps
    for name, v in (("P", -E.astype(np.float64)), ("Q", Q.astype(np.float64)),
                    ("absm", np.abs(M) / float(N))):
        err = binning_error(v, 40)
        assert abs(v.mean() - ex[name]) < 4 * err, (name, v.mean(), ex[name], err)
    # per-attempt visits: P(Q) of the visited states
    freq = hist / hist.sum()
    big = pq > 0.02
    se = np.sqrt(pq * (1 - pq) / (60000 / 20.0))  # ~20 sweeps of correlation, generous
    assert np.all(np.abs(freq[big] - pq[big]) < 4 * se[big]), (freq[big], pq[big])
    if beta > 1:
        assert ex["P"] < 0  # the pair weight -beta * H[1] favours anti-aligned pairs
