"""Pins of the oracle's weight-iteration schedules (test infrastructure) by
what the mathematics fixes: the converged Wang-Landau log-weights and the
iterated multicanonical weights are -ln g(E) + const, with g the exact density
of states -- brute force for the 4x4 Ising and Blume-Capel tori, Beale's data
(the reference's own ising2D_8x8.csv) for 8x8.  The Wang-Landau flatness
schedule has no counterpart in the reference, so these are its only pins
besides the GPU parity tests."""
import json
import math
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ISING, BC = 0, 1


def _rmse(lw, logg, ok):
    """RMSE of -lw against ln g over the bins `ok`, both shifted to mean 0."""
    est = -lw[ok] - np.mean(-lw[ok])
    ref = logg[ok] - np.mean(logg[ok])
    return math.sqrt(np.mean((est - ref) ** 2))


def _ising_logg(L):
    from oracle import exact
    N = L * L
    g = np.zeros(N + 1)
    for (E, _), c in exact.dos_ising_bruteforce(2, L).items():
        g[(E + 2 * N) // 4] += c
    with np.errstate(divide="ignore"):
        return np.where(g > 0, np.log(g), np.nan)


def _bc_logg(D=1.0):
    d = json.load(open(os.path.join(GOLD, "bc_4x4_dos.json")))
    N = 16
    g = np.zeros(4 * N + 1)
    for pair, _, Q, c in d["rows"]:
        g[int(-pair + D * Q) + N] += c
    with np.errstate(divide="ignore"):
        return np.where(g > 0, np.log(g), np.nan)


def _walkers(orc, model, L, n, seed):
    spins = [orc.init_random(model, 2, L, seed, k) for k in range(n)]
    return spins


def test_wl_schedule_converges_on_ising_4x4(orc):
    L, N, n = 4, 16, 4
    logg = _ising_logg(L)
    spins = _walkers(orc, ISING, L, n, 5)
    lw = [np.zeros(N + 1) for _ in range(n)]
    hists = [np.zeros(N + 1, np.int64) for _ in range(n)]
    reached = np.zeros(N + 1, bool)
    logf, sweep, upd, stages, steps, acc = orc.wl_iterate_group(
        ISING, 2, L, spins, lw, hists, 1.0, reached, 5, list(range(n)), 0, -2 * N, 4, 100, 5000,
        0.8, 1e-6, 0.5)
    assert logf < 1e-6 and stages == 20
    assert all(np.array_equal(lw[0], x) for x in lw)  # shared after every update
    ok = np.isfinite(logg)
    # the forbidden energies (+-(2N - 4)) were never reached
    assert np.array_equal(reached, ok)
    rmse = _rmse(lw[0], logg, ok)
    print(f"Ising 4x4: {upd} updates, rmse {rmse:.4f}")
    assert rmse < 0.1


def test_wl_schedule_converges_on_blume_capel_4x4(orc):
    N, n = 16, 4
    logg = _bc_logg(1.0)
    spins = _walkers(orc, BC, 4, n, 9)
    lw = [np.zeros(4 * N + 1) for _ in range(n)]
    hists = [np.zeros(4 * N + 1, np.int64) for _ in range(n)]
    reached = np.zeros(4 * N + 1, bool)
    logf, *_ = orc.wl_iterate_group(BC, 2, 4, spins, lw, hists, 1.0, reached, 9, list(range(n)),
                                   0, -N, 1, 100, 5000, 0.8, 1e-5, 0.5, D=1.0)
    assert logf < 1e-5
    ok = np.isfinite(logg)
    assert np.array_equal(reached, ok)
    rmse = _rmse(lw[0], logg, ok)
    print(f"Blume-Capel 4x4: rmse {rmse:.4f}")
    assert rmse < 0.08


def test_wl_schedule_converges_to_beale_8x8(orc):
    raw = np.loadtxt(os.path.join(GOLD, "ising2D_8x8.csv"), delimiter=",", skiprows=1)
    L, N, n = 8, 64, 4
    logg = np.full(N + 1, np.nan)
    logg[((raw[:, 0] + 2 * N) // 4).astype(int)] = raw[:, 1]
    spins = _walkers(orc, ISING, L, n, 13)
    lw = [np.zeros(N + 1) for _ in range(n)]
    hists = [np.zeros(N + 1, np.int64) for _ in range(n)]
    reached = np.zeros(N + 1, bool)
    logf, *_ = orc.wl_iterate_group(ISING, 2, L, spins, lw, hists, 1.0, reached, 13,
                                   list(range(n)), 0, -2 * N, 4, 200, 20000, 0.8, 1e-5, 0.5)
    assert logf < 1e-5
    ok = np.isfinite(logg)
    rmse = _rmse(lw[0], logg, ok)
    print(f"Ising 8x8: rmse {rmse:.4f}")
    assert rmse < 0.1


def test_muca_iteration_converges_on_ising_4x4(orc):
    L, N, n = 4, 16, 4
    logg = _ising_logg(L)
    spins = _walkers(orc, ISING, L, n, 21)
    lw = [np.zeros(N + 1) for _ in range(n)]
    hists = [np.zeros(N + 1, np.int64) for _ in range(n)]
    sweep = 0
    for _ in range(12):
        sweep, steps, acc = orc.muca_iterate_update(ISING, 2, L, spins, lw, hists, 21,
                                                    list(range(n)), sweep, -2 * N, 4, 2000, 10)
        assert all(np.array_equal(lw[0], x) for x in lw)
        # merge_histograms!: the root holds the sum, the others their own
        assert hists[0].sum() == sum(steps)
    ok = np.isfinite(logg)
    rmse = _rmse(lw[0], logg, ok)
    print(f"muca Ising 4x4: rmse {rmse:.4f}")
    assert rmse < 0.1


def test_group_stats_rule(orc):
    """The flatness statistics count a bin from the first evaluation that sees
    it non-zero on, also after the histograms are reset."""
    reached = np.zeros(5, bool)
    h = [np.array([0, 3, 5, 0, 2], np.int64), np.array([0, 1, 0, 0, 2], np.int64)]
    assert orc._gen_group_stats(h, reached) == (4.0, 13.0 / 3.0, 3)
    assert list(reached) == [False, True, True, False, True]
    h = [np.array([1, 0, 4, 0, 0], np.int64)]
    assert orc._gen_group_stats(h, reached) == (0.0, 5.0 / 4.0, 4)
