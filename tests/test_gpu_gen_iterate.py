"""The weight iteration on the device (mcx_gen_iterate) and its pieces against
the oracle's sequential restatement: the Wang-Landau flatness schedule bit for
bit (spins, weights, histograms, ln f, counters, reached bins), the parallel
multicanonical update (integers bit for bit, the update's device log within a
few ulp), the flatness statistics against a host computation, the schedule's
convergence to exact densities of states, and the binning rules of Ising
walkers (integer-typed vs Float64 binnings, steps that are not powers of two)."""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
ISING, BC = 0, 1


def _walkers(mcx, model, L, n, kind, bins, seed, lw0=None, logf=1.0, D=0.0, dim=2):
    """n walkers of one model; bins = one DiscreteBinning, or a list (one per
    walker: walkers of equal binnings form a group)."""
    from montecarlox.spinsystems import SpinSystemBatch
    Ls = L if isinstance(L, (list, tuple)) else [L] * n
    sys = SpinSystemBatch(model, [Ls[0]] * dim, n=n, L=np.array(Ls), D=D)
    algs = []
    for k in range(n):
        b = bins[k] if isinstance(bins, list) else bins
        bo = mcx.BinnedObject(b, 0.0)
        if lw0 is not None:
            bo.values[:] = lw0[k]
        if kind == "muca":
            algs.append(mcx.Multicanonical(mcx.Philox(seed, k), bo))
        else:
            algs.append(mcx.WangLandau(mcx.Philox(seed, k), bo, logf=logf))
    sys.init_("random", [a.rng for a in algs])
    return sys, algs


def _host_stats(h, reached):
    reached |= h > 0
    if not reached.any():
        return 0.0, 0.0, 0
    return float(h[reached].min()), float(h[reached].sum()) / reached.sum(), int(reached.sum())


@pytest.mark.parametrize("lw_scale", [0.0, 0.7])
def test_flatness_equals_host_statistics(mcx, lw_scale):
    """mcx_gen_flatness == (min, mean, bins) of the root histogram over the bins
    seen non-zero so far, for zero and for non-zero starting weights (which no
    longer make every bin count)."""
    from montecarlox.ensembles import DiscreteBinning
    N = 36
    rng = np.random.default_rng(3)
    lw0 = rng.normal(0, lw_scale, size=(2, N + 1))
    print(f"flatness lw_scale={lw_scale}", flush=True)
    sys, algs = _walkers(mcx, "ising", 6, 2, "muca", DiscreteBinning(-2 * N, 4, N + 1), 7, lw0)
    w = sys.walkers(algs)
    reached = np.zeros(N + 1, bool)
    for sweeps in (3, 40, 2):
        mcx.sweep_(sys, algs, sweeps)
        h = np.asarray(algs[0].ensemble.histogram.values, np.int64)
        exp = _host_stats(h, reached)
        assert w.flatness(0) == exp, (sweeps, exp)
        mcx.reset_(algs[0])  # histogram to zero: reached bins keep counting (as 0)
    assert w.flatness(0) == (0.0, 0.0, int(reached.sum()))
    assert 0 < reached.sum() < N + 1


def _wl_case(model, L, n, groups, D=0.0):
    """(bins per walker, group members) for `groups` groups of n walkers each,
    lattice L[g] per group."""
    from montecarlox.ensembles import DiscreteBinning
    bins, Ls, members = [], [], []
    for g in range(groups):
        N = L[g] ** 2
        b = DiscreteBinning(-2 * N, 4, N + 1) if model == "ising" else \
            DiscreteBinning(-float(N), 1.0, 4 * N + 1)
        members.append(list(range(len(bins), len(bins) + n)))
        bins += [b] * n
        Ls += [L[g]] * n
    return bins, Ls, members


@pytest.mark.parametrize("model,L,D", [("ising", (4, 6), 0.0), ("blume_capel", (4, 6), 1.0)])
def test_wang_landau_iteration_equals_oracle(mcx, orc, model, L, D):
    n, groups, seed, sweeps = 3, 2, 29, 30
    code = ISING if model == "ising" else BC
    bins, Ls, members = _wl_case(model, L, n, groups, D)
    print(f"WL iterate {model} L={L}", flush=True)
    sys, algs = _walkers(mcx, model, Ls, n * groups, "wl", bins, seed, D=D)
    init = [sys.batch.get_spins(k) for k in range(n * groups)]
    w = sys.walkers(algs)
    pr = w.iterate_(sweeps, 400, flatness=0.8, logf_final=1e-3, power=0.5)
    tot_updates = tot_stages = still = 0
    for g in range(groups):
        m = members[g]
        b = bins[m[0]]
        nb = b.num
        spins = [init[k].copy() for k in m]
        lw = [np.zeros(nb) for _ in m]
        hists = [np.zeros(nb, np.int64) for _ in m]
        reached = np.zeros(nb, bool)
        logf, sweep, upd, stages, steps, acc = orc.wl_iterate_group(
            code, 2, L[g], spins, lw, hists, 1.0, reached, seed, m, 0, b.start, b.step, sweeps,
            400, 0.8, 1e-3, 0.5, D=D)
        tot_updates = max(tot_updates, upd)
        tot_stages += stages
        still += logf >= 1e-3
        assert stages >= 5
        for j, k in enumerate(m):
            e = algs[k].ensemble
            assert np.array_equal(sys.batch.get_spins(k), spins[j]), f"walker {k} spins"
            assert np.array_equal(e.logweight.values, lw[j]), f"walker {k} weights"
            assert np.array_equal(np.asarray(e.device_histogram, np.int64), hists[j])
            assert e.logf == logf
            assert (algs[k].steps, algs[k].accepted) == (steps[j], acc[j])
        # the device's reached bins: flatness of the (zeroed) root histogram
        assert w.flatness(g)[2] == int(reached.sum())
    assert pr["updates"] == tot_updates and pr["stages"] == tot_stages
    assert pr["active_groups"] == still


def test_multicanonical_iteration_equals_oracle(mcx, orc):
    """Three updates of the parallel multicanonical iteration (thermalise,
    reset!, record, merge_histograms!, update! on the root,
    distribute_logweight!) on two groups, each update checked against the
    oracle continued from the device's weights."""
    from montecarlox.ensembles import DiscreteBinning
    n, seed = 3, 41
    b6, b8 = DiscreteBinning(-72, 4, 37), DiscreteBinning(-128, 4, 65)
    bins = [b6] * n + [b8] * n
    sys, algs = _walkers(mcx, "ising", [6] * n + [8] * n, 2 * n, "muca", bins, seed)
    w = sys.walkers(algs)
    spins = [sys.batch.get_spins(k) for k in range(2 * n)]
    sweep = [0, 0]
    for it in range(3):
        lw_dev = [a.ensemble.logweight.values.copy() for a in algs]
        print(f"muca iterate update {it}", flush=True)
        pr = w.iterate_(25, 1, therm_sweeps=5)
        assert pr["updates"] == 1 and pr["sweeps"] == 30
        for g, (L, b) in enumerate(((6, b6), (8, b8))):
            m = list(range(g * n, (g + 1) * n))
            lw = [lw_dev[k].copy() for k in m]
            hists = [np.zeros(b.num, np.int64) for _ in m]
            sp = [spins[k] for k in m]
            sweep[g], st, ac = orc.muca_iterate_update(ISING, 2, L, sp, lw, hists, seed, m,
                                                       sweep[g], b.start, b.step, 25, 5)
            for j, k in enumerate(m):
                e = algs[k].ensemble
                assert np.array_equal(sys.batch.get_spins(k), sp[j]), (it, k)
                assert np.array_equal(np.asarray(e.histogram.values, np.int64), hists[j]), (it, k)
                assert (algs[k].steps, algs[k].accepted) == (st[j], ac[j])
                got = e.logweight.values
                assert np.all(np.abs(got - lw[j]) <= 4 * np.spacing(np.abs(lw[j])) + 1e-300)
                assert np.array_equal(got, algs[m[0]].ensemble.logweight.values)


def test_wang_landau_iteration_converges_to_beale(mcx):
    """Ising 8x8, 16 walkers, ln f from 1 to below 1e-6 with flatness 0.8: the
    shared log-weights give the reference's exact Beale density of states."""
    from montecarlox.ensembles import DiscreteBinning
    raw = np.loadtxt(os.path.join(GOLD, "ising2D_8x8.csv"), delimiter=",", skiprows=1)
    L, N, n = 8, 64, 16
    exact = np.full(N + 1, np.nan)
    exact[((raw[:, 0] + 2 * N) // 4).astype(int)] = raw[:, 1]
    print("WL iterate Ising 8x8 to 1e-6", flush=True)
    sys, algs = _walkers(mcx, "ising", L, n, "wl", DiscreteBinning(-2 * N, 4, N + 1), 13)
    w = sys.walkers(algs)
    pr = w.iterate_(200, 2000, flatness=0.8, logf_final=1e-6, power=0.5)
    assert pr["active_groups"] == 0 and pr["stages"] == 20
    lw = algs[0].ensemble.logweight.values
    ok = np.isfinite(exact)
    est = -lw[ok] - np.mean(-lw[ok])
    ref = exact[ok] - np.mean(exact[ok])
    rmse = math.sqrt(np.mean((est - ref) ** 2))
    print(f"updates {pr['updates']}, rmse {rmse:.4f}")
    assert rmse < 0.1
    assert w.flatness(0)[2] == int(ok.sum())  # the two forbidden energies never reached


def test_wang_landau_iteration_converges_blume_capel_4x4(mcx):
    from montecarlox.ensembles import DiscreteBinning
    d = json.load(open(os.path.join(GOLD, "bc_4x4_dos.json")))
    N, n = 16, 8
    g = np.zeros(4 * N + 1)
    for pair, _, Q, c in d["rows"]:
        g[int(-pair + Q) + N] += c
    ok = g > 0
    print("WL iterate Blume-Capel 4x4", flush=True)
    sys, algs = _walkers(mcx, "blume_capel", 4, n, "wl", DiscreteBinning(-float(N), 1.0, 4 * N + 1),
                         17, D=1.0)
    w = sys.walkers(algs)
    pr = w.iterate_(100, 3000, flatness=0.8, logf_final=1e-6, power=0.5)
    assert pr["active_groups"] == 0
    lw = algs[0].ensemble.logweight.values
    est = -lw[ok] - np.mean(-lw[ok])
    ref = np.log(g[ok]) - np.mean(np.log(g[ok]))
    rmse = math.sqrt(np.mean((est - ref) ** 2))
    print(f"updates {pr['updates']}, rmse {rmse:.4f}")
    assert rmse < 0.1


def test_iterate_rejects_bad_schedules(mcx):
    from montecarlox.ensembles import DiscreteBinning
    sys, algs = _walkers(mcx, "ising", 4, 2, "wl", DiscreteBinning(-32, 4, 17), 1)
    w = sys.walkers(algs)
    for kw in ({"sweeps": 0, "max_updates": 1}, {"sweeps": 1, "max_updates": 0},
               {"sweeps": 1, "max_updates": 1, "flatness": 1.5},
               {"sweeps": 1, "max_updates": 1, "power": 1.0},
               {"sweeps": 1, "max_updates": 1, "logf_final": 0.0}):
        with pytest.raises(mcx.MCXArgumentError):
            w.iterate_(**kw)


@pytest.mark.parametrize("kind", ["muca", "wl"])
@pytest.mark.parametrize("start,step,integer", [(-73, 3, True), (72, -4, True), (-74.0, 4.0, False),
                                               (-72.0, 1.0, False), (-73.5, 3.0, False)])
def test_ising_binning_rules(mcx, orc, kind, start, step, integer):
    """_binindex on Ising energies: an integer-typed binning divides (truncating,
    also for steps that are not powers of two and negative steps); a Float64
    binning rounds half to even (-74.0 / -73.5: every energy sits half a step
    off a centre) -- from the walker's first attempt on."""
    from montecarlox.ensembles import DiscreteBinning
    L, N, n, seed, sweeps = 6, 36, 2, 23, 15
    energies = range(-2 * N, 2 * N + 1, 4)
    if integer:
        idx = [math.trunc((e - start) / step) for e in energies]
    else:
        idx = [round((e - start) / step) for e in energies]  # half to even
    assert min(idx) >= 0
    nb = max(idx) + 1
    bins = DiscreteBinning(start, step, nb)
    assert bins.integer == integer
    rng = np.random.default_rng(5)
    lw0 = rng.normal(0, 0.5, size=(n, nb))
    print(f"Ising binning start={start} step={step} integer={integer} {kind}", flush=True)
    sys, algs = _walkers(mcx, "ising", L, n, kind, bins, seed, lw0, logf=0.5)
    init = [sys.batch.get_spins(k) for k in range(n)]
    mcx.sweep_(sys, algs, sweeps)
    for k in range(n):
        s = init[k].copy()
        lw = lw0[k].copy()
        h = np.zeros(nb, np.int64)
        _, _, _, st, ac = orc.generalized_sweeps(ISING, 1 if kind == "wl" else 0, 2, L, lw, h, s,
                                                 seed, k, 0, sweeps, bin_start=start,
                                                 bin_step=step, logf=0.5, int_bins=integer)
        e = algs[k].ensemble
        hg = e.histogram.values if kind == "muca" else e.device_histogram
        assert np.array_equal(s, sys.batch.get_spins(k))
        assert np.array_equal(h, np.asarray(hg, np.int64))
        assert np.array_equal(lw, e.logweight.values)
        assert (algs[k].steps, algs[k].accepted) == (st, ac)
