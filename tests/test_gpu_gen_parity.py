"""Multicanonical / Wang-Landau walkers on the GPU vs the sequential oracle.

Integer outputs (spins, histograms, E/M/Q series, counters) and Wang-Landau
weights (exact FP64 subtractions) are compared bit for bit; the multicanonical
recursion lw -= log(h) is compared to 4 ulp (device log vs libm log).
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _walkers(mcx, model, dim, L, n, kind, start, step, nbins, seed, lw0=None, logf=1.0, D=0.0,
             cap=0):
    from montecarlox.spinsystems import SpinSystemBatch
    from montecarlox.ensembles import DiscreteBinning
    sys = SpinSystemBatch(model, [L] * dim, n=n, D=D, series_capacity=cap)
    algs = []
    for k in range(n):
        # a range domain start:step:stop carries its step (binned_object.jl:44-46)
        bo = mcx.BinnedObject(DiscreteBinning(start, step, nbins), 0.0)
        if lw0 is not None:
            bo.values[:] = lw0[k]
        if kind == "muca":
            algs.append(mcx.Multicanonical(mcx.Philox(seed, k), bo))
        else:
            algs.append(mcx.WangLandau(mcx.Philox(seed, k), bo, logf=logf))
    sys.init_("random", [a.rng for a in algs])
    return sys, algs


def _check(mcx, orc, model, dim, L, n, kind, start, step, nbins, seed, nsweeps, D=0.0, logf=1.0,
           lw_scale=0.3, runs=None, lw0=None):
    rng = np.random.default_rng(seed)
    if lw0 is None:
        lw0 = rng.normal(0, lw_scale, size=(n, nbins))
    sys, algs = _walkers(mcx, model, dim, L, n, kind, start, step, nbins, seed, lw0, logf, D,
                         cap=nsweeps)
    init = [sys.batch.get_spins(k) for k in range(n)]
    print(f"gen model={model} dim={dim} L={L} kind={kind} nbins={nbins}", flush=True)
    for p in (runs or [nsweeps]):
        mcx.sweep_(sys, algs, p, record=True)
    code = {"ising": 0, "blume_capel": 1}[model]
    for k in range(n):
        s = init[k].copy()
        lw = lw0[k].copy()
        h = np.zeros(nbins, np.int64)
        E, M, Q, st, ac = orc.generalized_sweeps(code, 1 if kind == "wl" else 0, dim, L, lw, h, s,
                                                 seed, k, 0, nsweeps, bin_start=start,
                                                 bin_step=step, logf=logf, D=D)
        e = algs[k].ensemble
        hg = e.histogram.values if kind == "muca" else e.device_histogram
        assert np.array_equal(s, sys.batch.get_spins(k)), f"walker {k} spins"
        assert np.array_equal(h, np.asarray(hg, np.int64)), f"walker {k} histogram"
        assert np.array_equal(lw, e.logweight.values, equal_nan=True), f"walker {k} log-weights"
        Eg, Mg, Qg = sys.batch.series(k)
        assert np.array_equal(E, Eg) and np.array_equal(M, Mg)
        if Qg is not None:
            assert np.array_equal(Q, Qg)
        assert algs[k].steps == st and algs[k].accepted == ac
    return sys, algs


@pytest.mark.parametrize("kind", ["muca", "wl"])
def test_ising_2d_walkers(mcx, orc, kind):
    N = 64
    _check(mcx, orc, "ising", 2, 8, 3, kind, -2 * N, 4, N + 1, 21, 12)


@pytest.mark.parametrize("kind", ["muca", "wl"])
@pytest.mark.parametrize("off,step", [(0, 1), (0, 2), (2, 4), (0, 8)])
def test_ising_integer_walk_bin_layouts(mcx, orc, kind, off, step):
    """Integer-energy walks whose bins follow the energy linearly (a step of
    1 or 2 dividing E - start: the walk carries the bin alone) and whose bins
    do not (step 4 with E - start = 2 mod 4; step 8)."""
    N = 64
    start = -2 * N - off
    nbins = (4 * N + off) // step + 2
    _check(mcx, orc, "ising", 2, 8, 2, kind, start, step, nbins, 23, 10)


@pytest.mark.parametrize("kind", ["muca", "wl"])
@pytest.mark.parametrize("model", ["ising", "blume_capel"])
def test_nan_weights_are_rejected(mcx, orc, kind, model):
    """Bins whose log-weight is NaN: a proposal into one fails both of the
    walk's fast compares (the multicanonical pass finds such an attempt
    afterwards from its records) and is settled exactly -- rejected, since
    lr > 0 || u < exp(lr) is false for lr = NaN -- and the walk goes on."""
    n, seed = 2, 29
    if model == "ising":
        L, N = 8, 64
        start, step, nbins, D = -2 * N, 4, N + 1, 0.0
    else:
        L, N = 6, 36
        start, step, D = -2.0 * N, 1.0, 1.0
        nbins = int(round(3 * N - start)) + 2
    lw0 = np.random.default_rng(seed).normal(0, 0.3, size=(n, nbins))
    lw0[:, 3::7] = np.nan
    # the walkers' starting bins keep a number (the same random start as _walkers)
    sys, _ = _walkers(mcx, model, 2, L, n, kind, start, step, nbins, seed, D=D)
    for k in range(n):
        s0 = sys.batch.get_spins(k).astype(np.int64).reshape(L, L)
        E = -(s0 * np.roll(s0, 1, 0)).sum() - (s0 * np.roll(s0, 1, 1)).sum() + D * (s0 ** 2).sum()
        b = int((E - start) // step)
        lw0[k, max(0, b - 2):b + 3] = 0.0
    sys.batch.close()
    _check(mcx, orc, model, 2, L, n, kind, start, step, nbins, seed, 8, D=D, logf=0.5, lw0=lw0)


@pytest.mark.parametrize("kind", ["muca", "wl"])
def test_ising_3d_walkers(mcx, orc, kind):
    N = 64
    _check(mcx, orc, "ising", 3, 4, 2, kind, -3 * N, 4, 3 * N // 2 + 1, 5, 6)


@pytest.mark.parametrize("kind", ["muca", "wl"])
@pytest.mark.parametrize("D", [1.0, 0.3])
def test_blume_capel_walkers(mcx, orc, kind, D):
    L, N = 6, 36
    # domain wide enough for every state: E in [-2N, 2N + D N]
    start, step = -2.0 * N, (1.0 if D == 1.0 else 0.1)
    nbins = int(round((2 * N + D * N - start) / step)) + 2
    _check(mcx, orc, "blume_capel", 2, L, 2, kind, start, step, nbins, 9, 5, D=D, logf=0.25)


def test_histogram_in_hbm_path(mcx, orc):
    # L = 64 Blume-Capel: 16385 bins -- the uint32 LDS histogram does not fit, int64 atomics in HBM
    L, N = 64, 4096
    _check(mcx, orc, "blume_capel", 2, L, 1, "muca", -N * 1.0, 1.0, 4 * N + 1, 3, 2, D=1.0)


@pytest.mark.parametrize("kind,lw_scale", [("wl", 0.0), ("muca", 0.0), ("muca", 1.5)])
def test_windowed_weights_blume_capel(mcx, orc, kind, lw_scale):
    """4097 bins (33 KB of weights): the walker keeps a window of 2048 bins in
    LDS and slides it as its energy moves (Wang-Landau with ln f = 1 wanders
    across the whole range); results equal the oracle's."""
    L, N = 32, 1024
    _check(mcx, orc, "blume_capel", 2, L, 2, kind, -1.0 * N, 1.0, 4 * N + 1, 13, 60, D=1.0,
           logf=1.0, lw_scale=lw_scale)


def test_windowed_weights_ising_and_fp64_bins(mcx, orc):
    # Ising L=64, step 4: 4097 bins (integer walk); Blume-Capel 3D with step 0.5:
    # 13825 bins, FP64 energies (rounded bins, the generic walk)
    _check(mcx, orc, "ising", 2, 64, 2, "wl", -2 * 4096, 4, 4097, 17, 30, logf=1.0, lw_scale=0.0)
    L, N = 12, 1728
    _check(mcx, orc, "blume_capel", 3, L, 1, "wl", -1.0 * N, 0.5, 8 * N + 1, 19, 25, D=0.5,
           logf=1.0, lw_scale=0.0)


def test_split_runs(mcx, orc):
    _check(mcx, orc, "ising", 2, 8, 2, "wl", -128, 4, 65, 4, 9, runs=[3, 6], logf=0.5)


def test_bounds_error(mcx):
    sys, algs = _walkers(mcx, "ising", 2, 4, 1, "muca", 0, 4, 3, 1)  # domain [0, 8]
    sys.batch.set_spins(0, np.ones(16, np.int8))  # E = -32: outside
    with pytest.raises(mcx.MCXBoundsError):
        mcx.sweep_(sys, algs, 1)
    assert algs[0].steps == 0 and algs[0].ensemble.histogram.values.sum() == 0


@pytest.mark.parametrize("model,kind,start,step,nbins,D", [
    # the oracle stops these after 195, 264, 5601 and 1772 attempts (chunks of 64)
    ("ising", "muca", -40, 4, 21, 0.0),            # integer walk
    ("ising", "wl", -60, 4, 31, 0.0),
    ("blume_capel", "muca", 10.0, 0.1, 801, 0.3),  # FP64 energies: the generic walk
    ("blume_capel", "wl", 10.0, 0.1, 801, 0.3),
])
def test_bounds_error_mid_run(mcx, orc, model, kind, start, step, nbins, D):
    """Flat weights accept every move, so the walker's energy wanders out of a
    narrow domain somewhere in a later chunk or sweep; the run stops at the
    same attempt as the oracle's (same site) with the same spins."""
    import re
    L, seed = 16, 21
    sys, algs = _walkers(mcx, model, 2, L, 1, kind, start, step, nbins, seed,
                         np.zeros((1, nbins)), 0.5, D)
    s = sys.batch.get_spins(0)
    print(f"gen bounds mid-run model={model} kind={kind}", flush=True)
    with pytest.raises(mcx.MCXBoundsError) as ei:
        mcx.sweep_(sys, algs, 40)
    h = np.zeros(nbins, np.int64)
    code = {"ising": 0, "blume_capel": 1}[model]
    with pytest.raises(orc.BoundsError) as eo:
        orc.generalized_sweeps(code, 1 if kind == "wl" else 0, 2, L, np.zeros(nbins), h, s, seed,
                               0, 0, 40, bin_start=start, bin_step=step, logf=0.5, D=D)
    site_dev = int(re.search(r"site (\d+)", str(ei.value)).group(1))
    site_orc = int(re.search(r"site (\d+)", str(eo.value)).group(1))
    assert site_dev == site_orc
    assert np.array_equal(s, sys.batch.get_spins(0))
    # the walk got past the first chunk (the error is not the initial state's)
    assert h.sum() > 64


def test_flat_weights_accept_everything(mcx):
    sys, algs = _walkers(mcx, "ising", 2, 6, 2, "muca", -72, 4, 37, 8)
    mcx.sweep_(sys, algs, 10)
    for a in algs:
        assert a.steps == 360 and mcx.acceptance_rate(a) == 1.0
        assert a.ensemble.histogram.values.sum() == a.steps
    mcx.reset_(algs[0])
    assert algs[0].steps == 0 and algs[0].ensemble.histogram.values.sum() == 0


def test_update_merge_distribute(mcx, orc):
    # parallel_multicanonical KATs (test_parallel_ensembles.jl:95-113) on device walkers
    sys, algs = _walkers(mcx, "ising", 2, 4, 2, "muca", -32, 4, 4, 2)
    w = sys.walkers(algs)
    algs[0].ensemble.histogram.values[:] = [1, 2, 3, 4]
    algs[1].ensemble.histogram.values[:] = [4, 3, 2, 1]
    pc = mcx.ParallelMulticanonical(mcx.ThreadsBackend(2), algs)
    mcx.merge_histograms_(pc)
    assert list(algs[0].ensemble.histogram.values) == [5, 5, 5, 5]
    assert list(algs[1].ensemble.histogram.values) == [4, 3, 2, 1]
    algs[0].ensemble.logweight.values[:] = [1.0, 2.0, 3.0, 4.0]
    mcx.distribute_logweight_(pc)
    assert list(algs[1].ensemble.logweight.values) == [1.0, 2.0, 3.0, 4.0]
    # update!: lw -= log(h) where h > 0, on the device
    algs[1].ensemble.histogram.values[:] = [0, 7, 1, 3]
    mcx.update_(algs[1].ensemble)
    ref = np.array([1.0, 2.0, 3.0, 4.0])
    orc.muca_update(ref, np.array([0.0, 7.0, 1.0, 3.0]))
    got = algs[1].ensemble.logweight.values
    assert np.all(np.abs(got - ref) <= 4 * np.spacing(np.abs(ref)))
    # unbound ensemble with a Float64 histogram (test_multicanonical.jl:53-73)
    e = mcx.MulticanonicalEnsemble(mcx.BinnedObject(np.arange(0.0, 5.0).tolist(), 0.0))  # 0.0:1.0:4.0
    e.histogram.values[:] = [0.2, 0.8, 1.1, 2.5]
    assert e.update_() is None
    exp = -np.log([0.2, 0.8, 1.1, 2.5])
    assert np.all(np.abs(e.logweight.values - exp) <= 4 * np.spacing(np.abs(exp)))


def test_wang_landau_converges_on_gpu(mcx):
    raw = np.loadtxt(os.path.join(GOLD, "ising2D_8x8.csv"), delimiter=",", skiprows=1)
    L, N, n = 8, 64, 8
    nb = N + 1
    exact = np.full(nb, np.nan)
    exact[((raw[:, 0] + 2 * N) // 4).astype(int)] = raw[:, 1]
    sys, algs = _walkers(mcx, "ising", 2, L, n, "wl", -2 * N, 4, nb, 13, logf=1.0)
    w = sys.walkers(algs)
    logf = 1.0
    while logf > 1e-4:
        for a in algs:
            a.ensemble.logf = logf
        mcx.sweep_(sys, algs, 2000)
        w.average_(0)
        logf *= 0.5
    lw = algs[0].ensemble.logweight.values
    ok = np.isfinite(exact)
    est = -(lw - lw[N // 2])
    ref = exact - exact[N // 2]
    rmse = math.sqrt(np.mean((est[ok] - ref[ok]) ** 2))
    print("rmse", rmse)
    assert rmse < 0.1


def _pair_spin2_walkers(mcx, dim, L, n, T, seed, lw0, start=0, cap=0):
    """Metropolis(rng, CustomEnsemble(BoltzmannEnsemble(T=T),
    MulticanonicalEnsemble(start:1:N))) per system, muca_BlumeCapel.jl:43-76."""
    from montecarlox.spinsystems import SpinSystemBatch
    N = L ** dim
    sys = SpinSystemBatch("blume_capel", [L] * dim, n=n, series_capacity=cap)
    algs = []
    for k in range(n):
        spin2 = mcx.MulticanonicalEnsemble(range(start, N + 1))
        spin2.logweight.values[:] = lw0[k]
        ens = mcx.CustomEnsemble(mcx.BoltzmannEnsemble(T=T), spin2, True)
        algs.append(mcx.Metropolis(mcx.Philox(seed, k), ens))
    sys.init_("random", [a.rng for a in algs])
    return sys, algs


@pytest.mark.parametrize("dim,L,n,T", [(2, 8, 3, 0.9), (2, 16, 2, 0.6), (3, 4, 2, 1.5),
                                       (2, 6, 2, 0.9), (2, 64, 1, 0.9)])
def test_pair_spin2_walkers_match_oracle(mcx, orc, dim, L, n, T):
    """The muca_BlumeCapel.jl iteration on device walkers -- sweeps, update! of
    the spin2 weights, reset!(alg) (counters only), reset_histogram!, more
    sweeps -- against the oracle's sequential restatement, bit for bit (the
    update's device log within 4 ulp, then continued from the device weights)."""
    N, seed, n1, n2 = L ** dim, 31, 25, 15
    rng = np.random.default_rng(L * 10 + dim)
    lw0 = rng.normal(0, 0.4, size=(n, N + 1))
    print(f"pair+spin2 walkers dim={dim} L={L} n={n} T={T}", flush=True)
    sys, algs = _pair_spin2_walkers(mcx, dim, L, n, T, seed, lw0, cap=n1 + n2)
    init = [sys.batch.get_spins(k) for k in range(n)]
    mcx.sweep_(sys, algs, n1, record=True)
    beta = 1.0 / T
    spins = [s.copy() for s in init]
    hists = [np.zeros(N + 1, np.int64) for _ in range(n)]
    ser = []
    for k in range(n):
        E, M, Q, st, ac = orc.pair_spin2_sweeps(dim, L, beta, lw0[k], hists[k], spins[k], seed, k,
                                                0, n1, 0, 1)
        ser.append((E, M, Q))
        assert np.array_equal(sys.batch.get_spins(k), spins[k]), f"walker {k} spins"
        assert np.array_equal(algs[k].ensemble.spin2.histogram.values, hists[k])
        assert (algs[k].steps, algs[k].accepted) == (st, ac)
    # update!(ensemble(alg).spin2), reset!(alg), reset_histogram!(alg)
    lw1 = []
    for k, a in enumerate(algs):
        mcx.update_(a.ensemble.spin2)
        ref = lw0[k].copy()
        orc.muca_update(ref, hists[k].astype(np.float64))
        got = a.ensemble.spin2.logweight.values.copy()
        assert np.all(np.abs(got - ref) <= 4 * np.spacing(np.abs(ref)) + 1e-300)
        lw1.append(got)
        mcx.reset_(a)
        assert a.steps == 0 and a.accepted == 0
        assert np.array_equal(a.ensemble.spin2.histogram.values, hists[k])  # kept by reset!
        mcx.reset_histogram_(a)
        assert not a.ensemble.spin2.histogram.values.any()
    mcx.sweep_(sys, algs, n2, record=True)
    for k in range(n):
        h = np.zeros(N + 1, np.int64)
        E, M, Q, st, ac = orc.pair_spin2_sweeps(dim, L, beta, lw1[k], h, spins[k], seed, k, n1, n2,
                                                0, 1)
        Eg, Mg, Qg = sys.batch.series(k)
        assert np.array_equal(Eg, np.concatenate([ser[k][0], E]))
        assert np.array_equal(Mg, np.concatenate([ser[k][1], M]))
        assert np.array_equal(Qg, np.concatenate([ser[k][2], Q]))
        assert np.array_equal(sys.batch.get_spins(k), spins[k])
        assert np.array_equal(algs[k].ensemble.spin2.histogram.values, h)
        assert (algs[k].steps, algs[k].accepted) == (st, ac)


def test_pair_spin2_bounds_and_rules(mcx, orc):
    """sum s^2 leaving a restricted domain stops the walker as the reference's
    BoundsError would; invalid setups are refused."""
    print("pair+spin2 bounds", flush=True)
    L, N = 4, 16
    lw0 = np.zeros((1, N + 1 - 12))
    sys, algs = _pair_spin2_walkers(mcx, 2, L, 1, 0.9, 5, lw0, start=12)
    s = sys.batch.get_spins(0)
    with pytest.raises(mcx.MCXBoundsError):
        mcx.sweep_(sys, algs, 50)
    ref = s.copy()
    h = np.zeros(N + 1 - 12, np.int64)
    with pytest.raises(orc.BoundsError):
        orc.pair_spin2_sweeps(2, L, 1 / 0.9, lw0[0], h, ref, 5, 0, 0, 50, 12, 1)
    # float bins with step 2 would round where the integer range divides
    from montecarlox.spinsystems import SpinSystemBatch
    sys2 = SpinSystemBatch("blume_capel", [4, 4], n=1)
    e = mcx.MulticanonicalEnsemble(mcx.BinnedObject(np.arange(0.0, 17.0, 2.0).tolist(), 0.0))
    a = mcx.Metropolis(mcx.Philox(1), mcx.CustomEnsemble(mcx.BoltzmannEnsemble(beta=1.0), e))
    sys2.init_("random", [a.rng])
    with pytest.raises(mcx.MCXArgumentError):
        mcx.sweep_(sys2, [a], 1)
    # Ising: sum s^2 = N is constant
    from montecarlox.spinsystems import SpinSystemBatch as SSB
    sys3 = SSB("ising", [4, 4], n=1)
    a3 = mcx.Metropolis(mcx.Philox(1), mcx.CustomEnsemble(
        mcx.BoltzmannEnsemble(beta=1.0), mcx.MulticanonicalEnsemble(range(0, 17))))
    sys3.init_("random", [a3.rng])
    with pytest.raises(mcx.MCXArgumentError):
        mcx.sweep_(sys3, [a3], 1)
