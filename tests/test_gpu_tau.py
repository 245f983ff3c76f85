"""Adaptive replica-exchange tooling on the device.

* mcx_batch_tau_int equals the oracle's tau_int (exact lag sums) on the same
  recorded series, bit for bit, for every observable, window and cap;
* temperature-ordered recording equals the oracle's replica-exchange run with
  its per-sweep energies regrouped by ladder index (the (index(pt), E)
  samples of pt_Ising2D_fullpower_mpi.jl:88-98 grouped as in
  parallel_tempering.jl:53-70), series and FP64 moments bit for bit;
* optimize_exchange_interval! on the device series equals the retune of the
  oracle's taus; a checkpoint keeps the order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _seq_moments(e, m):
    """The device's FP64 running sums, in sweep order."""
    s = [0.0] * 6
    for ev, mv in zip(e, m):
        ev, mv = float(ev), float(mv)
        s[0] += 1.0
        s[1] += ev
        s[2] += ev * ev
        s[3] += abs(mv)
        s[4] += mv * mv
        s[5] += (mv * mv) * (mv * mv)
    return np.array(s)


@pytest.mark.parametrize("model,dim,Ls,D", [
    ("ising", 2, [8, 16, 32, 10], 0.0),
    ("blume_capel", 2, [8, 12, 16], 1.25),
    ("blume_capel", 3, [4, 6], 0.0),
])
def test_tau_int_equals_oracle(mcx, orc, model, dim, Ls, D):
    from montecarlox.spinsystems import SpinSystemBatch
    print(f"tau model={model} dim={dim} Ls={Ls}", flush=True)
    nsw = 3000
    sys = SpinSystemBatch(model, [Ls[0]] * dim, n=len(Ls), L=Ls, D=D, series_capacity=nsw)
    betas = np.linspace(0.25, 0.6, len(Ls))
    algs = [mcx.Metropolis(mcx.Philox(21, k), beta=float(b)) for k, b in enumerate(betas)]
    sys.init_("random", [a.rng for a in algs])
    mcx.sweep_(sys, algs, nsw, record=True)
    b = sys.batch
    series = [b.series(k) for k in range(b.n)]
    obs = ["energy", "magnetization", "abs_magnetization"] + (["q"] if model != "ising" else [])
    for o in obs:
        for first, count, max_lag, c, mp in [(0, None, None, 5.0, 2), (100, 1500, 200, 5.0, 400),
                                             (0, 64, 7, 2.5, 2), (500, 2000, 3000, 6.0, 2)]:
            got = b.tau_int(o, first=first, count=count, max_lag=max_lag, c=c, min_points=mp)
            cnt = nsw - first if count is None else count
            for k, (E, M, Q) in enumerate(series):
                sl = slice(first, first + cnt)
                if o == "energy":
                    want = orc.tau_int_times([E[sl]], ys=[Q[sl]] if Q is not None else None,
                                             D=[D], min_points=mp, max_lag=max_lag, c=c)[0]
                elif o == "q":
                    want = orc.tau_int_times([Q[sl]], min_points=mp, max_lag=max_lag, c=c)[0]
                else:
                    want = orc.tau_int_times([M[sl]], absx=o == "abs_magnetization",
                                             min_points=mp, max_lag=max_lag, c=c)[0]
                assert got[k] == want or (np.isnan(got[k]) and np.isnan(want)), \
                    (o, first, count, max_lag, c, k, got[k], want)
    # empty windows give NaN, as integrated_autocorrelation_times does
    assert np.isnan(b.tau_int("energy", count=100, min_points=400)).all()
    assert np.isnan(b.tau_int("energy", max_lag=0)).all()
    with pytest.raises(mcx.MCXBoundsError):
        b.tau_int("energy", first=nsw - 10, count=20)
    if model == "ising":
        with pytest.raises(mcx.MCXArgumentError):
            b.tau_int("q")


def _oracle_pt_series(orc, dim, Ls, n_sets, R, ladder, seed, keys, rounds, sb, init_spins):
    """Replica exchange on the oracle with each sweep's (E, M) filed under the
    ladder slot that its replica held during that sweep."""
    n = n_sets * R
    spins = [s.copy() for s in init_spins]
    idx = np.tile(np.arange(1, R + 1, dtype=np.int32), (n_sets, 1))
    steps = np.zeros((n_sets, R - 1), np.int32)
    acc = np.zeros((n_sets, R - 1), np.int32)
    stage = np.zeros(n_sets, np.int32)
    sE = [[] for _ in range(n)]
    sM = [[] for _ in range(n)]
    for rnd in range(rounds):
        x = np.zeros(n)
        for c in range(n):
            s_, r_ = divmod(c, R)
            slot = s_ * R + idx[s_, r_] - 1
            E, M, _, _ = orc.canonical_sweeps(0, 0, dim, Ls[c], float(ladder[slot]), spins[c],
                                              seed, int(keys[c]), rnd * sb, sb)
            sE[slot].extend(E.tolist())
            sM[slot].extend(M.tolist())
            x[c] = E[-1]
        for s_ in range(n_sets):
            sl = slice(s_ * R, (s_ + 1) * R)
            row = idx[s_].copy()
            st_, ac_ = steps[s_].copy(), acc[s_].copy()
            stage[s_] = orc.pt_round(ladder[sl], x[sl], row, st_, ac_, int(stage[s_]), seed,
                                     keys[sl].astype(np.uint32), rnd)
            idx[s_], steps[s_], acc[s_] = row, st_, ac_
    return [np.array(e, np.int32) for e in sE], [np.array(m, np.int32) for m in sM], idx


@pytest.mark.parametrize("graph", [False, True])
def test_temperature_ordered_series(mcx, orc, tmp_path, graph):
    from montecarlox.spinsystems import SpinSystemBatch
    print(f"pt temperature order graph={graph}", flush=True)
    n_sets, R, L, seed, rounds, sb = 2, 5, 8, 4242, 24, 10
    ladder = np.concatenate([mcx.set_betas(R, 0.25, 0.55, "uniform"),
                             mcx.set_betas(R, 0.3, 0.5, "geometric")])
    Ls = [L] * (n_sets * R)
    sys = SpinSystemBatch("ising", [L, L], n=n_sets * R, series_capacity=rounds * sb)
    pt = mcx.ParallelTempering(sys, ladder, n_sets=n_sets, seed=seed)
    sys.init_("random", [a.rng for a in pt.algs])
    init = [sys.batch.get_spins(k) for k in range(len(Ls))]
    pt.set_series_order("temperature")
    half = rounds // 2
    pt.run_(half, sb, record=True, graph=graph)
    ckpt = mcx.init_checkpoint(str(tmp_path / "pt.mcx"), {"sys": sys, "pt": pt})
    pt.run_(rounds - half, sb, record=True, graph=graph)
    keys = np.arange(1, n_sets * R + 1)
    sE, sM, idx = _oracle_pt_series(orc, 2, Ls, n_sets, R, ladder, seed, keys, rounds, sb, init)
    assert np.array_equal(pt.indices, idx)
    b = sys.batch
    mom = b.moments()
    assert b.series_length() == rounds * sb
    for slot in range(n_sets * R):
        E, M, _ = b.series(slot)
        assert np.array_equal(E, sE[slot]) and np.array_equal(M, sM[slot]), f"slot {slot}"
        assert np.array_equal(mom[slot], _seq_moments(sE[slot], sM[slot])), f"moments {slot}"
    # tau_int per ladder index and the retune from it
    taus = pt.tau_int(max_lag=60)
    want = np.array([orc.tau_int(sE[k], max_lag=60) for k in range(n_sets * R)])
    assert np.array_equal(taus.ravel(), want)
    from montecarlox.pt import _retune_exchange_sweeps_
    sw = np.full((n_sets, R), 50, np.int64)
    got = mcx.optimize_exchange_interval_(pt, None, sw, base_sweeps=50, min_sweeps=5,
                                          max_sweeps=200, min_points=100, max_lag=60)
    ref = np.full((n_sets, R), 50, np.int64)
    for s_ in range(n_sets):
        _retune_exchange_sweeps_(ref[s_], want[s_ * R:(s_ + 1) * R], 50, 5, 200)
    assert np.array_equal(sw, ref)
    assert np.array_equal(got, np.take_along_axis(ref, pt.indices - 1, axis=1))
    # a restore keeps the order and continues the same per-temperature series
    c2 = mcx.restore_checkpoint(ckpt.file)
    assert c2.pt.series_order == "temperature"
    c2.pt.run_(rounds - half, sb, record=True, graph=graph)
    for slot in range(n_sets * R):
        assert np.array_equal(c2.sys.batch.series(slot)[0], sE[slot])
    assert np.array_equal(c2.sys.batch.moments(), mom)
    mcx.finalize_(ckpt)


def test_series_order_rules(mcx):
    from montecarlox.spinsystems import SpinSystemBatch
    print("pt series order rules", flush=True)
    sys = SpinSystemBatch("ising", [8, 8], n=4, series_capacity=20)
    pt = mcx.ParallelTempering(sys, mcx.set_betas(4, 0.3, 0.5), seed=3)
    sys.init_("random", [a.rng for a in pt.algs])
    pt.run_(1, 5, record=True)
    with pytest.raises(mcx.MCXStateError):
        pt.tau_int()  # recorded per replica
    pt.set_series_order("temperature")
    assert sys.batch.series_length() == 0 and not sys.batch.moments().any()  # cleared
    with pytest.raises(mcx.MCXStateError):
        sys.batch.run("metropolis", 2, record=True, plain=True)
    pt.run_(2, 5, record=True)
    assert pt.tau_int().shape == (4,)
    with pytest.raises(mcx.MCXArgumentError):
        pt.set_series_order("sideways")
    pt.set_series_order("replica")
    sys.batch.run("metropolis", 2, record=True, plain=True)  # allowed again
