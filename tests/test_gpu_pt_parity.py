"""Replica exchange on the GPU vs the oracle: ladders, counters, stages and
spins after several rounds, bit for bit; graph replay equals eager launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_pt(orc, dim, Ls, n_sets, R, ladder, seed, keys, rounds, sb, alg=0, init_spins=None):
    n = n_sets * R
    spins = [s.copy() for s in init_spins]
    idx = np.tile(np.arange(1, R + 1, dtype=np.int32), (n_sets, 1))
    steps = np.zeros((n_sets, R - 1), np.int32)
    acc = np.zeros((n_sets, R - 1), np.int32)
    stage = np.zeros(n_sets, np.int32)
    for rnd in range(rounds):
        x = np.zeros(n)
        for c in range(n):
            s_, r_ = divmod(c, R)
            beta = ladder[s_ * R + idx[s_, r_] - 1]
            E, _, _, _ = orc.canonical_sweeps(0, alg, dim, Ls[c], float(beta), spins[c], seed,
                                              int(keys[c]), rnd * sb, sb)
            x[c] = E[-1]
        for s_ in range(n_sets):
            sl = slice(s_ * R, (s_ + 1) * R)
            row = idx[s_].copy()
            st_, ac_ = steps[s_].copy(), acc[s_].copy()
            stage[s_] = orc.pt_round(ladder[sl], x[sl], row, st_, ac_, int(stage[s_]), seed,
                                     keys[sl].astype(np.uint32), rnd)
            idx[s_], steps[s_], acc[s_] = row, st_, ac_
    return spins, idx, steps, acc, stage


def _gpu_pt(mcx, dim, Ls, n_sets, R, ladder, seed, rounds, sb, graph=False, alg="metropolis"):
    from montecarlox.spinsystems import SpinSystemBatch
    sys = SpinSystemBatch("ising", [Ls[0]] * dim, n=len(Ls), L=Ls)
    pt = mcx.ParallelTempering(sys, ladder, n_sets=n_sets, seed=seed, alg=alg)
    sys.init_("random", [a.rng for a in pt.algs])
    init = [sys.batch.get_spins(k) for k in range(len(Ls))]
    pt.run_(rounds, sb, graph=graph)
    out = [sys.batch.get_spins(k) for k in range(len(Ls))]
    return pt, init, out


@pytest.mark.parametrize("graph", [False, True])
def test_pt_matches_oracle_2d(mcx, orc, graph):
    n_sets, R, L, seed, rounds, sb = 2, 6, 8, 1000, 7, 3
    ladder = np.concatenate([mcx.set_betas(R, 1 / 3.0, 1 / 1.5, "uniform"),
                             mcx.set_betas(R, 0.3, 0.5, "geometric")])
    Ls = [L] * (n_sets * R)
    print(f"pt 2d graph={graph}", flush=True)
    pt, init, out = _gpu_pt(mcx, 2, Ls, n_sets, R, ladder, seed, rounds, sb, graph=graph)
    keys = np.arange(1, n_sets * R + 1)
    spins, idx, steps, acc, stage = _oracle_pt(orc, 2, Ls, n_sets, R, ladder, seed, keys, rounds,
                                               sb, init_spins=init)
    assert np.array_equal(pt.indices, idx)
    assert np.array_equal(pt.steps, steps) and np.array_equal(pt.accepted, acc)
    assert np.array_equal(pt.stage, stage)
    for k in range(len(Ls)):
        assert np.array_equal(out[k], spins[k]), f"replica {k}"
    assert acc.sum() > 0
    assert list(pt.rounds()) == [rounds] * n_sets


def test_pt_uneven_sets_3d(mcx, orc):
    # sets of different lattice sizes in one launch (3D, as in the 256-set workload)
    n_sets, R, seed, rounds, sb = 3, 4, 11, 5, 2
    set_L = [4, 8, 6]
    Ls = [set_L[s] for s in range(n_sets) for _ in range(R)]
    ladder = np.tile(mcx.set_betas(R, 0.20, 0.24, "geometric"), n_sets)
    print("pt 3d uneven", flush=True)
    pt, init, out = _gpu_pt(mcx, 3, Ls, n_sets, R, ladder, seed, rounds, sb)
    keys = np.arange(1, n_sets * R + 1)
    spins, idx, steps, acc, stage = _oracle_pt(orc, 3, Ls, n_sets, R, ladder, seed, keys, rounds,
                                               sb, init_spins=init)
    assert np.array_equal(pt.indices, idx) and np.array_equal(pt.accepted, acc)
    for k in range(len(Ls)):
        assert np.array_equal(out[k], spins[k])


def test_pt_heatbath(mcx, orc):
    n_sets, R, L, seed, rounds, sb = 1, 4, 8, 5, 6, 2
    ladder = mcx.set_betas(R, 0.3, 0.6, "uniform")
    Ls = [L] * R
    pt, init, out = _gpu_pt(mcx, 2, Ls, n_sets, R, ladder, seed, rounds, sb, alg="heatbath")
    spins, idx, steps, acc, stage = _oracle_pt(orc, 2, Ls, n_sets, R, ladder, seed,
                                               np.arange(1, R + 1), rounds, sb, alg=1,
                                               init_spins=init)
    assert np.array_equal(pt.indices, idx) and np.array_equal(pt.accepted, acc)
    for k in range(R):
        assert np.array_equal(out[k], spins[k])


def test_pt_glauber(mcx, orc):
    n_sets, R, L, seed, rounds, sb = 2, 4, 8, 15, 6, 3
    ladder = mcx.set_betas(R, 0.3, 0.6, "uniform")
    Ls = [L] * (n_sets * R)
    pt, init, out = _gpu_pt(mcx, 2, Ls, n_sets, R, np.tile(ladder, n_sets), seed, rounds, sb,
                            alg="glauber")
    spins, idx, steps, acc, stage = _oracle_pt(orc, 2, Ls, n_sets, R, np.tile(ladder, n_sets),
                                               seed, np.arange(1, n_sets * R + 1), rounds, sb,
                                               alg=2, init_spins=init)
    assert np.array_equal(pt.indices, idx) and np.array_equal(pt.accepted, acc)
    for k in range(n_sets * R):
        assert np.array_equal(out[k], spins[k])


def test_update_with_given_energies(mcx):
    # equal ladder betas: log ratio 0, every attempted pair swaps (replica_exchange.jl:158-178)
    from montecarlox.spinsystems import SpinSystemBatch
    sys = SpinSystemBatch("ising", [8, 8], n=4)
    pt = mcx.ParallelTempering(sys, [0.5] * 4, seed=3)
    pt.update_([-10.0, -8.0, -6.0, 0.0])
    assert list(mcx.index(pt)) == [2, 1, 4, 3] and list(pt.stage) == [1]
    pt.update_([-10.0, -8.0, -6.0, 0.0])
    assert list(mcx.index(pt)) == [3, 1, 4, 2] and list(pt.steps[0]) == [1, 1, 1]
    assert mcx.acceptance_rate(pt) == 1.0
    assert list(mcx.acceptance_rates(pt)) == [1.0, 1.0, 1.0]
    pt.reset_()
    assert list(mcx.index(pt)) == [1, 2, 3, 4] and list(pt.stage) == [0]
    # the ensembles travel with the ladder index
    pt2 = mcx.ParallelTempering(SpinSystemBatch("ising", [8, 8], n=2), [1.0, 0.5], seed=3)
    pt2.update_([0.0, -5.0])  # lr = 2.5 > 0: accepted
    assert pt2.algs[0].ensemble.beta == 0.5 and pt2.algs[1].ensemble.beta == 1.0


def test_graph_replays_equal_eager_runs(mcx, orc):
    """The instantiated graph is kept: later runs of the same shape replay it
    (and recording runs update it in place).  Every replay, with eager work
    and synchronisations between them, equals the eager run and the oracle."""
    from montecarlox.spinsystems import SpinSystemBatch
    n_sets, R, L, seed, rounds, sb = 2, 4, 8, 77, 3, 2
    ladder = np.concatenate([mcx.set_betas(R, 0.3, 0.5, "uniform"),
                             mcx.set_betas(R, 0.25, 0.45, "geometric")])
    Ls = [L] * (n_sets * R)
    runs = {}
    for graph in (False, True):
        sys_ = SpinSystemBatch("ising", [L, L], n=len(Ls), L=Ls, series_capacity=4 * rounds * sb)
        pt = mcx.ParallelTempering(sys_, ladder, n_sets=n_sets, seed=seed)
        sys_.init_("random", [a.rng for a in pt.algs])
        init = [sys_.batch.get_spins(k) for k in range(len(Ls))]
        snaps = []
        for rep in range(4):
            print(f"graph={graph} replay {rep}", flush=True)
            pt.run_(rounds, sb, graph=graph, record=rep >= 2)  # reps 2, 3: update path
            snaps.append((pt.indices.copy(), pt.accepted.copy(), pt.stage.copy(),
                          [sys_.batch.get_spins(k) for k in range(len(Ls))]))
            sys_.batch.observables()  # eager device work between the replays
        E = [sys_.batch.series(k)[0] for k in range(len(Ls))]
        runs[graph] = (init, snaps, E)
        pt.close()
        sys_.batch.close()
    (init, eager, Ee), (_, replays, Eg) = runs[False], runs[True]
    for rep, (a, b) in enumerate(zip(eager, replays)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), rep
        assert np.array_equal(a[2], b[2]), rep
        assert all(np.array_equal(x, y) for x, y in zip(a[3], b[3])), rep
    assert all(np.array_equal(x, y) for x, y in zip(Ee, Eg))
    keys = np.arange(1, n_sets * R + 1)
    spins, idx, steps, acc, stage = _oracle_pt(orc, 2, Ls, n_sets, R, ladder, seed, keys,
                                               4 * rounds, sb, init_spins=init)
    assert np.array_equal(replays[-1][0], idx) and np.array_equal(replays[-1][1], acc)
    assert all(np.array_equal(x, y) for x, y in zip(replays[-1][3], spins))
