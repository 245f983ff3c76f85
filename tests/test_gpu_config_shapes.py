"""Parity at the benchmark configurations' own shapes.

* 4096 independent 2D Ising chains, L=64, T_k = 1.5 + 2k/4095, Philox key
  (7, k): every chain bit for bit against the oracle for the first sweeps, and
  a strided sample of 33 chains across the temperature range for 1000 sweeps
  (spins, per-sweep E/M series, accepted counts).
* Replica exchange in 3D: one 64-replica set_betas(64, 0.20, 0.24, :geometric)
  ladder for each L in {16, 24, 32, 48}, all four sets in one launch, several
  rounds of 10 Metropolis sweeps + an exchange: ladders, counters, stages,
  spins bit for bit.
* 4096 independent 2D Blume-Capel chains, L=64, D=1, T_k = 1.5 + 2k/4095,
  Philox key (5, k): Metropolis, heat bath and Glauber, every chain bit for bit
  against the oracle over a few sweeps.  At 2^24 site updates per sweep the
  decisions that need a site's 16-bit tail (its prefix equal to the
  threshold's leading bits, about 2^-15 per Metropolis site and 2^-16 per
  heat-bath comparison) occur hundreds of times per sweep.
* One 2D Ising chain, L=32, beta = 0.4406868, seed 42 / key (42, 0): the plain
  one-chain-per-launch path against the oracle, and the batched path against
  the plain path over the full 10^4 + 10^5 sweeps.

The oracle's chains run in a thread pool (its C calls release the GIL).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))


def _pool(fn, items):
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(fn, items))


def test_4096_chains_l64_bit_exact(mcx, orc):
    n, L, seed, first, more = 4096, 64, 7, 3, 997
    betas = 1.0 / (1.5 + 2.0 * np.arange(n) / (n - 1))
    print(f"{n} chains L={L}: {first} sweeps all chains, {first + more} sweeps strided", flush=True)
    b = mcx.LatticeBatch("ising", 2, L, n_chains=n, seed=seed, series_capacity=first + more)
    b.set_betas(betas)
    b.init("random")
    init = [b.get_spins(k) for k in range(n)]
    b.run("metropolis", first, record=True)
    mid = [b.get_spins(k) for k in range(n)]
    _, acc_mid = b.counters()
    b.run("metropolis", more, record=True)
    _, acc_end = b.counters()

    def first_sweeps(k):
        s = init[k].copy()
        E, M, _, a = orc.canonical_sweeps(0, 0, 2, L, float(betas[k]), s, seed, k, 0, first)
        return s, E, M, a

    res = _pool(first_sweeps, range(n))
    for k in range(n):
        s, E, M, a = res[k]
        assert np.array_equal(s, mid[k]), f"chain {k} spins after {first} sweeps"
        Eg, Mg, _ = b.series(k, 0, first)
        assert np.array_equal(E, Eg) and np.array_equal(M, Mg), f"chain {k} series"
        assert a == acc_mid[k], f"chain {k} accepted"
    strided = sorted(set(range(0, n, 128)) | {n - 1})
    assert len(strided) >= 32

    def rest(k):
        s = res[k][0].copy()
        E, M, _, a = orc.canonical_sweeps(0, 0, 2, L, float(betas[k]), s, seed, k, first, more)
        return s, E, M, a

    for k, (s, E, M, a) in zip(strided, _pool(rest, strided)):
        assert np.array_equal(s, b.get_spins(k)), f"chain {k} (T={1 / betas[k]:.4f}) spins"
        Eg, Mg, _ = b.series(k, first, more)
        assert np.array_equal(E, Eg) and np.array_equal(M, Mg), f"chain {k} series"
        assert res[k][3] + a == acc_end[k], f"chain {k} accepted"
    b.close()


@pytest.mark.parametrize("alg", ["metropolis", "heatbath", "glauber"])
def test_bc_4096_chains_l64_bit_exact(mcx, orc, alg):
    n, L, seed, D, sweeps = 4096, 64, 5, 1.0, 3
    code = {"metropolis": 0, "heatbath": 1, "glauber": 2}[alg]
    betas = 1.0 / (1.5 + 2.0 * np.arange(n) / (n - 1))
    print(f"{n} Blume-Capel chains L={L} {alg}: {sweeps} sweeps", flush=True)
    b = mcx.LatticeBatch("blume_capel", 2, L, n_chains=n, seed=seed, D=D, series_capacity=sweeps)
    b.set_betas(betas)
    b.init("random")
    init = [b.get_spins(k) for k in range(n)]
    b.run(alg, sweeps, record=True)
    _, acc = b.counters()

    def oracle_chain(k):
        s = init[k].copy()
        E, M, Q, a = orc.canonical_sweeps(1, code, 2, L, float(betas[k]), s, seed, k, 0, sweeps,
                                          D=D)
        return s, E, M, Q, a

    for k, (s, E, M, Q, a) in enumerate(_pool(oracle_chain, range(n))):
        assert np.array_equal(s, b.get_spins(k)), f"chain {k} spins"
        Eg, Mg, Qg = b.series(k)
        assert np.array_equal(E, Eg) and np.array_equal(M, Mg) and np.array_equal(Q, Qg), \
            f"chain {k} series"
        assert a == acc[k], f"chain {k} accepted"
    b.close()


def test_replica_exchange_3d_l16_to_48_full_ladders(mcx, orc):
    R, Ls_set, seed, rounds, sb = 64, [16, 24, 32, 48], 11, 3, 10
    S = len(Ls_set)
    ladder = np.tile(mcx.set_betas(R, 0.20, 0.24, "geometric"), S)
    Ls = np.repeat(Ls_set, R)
    print(f"replica exchange 3D: {S} sets x {R} replicas, L={Ls_set}, {rounds} x {sb} sweeps",
          flush=True)
    from montecarlox.spinsystems import SpinSystemBatch
    sys_ = SpinSystemBatch("ising", [16] * 3, n=len(Ls), L=Ls)
    pt = mcx.ParallelTempering(sys_, ladder, n_sets=S, seed=seed)
    sys_.init_("random", [a.rng for a in pt.algs])
    spins = [sys_.batch.get_spins(k) for k in range(len(Ls))]
    pt.run_(rounds, sb)
    keys = np.array([a.rng.key for a in pt.algs], np.uint32)
    idx = np.tile(np.arange(1, R + 1, dtype=np.int32), (S, 1))
    steps = np.zeros((S, R - 1), np.int32)
    acc = np.zeros((S, R - 1), np.int32)
    stage = np.zeros(S, np.int32)
    for rnd in range(rounds):
        def sweep(c):
            s_, r_ = divmod(c, R)
            beta = ladder[s_ * R + idx[s_, r_] - 1]
            E, _, _, _ = orc.canonical_sweeps(0, 0, 3, int(Ls[c]), float(beta), spins[c], seed,
                                              int(keys[c]), rnd * sb, sb)
            return float(E[-1])
        x = np.array(_pool(sweep, range(len(Ls))))
        for s_ in range(S):
            sl = slice(s_ * R, (s_ + 1) * R)
            row, st_, ac_ = idx[s_].copy(), steps[s_].copy(), acc[s_].copy()
            stage[s_] = orc.pt_round(ladder[sl], x[sl], row, st_, ac_, int(stage[s_]), seed,
                                     keys[sl], rnd)
            idx[s_], steps[s_], acc[s_] = row, st_, ac_
    assert np.array_equal(pt.indices, idx)
    assert np.array_equal(pt.steps, steps) and np.array_equal(pt.accepted, acc)
    assert np.array_equal(pt.stage, stage)
    for k in range(len(Ls)):
        assert np.array_equal(sys_.batch.get_spins(k), spins[k]), f"replica {k} (L={Ls[k]})"
    pt.close()
    sys_.batch.close()


def test_one_chain_l32_critical_plain_path(mcx, orc):
    L, beta, seed, therm, meas, checked = 32, 0.4406868, 42, 10_000, 100_000, 10_000
    total = therm + meas
    print(f"one chain L={L} beta={beta}: plain path {total} sweeps, oracle {checked}", flush=True)
    runs = {}
    for plain in (True, False):
        b = mcx.LatticeBatch("ising", 2, L, n_chains=1, seed=seed, keys=[0],
                             series_capacity=total)
        b.set_betas([beta])
        b.init("random")
        if plain:
            init = b.get_spins(0)
        b.run("metropolis", therm, record=True, plain=plain)
        b.run("metropolis", meas, record=True, plain=plain)
        E, M, _ = b.series(0)
        runs[plain] = (b.get_spins(0), E, M, b.counters()[1][0], b.moments()[0].copy())
        b.close()
    s = init.copy()
    E, M, _, a = orc.canonical_sweeps(0, 0, 2, L, beta, s, seed, 0, 0, checked)
    sp, Ep, Mp, ap, _ = runs[True]
    assert np.array_equal(E, Ep[:checked]) and np.array_equal(M, Mp[:checked])
    # the plain path == the batched path over the whole run, moments included
    sb, Eb, Mb, ab, mb = runs[False]
    assert np.array_equal(sp, sb) and ap == ab
    assert np.array_equal(Ep, Eb) and np.array_equal(Mp, Mb)
    assert np.array_equal(runs[True][4], mb)
