"""Bit-exact parity of the HIP Blume-Capel kernels (bit-plane LDS batched and
plain one-chain-per-launch) with the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BC = 1


def _run(mcx, orc, dim, alg, Ls, betas, Ds, nsweeps, seed=3, plain=False, runs=None):
    n = len(Ls)
    print(f"bc dim={dim} alg={alg} L={Ls} plain={plain}", flush=True)
    b = mcx.LatticeBatch("blume_capel", dim, list(Ls), seed=seed, D=Ds, series_capacity=nsweeps)
    b.set_betas(betas)
    b.init("random")
    start = [b.get_spins(k) for k in range(n)]
    for p in (runs or [nsweeps]):
        b.run({0: "metropolis", 1: "heatbath", 2: "glauber"}[alg], p, record=True, plain=plain)
    _, acc = b.counters()
    mom = b.moments()
    for k in range(n):
        s0 = orc.init_random(BC, dim, Ls[k], seed, k, 0)
        assert np.array_equal(s0, start[k])
        s = s0.copy()
        E, M, Q, a = orc.canonical_sweeps(BC, alg, dim, Ls[k], float(betas[k]), s, seed, k, 0,
                                          nsweeps, D=float(Ds[k]))
        Eg, Mg, Qg = b.series(k)
        assert np.array_equal(s, b.get_spins(k)), f"chain {k} spins"
        assert np.array_equal(E, Eg) and np.array_equal(M, Mg) and np.array_equal(Q, Qg)
        assert acc[k] == a
        e = E.astype(np.float64) + Ds[k] * Q
        assert abs(mom[k][1] - e.sum()) <= 1e-9 * max(1.0, abs(e.sum()))
    return b


@pytest.mark.parametrize("L", [4, 6, 16, 34, 66])
def test_bc_2d_metropolis(mcx, orc, L):
    _run(mcx, orc, 2, 0, [L] * 3, [0.3, 0.7, 1.4], [0.0, 1.0, 1.9], 8)


@pytest.mark.parametrize("L", [4, 10, 16])
def test_bc_2d_heatbath(mcx, orc, L):
    _run(mcx, orc, 2, 1, [L] * 3, [0.3, 0.7, 1.4], [0.5, 1.0, 1.7], 8)


@pytest.mark.parametrize("alg", [0, 1])
def test_bc_3d(mcx, orc, alg):
    _run(mcx, orc, 3, alg, [4, 6], [0.4, 0.9], [0.2, 1.1], 5)


@pytest.mark.parametrize("L", [4, 16, 34])
def test_bc_2d_glauber(mcx, orc, L):
    # Glauber on Blume-Capel: the Metropolis proposal with logistic acceptance
    _run(mcx, orc, 2, 2, [L] * 3, [0.3, 0.7, 1.4], [0.0, 1.0, 1.9], 8)


def test_bc_glauber_3d_plain_extreme(mcx, orc):
    _run(mcx, orc, 3, 2, [4, 6], [0.4, 0.9], [0.2, 1.1], 5)
    _run(mcx, orc, 2, 2, [16, 6], [0.6, 1.0], [1.0, 0.3], 6, plain=True)
    _run(mcx, orc, 2, 2, [8, 8, 8], [0.0, 5.0, 60.0], [1.0, 1.0, 1.9], 6)


def test_bc_plain_and_split_runs(mcx, orc):
    _run(mcx, orc, 2, 0, [16, 6], [0.6, 1.0], [1.0, 0.3], 6, plain=True)
    _run(mcx, orc, 2, 1, [12], [0.8], [0.9], 6, plain=True)
    _run(mcx, orc, 2, 0, [16, 16], [0.6, 1.0], [1.0, 0.3], 9, runs=[4, 5])


def test_bc_extreme(mcx, orc):
    _run(mcx, orc, 2, 0, [8, 8, 8], [0.0, 5.0, 60.0], [1.0, 1.0, 1.9], 6)
    _run(mcx, orc, 2, 1, [8, 8], [0.0, 30.0], [1.0, 0.2], 6)
