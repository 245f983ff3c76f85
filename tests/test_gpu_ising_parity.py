"""Bit-exact parity of the HIP Ising kernels with the CPU oracle.

Every case runs the batched (LDS, bit-plane) kernel and / or the plain
one-chain-per-launch kernel through the C ABI and compares final spins, the
per-sweep E/M series and the accepted-move counters with the oracle's
sequential checkerboard restatement on the same seeded inputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ISING, METRO, HB, GLAUBER = 0, 0, 1, 2
NAMES = {METRO: "metropolis", HB: "heatbath", GLAUBER: "glauber"}


def _run_case(mcx, orc, dim, alg, Ls, betas, nsweeps, seed=7, plain=False, keys=None,
              runs=(None,)):
    n = len(Ls)
    print(f"case dim={dim} alg={alg} L={list(Ls)} nsweeps={nsweeps} plain={plain}", flush=True)
    b = mcx.LatticeBatch("ising", dim, list(Ls), seed=seed, keys=keys, series_capacity=nsweeps)
    b.set_betas(betas)
    b.init("random")
    start = [b.get_spins(k) for k in range(n)]
    parts = [nsweeps] if runs == (None,) else list(runs)
    assert sum(parts) == nsweeps
    for p in parts:
        b.run(NAMES[alg], p, record=True, plain=plain)
    steps, acc = b.counters()
    for k in range(n):
        key = k if keys is None else keys[k]
        s0 = orc.init_random(0, dim, Ls[k], seed, key, 0)
        assert np.array_equal(s0, start[k]), f"chain {k}: initial state differs"
        s = s0.copy()
        E, M, _, a = orc.canonical_sweeps(ISING, alg, dim, Ls[k], float(betas[k]), s, seed, key, 0,
                                          nsweeps)
        Eg, Mg, _ = b.series(k)
        assert np.array_equal(s, b.get_spins(k)), f"chain {k} (L={Ls[k]}): spins differ"
        assert np.array_equal(E, Eg), f"chain {k}: E series differs"
        assert np.array_equal(M, Mg), f"chain {k}: M series differs"
        assert acc[k] == a, f"chain {k}: accepted {acc[k]} != {a}"
        assert steps[k] == Ls[k] ** dim * nsweeps
    return b


@pytest.mark.parametrize("L", [4, 6, 16, 18, 32, 64, 66, 130])
def test_2d_metropolis_sizes(mcx, orc, L):
    betas = [0.2, 0.4406868, 0.7, 1.0]
    _run_case(mcx, orc, 2, METRO, [L] * 4, betas, 12)


@pytest.mark.parametrize("L", [4, 10, 16, 66])
def test_2d_heatbath_sizes(mcx, orc, L):
    _run_case(mcx, orc, 2, HB, [L] * 3, [0.1, 0.4406868, 0.9], 10)


@pytest.mark.parametrize("L", [4, 8, 12, 16])
def test_3d_metropolis(mcx, orc, L):
    _run_case(mcx, orc, 3, METRO, [L] * 3, [0.15, 0.2216544, 0.35], 6)


@pytest.mark.parametrize("L", [4, 8, 14])
def test_3d_heatbath(mcx, orc, L):
    _run_case(mcx, orc, 3, HB, [L] * 2, [0.1, 0.2216544], 6)


@pytest.mark.parametrize("L", [4, 16, 66, 130])
def test_2d_glauber_sizes(mcx, orc, L):
    # Glauber(rng; beta): logistic acceptance (metropolis.jl:108-127)
    _run_case(mcx, orc, 2, GLAUBER, [L] * 4, [0.2, 0.4406868, 0.7, 1.0], 10)


@pytest.mark.parametrize("L", [4, 8, 14])
def test_3d_glauber(mcx, orc, L):
    _run_case(mcx, orc, 3, GLAUBER, [L] * 3, [0.15, 0.2216544, 0.35], 6)


@pytest.mark.parametrize("dim,L", [(2, 8), (2, 24), (2, 40), (2, 56), (3, 24), (3, 32), (3, 40)])
def test_flat_layout_sizes(mcx, orc, dim, L):
    """L < 64 with L % 8 == 0: each colour's sites packed flat, 32 to a word
    (rows split across words); every algorithm, batched and plain."""
    betas = [0.3, 0.4406868, 0.7] if dim == 2 else [0.15, 0.2216544, 0.35]
    for alg in (METRO, HB, GLAUBER):
        _run_case(mcx, orc, dim, alg, [L] * 3, betas, 4)
    _run_case(mcx, orc, dim, METRO, [L], betas[1:2], 2, plain=True)


@pytest.mark.parametrize("dim", [2, 3])
def test_one_word_rows_sizes(mcx, orc, dim):
    """L = 64: rows of exactly one word, run by the flat word loop with row
    neighbours (its own instantiation); every algorithm, 2D and 3D."""
    betas = [0.3, 0.4406868] if dim == 2 else [0.2216544]
    for alg in (METRO, HB, GLAUBER):
        _run_case(mcx, orc, dim, alg, [64] * len(betas), betas, 3 if dim == 2 else 2)


@pytest.mark.parametrize("dim,L,alg", [(2, 10, METRO), (2, 10, HB), (2, 16, GLAUBER),
                                       (3, 6, METRO)])
def test_launches_past_the_small_launch_threshold(mcx, orc, dim, L, alg):
    """1100 one-wave chains: at least as many waves as the device has SIMDs,
    so the launch runs one lane per word (a smaller one shares each word among
    a quad of lanes); row planes (L = 10, 3D L = 6) and flat planes (L = 16)."""
    rng = np.random.default_rng(3)
    betas = rng.uniform(0.1, 1.0, 1100) if dim == 2 else rng.uniform(0.1, 0.5, 1100)
    _run_case(mcx, orc, dim, alg, [L] * 1100, betas, 3)


def test_glauber_extremes_plain_and_split(mcx, orc):
    # beta = 0: p = 1/2 everywhere; beta = 100: logistic rounds to 0 and 1
    _run_case(mcx, orc, 2, GLAUBER, [16] * 4, [0.0, 1e-12, 5.0, 100.0], 8)
    _run_case(mcx, orc, 2, GLAUBER, [32, 18], [0.44, 0.3], 5, plain=True)
    _run_case(mcx, orc, 3, GLAUBER, [8], [0.22], 3, plain=True)
    _run_case(mcx, orc, 2, GLAUBER, [32, 32], [0.44, 0.5], 13, runs=(5, 8))


def test_extreme_betas(mcx, orc):
    # beta = 0: every move accepted (threshold 2^32); beta = 100: exp underflows to 0
    _run_case(mcx, orc, 2, METRO, [16] * 4, [0.0, 1e-12, 5.0, 100.0], 8)
    _run_case(mcx, orc, 2, HB, [16] * 3, [0.0, 5.0, 100.0], 8)


def test_uneven_sizes_one_launch(mcx, orc):
    Ls = [64, 16, 32, 48, 8, 130, 64, 4]
    betas = np.linspace(0.25, 0.8, len(Ls))
    _run_case(mcx, orc, 2, METRO, Ls, betas, 9)


def test_plain_equals_batched_equals_oracle(mcx, orc):
    for alg in (METRO, HB):
        _run_case(mcx, orc, 2, alg, [32, 18], [0.44, 0.3], 5, plain=True)
        _run_case(mcx, orc, 3, alg, [8], [0.22], 3, plain=True)


def test_split_runs_resume(mcx, orc):
    # n + m sweeps in separate calls equal one call of n + m (counter-based streams)
    _run_case(mcx, orc, 2, METRO, [32, 32], [0.44, 0.5], 17, runs=(5, 1, 11))
    _run_case(mcx, orc, 2, HB, [16], [0.44], 9, runs=(4, 5), plain=False)


def test_custom_keys_and_seed(mcx, orc):
    _run_case(mcx, orc, 2, METRO, [16, 16, 16], [0.4, 0.4, 0.4], 6, seed=(1 << 40) + 3,
              keys=[7, 123456, 2 ** 32 - 1])


def test_record_flag_does_not_change_dynamics(mcx):
    outs = []
    for record in (False, True):
        b = mcx.LatticeBatch("ising", 2, 64, n_chains=8, seed=11, series_capacity=20)
        b.set_betas(np.linspace(0.3, 0.6, 8))
        b.init("random")
        b.run("metropolis", 20, record=record)
        outs.append([b.get_spins(k) for k in range(8)])
        assert b.counters()[1].sum() > 0
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


def test_observables_match_full_recompute(mcx, orc):
    b = mcx.LatticeBatch("ising", 3, [8, 12], seed=3)
    b.set_betas([0.2, 0.3])
    b.init("random")
    b.run("metropolis", 7)
    E, M, Q = b.observables()
    for k, L in enumerate([8, 12]):
        s = b.get_spins(k)
        assert E[k] == -orc.pair_sum(3, L, s)
        assert M[k] == orc.magnetization(3, L, s)
        assert Q[k] == L ** 3


def test_moments_match_series(mcx):
    b = mcx.LatticeBatch("ising", 2, 32, n_chains=3, seed=5, series_capacity=40)
    b.set_betas([0.3, 0.44, 0.6])
    b.init("random")
    b.run("metropolis", 25, record=True)
    b.run("metropolis", 15, record=True)
    mom = b.moments()
    for k in range(3):
        E, M, _ = b.series(k)
        e, m = E.astype(np.float64), M.astype(np.float64)
        np.testing.assert_allclose(mom[k], [40, e.sum(), (e * e).sum(), np.abs(m).sum(),
                                            (m * m).sum(), (m ** 4).sum()], rtol=1e-12)
