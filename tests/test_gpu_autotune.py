"""Launch shapes (workgroup size, chains per workgroup, Blume-Capel lanes per
plane word) change only the speed: every candidate the autotuner measures,
and every committed tuned entry, reproduces the built-in shape's spins,
observables, counters and moments bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model,dim,Ls,alg", [
    ("ising", 2, [16] * 24, "metropolis"),
    ("ising", 2, [32] * 8 + [16] * 8, "heatbath"),
    ("ising", 3, [8] * 6 + [12] * 6, "metropolis"),
    ("ising", 2, [64] * 6, "glauber"),
    ("blume_capel", 2, [16] * 12, "metropolis"),
    ("blume_capel", 2, [64] * 6, "heatbath"),
    ("blume_capel", 3, [8] * 4, "glauber"),
])
def test_every_candidate_shape_is_bit_identical(mcx, model, dim, Ls, alg):
    from montecarlox import autotune
    print(f"shapes {model} dim={dim} L={sorted(set(Ls))} {alg}", flush=True)
    e = autotune.tune_batched(model, dim, Ls, alg, sweeps=12, reps=1, log=None)
    assert all(t["identical"] for t in e["tried"]) and len(e["tried"]) >= 6


def test_committed_tuned_shapes_are_bit_identical(mcx):
    from montecarlox import autotune
    tab = autotune.load_table()["batched"]
    assert tab, "no tuned shapes committed"
    for k, e in tab.items():
        model, dim, sizes, alg = k.split("/")
        sizes = [int(x) for x in sizes.split(",")]
        Ls = [L for L in sizes for _ in range(3)]  # a few chains of every size
        shape = {f: e[f] for f in ("threads", "chains_per_group", "lanes_per_word")}
        print(f"tuned {k}: {shape}", flush=True)
        r = autotune.tune_batched(model, int(dim), Ls, alg, sweeps=10, reps=1, shapes=[shape],
                                  log=None)
        assert r["tried"][0]["identical"]


def test_batches_take_the_tuned_shape(mcx):
    from montecarlox import autotune
    tab = autotune.load_table()["batched"]
    k = "ising/2/64/metropolis"
    if k not in tab:
        pytest.skip("no tuned entry for the 4096-chain workload")
    b = mcx.LatticeBatch("ising", 2, 64, n_chains=2)
    exp = {f: tab[k][f] for f in ("threads", "chains_per_group", "lanes_per_word")}
    assert b.launch_shape("metropolis") == exp == b.tuned_shapes["metropolis"]
    b2 = mcx.LatticeBatch("ising", 2, 64, n_chains=2, tuned=False)
    assert b2.launch_shape("metropolis") == {"threads": 0, "chains_per_group": 0,
                                             "lanes_per_word": 0}
    with pytest.raises(mcx.MCXArgumentError):
        b2.set_launch_shape("metropolis", threads=96)
    with pytest.raises(mcx.MCXArgumentError):
        b2.set_launch_shape("metropolis", lanes_per_word=2)  # Ising has no lanes per word
    b.close()
    b2.close()
