"""The committed autotuner table: well-formed entries, lookups by (model, dim,
lattice sizes, algorithm), and the server settings (no GPU needed)."""
import numpy as np


def test_table_entries_are_valid():
    from montecarlox import autotune
    tab = autotune.load_table()
    assert tab["batched"], "no tuned launch shapes"
    for k, e in tab["batched"].items():
        model, dim, sizes, alg = k.split("/")
        assert model in ("ising", "blume_capel") and dim in ("2", "3")
        assert alg in ("metropolis", "heatbath", "glauber")
        assert all(int(x) % 2 == 0 and int(x) >= 4 for x in sizes.split(","))
        assert e["threads"] % 64 == 0 and 0 <= e["threads"] <= 1024
        assert e["chains_per_group"] >= 0
        assert e["lanes_per_word"] in (-1, 0, 1, 2, 4, 8)
        if model == "ising":
            assert e["lanes_per_word"] == 0
        # the chosen shape is the fastest measured one, and every candidate
        # reproduced the built-in shape's results
        assert e["ms"] <= e["default_ms"]
        assert all(t["identical"] for t in e["tried"])
        assert e["ms"] == min([e["default_ms"]] + [t["ms"] for t in e["tried"]])
    s = tab["server"]
    assert s["n_streams"] >= 1 and s["quantum_ms"] > 0
    assert s["p99_ms"] == min(t["p99_ms"] for t in s["tried"])


def test_lookup_by_sizes():
    from montecarlox import autotune
    assert autotune.key("ising", 2, [64, 64], "metropolis") == "ising/2/64/metropolis"
    assert autotune.key(1, 3, np.array([48, 16, 24, 16]), "glauber") == \
        "blume_capel/3/16,24,48/glauber"
    got = autotune.lookup(0, 2, [64] * 4096)
    assert set(got) <= {"metropolis", "heatbath", "glauber"}
    assert "metropolis" in got and set(got["metropolis"]) == {"threads", "chains_per_group",
                                                             "lanes_per_word"}
    assert autotune.lookup(0, 2, [62]) == {}
    assert set(autotune.server_settings()) <= {"n_streams", "quantum_ms", "band_sweeps"}
