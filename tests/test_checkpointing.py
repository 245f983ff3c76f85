"""Checkpoint / restore of host objects, after test/test_checkpointing.jl:11-147:
algorithm counters and streams, multicanonical and Wang-Landau ensembles,
keyword overrides, finalize!, and a deterministic restart (here of a
counter-based chain of the CPU oracle: n sweeps, checkpoint, restore, m sweeps
equals n + m sweeps).  The format is checked to load without pickle."""
import os

import numpy as np
import pytest


def _file(tmp_path, prefix):
    return str(tmp_path / f"{prefix}.ckpt")


@pytest.mark.parametrize("name", ["Metropolis", "Glauber", "HeatBath"])
def test_checkpoint_restore_importance_sampling(mcx, tmp_path, name):
    rng = mcx.Philox(2 ** 40 + 3, 17)
    alg = {"Metropolis": lambda: mcx.Metropolis(rng, beta=0.5),
           "Glauber": lambda: mcx.Glauber(rng, beta=0.25),
           "HeatBath": lambda: mcx.HeatBath(rng, beta=0.5)}[name]()
    alg.steps = 100
    if hasattr(alg, "accepted"):
        alg.accepted = 42
    f = _file(tmp_path, name)
    ckpt = mcx.init_checkpoint(f, {"alg": alg, "sweep": 100})
    assert os.path.isfile(f) and not os.path.exists(f + ".tmp")
    c2 = mcx.restore_checkpoint(f)
    assert c2.sweep == 100
    assert type(c2.alg) is type(alg)
    assert c2.alg.rng == rng and c2.alg.steps == 100
    assert getattr(c2.alg, "accepted", None) == getattr(alg, "accepted", None)
    assert mcx.algorithms.algorithm_beta(c2.alg) == mcx.algorithms.algorithm_beta(alg)
    mcx.finalize_(ckpt)
    assert not os.path.isfile(f)


def test_checkpoint_restore_ensembles(mcx, tmp_path):
    alg = mcx.Multicanonical(mcx.Philox(2), range(0, 11))
    ens = mcx.ensemble(alg)
    ens.logweight.values[:] = np.arange(11)
    ens.histogram.values[:] = 2 * np.arange(11)
    f = _file(tmp_path, "muca")
    ckpt = mcx.init_checkpoint(f, {"alg": alg, "sweep": 0})
    c2 = mcx.restore_checkpoint(f)
    assert mcx.ensemble(c2.alg) == ens
    assert c2.alg.ensemble.logweight(7) == 7.0
    mcx.finalize_(ckpt)

    alg = mcx.WangLandau(mcx.Philox(3), range(0, 6), logf=2.0)
    ens = mcx.ensemble(alg)
    ens.logweight.values[:] = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    ens.logf = 1.5
    ckpt = mcx.init_checkpoint(f, {"alg": alg, "sweep": 0})
    c2 = mcx.restore_checkpoint(f)
    assert mcx.ensemble(c2.alg) == ens and c2.alg.ensemble.logf == 1.5
    mcx.finalize_(ckpt)


def test_checkpoint_override(mcx, tmp_path):
    alg = mcx.Metropolis(mcx.Philox(7), beta=1.0)
    f = _file(tmp_path, "override")
    ckpt = mcx.init_checkpoint(f, {"alg": alg, "sweep": 0})
    mcx.checkpoint_(ckpt, sweep=1)
    assert mcx.restore_checkpoint(f).sweep == 1
    assert ckpt.sweep == 0  # the override goes to the file, not the session
    mcx.finalize_(ckpt)
    mcx.finalize_(ckpt)  # a second finalize is a no-op


def test_checkpoint_defaults_and_values(mcx, tmp_path):
    shared = mcx.Philox(5, 6)
    state = {"a": np.arange(6, dtype=np.int16).reshape(2, 3), "x": np.float64(0.1),
             "n": 2 ** 62 + 1, "t": (1, "two", None, True), "d": {"k": [1.5, float("inf")]},
             "r1": shared, "r2": [shared], "u": np.uint64(2 ** 64 - 1)}
    ckpt = mcx.init_checkpoint(state, sweep=3)  # file in a fresh temporary directory
    assert os.path.basename(ckpt.file) == "ckpt.mcx" and os.path.isfile(ckpt.file)
    c2 = mcx.restore_checkpoint(ckpt.file)
    assert c2.sweep == 3 and c2.n == 2 ** 62 + 1 and c2.t == (1, "two", None, True)
    assert c2.a.dtype == np.int16 and np.array_equal(c2.a, state["a"])
    assert c2.x == 0.1 and c2.u == np.uint64(2 ** 64 - 1) and c2.d == {"k": [1.5, float("inf")]}
    assert c2.r1 is c2.r2[0] and c2.r1 == shared  # shared objects stay shared
    assert set(c2.keys()) == set(state) | {"sweep"}
    with pytest.raises(AttributeError):
        c2.sweep = 4
    with np.load(ckpt.file, allow_pickle=False) as z:  # loads with pickle off
        assert "meta" in z.files
    mcx.finalize_(ckpt)
    empty = mcx.init_checkpoint(str(tmp_path / "sub" / "dir" / "c.mcx"))
    assert os.path.isfile(empty.file) and list(mcx.restore_checkpoint(empty.file).keys()) == []


def test_checkpoint_refuses_unknown_objects(mcx, tmp_path):
    f = _file(tmp_path, "bad")
    with pytest.raises(mcx.MCXArgumentError):
        mcx.init_checkpoint(f, {"f": lambda x: -0.5 * x * x})
    with pytest.raises(mcx.MCXArgumentError):
        mcx.init_checkpoint(f, {"o": np.array([object()], dtype=object)})
    assert not os.path.exists(f)
    with open(f, "wb") as fh:
        np.savez(fh, meta=np.frombuffer(b'{"format": "other"}', np.uint8))
    with pytest.raises(mcx.MCXArgumentError):
        mcx.restore_checkpoint(f)


def test_deterministic_restart_oracle_chain(mcx, orc, tmp_path):
    """The counter-based stream needs only (seed, key, sweep) and the spins to
    continue: n + m sweeps in one go equal n sweeps, a checkpoint, a restore
    into fresh objects and m sweeps."""
    L, beta, n, m = 12, 0.44, 40, 35
    rng = mcx.Philox(2 ** 35 + 11, 3)
    s_ref = orc.init_random(0, 2, L, rng.seed, rng.key, 0)
    E_ref, M_ref, _, acc_ref = orc.canonical_sweeps(0, 0, 2, L, beta, s_ref, rng.seed, rng.key,
                                                    0, n + m)
    alg = mcx.Metropolis(rng, beta=beta)
    s = orc.init_random(0, 2, L, rng.seed, rng.key, 0)
    E1, M1, _, acc1 = orc.canonical_sweeps(0, 0, 2, L, beta, s, rng.seed, rng.key, 0, n)
    alg.steps, alg.accepted = n * L * L, acc1
    f = _file(tmp_path, "det")
    ckpt = mcx.init_checkpoint(f, {"alg": alg, "spins": s.copy(), "sweep": n})
    c2 = mcx.restore_checkpoint(f)
    s2 = np.ascontiguousarray(c2.spins)
    E2, M2, _, acc2 = orc.canonical_sweeps(0, 0, 2, L, c2.alg.beta, s2, c2.alg.rng.seed,
                                           c2.alg.rng.key, c2.sweep, m)
    assert np.array_equal(np.concatenate([E1, E2]), E_ref)
    assert np.array_equal(np.concatenate([M1, M2]), M_ref)
    assert np.array_equal(s2, s_ref) and c2.alg.accepted + acc2 == acc_ref
    mcx.finalize_(ckpt)
