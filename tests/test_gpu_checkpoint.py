"""Deterministic restart on the device (test/test_checkpointing.jl:100-139 for
the spin-lattice path): n sweeps, checkpoint, m sweeps equals a restore into
fresh objects and the same m sweeps, bit for bit -- spins, recorded series,
moments, counters, sweep counters, replica-exchange ladders and walker
weights / histograms.  Snapshots of another shape are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state_of(sys):
    b = sys.batch
    spins = [b.get_spins(k) for k in range(b.n)]
    series = [b.series(k) for k in range(b.n)] if b.capacity else []
    st, ac = b.counters()
    return spins, series, b.moments(), st, ac, b.sweeps(), b.observables()


def _same(a, b):
    sa, sera, ma, sta, aca, swa, oa = a
    sb, serb, mb, stb, acb, swb, ob = b
    assert all(np.array_equal(x, y) for x, y in zip(sa, sb))
    for (e1, m1, q1), (e2, m2, q2) in zip(sera, serb):
        assert np.array_equal(e1, e2) and np.array_equal(m1, m2)
        assert (q1 is None and q2 is None) or np.array_equal(q1, q2)
    assert np.array_equal(ma, mb)  # FP64 moments continue bit-exactly
    assert np.array_equal(sta, stb) and np.array_equal(aca, acb) and np.array_equal(swa, swb)
    assert all(np.array_equal(x, y) for x, y in zip(oa, ob))


@pytest.mark.parametrize("model,dim,Ls,alg,D", [
    ("ising", 2, [8, 16, 34, 64, 6], "metropolis", 0.0),
    ("blume_capel", 3, [4, 6, 10], "heatbath", 0.7),
    ("blume_capel", 2, [10, 64], "metropolis", 1.3),
])
def test_restart_canonical_batch(mcx, tmp_path, model, dim, Ls, alg, D):
    from montecarlox.spinsystems import SpinSystemBatch
    print(f"ckpt canonical model={model} dim={dim} Ls={Ls} alg={alg}", flush=True)
    n, m = 30, 25
    sys = SpinSystemBatch(model, [Ls[0]] * dim, n=len(Ls), L=Ls, D=D, series_capacity=n + m)
    cls = mcx.Metropolis if alg == "metropolis" else mcx.HeatBath
    betas = np.linspace(0.2, 0.7, len(Ls))
    algs = [cls(mcx.Philox(2 ** 33 + 9, 40 + k), beta=float(betas[k])) for k in range(len(Ls))]
    sys.init_("random", [a.rng for a in algs])
    mcx.sweep_(sys, algs, n, record=True)
    f = str(tmp_path / "c.mcx")
    ckpt = mcx.init_checkpoint(f, {"sys": sys, "algs": algs, "sweep": n})
    mcx.sweep_(sys, algs, m, record=True)
    ref = _state_of(sys)
    ref_counts = [(a.steps, getattr(a, "accepted", None)) for a in algs]

    c2 = mcx.restore_checkpoint(f)
    assert c2.sweep == n and c2.sys is not sys and c2.sys.batch.handle.value != sys.batch.handle.value
    assert c2.sys.batch.series_length() == n
    mcx.sweep_(c2.sys, c2.algs, m, record=True)
    _same(_state_of(c2.sys), ref)
    assert [(a.steps, getattr(a, "accepted", None)) for a in c2.algs] == ref_counts
    mcx.finalize_(ckpt)


def test_restart_before_first_sweep_and_views(mcx, tmp_path):
    """A checkpoint before any sweep (no tables yet) and a single-system view."""
    print("ckpt single Ising view", flush=True)
    sys = mcx.Ising([16, 16])
    alg = mcx.Metropolis(mcx.Philox(4, 1), beta=0.44)
    mcx.init_(sys, "random", rng=alg.rng)
    f = str(tmp_path / "v.mcx")
    ckpt = mcx.init_checkpoint(f, {"sys": sys, "alg": alg})
    mcx.sweep_(sys, alg, 40)
    c2 = mcx.restore_checkpoint(f)
    assert isinstance(c2.sys, mcx.Ising)
    mcx.sweep_(c2.sys, c2.alg, 40)
    assert np.array_equal(c2.sys.spins, sys.spins) and c2.sys.energy() == sys.energy()
    assert (c2.alg.steps, c2.alg.accepted) == (alg.steps, alg.accepted)
    mcx.finalize_(ckpt)


@pytest.mark.parametrize("graph", [False, True])
def test_restart_parallel_tempering(mcx, tmp_path, graph):
    from montecarlox.spinsystems import SpinSystemBatch
    print(f"ckpt PT graph={graph}", flush=True)
    S, R, L = 2, 4, 6
    sys = SpinSystemBatch("ising", [L] * 3, n=S * R, series_capacity=40)
    pt = mcx.ParallelTempering(sys, mcx.set_betas(R, 0.15, 0.4), n_sets=S, seed=77)
    sys.init_("random", [a.rng for a in pt.algs])
    pt.run_(6, 3, record=True, graph=graph)
    f = str(tmp_path / "pt.mcx")
    ckpt = mcx.init_checkpoint(f, {"sys": sys, "pt": pt, "round": 6})
    pt.run_(5, 3, record=True, graph=graph)
    ref = _state_of(sys)
    ref_pt = (pt.indices.copy(), pt.steps.copy(), pt.accepted.copy(), pt.stage.copy(), pt.rounds())

    c2 = mcx.restore_checkpoint(f)
    assert c2.pt is c2.sys._pt and c2.round == 6
    c2.pt.run_(5, 3, record=True, graph=graph)
    _same(_state_of(c2.sys), ref)
    got = (c2.pt.indices, c2.pt.steps, c2.pt.accepted, c2.pt.stage, c2.pt.rounds())
    assert all(np.array_equal(x, y) for x, y in zip(got, ref_pt))
    assert [a.ensemble.beta for a in c2.pt.algs] == [a.ensemble.beta for a in pt.algs]
    assert [(a.steps, a.accepted) for a in c2.pt.algs] == [(a.steps, a.accepted) for a in pt.algs]
    mcx.finalize_(ckpt)


@pytest.mark.parametrize("kind", ["muca", "wl"])
def test_restart_walkers(mcx, tmp_path, kind):
    from montecarlox.ensembles import DiscreteBinning
    from montecarlox.spinsystems import SpinSystemBatch
    print(f"ckpt walkers kind={kind}", flush=True)
    n, L = 4, 6
    sys = SpinSystemBatch("ising", [L, L], n=n)
    rng = np.random.default_rng(5)
    algs = []
    for k in range(n):
        bo = mcx.BinnedObject(DiscreteBinning(-2.0 * L * L, 4.0, L * L + 1), 0.0)
        bo.values[:] = rng.normal(0, 0.3, bo.values.size)
        if kind == "muca":
            algs.append(mcx.Multicanonical(mcx.Philox(12, k), bo))
        else:
            algs.append(mcx.WangLandau(mcx.Philox(12, k), bo, logf=0.5))
    sys.init_("random", [a.rng for a in algs])
    mcx.sweep_(sys, algs, 50)
    for a in algs:
        mcx.update_(a.ensemble)
    f = str(tmp_path / "w.mcx")
    ckpt = mcx.init_checkpoint(f, {"sys": sys, "algs": algs})
    mcx.sweep_(sys, algs, 60)
    ref = _state_of(sys)
    ref_w = [(a.ensemble.logweight.values.copy(), a.ensemble.device_histogram.copy(), a.steps,
              a.accepted) for a in algs]

    c2 = mcx.restore_checkpoint(f)
    assert c2.sys._walkers is not None and c2.sys._walkers.algs[0] is c2.algs[0]
    mcx.sweep_(c2.sys, c2.algs, 60)
    assert c2.sys._walkers.algs[0] is c2.algs[0]  # no rebinding: the restored walkers ran
    _same(_state_of(c2.sys), ref)
    for a, (lw, h, st, ac) in zip(c2.algs, ref_w):
        assert np.array_equal(a.ensemble.logweight.values, lw)
        assert np.array_equal(a.ensemble.device_histogram, h)
        assert (a.steps, a.accepted) == (st, ac)
        if kind == "wl":
            assert a.ensemble.logf == 0.5 * 0.5
    mcx.finalize_(ckpt)


def test_snapshot_shape_mismatch(mcx):
    print("ckpt shape mismatch", flush=True)
    a = mcx.LatticeBatch("ising", 2, [8, 16], seed=1)
    b = mcx.LatticeBatch("ising", 2, [8, 12], seed=1)
    c = mcx.LatticeBatch("blume_capel", 2, [8, 16], seed=1)
    blob = a.save_state()
    for other in (b, c):
        with pytest.raises(mcx.MCXArgumentError):
            other.load_state(blob)
    with pytest.raises(mcx.MCXArgumentError):
        a.load_state(blob[: blob.size // 2])
    with pytest.raises(mcx.MCXArgumentError):
        a.load_state(np.zeros(64, np.uint8))
    a.load_state(blob)  # its own snapshot loads


# byte layout of a snapshot: an 80-byte header, then (batch) the 56-byte chain
# records, whose table field sits at +40 and sweep counter at +16
_HDR, _REC, _REC_TABLE, _REC_SWEEP = 80, 56, 40, 16


def _refused(mcx, load, blob):
    with pytest.raises(mcx.MCXArgumentError):
        load(blob)


def test_corrupted_batch_snapshots_are_refused(mcx):
    """A snapshot whose chain records point outside the tables, whose sweep
    counter overflows the Philox counter, or whose spins leave the model's
    states is refused before anything reaches the device; the batch keeps
    working and still loads the intact snapshot."""
    print("corrupted batch snapshots", flush=True)
    for model, bad_spin in (("ising", [5, 0]), ("blume_capel", [2, -3])):
        b = mcx.LatticeBatch(model, 2, 8, n_chains=4, seed=3, D=0.5)
        b.set_betas(np.linspace(0.3, 0.5, 4))
        b.init("random")
        b.run("metropolis", 5)
        good = b.save_state(series=False)
        for chain in (0, 3):
            blob = good.copy()
            at = _HDR + chain * _REC + _REC_TABLE
            blob[at:at + 4] = np.frombuffer(np.int32(4).tobytes(), np.uint8)  # table == n_chains
            _refused(mcx, b.load_state, blob)
            blob = good.copy()
            at = _HDR + chain * _REC + _REC_TABLE
            blob[at:at + 4] = np.frombuffer(np.int32(-1).tobytes(), np.uint8)
            _refused(mcx, b.load_state, blob)
            blob = good.copy()
            at = _HDR + chain * _REC + _REC_SWEEP
            blob[at:at + 8] = np.frombuffer(np.uint64(1 << 48).tobytes(), np.uint8)
            _refused(mcx, b.load_state, blob)
        pool = 4 * 256  # four 8x8 lattices, 256-byte aligned, at the end of the blob
        for v in bad_spin:
            blob = good.copy()
            blob[len(blob) - pool + 256 + 17] = np.uint8(v & 0xFF)  # chain 1, site 17
            _refused(mcx, b.load_state, blob)
        _refused(mcx, b.load_state, good[:-1])
        b.run("metropolis", 3)
        b.load_state(good)
        s0 = b.get_spins(1)
        b.run("metropolis", 2)
        b.load_state(good)
        assert np.array_equal(b.get_spins(1), s0)
        b.close()


def test_corrupted_ladder_snapshots_are_refused(mcx):
    from montecarlox.spinsystems import SpinSystemBatch
    print("corrupted replica-exchange snapshots", flush=True)
    S, R = 2, 4
    sys_ = SpinSystemBatch("ising", [8, 8], n=S * R)
    pt = mcx.ParallelTempering(sys_, np.tile(mcx.set_betas(R, 0.3, 0.5, "uniform"), S),
                               n_sets=S, seed=5)
    sys_.init_("random", [a.rng for a in pt.algs])
    pt.run_(6, 2)
    good = pt.save_state()
    n = S * R
    idx_at, pos_at = _HDR + 8 * n, _HDR + 12 * n
    stage_at = pos_at + 4 * n + 8 * S * (R - 1)
    ints = lambda a, k: np.frombuffer(a[k:k + 4 * n].tobytes(), np.int32)
    # pos no longer the inverse of idx (two ladder positions swapped)
    blob = good.copy()
    p = ints(blob, pos_at).copy()
    p[[0, 1]] = p[[1, 0]]
    blob[pos_at:pos_at + 4 * n] = np.frombuffer(p.tobytes(), np.uint8)
    _refused(mcx, pt.load_state, blob)
    # a ladder index outside 1..R
    blob = good.copy()
    i = ints(blob, idx_at).copy()
    i[5] = R + 1
    blob[idx_at:idx_at + 4 * n] = np.frombuffer(i.tobytes(), np.uint8)
    _refused(mcx, pt.load_state, blob)
    # a position outside the set
    blob = good.copy()
    p = ints(blob, pos_at).copy()
    p[2] = -7
    blob[pos_at:pos_at + 4 * n] = np.frombuffer(p.tobytes(), np.uint8)
    _refused(mcx, pt.load_state, blob)
    # a stage that is neither 0 nor 1
    blob = good.copy()
    blob[stage_at:stage_at + 4] = np.frombuffer(np.int32(2).tobytes(), np.uint8)
    _refused(mcx, pt.load_state, blob)
    pt.load_state(good)
    pt.run_(2, 2)
    pt.close()
    sys_.batch.close()
