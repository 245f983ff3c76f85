"""The continuous-batching server: every job's result equals the CPU oracle's
run of the same (model, L, beta, D, seed, key) from the same random start, is
bit-identical alone and in a mix that forces many quanta and admissions under
a tight HBM budget, and equals a one-chain LatticeBatch run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODEL = {"ising": 0, "blume_capel": 1}
ALG = {"metropolis": 0, "heatbath": 1, "glauber": 2}


def _jobs(mcx):
    J = mcx.Job
    return [
        J("ising", 16, 0.4406868, 200, "metropolis", seed=5, key=0),
        J("ising", 32, 0.3, 150, "heatbath", seed=5, key=1),
        J("blume_capel", 20, 0.7, 120, "metropolis", D=1.2, seed=5, key=2),
        J("blume_capel", 64, 0.9, 60, "heatbath", D=0.5, seed=5, key=3),
        J("ising", 8, 0.22, 100, "metropolis", dim=3, seed=5, key=4),
        J("ising", 6, 1.0, 300, "metropolis", seed=9, key=5),
        J("blume_capel", 6, 0.5, 90, "heatbath", dim=3, D=0.3, seed=9, key=6),
        J("ising", 128, 0.4406868, 30, "metropolis", seed=2**33 + 7, key=7),
        J("ising", 66, 0.5, 40, "heatbath", seed=5, key=8),
        J("blume_capel", 4, 2.0, 500, "metropolis", D=1.9, seed=5, key=9),
        J("ising", 256, 0.45, 6, "metropolis", seed=5, key=10),        # 1024-thread class
        J("blume_capel", 130, 0.6, 8, "heatbath", D=0.7, seed=5, key=11),
        J("ising", 24, 0.4406868, 90, "glauber", seed=5, key=12),
        J("blume_capel", 18, 0.8, 70, "glauber", D=0.8, seed=5, key=13),
        J("ising", 6, 0.25, 120, "glauber", dim=3, seed=9, key=14),
        J("blume_capel", 8, 0.6, 40, "glauber", dim=3, D=0.6, seed=9, key=15),
    ]


def _oracle(orc, job):
    m, a = MODEL[job.model], ALG[job.alg]
    s = orc.init_random(m, job.dim, job.L, job.seed, job.key, 0)
    E, M, Q, acc = orc.canonical_sweeps(m, a, job.dim, job.L, job.beta, s, job.seed, job.key, 0,
                                        job.nsweeps, D=job.D)
    e = E.astype(np.float64) + (job.D * Q if m == 1 else 0.0)
    mf = M.astype(np.float64)
    mom = np.array([job.nsweeps, e.sum(), (e * e).sum(), np.abs(mf).sum(), (mf * mf).sum(),
                    ((mf * mf) ** 2).sum()])
    return int(E[-1]), int(M[-1]), int(Q[-1]), acc, mom


def _same(r1, r2):
    for f in ("E_pair", "M", "Q", "accepted", "nsweeps", "L"):
        assert r1[f] == r2[f], f
    assert np.array_equal(r1["moments"], r2["moments"])


# short quanta and a budget of four 256^2 lattices: many quanta, admissions wait
SERVER = dict(hbm_budget=4 * 256 * 256, max_jobs=4, quantum_ms=0.05)


def test_server_matches_oracle_alone_and_in_mix(mcx, orc):
    jobs = _jobs(mcx)
    print("serving: mix of", len(jobs), "jobs", flush=True)
    with mcx.Server(**SERVER) as srv:
        mix = srv.run(jobs, timeout_ms=120000)
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["quanta"] > 20
    for job, r in zip(jobs, mix):
        assert r["tag"] == job.tag and r["status"] == 0
        E, M, Q, acc, mom = _oracle(orc, job)
        assert (r["E_pair"], r["M"], r["Q"], r["accepted"]) == (E, M, Q, acc), job
        assert np.allclose(r["moments"], mom, rtol=1e-12, atol=0), job
    print("serving: each job alone", flush=True)
    with mcx.Server(**SERVER) as srv:
        for job, r in zip(jobs, mix):
            alone = srv.run([job], timeout_ms=60000)[0]
            _same(alone, r)


def test_large_mixed_layout_class(mcx, orc):
    """1100 jobs in one class, flat (L = 8) and row (L = 10) planes mixed: a
    launch of at least as many waves as the device has SIMDs, one lane per
    word (a smaller class shares each word among a quad of lanes)."""
    # 10^4 sweeps each: the jobs submitted first are still running when the
    # last arrive, so later quanta launch the whole class at once (earlier,
    # smaller ones share words among quads: a job crosses both kinds)
    J = mcx.Job
    jobs = [J("ising", 8 if k % 2 else 10, 0.2 + 0.0005 * k, 10000, "metropolis", seed=11,
              key=k) for k in range(1100)]
    print("serving: one class of", len(jobs), "jobs", flush=True)
    with mcx.Server(hbm_budget=1 << 24, max_jobs=2048) as srv:
        res = srv.run(jobs, timeout_ms=120000)
    assert all(r["status"] == 0 and r["nsweeps"] == 10000 for r in res)
    for k in list(range(0, 1100, 10)) + [1097, 1098, 1099]:
        E, M, Q, acc, mom = _oracle(orc, jobs[k])
        r = res[k]
        assert (r["E_pair"], r["M"], r["accepted"]) == (E, M, acc), jobs[k]
        assert np.allclose(r["moments"], mom, rtol=1e-12, atol=0), jobs[k]


@pytest.mark.parametrize("band_sweeps", [0, 3, -1])
def test_tiled_path_equals_lds_path_and_oracle(mcx, orc, band_sweeps):
    """Every lattice on the multi-workgroup path (tile_sites=16: band kernels
    with the built-in depths, bands of 3 sweeps, and the per-colour kernels
    without bands) against every lattice on the LDS-resident path
    (tile_sites=-1): bit-identical results, and both equal to the oracle."""
    jobs = _jobs(mcx)
    print(f"serving: tiled (band_sweeps={band_sweeps}) vs LDS path", flush=True)
    with mcx.Server(hbm_budget=1 << 22, max_jobs=6, quantum_ms=0.2, tile_sites=16,
                    band_sweeps=band_sweeps) as srv:
        tiled = srv.run(jobs, timeout_ms=120000)
    with mcx.Server(hbm_budget=1 << 22, max_jobs=6, quantum_ms=0.2, tile_sites=-1) as srv:
        lds = srv.run(jobs, timeout_ms=120000)
    for job, a, b in zip(jobs, tiled, lds):
        _same(a, b)
        E, M, Q, acc, mom = _oracle(orc, job)
        assert (a["E_pair"], a["M"], a["Q"], a["accepted"]) == (E, M, Q, acc), job
        assert np.allclose(a["moments"], mom, rtol=1e-12, atol=0), job


def test_band_kernels_small_and_odd_lattices(mcx, orc):
    """The multi-workgroup band kernels (T sweeps per launch, halo of 2T slabs)
    on lattices smaller than their halo, bands that do not divide L, and 3D
    z-plane bands, at temperatures where most sites keep flipping: equal to
    the oracle."""
    J = mcx.Job
    jobs = [J("ising", 4, 0.3, 37, "metropolis", seed=3, key=0),
            J("ising", 6, 0.25, 41, "heatbath", seed=3, key=1),
            J("ising", 10, 0.35, 29, "glauber", seed=3, key=2),
            J("ising", 34, 0.3, 23, "metropolis", seed=3, key=3),
            J("ising", 70, 0.4, 19, "metropolis", seed=3, key=4),
            J("blume_capel", 4, 0.3, 33, "metropolis", D=0.2, seed=3, key=5),
            J("blume_capel", 6, 0.4, 27, "heatbath", D=0.5, seed=3, key=6),
            J("blume_capel", 38, 0.35, 17, "glauber", D=1.0, seed=3, key=7),
            J("ising", 4, 0.15, 21, "metropolis", dim=3, seed=3, key=8),
            J("ising", 10, 0.2, 13, "heatbath", dim=3, seed=3, key=9),
            J("blume_capel", 6, 0.3, 15, "metropolis", dim=3, D=0.4, seed=3, key=10),
            J("blume_capel", 12, 0.25, 9, "heatbath", dim=3, D=0.8, seed=3, key=11),
            # one class, its band shape set by L=200 (several 8-lane bands),
            # applied to a lattice smaller than one band
            J("blume_capel", 200, 0.35, 7, "metropolis", D=0.7, seed=3, key=12),
            J("blume_capel", 18, 0.3, 11, "metropolis", D=0.3, seed=3, key=13)]
    print("serving: band kernels on small / odd lattices", flush=True)
    with mcx.Server(hbm_budget=1 << 24, max_jobs=14, quantum_ms=0.3, tile_sites=16) as srv:
        out = srv.run(jobs, timeout_ms=120000)
    for job, r in zip(jobs, out):
        E, M, Q, acc, mom = _oracle(orc, job)
        assert (r["E_pair"], r["M"], r["Q"], r["accepted"]) == (E, M, Q, acc), job
        assert np.allclose(r["moments"], mom, rtol=1e-12, atol=0), job


def test_tiled_path_beyond_lds(mcx, orc):
    """A 3D lattice too big for one CU's LDS runs on the tiled path."""
    job = mcx.Job("blume_capel", 96, 0.5, 3, "heatbath", dim=3, D=0.4, seed=8, key=1)
    print("serving: 3D L=96 Blume-Capel on the tiled path", flush=True)
    with mcx.Server(hbm_budget=1 << 22, max_jobs=2, tile_sites=1 << 14) as srv:
        r = srv.run([job], timeout_ms=60000)[0]
    with pytest.raises(mcx.MCXArgumentError):
        with mcx.Server(hbm_budget=1 << 22, max_jobs=2, tile_sites=-1) as srv:
            srv.submit(job)
    E, M, Q, acc, mom = _oracle(orc, job)
    assert (r["E_pair"], r["M"], r["Q"], r["accepted"]) == (E, M, Q, acc)
    assert np.allclose(r["moments"], mom, rtol=1e-12, atol=0)


def test_server_quantum_size_does_not_change_the_chain(mcx, orc):
    jobs = _jobs(mcx)[:6]
    print("serving: quantum sizes", flush=True)
    with mcx.Server(hbm_budget=1 << 20, max_jobs=16, quantum_ms=50.0) as srv:
        a = srv.run(jobs, timeout_ms=60000)
        assert srv.stats()["quanta"] < 10
    with mcx.Server(hbm_budget=1 << 20, max_jobs=2, quantum_ms=0.02, n_streams=1) as srv:
        b = srv.run(jobs, timeout_ms=60000)
        assert srv.stats()["quanta"] > 20
    for r1, r2 in zip(a, b):
        _same(r1, r2)


def test_server_matches_lattice_batch(mcx):
    print("serving: vs LatticeBatch", flush=True)
    job = mcx.Job("blume_capel", 24, 0.8, 77, "heatbath", D=0.9, seed=123, key=31)
    with mcx.Server(hbm_budget=1 << 16, max_jobs=2, quantum_ms=0.05) as srv:
        r = srv.run([job])[0]
    b = mcx.LatticeBatch("blume_capel", 2, [24], seed=123, keys=[31], D=[0.9], series_capacity=77)
    b.set_betas([0.8])
    b.init("random")
    b.run("heatbath", 77, record=True)
    E, M, Q = b.observables()
    _, acc = b.counters()
    assert (r["E_pair"], r["M"], r["Q"], r["accepted"]) == (E[0], M[0], Q[0], acc[0])
    # the kernels continue the moment sums sweep by sweep: bit-identical
    assert np.array_equal(r["moments"], b.moments()[0])
    b.close()


def test_server_poisson_stream_small(mcx, orc):
    """A scaled-down serving stream (sweeps / 100): every result checks against
    the oracle's last-sweep observables and acceptance count."""
    jobs = mcx.poisson_workload(24, sweep_scale=0.01, L_range=(16, 96))
    print("serving: poisson stream of", len(jobs), flush=True)
    with mcx.Server(hbm_budget=64 * 96 * 96, max_jobs=64) as srv:
        res = srv.run(jobs, timeout_ms=120000)
        st = srv.stats()
    assert st["completed"] == len(jobs)
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    for job, r in zip(jobs, res):
        E, M, Q, acc, _ = _oracle(orc, job)
        assert (r["E_pair"], r["M"], r["Q"], r["accepted"]) == (E, M, Q, acc), job


def test_server_rejects_bad_jobs(mcx):
    with mcx.Server(hbm_budget=4096, max_jobs=2) as srv:
        with pytest.raises(mcx.MCXArgumentError):
            srv.submit(mcx.Job("ising", 17, 0.4, 10))      # odd L
        with pytest.raises(mcx.MCXArgumentError):
            srv.submit(mcx.Job("ising", 66, 0.4, 10))      # over the budget
        with pytest.raises(mcx.MCXArgumentError):
            srv.submit(mcx.Job("ising", 16, 0.4, 0))       # no sweeps
        with pytest.raises(mcx.MCXArgumentError):
            srv.submit(mcx.Job("potts", 16, 0.4, 10))
        assert srv.stats()["submitted"] == 0


def test_server_destroy_with_work_in_flight(mcx):
    srv = mcx.Server(hbm_budget=1 << 16, max_jobs=2, quantum_ms=0.05)
    for k in range(6):
        srv.submit(mcx.Job("ising", 32, 0.44, 400, seed=1, key=k))
    srv.close()  # stops after the current quantum; must not hang


def test_slo_report(mcx):
    """Deadlines of slo_ms + attempts / rate: a loose SLO is met by every job, an
    impossible one by none; per-job flags, per-model counts and the latency
    quantiles agree with the results and with the server's own statistics."""
    jobs = _jobs(mcx)
    print("SLO report", flush=True)
    with mcx.Server(hbm_budget=1 << 26, max_jobs=8) as srv:
        r0 = srv.slo_report()
        assert r0["jobs"] == 0 and r0["attainment"] == 1.0
        res = srv.run(jobs[:4])              # before any SLO: no deadline
        assert all(r["slo_met"] == -1 for r in res)
        srv.set_slo(60_000.0)                # one minute: always met
        res = srv.run(jobs)
        assert all(r["slo_met"] == 1 for r in res)
        r1 = srv.slo_report()
        assert (r1["jobs"], r1["met"], r1["missed"]) == (len(jobs), len(jobs), 0)
        assert r1["worst_lateness_ms"] < 0 and 0 < r1["mean_deadline_frac"] < 1
        n_bc = sum(j.model == "blume_capel" for j in jobs)
        assert r1["jobs_by_model"] == {"ising": len(jobs) - n_bc, "blume_capel": n_bc}
        srv.set_slo(0.0, 1e9)                # a billion attempts per ns: never met
        res = srv.run(jobs[:5])
        assert all(r["slo_met"] == 0 for r in res)
        r2 = srv.slo_report()
        assert r2["jobs"] == len(jobs) + 5 and r2["missed"] == 5 and r2["worst_lateness_ms"] > 0
        assert r2["attainment"] == len(jobs) / (len(jobs) + 5)
        st = srv.stats()
        assert abs(r2["p50_ms"] - st["p50_ms"]) < 1e-9 and abs(r2["p99_ms"] - st["p99_ms"]) < 1e-9
        assert r2["p50_ms"] <= r2["p90_ms"] <= r2["p99_ms"] <= r2["p999_ms"] <= r2["max_ms"]
        with pytest.raises(mcx.MCXArgumentError):
            srv.set_slo(float("nan"))
