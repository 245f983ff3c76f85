"""Host-side object model of the reference-API mirror (no GPU): binned
objects, ensembles, algorithm constructors and the tau_int estimator, checked
against the reference's own known answers."""
import math

import numpy as np
import pytest


def test_binned_object_known_answers(mcx):
    # test/test_binned_objects.jl:8-51
    bo = mcx.BinnedObject(range(0, 11, 2), 0.0)
    assert bo.size() == (6,) and bo.bins.start == 0 and bo.bins.step == 2
    assert mcx.get_centers(bo) == [0, 2, 4, 6, 8, 10]
    assert mcx.get_edges(bo) == list(range(-1, 12, 2))
    bo[4] = 1.5
    assert bo[4] == 1.5 and bo(4) == bo[4]
    with pytest.raises(mcx.MCXArgumentError):
        mcx.BinnedObject([1], 0.0)
    with pytest.raises(mcx.MCXArgumentError):
        mcx.BinnedObject([0, 1, 3], 0.0)
    b = mcx.BinnedObject([0.0, 2.0, 4.0, 6.0], 0.0, interpretation="discrete")
    assert b.bins.index(4.0) == 2  # _binindex 3 (1-based)
    c = mcx.BinnedObject(np.arange(0.0, 4.5, 1.0).tolist(), 0.0)
    assert c.size() == (4,) and c.bins.centers() == [0.5, 1.5, 2.5, 3.5]
    with pytest.raises(mcx.MCXBoundsError):
        bo[100]


def test_float_binning_rounds_half_to_even(mcx):
    bo = mcx.BinnedObject([0.0, 1.0, 2.0, 3.0], 0.0, interpretation="discrete")
    assert bo.bins.index(0.5) == 0 and bo.bins.index(1.5) == 2 and bo.bins.index(2.5) == 2


def test_set_known_answers(mcx):
    # test/test_multicanonical.jl:82-107
    alg = mcx.Multicanonical(mcx.MulticanonicalEnsemble(np.arange(0.0, 7.0).tolist()))
    lw = alg.ensemble.logweight
    lw.values[:] = 0
    assert mcx.set_(lw, (1.0, 4.0), lambda x: 10.0 + x) is None
    assert np.allclose(lw.values, [0.0, 11.5, 12.5, 13.5, 0.0, 0.0])
    mcx.set_(lw, np.arange(0.0, 7.0), lambda x: -x * x)
    assert np.allclose(lw.values, -np.array(mcx.get_centers(alg.ensemble.histogram)) ** 2)
    with pytest.raises(mcx.MCXArgumentError):
        mcx.set_(lw, (100.0, 200.0), lambda x: x)


def test_boltzmann_known_answers(mcx):
    # test/test_ensembles.jl:20-54
    assert mcx.BoltzmannEnsemble(beta=2.0).beta == 2.0
    assert mcx.BoltzmannEnsemble(**{"β": 3.0}).beta == 3.0
    assert math.isclose(mcx.BoltzmannEnsemble(T=1.0).beta, 1.0)
    for kw in ({"beta": 1.0, "β": 1.0}, {"beta": 1.0, "T": 1.0}, {}):
        with pytest.raises(mcx.MCXArgumentError):
            mcx.BoltzmannEnsemble(**kw)
    e = mcx.BoltzmannEnsemble(beta=2.0)
    assert e(1.0) == -2.0 and e([1.0, 2.0, 3.0]) == -12.0


def test_generalized_ensemble_constructors(mcx):
    lw = mcx.BinnedObject(range(0, 4), 0.0)
    e = mcx.MulticanonicalEnsemble(lw)
    assert e.logweight is lw and np.all(e.histogram.values == 0)
    e3 = mcx.MulticanonicalEnsemble(range(0, 4), init=-1.0, record_visits=False)
    assert not e3.record_visits and np.allclose(e3.logweight.values, -1.0)
    with pytest.raises(mcx.MCXAssertionError):
        mcx.MulticanonicalEnsemble(lw, histogram=mcx.BinnedObject(range(0, 5), 0.0))
    w = mcx.WangLandauEnsemble(range(0, 4), logf=0.5)
    assert w.logf == 0.5
    w.update_()
    assert w.logf == 0.25
    wl = mcx.WangLandau(mcx.Philox(780), lw, logf=math.log(2.0))
    assert wl.ensemble.logf == math.log(2.0)


def test_algorithm_constructors_and_counters(mcx):
    a = mcx.Metropolis(mcx.Philox(1, 2), beta=0.4)
    assert a.steps == 0 and mcx.acceptance_rate(a) == 0.0 and a.ensemble.beta == 0.4
    a.steps, a.accepted = 10, 4
    assert mcx.acceptance_rate(a) == 0.4
    mcx.reset_(a)
    assert a.steps == 0 and a.accepted == 0
    h = mcx.HeatBath(mcx.Philox(1), beta=0.3)
    assert h.steps == 0 and h.beta == 0.3
    with pytest.raises(mcx.MCXArgumentError):
        mcx.Philox(1, 2 ** 32)


def test_tau_int_known_answers(mcx):
    # test/test_measurements.jl:151-187
    from montecarlox.measurements import integrated_autocorrelation_time, integrated_autocorrelation_times
    assert integrated_autocorrelation_time(np.ones(64)) == 0.5
    rng = np.random.default_rng(123)
    assert 0.5 <= integrated_autocorrelation_time(rng.standard_normal(4000)) <= 2.0
    phi, n = 0.8, 6000
    ar = np.empty(n)
    ar[0] = rng.standard_normal()
    for t in range(1, n):
        ar[t] = phi * ar[t - 1] + math.sqrt(1 - phi ** 2) * rng.standard_normal()
    assert 2.5 <= integrated_autocorrelation_time(ar) <= 7.0
    assert integrated_autocorrelation_time(ar, max_lag=50) >= 0.5
    with pytest.raises(mcx.MCXArgumentError):
        integrated_autocorrelation_time(rng.standard_normal(20), max_lag=11)
    t = integrated_autocorrelation_times([rng.standard_normal(20)], min_points=2, max_lag=0)
    assert len(t) == 1 and np.isnan(t[0])


def test_binning_error_on_ar1(mcx):
    # AR(1): tau_int = (1 + phi) / (2 (1 - phi)) = 4.5 for phi = 0.8;
    # error of the mean = sigma * sqrt(2 tau / n)
    from montecarlox.measurements import binning_error
    rng = np.random.default_rng(5)
    phi, n = 0.8, 400000
    x = np.empty(n)
    x[0] = 0
    eps = rng.standard_normal(n) * math.sqrt(1 - phi ** 2)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + eps[t]
    expect = math.sqrt(2 * 4.5 / n)
    assert abs(binning_error(x, 40) / expect - 1) < 0.35


def test_poisson_workload_matches_the_stream_definition(mcx):
    """The serving stream: deterministic from its seed, even L in [16, 512],
    about 70 % Ising, mean offered work near 8.1e8 attempts per job and mean
    inter-arrival 1/200 s."""
    a = mcx.poisson_workload(2000)
    b = mcx.poisson_workload(2000)
    assert [(j.L, j.nsweeps, j.beta, j.D, j.model) for j in a] == \
           [(j.L, j.nsweeps, j.beta, j.D, j.model) for j in b]
    Ls = np.array([j.L for j in a])
    assert Ls.min() >= 16 and Ls.max() <= 512 and np.all(Ls % 2 == 0)
    ising = np.array([j.model == "ising" for j in a])
    assert abs(ising.mean() - 0.7) < 0.04
    assert all(j.D == 0.0 for j in a if j.model == "ising")
    assert all(0.0 <= j.D <= 1.9 for j in a if j.model != "ising")
    assert all(1.5 <= 1.0 / j.beta <= 3.5 for j in a)
    sw = np.array([j.nsweeps for j in a])
    assert sw.min() >= 1000 and sw.max() <= 100000
    work = np.mean([j.attempts for j in a])
    assert 6.5e8 < work < 9.8e8
    t = np.array([j.arrival_s for j in a])
    assert np.all(np.diff(t) > 0) and abs(t[-1] / len(a) - 1 / 200) < 0.0005
    assert [j.key for j in a] == list(range(2000))


def test_job_descriptor_validation_on_host(mcx):
    with pytest.raises(mcx.MCXArgumentError):
        mcx.Job("potts", 16, 0.4, 10).desc()
    with pytest.raises(mcx.MCXArgumentError):
        mcx.Job("ising", 16, 0.4, 10, alg="kawasaki").desc()
    assert mcx.Job("ising", 16, 0.4, 10, alg="glauber").desc().alg == 4  # MCX_GLAUBER
    d = mcx.Job("blume_capel", 16, 0.4, 10, alg="heatbath", D=1.5, seed=2**40 + 3, key=7).desc()
    assert (d.model, d.alg, d.seed, d.key, d.D) == (1, 1, 2**40 + 3, 7, 1.5)
