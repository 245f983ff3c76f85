"""mcx_batch_tau_int_noalloc (static device staging, no allocation) against
mcx_batch_tau_int (per-call device buffer), through the C ABI: two calls in a
row, the second on a smaller batch, bit for bit, NaN rules included."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBS = {"energy": 0, "magnetization": 1, "abs_magnetization": 2, "q": 3}


def _both(mcx, b, obs, first, count, min_points, max_lag, c):
    from montecarlox import _lib
    lib = _lib.load()
    out = {}
    for name in ("mcx_batch_tau_int_noalloc", "mcx_batch_tau_int"):
        taus = np.full(b.n, -1.0)
        st = getattr(lib, name)(b._h, OBS[obs], first, count, min_points, max_lag, c,
                                taus.ctypes.data_as(C.POINTER(C.c_double)))
        assert st == 0, (name, lib.mcx_last_error())
        out[name] = taus
    return out["mcx_batch_tau_int_noalloc"], out["mcx_batch_tau_int"]


def _batch(mcx, model, n, L, sweeps, seed, D=0.0):
    b = mcx.LatticeBatch(model, 2, L, n_chains=n, seed=seed, D=D, series_capacity=sweeps)
    b.set_betas(np.linspace(0.25, 0.6, n))
    b.init("random")
    b.run("metropolis", sweeps, record=True)
    return b


def test_noalloc_tau_equals_allocating_tau_twice_in_a_row(mcx):
    print("no-alloc tau: 512 Ising chains L=8 x 600 sweeps, then 24 Blume-Capel chains L=6",
          flush=True)
    b1 = _batch(mcx, "ising", 512, 8, 600, 5)
    for obs, first, count, mp, ml, c in (("energy", 0, 600, 100, -1, 5.0),
                                         ("abs_magnetization", 50, 500, 100, 40, 4.0),
                                         ("magnetization", 0, 600, 700, -1, 5.0)):
        a, r = _both(mcx, b1, obs, first, count, mp, ml, c)
        assert a.view(np.uint64).tolist() == r.view(np.uint64).tolist(), obs
    assert np.all(np.isnan(a))  # min_points > count: NaN everywhere, on both paths
    # second call: a smaller batch (fewer slots, another model) on the same GPU array
    b2 = _batch(mcx, "blume_capel", 24, 6, 300, 9, D=0.7)
    for obs in ("energy", "q", "abs_magnetization"):
        a, r = _both(mcx, b2, obs, 10, 290, 50, -1, 5.0)
        assert a.view(np.uint64).tolist() == r.view(np.uint64).tolist(), obs
        assert np.isfinite(a).all() and (a >= 0.5).all()
    b2.close()
    b1.close()


def test_noalloc_tau_rejects_more_slots_than_its_array(mcx):
    from montecarlox import _lib
    lib = _lib.load()
    limit = 65536
    b = mcx.LatticeBatch("ising", 2, 4, n_chains=limit + 1, seed=1, series_capacity=4)
    taus = np.empty(limit + 1)
    st = lib.mcx_batch_tau_int_noalloc(b._h, 0, 0, 0, 2, -1, 5.0,
                                       taus.ctypes.data_as(C.POINTER(C.c_double)))
    assert st == _lib.MCX_EINVAL
    assert b"static limit" in lib.mcx_last_error()
    b.close()
