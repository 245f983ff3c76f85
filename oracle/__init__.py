"""CPU oracle for the MonteCarloX spin-lattice path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper around ``oracle/liboracle.so`` (built from ``mcx_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product never
does.  It is the checker, never the thing measured or shipped.

Pinning status of each function (what checks it besides the GPU parity tests):
  philox4x32_10        Random123 known-answer vectors + rocRAND's host engine
  canonical_sweeps     reference bookkeeping KATs, cached == full invariants,
                       brute-force 4x4 / 3x3 enumerations, Kaufman exact Z (L=16),
                       Beale exact DOS (L=8), detailed balance on 2x... enumerations
  generalized_sweeps   reference muca/WL KATs (flat weights accept all, histogram
                       total == steps, WL decrement), flat-histogram convergence
                       to the Beale DOS at L=8
  wl_iterate_group     (the Wang-Landau flatness schedule has no reference
                       counterpart) convergence to the exact density of states:
                       brute-force Ising 4x4 and Blume-Capel 4x4, Beale Ising 8x8
  muca_iterate_update  the reference's merge / update! / distribute KATs and the
                       convergence of the iterated weights to the exact 4x4 DOS
  pt_round             reference exchange KATs (lr = 2.5, u=0 accepts, u=1 rejects)
  refpath_ising2d      timing only (CPU baseline); detailed-balance statistics
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

P_METRO, P_HB, P_INIT, P_BC_METRO, P_BC_HB, P_GEN, P_PT = 1, 2, 3, 4, 5, 6, 8

_lib = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _NullablePtr:
    """ndpointer that also accepts None."""

    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        if obj is None:
            return None
        return self.base.from_param(obj)


def _declare(L):
    L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.orc_site_group.argtypes = [C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_ising_group.argtypes = [C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_bitsliced_word.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint64, C.c_int]
    L.orc_bitsliced_word.restype = C.c_uint32
    L.orc_site_word.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint64, C.c_int]
    L.orc_site_word.restype = C.c_uint32
    L.orc_pt_word.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
    L.orc_pt_word.restype = C.c_uint32
    L.orc_bc_site_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint64,
                                      C.c_int, C.POINTER(C.c_int)]
    L.orc_bc_site_uniform.restype = C.c_uint32
    L.orc_logistic.argtypes = [C.c_double]
    L.orc_logistic.restype = C.c_double
    L.orc_threshold32.argtypes = [C.c_double]
    L.orc_threshold32.restype = C.c_uint64
    for f in ("orc_pair_sum", "orc_magnetization", "orc_sum_sq"):
        getattr(L, f).argtypes = [C.c_int, C.c_int, _i8p]
        getattr(L, f).restype = C.c_int64
    L.orc_init_random.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint64, _i8p]
    n32 = _NullablePtr(_i32p)
    L.orc_canonical_sweeps.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64, _i8p,
                                       n32, n32, n32, C.POINTER(C.c_int64)]
    L.orc_generalized_sweeps.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_double, C.c_int64, C.c_int, _f64p, _i64p,
                                         C.c_double,
                                         C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64, _i8p,
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                         n32, n32, n32, C.POINTER(C.c_int64)]
    L.orc_generalized_sweeps.restype = C.c_int
    L.orc_pt_round.argtypes = [C.c_int, _f64p, _f64p, _i32p, _i32p, _i32p, C.POINTER(C.c_int32),
                               C.c_uint64, _u32p, C.c_uint64]
    L.orc_muca_update.argtypes = [C.c_int64, _f64p, _f64p]
    L.orc_exchange_pair.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_double)]
    L.orc_exchange_pair.restype = C.c_int
    L.orc_refpath_series_ising2d.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_int64, _i8p,
                                             _i32p, _i32p, C.POINTER(C.c_int64)]
    L.orc_refpath_ising2d.argtypes = [C.c_int, C.c_int, _f64p, C.c_uint64, C.c_int64, C.c_int,
                                      _NullablePtr(_i8p)]
    L.orc_refpath_ising2d.restype = C.c_int64
    nf64 = _NullablePtr(_f64p)
    L.orc_refpath_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _f64p, nf64, nf64,
                                  C.c_int64, C.c_double, C.c_double, C.c_uint64, C.c_int64, C.c_int,
                                  C.c_double]
    L.orc_refpath_run.restype = C.c_int64
    L.orc_pair_spin2_sweeps.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_int64, _f64p, _i64p, C.c_uint64, C.c_uint32,
                                        C.c_uint64, C.c_int64, _i8p, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), n32, n32, n32,
                                        C.POINTER(C.c_int64)]
    L.orc_pair_spin2_sweeps.restype = C.c_int
    L.orc_tau_int.argtypes = [_i32p, n32, C.c_int64, C.c_double, C.c_int, C.c_int32, C.c_double]
    L.orc_tau_int.restype = C.c_double
    L.orc_tau_int_centred.argtypes = [_f64p, C.c_int64, C.c_int32, C.c_double]
    L.orc_tau_int_centred.restype = C.c_double


# --------------------------------------------------------------------------
# thin Python API
# --------------------------------------------------------------------------

def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def site_group(dim, L, site):
    g, b, c = C.c_uint32(), C.c_int(), C.c_int()
    lib().orc_site_group(dim, L, site, C.byref(g), C.byref(b), C.byref(c))
    return g.value, b.value, c.value


def ising_group(dim, L, site):
    """(group, bit, colour) of an Ising site's bit-sliced sweep stream."""
    g, b, c = C.c_uint32(), C.c_int(), C.c_int()
    lib().orc_ising_group(dim, L, site, C.byref(g), C.byref(b), C.byref(c))
    return g.value, b.value, c.value


def bitsliced_word(seed, chain_key, group, bit, sweep, purpose):
    return lib().orc_bitsliced_word(seed, chain_key, group, bit, sweep, purpose)


def site_word(seed, chain_key, group, bit, sweep, purpose):
    return lib().orc_site_word(seed, chain_key, group, bit, sweep, purpose)


def bc_site_uniform(seed, chain_key, group, bit, sweep, purpose):
    """(U, alt) of a Blume-Capel site: U = u16 * 2^16 + t16, alt = the proposal bit."""
    alt = C.c_int()
    U = lib().orc_bc_site_uniform(seed, chain_key, group, bit, sweep, purpose, C.byref(alt))
    return U, alt.value


def pt_word(seed, chain_key, rnd):
    return lib().orc_pt_word(seed, chain_key, rnd)


def logistic(x):
    return lib().orc_logistic(x)


def threshold32(p):
    return lib().orc_threshold32(p)


def pair_sum(dim, L, spins):
    return lib().orc_pair_sum(dim, L, np.ascontiguousarray(spins, np.int8))


def magnetization(dim, L, spins):
    return lib().orc_magnetization(dim, L, np.ascontiguousarray(spins, np.int8))


def sum_sq(dim, L, spins):
    return lib().orc_sum_sq(dim, L, np.ascontiguousarray(spins, np.int8))


def init_random(model, dim, L, seed, chain_key, sweep=0):
    N = L ** dim
    s = np.zeros(N, np.int8)
    lib().orc_init_random(model, dim, L, seed, chain_key, sweep, s)
    return s


def canonical_sweeps(model, alg, dim, L, beta, spins, seed, chain_key, sweep0, nsweeps, D=0.0):
    """Run nsweeps in place on `spins` (int8, modified).  Returns (E_pair, M, Q, accepted)."""
    E = np.zeros(nsweeps, np.int32)
    M = np.zeros(nsweeps, np.int32)
    Q = np.zeros(nsweeps, np.int32)
    acc = C.c_int64(0)
    lib().orc_canonical_sweeps(model, alg, dim, L, beta, D, seed, chain_key, sweep0, nsweeps,
                               spins, E, M, Q, C.byref(acc))
    return E, M, Q, acc.value


class BoundsError(IndexError):
    pass


def generalized_sweeps(model, wl, dim, L, lw, hist, spins, seed, chain_key, sweep0, nsweeps,
                       bin_start, bin_step, logf=0.0, D=0.0, int_bins=None):
    """Sequential multicanonical (wl=0) / Wang-Landau (wl=1) sweeps, in place on
    spins, lw (float64) and hist (int64).  int_bins: Ising energies on an
    integer-typed binning (truncating div; default for Ising), else rounded.
    Returns (E_pair, M, Q, steps, accepted)."""
    if int_bins is None:
        int_bins = model == 0
    E = np.zeros(nsweeps, np.int32)
    M = np.zeros(nsweeps, np.int32)
    Q = np.zeros(nsweeps, np.int32)
    steps, acc, err = C.c_int64(0), C.c_int64(0), C.c_int64(-1)
    rc = lib().orc_generalized_sweeps(model, wl, dim, L, D, float(bin_start), float(bin_step),
                                      len(lw), 1 if int_bins else 0, lw, hist, logf, seed,
                                      chain_key, sweep0, nsweeps,
                                      spins, C.byref(steps), C.byref(acc), E, M, Q, C.byref(err))
    if rc == 2:
        raise BoundsError(f"energy bin out of range at site {err.value}")
    return E, M, Q, steps.value, acc.value


def pair_spin2_sweeps(dim, L, beta, lw, hist, spins, seed, chain_key, sweep0, nsweeps,
                      bin_start, bin_step):
    """Blume-Capel under the CustomEnsemble (Boltzmann pair term, multicanonical
    sum s^2) of muca_BlumeCapel.jl, in place on spins and hist (int64); lw is
    read only.  Returns (E_pair, M, Q, steps, accepted)."""
    E = np.zeros(nsweeps, np.int32)
    M = np.zeros(nsweeps, np.int32)
    Q = np.zeros(nsweeps, np.int32)
    steps, acc, err = C.c_int64(0), C.c_int64(0), C.c_int64(-1)
    rc = lib().orc_pair_spin2_sweeps(dim, L, float(beta), float(bin_start), float(bin_step),
                                     len(lw), np.ascontiguousarray(lw, np.float64), hist, seed,
                                     chain_key, sweep0, nsweeps, spins, C.byref(steps),
                                     C.byref(acc), E, M, Q, C.byref(err))
    if rc == 2:
        raise BoundsError(f"sum s^2 bin out of range at site {err.value}")
    return E, M, Q, steps.value, acc.value


def pt_round(betas_of_ladder, x, indices, steps, accepted, stage, seed, chain_keys, rnd):
    """One exchange round; mutates indices/steps/accepted, returns new stage."""
    st = C.c_int32(stage)
    lib().orc_pt_round(len(x), np.ascontiguousarray(betas_of_ladder, np.float64),
                       np.ascontiguousarray(x, np.float64), indices, steps, accepted,
                       C.byref(st), seed, np.ascontiguousarray(chain_keys, np.uint32), rnd)
    return st.value


def exchange_pair(beta_i, beta_j, x_i, x_j, u):
    """(accepted, log ratio) of one replica-exchange attempt with shared u."""
    lr = C.c_double()
    ok = lib().orc_exchange_pair(beta_i, beta_j, x_i, x_j, u, C.byref(lr))
    return bool(ok), lr.value


def refpath_series_ising2d(L, beta, seed, nsweeps, spins):
    """Reference CPU path (random site, Xoshiro256++) on one chain; spins in place.
    Returns (E_pair, M, accepted) per sweep."""
    E = np.zeros(nsweeps, np.int32)
    M = np.zeros(nsweeps, np.int32)
    acc = C.c_int64(0)
    lib().orc_refpath_series_ising2d(L, beta, seed, nsweeps, spins, E, M, C.byref(acc))
    return E, M, acc.value


def muca_update(lw, h):
    lib().orc_muca_update(len(lw), lw, np.ascontiguousarray(h, np.float64))


def refpath_ising2d(L, betas, seed, nsweeps, nthreads, keep_spins=False):
    betas = np.ascontiguousarray(betas, np.float64)
    out = np.zeros(len(betas) * L * L, np.int8) if keep_spins else None
    n = lib().orc_refpath_ising2d(L, len(betas), betas, seed, nsweeps, nthreads, out)
    return n, out


def refpath_run(model, dim, Ls, betas, seed, nsweeps, nthreads, D=None, lw=None,
                bin_start=0.0, bin_step=1.0, logf=None):
    """General reference CPU path (random site, Xoshiro256++): Ising or
    Blume-Capel (model 0 / 1), 2D or 3D, per-chain L / beta / D; canonical
    Metropolis, or multicanonical with weights lw over DiscreteBinning(bin_start,
    bin_step, len(lw)), or (logf given) Wang-Landau walkers starting from lw.
    Returns the attempts made (timing baseline only)."""
    Ls = np.ascontiguousarray(Ls, np.int32)
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float64), Ls.shape))
    Dv = None if D is None else np.ascontiguousarray(np.broadcast_to(np.asarray(D, np.float64),
                                                                      Ls.shape))
    lwv = None if lw is None else np.ascontiguousarray(lw, np.float64)
    mode = 0 if lwv is None else (2 if logf is not None else 1)
    return lib().orc_refpath_run(model, dim, mode, len(Ls), Ls, betas, Dv, lwv,
                                 0 if lwv is None else len(lwv), float(bin_start), float(bin_step),
                                 seed, nsweeps, nthreads, 0.0 if logf is None else float(logf))


def _lag_cap(n, max_lag):
    return n // 2 if max_lag is None else min(int(max_lag), n // 2)


def tau_int(x, y=None, D=0.0, absx=False, max_lag=None, c=5.0):
    """integrated_autocorrelation_time of an integer series (exact lag sums);
    with y: of the energy x + D*y."""
    x = np.ascontiguousarray(x, np.int32)
    yy = None if y is None else np.ascontiguousarray(y, np.int32)
    return lib().orc_tau_int(x, yy, x.size, float(D), int(absx), _lag_cap(x.size, max_lag), c)


def tau_int_centred(x, max_lag=None, c=5.0):
    """integrated_autocorrelation_time as the reference evaluates it (centred FP64)."""
    x = np.ascontiguousarray(x, np.float64)
    return lib().orc_tau_int_centred(x, x.size, _lag_cap(x.size, max_lag), c)


def tau_int_times(traces, ys=None, D=None, absx=False, min_points=2, max_lag=None, c=5.0):
    """integrated_autocorrelation_times (autocorrelations.jl:84-110) over integer
    traces: NaN below min_points or for an empty lag window."""
    out = np.full(len(traces), np.nan)
    for k, x in enumerate(traces):
        n = len(x)
        if n < min_points:
            continue
        cap = _lag_cap(n, max_lag)
        if cap < 1:
            continue
        out[k] = tau_int(x, None if ys is None else ys[k], 0.0 if D is None else D[k], absx,
                         cap, c)
    return out


def _gen_group_stats(hists, reached):
    """{min, mean, bins} of the summed histograms over the reached bins, the
    mask OR-ed with sum > 0 first (the flatness rule of the walker schedule)."""
    h = np.zeros_like(hists[0])
    for x in hists:
        h = h + x
    reached |= h > 0
    if not reached.any():
        return 0.0, 0.0, 0
    hr = h[reached]
    return float(hr.min()), float(hr.sum()) / float(hr.size), int(hr.size)


def wl_iterate_group(model, dim, L, spins, lw, hists, logf, reached, seed, keys, sweep0,
                     bin_start, bin_step, sweeps, max_updates, flatness, logf_final, power,
                     D=0.0, int_bins=None):
    """The Wang-Landau schedule of one binning group of walkers, restated
    sequentially (test infrastructure): per update, `sweeps` sweeps of every
    walker (orc_generalized_sweeps, wl = 1, ln f = logf), the walkers' mean
    log-weight summed in walker order and copied to every walker, and, when the
    summed histogram is flat over the reached bins (min >= flatness * mean),
    logf *= power and the histograms are zeroed; stops once logf < logf_final
    or after max_updates updates.  update!(::WangLandauEnsemble; power) is
    ensembles/wang_landau.jl:23; the accept rule algorithms/wang_landau.jl:29-37.
    spins / lw / hists (lists, one per walker) and reached are updated in
    place.  Returns (logf, sweep counter, updates, stages, steps, accepted)."""
    n = len(spins)
    updates = stages = 0
    steps = [0] * n
    acc = [0] * n
    sweep = sweep0
    while updates < max_updates and logf >= logf_final:
        for k in range(n):
            _, _, _, st, ac = generalized_sweeps(model, 1, dim, L, lw[k], hists[k], spins[k], seed,
                                                 keys[k], sweep, sweeps, bin_start, bin_step,
                                                 logf=logf, D=D, int_bins=int_bins)
            steps[k] += st
            acc[k] += ac
        sweep += sweeps
        updates += 1
        mean = np.zeros_like(lw[0])
        for k in range(n):
            mean = mean + lw[k]
        mean = mean / float(n)
        for k in range(n):
            lw[k][:] = mean
        mn, avg, cnt = _gen_group_stats(hists, reached)
        if cnt > 0 and mn >= flatness * avg:
            logf = logf * power
            stages += 1
            for h in hists:
                h[:] = 0
    return logf, sweep, updates, stages, steps, acc


def muca_iterate_update(model, dim, L, spins, lw, hists, seed, keys, sweep0, bin_start, bin_step,
                        sweeps, therm_sweeps, D=0.0, int_bins=None):
    """One update of the parallel multicanonical iteration of one group,
    restated sequentially (test infrastructure): therm_sweeps sweeps, reset!
    (histograms and counters, algorithms/multicanonical.jl:27-33), `sweeps`
    sweeps, merge_histograms! into the root (parallel_multicanonical.jl:38-47),
    update! of the root (ensembles/multicanonical.jl:32-44, libm log), and
    distribute_logweight! (:59-68).  In place; returns (sweep counter, steps,
    accepted) of the recording sweeps."""
    n = len(spins)
    sweep = sweep0
    if therm_sweeps > 0:
        for k in range(n):
            generalized_sweeps(model, 0, dim, L, lw[k], hists[k], spins[k], seed, keys[k], sweep,
                               therm_sweeps, bin_start, bin_step, D=D, int_bins=int_bins)
        sweep += therm_sweeps
    steps, acc = [0] * n, [0] * n
    for k in range(n):
        hists[k][:] = 0
        _, _, _, steps[k], acc[k] = generalized_sweeps(model, 0, dim, L, lw[k], hists[k], spins[k],
                                                       seed, keys[k], sweep, sweeps, bin_start,
                                                       bin_step, D=D, int_bins=int_bins)
    sweep += sweeps
    for k in range(1, n):
        hists[0] += hists[k]
    muca_update(lw[0], hists[0].astype(np.float64))
    for k in range(1, n):
        lw[k][:] = lw[0]
    return sweep, steps, acc
