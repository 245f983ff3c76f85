"""Multicanonical / Wang-Landau walkers on the GPU.

  Multicanonical(rng, bins), WangLandau(rng, bins; logf)   src/algorithms/multicanonical.jl,
                                                           src/algorithms/wang_landau.jl
  update!(::MulticanonicalEnsemble), update!(::WangLandauEnsemble)
                                                           src/ensembles/{multicanonical,wang_landau}.jl
  ParallelMulticanonical, merge_histograms!, distribute_logweight!
                                                           src/algorithms/parallel_multicanonical.jl
  Metropolis(rng, CustomEnsemble(pair, spin2)) on Blume-Capel, reset_histogram!
                                                           docs/src/examples/spin_systems/muca_BlumeCapel.jl:43-89

GeneralizedWalkers binds the ensembles of a list of ImportanceSampling
algorithms to device walkers (mcx_gen).  While bound, the device holds the
log-weights and histograms; the host BinnedObject arrays are refreshed after
every device call and uploaded again when the caller changed them.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import MCXArgumentError, call
from .algorithms import ImportanceSampling, Metropolis
from .ensembles import (BinnedObject, CustomEnsemble, DiscreteBinning, MulticanonicalEnsemble,
                        WangLandauEnsemble)


class GenDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_groups", C.c_int32),
                ("group_of_chain", C.POINTER(C.c_int32)), ("bin_start", C.POINTER(C.c_double)),
                ("bin_step", C.POINTER(C.c_double)), ("n_bins", C.POINTER(C.c_int64)),
                ("pair_beta", C.POINTER(C.c_double)), ("bin_integer", C.POINTER(C.c_int32))]


class GenSchedule(C.Structure):
    _fields_ = [("sweeps", C.c_int64), ("therm_sweeps", C.c_int64), ("max_updates", C.c_int64),
                ("flatness", C.c_double), ("logf_final", C.c_double), ("power", C.c_double),
                ("flags", C.c_uint32), ("pad", C.c_int32)]


class GenProgress(C.Structure):
    _fields_ = [("updates", C.c_int64), ("stages", C.c_int64), ("sweeps", C.c_int64),
                ("active", C.c_int32), ("pad", C.c_int32), ("device_ms", C.c_double)]


_P = C.c_void_p
_pd, _pi64, _pi32 = C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
for _n, _a in {
    "mcx_gen_create": [_P, C.POINTER(GenDesc), C.POINTER(_P)],
    "mcx_gen_destroy": [_P],
    "mcx_gen_set_logweight": [_P, C.c_int32, _pd, C.c_int64],
    "mcx_gen_get_logweight": [_P, C.c_int32, _pd, C.c_int64],
    "mcx_gen_get_histogram": [_P, C.c_int32, _pi64, C.c_int64],
    "mcx_gen_set_histogram": [_P, C.c_int32, _pi64, C.c_int64],
    "mcx_gen_set_logf": [_P, C.c_int32, C.c_double],
    "mcx_gen_get_logf": [_P, _pd],
    "mcx_gen_run": [_P, C.c_int64, C.c_uint32],
    "mcx_gen_reset": [_P, C.c_int32],
    "mcx_gen_update": [_P, C.c_int32],
    "mcx_gen_merge": [_P, C.c_int32],
    "mcx_gen_distribute": [_P, C.c_int32],
    "mcx_gen_average_logweight": [_P, C.c_int32],
    "mcx_gen_flatness": [_P, C.c_int32, _pd],
    "mcx_gen_get_counters": [_P, _pi64, _pi64],
    "mcx_gen_group_info": [_P, C.c_int32, _pi32, _pi32, _pi64],
    "mcx_muca_update_values": [_P, _pd, _pd, C.c_int64],
    "mcx_gen_reset_histogram": [_P, C.c_int32],
    "mcx_gen_set_pair_beta": [_P, C.c_int32, C.c_double],
    "mcx_gen_state_size": [_P, C.POINTER(C.c_size_t)],
    "mcx_gen_save_state": [_P, _P, C.c_size_t],
    "mcx_gen_load_state": [_P, _P, C.c_size_t],
    "mcx_gen_iterate": [_P, C.POINTER(GenSchedule), C.POINTER(GenProgress)],
}.items():
    _lib.declare(_n, _a)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def muca_update_values(lw: np.ndarray, h: np.ndarray, ctx=None):
    """lw[i] -= log(h[i]) where h[i] > 0, on the device, in place."""
    ctx = ctx or _lib.default_context()
    if lw.dtype != np.float64 or not lw.flags.c_contiguous:
        raise MCXArgumentError("logweight values must be a contiguous float64 array")
    hh = np.ascontiguousarray(h, np.float64)
    call("mcx_muca_update_values", ctx.handle, _ptr(lw, C.c_double), _ptr(hh, C.c_double), lw.size)


def is_walker_alg(a) -> bool:
    """Algorithms that run as device walkers: multicanonical / Wang-Landau
    ImportanceSampling, or Metropolis over a CustomEnsemble."""
    return isinstance(a, ImportanceSampling) or (type(a) is Metropolis and
                                                 isinstance(a.ensemble, CustomEnsemble))


def _binned(a):
    """The ensemble holding the walker's log-weights and histogram."""
    e = a.ensemble
    return e.spin2 if isinstance(e, CustomEnsemble) else e


def _binning(ens):
    """(start, step, num, integer-typed) of the ensemble's DiscreteBinning."""
    b = ens.logweight.bins
    if not isinstance(b, DiscreteBinning):
        raise MCXArgumentError("device walkers need discrete (start, step, num) binning")
    return (float(b.start), float(b.step), int(b.num), bool(b.integer))


class GeneralizedWalkers:
    """Device walkers for the ImportanceSampling algorithms `algs` driving the
    systems of a SpinSystemBatch (one algorithm per system).  Walkers whose
    ensembles share a binning form a group; the first of a group is its root."""

    def __init__(self, systems, algs: Sequence[ImportanceSampling]):
        algs = list(algs)
        if len(algs) != len(systems):
            raise MCXArgumentError("one algorithm per system")
        kinds = {type(a.ensemble) for a in algs}
        if kinds == {MulticanonicalEnsemble} and all(isinstance(a, ImportanceSampling) for a in algs):
            self.kind = _lib.MULTICANONICAL
        elif kinds == {WangLandauEnsemble} and all(isinstance(a, ImportanceSampling) for a in algs):
            self.kind = _lib.WANGLANDAU
        elif kinds == {CustomEnsemble} and all(type(a) is Metropolis for a in algs):
            self.kind = _lib.MUCA_PAIR_SPIN2
            for a in algs:
                b = a.ensemble.spin2.logweight.bins
                if not isinstance(b, DiscreteBinning) or b.step < 1 or \
                        float(b.start) != int(b.start) or float(b.step) != int(b.step) or \
                        (b.step != 1 and not isinstance(b.step, (int, np.integer))):
                    raise MCXArgumentError("sum s^2 is an integer: bin it with an integer range "
                                           "start:step:stop")
                if not a.ensemble.record_visits:
                    raise MCXArgumentError("device walkers record every visit (record_visits=true)")
        else:
            raise MCXArgumentError("all walkers use MulticanonicalEnsemble, all WangLandauEnsemble, "
                                   "or all Metropolis over a CustomEnsemble")
        self.systems, self.algs = systems, algs
        keys, groups, self.group_of = [], [], []
        for a in algs:
            bk = _binning(_binned(a))
            if bk not in keys:
                keys.append(bk)
                groups.append([])
            g = keys.index(bk)
            groups[g].append(len(self.group_of))
            self.group_of.append(g)
        self.groups = groups
        start = np.array([k[0] for k in keys])
        step = np.array([k[1] for k in keys])
        nb = np.array([k[2] for k in keys], np.int64)
        integer = np.array([1 if k[3] else 0 for k in keys], np.int32)
        gof = np.array(self.group_of, np.int32)
        systems.bind_streams([a.rng for a in algs])
        pb = None
        if self.kind == _lib.MUCA_PAIR_SPIN2:
            pb = np.array([float(a.ensemble.pair.beta) for a in algs], np.float64)
        d = GenDesc(self.kind, len(keys), _ptr(gof, C.c_int32), _ptr(start, C.c_double),
                    _ptr(step, C.c_double), _ptr(nb, C.c_int64),
                    _ptr(pb, C.c_double) if pb is not None else None, _ptr(integer, C.c_int32))
        self._h = C.c_void_p()
        call("mcx_gen_create", systems.batch.handle, C.byref(d), C.byref(self._h))
        self._last_lw = [None] * len(algs)
        self._last_h = [None] * len(algs)
        self._last_logf = [None] * len(algs)
        self._last_pb = None if pb is None else pb.copy()
        for k, a in enumerate(algs):
            a.ensemble._dev = (self, k)
            if self.kind == _lib.MUCA_PAIR_SPIN2:
                a.ensemble.spin2._dev = (self, k)
        self.sync_in()

    # ---- host <-> device ---------------------------------------------------------
    def sync_in(self):
        if self._last_pb is not None:
            for k, a in enumerate(self.algs):
                b = float(a.ensemble.pair.beta)
                if b != self._last_pb[k]:
                    call("mcx_gen_set_pair_beta", self._h, k, b)
                    self._last_pb[k] = b
        for k, a in enumerate(self.algs):
            e = _binned(a)
            lw = np.ascontiguousarray(e.logweight.values, np.float64)
            if self._last_lw[k] is None or not np.array_equal(lw, self._last_lw[k]):
                call("mcx_gen_set_logweight", self._h, k, _ptr(lw, C.c_double), lw.size)
                self._last_lw[k] = lw.copy()
            if isinstance(e, MulticanonicalEnsemble):
                h = e.histogram.values
                if self._last_h[k] is None or not np.array_equal(h, self._last_h[k]):
                    hi = np.ascontiguousarray(np.rint(h), np.int64)
                    if not np.array_equal(hi, h):
                        raise MCXArgumentError("device histograms hold integer counts")
                    call("mcx_gen_set_histogram", self._h, k, _ptr(hi, C.c_int64), hi.size)
                    self._last_h[k] = np.asarray(h, np.float64).copy()
            else:
                if self._last_logf[k] != e.logf:
                    call("mcx_gen_set_logf", self._h, k, float(e.logf))
                    self._last_logf[k] = e.logf

    def sync_out(self):
        n = len(self.algs)
        st = np.empty(n, np.int64)
        ac = np.empty(n, np.int64)
        call("mcx_gen_get_counters", self._h, _ptr(st, C.c_int64), _ptr(ac, C.c_int64))
        for k, a in enumerate(self.algs):
            e = _binned(a)
            lw = np.empty(len(e.logweight.values), np.float64)
            call("mcx_gen_get_logweight", self._h, k, _ptr(lw, C.c_double), lw.size)
            e.logweight.values[:] = lw
            self._last_lw[k] = lw
            h = np.empty(len(lw), np.int64)
            call("mcx_gen_get_histogram", self._h, k, _ptr(h, C.c_int64), h.size)
            if isinstance(e, MulticanonicalEnsemble):
                e.histogram.values[:] = h
                self._last_h[k] = h.astype(np.float64)
            e.device_histogram = h
            a.steps, a.accepted = int(st[k]), int(ac[k])

    # ---- operations ----------------------------------------------------------------
    def run_(self, nsweeps: int, record: bool = False) -> float:
        self.sync_in()
        try:
            call("mcx_gen_run", self._h, int(nsweeps), _lib.RUN_RECORD if record else 0)
        finally:
            self.sync_out()
        return self.systems.batch.last_run_ms()

    def reset_(self, k: int = -1):
        self.sync_in()
        call("mcx_gen_reset", self._h, int(k))
        self.sync_out()

    def reset_histogram_(self, k: int = -1):
        self.sync_in()
        call("mcx_gen_reset_histogram", self._h, int(k))
        self.sync_out()

    def update_(self, k: int = -1):
        self.sync_in()
        call("mcx_gen_update", self._h, int(k))
        self.sync_out()

    def merge_(self, group: int = 0):
        self.sync_in()
        call("mcx_gen_merge", self._h, int(group))
        self.sync_out()

    def distribute_(self, group: int = 0):
        self.sync_in()
        call("mcx_gen_distribute", self._h, int(group))
        self.sync_out()

    def average_(self, group: int = 0):
        self.sync_in()
        call("mcx_gen_average_logweight", self._h, int(group))
        self.sync_out()

    def flatness(self, group: int = 0):
        """(min, mean, bins) of the group root's histogram over the group's
        reached bins (every bin any flatness evaluation has seen non-zero)."""
        self.sync_in()
        out = np.empty(3, np.float64)
        call("mcx_gen_flatness", self._h, int(group), _ptr(out, C.c_double))
        return float(out[0]), float(out[1]), int(out[2])

    def iterate_(self, sweeps: int, max_updates: int, therm_sweeps: int = 0,
                 flatness: float = 0.8, logf_final: float = 1e-6, power: float = 0.5) -> dict:
        """The weight iteration on the device (mcx_gen_iterate), per binning
        group.  Multicanonical walkers: per update, therm_sweeps sweeps,
        reset!, `sweeps` sweeps, merge_histograms!, update! of the root,
        distribute_logweight! (the ParallelMulticanonical loop of
        parallel_chains_threads.jl:180-201).  Wang-Landau walkers: per update,
        `sweeps` sweeps, the group's mean log-weight to every walker, and when
        the summed histogram is flat (min >= flatness * mean over the reached
        bins) ln f <- power * ln f and the histograms start again; a group stops
        below logf_final.  Returns the progress counters."""
        self.sync_in()
        sc = GenSchedule(int(sweeps), int(therm_sweeps), int(max_updates), float(flatness),
                         float(logf_final), float(power), 0, 0)
        pr = GenProgress()
        try:
            call("mcx_gen_iterate", self._h, C.byref(sc), C.byref(pr))
        finally:
            self.sync_out()
            if self.kind == _lib.WANGLANDAU:
                lf = np.empty(len(self.algs), np.float64)
                call("mcx_gen_get_logf", self._h, _ptr(lf, C.c_double))
                for k, a in enumerate(self.algs):
                    a.ensemble.logf = float(lf[k])
                    self._last_logf[k] = a.ensemble.logf
        return {"updates": pr.updates, "stages": pr.stages, "sweeps": pr.sweeps,
                "active_groups": pr.active, "device_ms": pr.device_ms}

    def save_state(self) -> np.ndarray:
        """Log-weights, histograms, ln f and counters of every walker (uint8);
        host-side changes to the ensembles are uploaded first."""
        self.sync_in()
        n = C.c_size_t()
        call("mcx_gen_state_size", self._h, C.byref(n))
        buf = np.empty(n.value, np.uint8)
        call("mcx_gen_save_state", self._h, buf.ctypes.data_as(C.c_void_p), buf.size)
        return buf

    def load_state(self, blob):
        """Restore a snapshot of walkers of the same kind and bin layout; the
        host ensembles and counters are refreshed from it."""
        b = np.ascontiguousarray(blob, np.uint8)
        call("mcx_gen_load_state", self._h, b.ctypes.data_as(C.c_void_p), b.size)
        self.sync_out()

    def close(self):
        if getattr(self, "_h", None):
            for a in self.algs:
                a.ensemble._dev = None
                if isinstance(a.ensemble, CustomEnsemble):
                    a.ensemble.spin2._dev = None
            call("mcx_gen_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ThreadsBackend:
    """ThreadsBackend(n) (src/infrastructure/parallel_backends.jl:25-33): the
    chains of one device batch."""

    def __init__(self, n: int):
        self.n = int(n)

    def size(self):
        return self.n


class ParallelChains:
    """ParallelChains(backend, algs) (src/infrastructure/parallel_chains.jl:14-128)."""

    def __init__(self, backend: ThreadsBackend, algs: Sequence):
        if backend.n != len(algs):
            raise MCXArgumentError("size(backend) must equal length(algs)")
        self.backend, self.alg = backend, list(algs)

    def size(self):
        return len(self.alg)

    def rank(self):
        return 0

    def is_root(self):
        return True

    def algorithm(self, i: int):
        return self.alg[i - 1]


def ParallelMulticanonical(backend: ThreadsBackend, algs: Sequence[ImportanceSampling]):
    for a in algs:
        if not isinstance(a.ensemble, MulticanonicalEnsemble):
            raise MCXArgumentError("ParallelMulticanonical needs MulticanonicalEnsemble algorithms")
    return ParallelChains(backend, algs)


def _walkers_of(pc: ParallelChains):
    devs = {getattr(a.ensemble, "_dev", None) and id(a.ensemble._dev[0]) for a in pc.alg}
    if len(devs) == 1 and None not in devs:
        w = pc.alg[0].ensemble._dev[0]
        return w, w.group_of[pc.alg[0].ensemble._dev[1]]
    return None, None


def merge_histograms_(pc: ParallelChains):
    """Sum every chain's histogram into the root chain's (parallel_multicanonical.jl:38-47)."""
    w, g = _walkers_of(pc)
    if w is None:
        raise MCXArgumentError("merge_histograms! runs on device walkers: run the chains first")
    w.merge_(g)


def distribute_logweight_(pc: ParallelChains):
    """Copy the root's log-weights to every chain (parallel_multicanonical.jl:59-68)."""
    w, g = _walkers_of(pc)
    if w is None:
        raise MCXArgumentError("distribute_logweight! runs on device walkers: run the chains first")
    w.distribute_(g)
