"""Replica exchange / parallel tempering on the GPU.

  ReplicaExchange fields indices/steps/accepted/stage  src/algorithms/replica_exchange.jl:13-19
  update!(rx, xs)                                      replica_exchange.jl:158-178
  ParallelTempering(betas; seed, rng)                  src/algorithms/parallel_tempering.jl:24-53
  set_betas(n, bmin, bmax, mode)                       parallel_tempering.jl:146-162
  acceptance_rates / acceptance_rate / index / reset!  replica_exchange.jl:48-101
  optimize_exchange_interval!, _retune_exchange_sweeps!
                                                       parallel_tempering.jl:53-135

The exchange itself runs in libmcx_hip.so (one thread per ladder); this class
mirrors the device state after every call.
"""
from __future__ import annotations

import ctypes as C
import math
from fractions import Fraction
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import MCXArgumentError, call
from .algorithms import Glauber, HeatBath, Metropolis, Philox, algorithm_code
from .ensembles import BoltzmannEnsemble


class PTDesc(C.Structure):
    _fields_ = [("n_sets", C.c_int32), ("replicas", C.c_int32),
                ("betas", C.POINTER(C.c_double))]


_P = C.c_void_p
for _n, _a in {
    "mcx_pt_create": [_P, C.POINTER(PTDesc), C.POINTER(_P)],
    "mcx_pt_destroy": [_P],
    "mcx_pt_reset": [_P],
    "mcx_pt_exchange": [_P, C.POINTER(C.c_double)],
    "mcx_pt_run": [_P, C.c_int32, C.c_int64, C.c_int64, C.c_uint32],
    "mcx_pt_get_ladder": [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "mcx_pt_get_counts": [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "mcx_pt_get_rounds": [_P, C.POINTER(C.c_uint64)],
    "mcx_pt_set_series_order": [_P, C.c_int32],
    "mcx_pt_state_size": [_P, C.POINTER(C.c_size_t)],
    "mcx_pt_save_state": [_P, _P, C.c_size_t],
    "mcx_pt_load_state": [_P, _P, C.c_size_t],
}.items():
    _lib.declare(_n, _a)

RUN_GRAPH = 4
_ORDERS = {"replica": 0, "temperature": 1}


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def set_betas(nreplicas: int, bmin: float, bmax: float, mode: str = "uniform") -> np.ndarray:
    """Beta ladder from bmax down to bmin.  uniform: exact linear interpolation
    of the end points, correctly rounded; geometric: exp of the uniform ladder
    of the logs."""
    mode = mode.lstrip(":")
    if nreplicas < 2:
        raise MCXArgumentError("nreplicas must be >= 2")

    def lin(a, b):
        fa, fb = Fraction(a), Fraction(b)
        return [float(fa + (fb - fa) * Fraction(k, nreplicas - 1)) for k in range(nreplicas)]

    if mode == "uniform":
        return np.array(lin(float(bmax), float(bmin)))
    if mode == "geometric":
        return np.exp(np.array(lin(math.log(float(bmax)), math.log(float(bmin)))))
    raise MCXArgumentError(f"unknown beta mode {mode}; use :uniform or :geometric")


class ParallelTempering:
    """n_sets independent replica-exchange ladders of R replicas over a
    SpinSystemBatch of n_sets*R systems (replica r of set s = system s*R + r).

    betas: ladder of one set (broadcast) or n_sets*R values.  Replica i
    (1-based, as in the reference) draws from Philox(seed, key = s*R + i),
    the counterpart of rng(seed + i)."""

    def __init__(self, systems, betas, n_sets: int = 1, seed: int = 1000, algs=None,
                 alg: str = "metropolis"):
        from .spinsystems import SpinSystemBatch
        if not isinstance(systems, SpinSystemBatch):
            raise MCXArgumentError("ParallelTempering needs a SpinSystemBatch")
        n = len(systems)
        if n % n_sets:
            raise MCXArgumentError("number of systems must be a multiple of n_sets")
        R = n // n_sets
        if R < 2:
            raise MCXArgumentError("need at least 2 replicas")
        b = np.asarray(betas, np.float64).ravel()
        if b.size == R:
            b = np.tile(b, n_sets)
        if b.size != n:
            raise MCXArgumentError(f"size(betas) must be R={R} or n_sets*R={n}")
        self.systems, self.n_sets, self.R, self.ladder = systems, n_sets, R, b
        cls = {"heatbath": HeatBath, "glauber": Glauber}.get(alg, Metropolis)
        if algs is None:
            algs = [cls(Philox(seed, s * R + r + 1), beta=float(b[s * R + r]))
                    for s in range(n_sets) for r in range(R)]
        self.algs = list(algs)
        systems.bind(self.algs)
        systems._tables = None  # the ladder now owns the tables
        d = PTDesc(n_sets, R, _ptr(self.ladder, C.c_double))
        self._h = C.c_void_p()
        call("mcx_pt_create", systems.batch.handle, C.byref(d), C.byref(self._h))
        self._alg = algorithm_code(self.algs[0])
        self.series_order = "replica"
        systems._pt = self
        self._sync()

    # ---- mirrored state -----------------------------------------------------
    def _sync(self):
        n, S, R = self.n_sets * self.R, self.n_sets, self.R
        idx = np.empty(n, np.int32)
        st = np.empty(S, np.int32)
        call("mcx_pt_get_ladder", self._h, _ptr(idx, C.c_int32), _ptr(st, C.c_int32))
        steps = np.empty(S * (R - 1), np.int32)
        acc = np.empty(S * (R - 1), np.int32)
        call("mcx_pt_get_counts", self._h, _ptr(steps, C.c_int32), _ptr(acc, C.c_int32))
        self.indices = idx.reshape(S, R)
        self.stage = st
        self.steps = steps.reshape(S, R - 1)
        self.accepted = acc.reshape(S, R - 1)
        # the ensembles move with the ladder index, as attempt_exchange_pair! swaps them
        for s in range(S):
            for r in range(R):
                a = self.algs[s * R + r]
                beta = float(self.ladder[s * R + self.indices[s, r] - 1])
                if isinstance(a, HeatBath):
                    a.beta = beta
                else:
                    a.ensemble = BoltzmannEnsemble(beta=beta)

    def rounds(self) -> np.ndarray:
        r = np.empty(self.n_sets, np.uint64)
        call("mcx_pt_get_rounds", self._h, _ptr(r, C.c_uint64))
        return r

    def update_(self, xs: Optional[Sequence[float]] = None):
        """update!(rx, xs): one exchange round of every set; xs defaults to the
        replicas' current energies on the device."""
        if xs is None:
            call("mcx_pt_exchange", self._h, None)
        else:
            x = np.ascontiguousarray(xs, np.float64).ravel()
            if x.size != self.n_sets * self.R:
                raise MCXArgumentError("xs must have length size(rx)")
            call("mcx_pt_exchange", self._h, _ptr(x, C.c_double))
        self._sync()

    def run_(self, n_rounds: int, sweeps_between: int, record: bool = False,
             graph: bool = False) -> float:
        """n_rounds x (sweeps_between sweeps, exchange); returns device ms."""
        flags = (_lib.RUN_RECORD if record else 0) | (RUN_GRAPH if graph else 0)
        _, acc0 = self.systems.batch.counters()
        call("mcx_pt_run", self._h, self._alg, int(n_rounds), int(sweeps_between), flags)
        _, acc1 = self.systems.batch.counters()
        for k, a in enumerate(self.algs):
            a.steps += int(self.systems.batch.nsites[k]) * int(n_rounds * sweeps_between)
            if hasattr(a, "accepted"):
                a.accepted += int(acc1[k] - acc0[k])
        self._sync()
        return self.systems.batch.last_run_ms()

    def set_series_order(self, order: str):
        """"temperature": record series and moments per ladder slot (slot
        s*R + i - 1 holds whichever replica of set s is at ladder index i);
        "replica": per chain.  Clears the recorded series and moments."""
        o = order.lstrip(":").lower()
        if o not in _ORDERS:
            raise MCXArgumentError(f"unknown series order {order!r}; use :replica or :temperature")
        call("mcx_pt_set_series_order", self._h, _ORDERS[o])
        self.series_order = o
        return self

    def tau_int(self, observable: str = "energy", first: int = 0, count: Optional[int] = None,
                min_points: int = 2, max_lag: Optional[int] = None, c: float = 5.0) -> np.ndarray:
        """tau_int per ladder index from the temperature-ordered device series:
        shape (n_sets, R), or (R,) for one set."""
        if self.series_order != "temperature":
            raise _lib.MCXStateError("tau_int per ladder index needs set_series_order(:temperature) "
                                     "before the recording")
        t = self.systems.batch.tau_int(observable, first, count, min_points, max_lag, c)
        t = t.reshape(self.n_sets, self.R)
        return t if self.n_sets > 1 else t[0]

    def save_state(self) -> np.ndarray:
        """Ladder positions, exchange counters, stages and round counters (uint8)."""
        n = C.c_size_t()
        call("mcx_pt_state_size", self._h, C.byref(n))
        buf = np.empty(n.value, np.uint8)
        call("mcx_pt_save_state", self._h, buf.ctypes.data_as(C.c_void_p), buf.size)
        return buf

    def load_state(self, blob):
        """Restore a snapshot of a ladder of the same shape and betas."""
        b = np.ascontiguousarray(blob, np.uint8)
        call("mcx_pt_load_state", self._h, b.ctypes.data_as(C.c_void_p), b.size)
        self._sync()

    def reset_(self):
        call("mcx_pt_reset", self._h)
        self._sync()
        return self

    def close(self):
        if getattr(self, "_h", None):
            call("mcx_pt_destroy", self._h)
            self._h = C.c_void_p()
            self.systems._pt = None
            self.systems._tables = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return self.n_sets * self.R


ReplicaExchange = ParallelTempering


def index(rx, i: Optional[int] = None):
    """index(rx) / index(rx, i) (1-based replica i; single-set ladders)."""
    if i is None:
        return rx.indices if rx.n_sets > 1 else rx.indices[0]
    return int(rx.indices.ravel()[i - 1])


def acceptance_rates(rx) -> np.ndarray:
    s = rx.steps.astype(np.float64)
    out = np.where(s > 0, rx.accepted / np.maximum(s, 1), 0.0)
    return out if rx.n_sets > 1 else out[0]


def acceptance_rate_pt(rx) -> float:
    tot = rx.steps.sum()
    return float(rx.accepted.sum() / tot) if tot > 0 else 0.0


def _group_samples(samples, n: int):
    """Per-ladder-index traces from pre-grouped traces or flat (index, value)
    tuples (parallel_tempering.jl:53-70)."""
    samples = list(samples)
    if samples and all(isinstance(s, tuple) and len(s) == 2 for s in samples):
        grouped = [[] for _ in range(n)]
        for idx, e in samples:
            if not 1 <= int(idx) <= n:
                raise MCXArgumentError(f"sample index {idx} out of bounds for {n} ladders")
            grouped[int(idx) - 1].append(float(e))
        return grouped
    return samples


def _retune_exchange_sweeps_(sweeps, taus, base_sweeps: int, min_sweeps: int, max_sweeps: int):
    """sweeps[i] = clamp(round(base * tau_i / median(finite taus)), min, max);
    a non-finite tau keeps scale 1 (parallel_tempering.jl:123-135).  Rounding
    is half to even, as Julia's round(Int, x)."""
    taus = np.asarray(taus, np.float64)
    finite = taus[np.isfinite(taus)]
    tau_ref = float(np.median(finite)) if finite.size else 1.0
    for i in range(len(sweeps)):
        scale = taus[i] / tau_ref if np.isfinite(taus[i]) else 1.0
        target = int(np.rint(base_sweeps * scale))
        sweeps[i] = min(max(target, min_sweeps), max_sweeps)
    return sweeps


def optimize_exchange_interval_(pt, local_samples, sweeps_after_exchange, base_sweeps: int,
                                min_sweeps: int = 1, max_sweeps: int = 2 ** 63 - 1,
                                min_points: int = 400, max_lag: Optional[int] = 200,
                                observable: str = "energy", first: int = 0,
                                count: Optional[int] = None):
    """optimize_exchange_interval!(pt, local_samples, sweeps_after_exchange;
    base_sweeps, min_sweeps, max_sweeps, min_points, max_lag)
    (parallel_tempering.jl:83-121): retune the post-exchange sweeps per ladder
    index from tau_int of the energy at each index; returns
    sweeps_after_exchange[index(pt)] (one value per replica).

    local_samples: per-index traces or flat (index, value) tuples as in the
    reference, or None for the temperature-ordered series recorded on the
    device (set_series_order(:temperature); sweeps [first, first + count)),
    whose tau_int is computed there.  sweeps_after_exchange (numpy int array,
    updated in place) has length R, or shape (n_sets, R) with device samples:
    every set is retuned from its own ladder."""
    R, S = pt.R, pt.n_sets
    sw = sweeps_after_exchange
    if local_samples is None:
        taus = np.asarray(pt.tau_int(observable, first, count, min_points, max_lag)).reshape(S, R)
        rows = np.asarray(sw).reshape(S, R) if np.ndim(sw) == 2 else None
        if rows is None and (S != 1 or len(sw) != R):
            raise MCXArgumentError("sweeps_after_exchange must have length size(pt) (n_sets x R)")
        for s in range(S):
            row = sw[s] if rows is not None else sw
            _retune_exchange_sweeps_(row, taus[s], base_sweeps, min_sweeps, max_sweeps)
    else:
        if S != 1:
            raise MCXArgumentError("host samples describe one ladder; use a one-set object or the "
                                   "device series")
        if len(sw) != R:
            raise MCXArgumentError("sweeps_after_exchange must have length size(pt)")
        from .measurements import integrated_autocorrelation_times
        grouped = _group_samples(local_samples, R)
        taus = integrated_autocorrelation_times(grouped, min_points=min_points, max_lag=max_lag)
        _retune_exchange_sweeps_(sw, taus, base_sweeps, min_sweeps, max_sweeps)
    out = np.asarray(sw).reshape(S, R)
    sel = np.take_along_axis(out, pt.indices.reshape(S, R) - 1, axis=1)
    return sel if S > 1 else sel[0]
