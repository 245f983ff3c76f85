"""LatticeBatch: a set of chains resident on the GPU (mcx_batch handle).

Thin numpy-facing wrapper over the C ABI; all computation happens in the HIP
library.  The reference-API mirror (spinsystems.py, algorithms.py, ...) is
built on top of this class.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Union

import numpy as np

from . import _lib
from ._lib import call

_MODELS = {"ising": _lib.ISING, "blume_capel": _lib.BLUME_CAPEL, "blumecapel": _lib.BLUME_CAPEL}
_ALGS = {"metropolis": _lib.METROPOLIS, "heatbath": _lib.HEATBATH, "heat_bath": _lib.HEATBATH,
         "glauber": _lib.GLAUBER}
_OBS = {"energy": 0, "e": 0, "magnetization": 1, "m": 1, "abs_magnetization": 2, "|m|": 2,
        "q": 3, "sum_sq": 3}
_INITS = {"up": _lib.INIT_UP, "down": _lib.INIT_DOWN, "zero": _lib.INIT_ZERO,
          "random": _lib.INIT_RANDOM}


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def alg_code(alg) -> int:
    if isinstance(alg, str):
        try:
            return _ALGS[alg.lower()]
        except KeyError:
            raise _lib.MCXArgumentError(f"unknown algorithm {alg!r}") from None
    return int(alg)


class LatticeBatch:
    """n_chains independent lattices (Ising or Blume-Capel, 2D or 3D) on one GPU."""

    def __init__(self, model: Union[str, int] = "ising", dim: int = 2,
                 L: Union[int, Sequence[int]] = 16, n_chains: Optional[int] = None,
                 seed: int = 0, keys: Optional[Sequence[int]] = None,
                 D: Union[float, Sequence[float]] = 0.0, series_capacity: int = 0,
                 ctx: Optional[_lib.Context] = None, tuned: bool = True):
        self.ctx = ctx or _lib.default_context()
        self.model = _MODELS[model.lower()] if isinstance(model, str) else int(model)
        self.dim = int(dim)
        if np.ndim(L) == 0:
            if n_chains is None:
                n_chains = 1
            self.L = np.full(n_chains, int(L), np.int32)
        else:
            self.L = np.ascontiguousarray(L, np.int32)
            if n_chains is not None and n_chains != len(self.L):
                raise _lib.MCXArgumentError("len(L) must equal n_chains")
        self.n = len(self.L)
        self.seed = int(seed)
        self.keys = None if keys is None else np.ascontiguousarray(keys, np.uint32)
        if self.keys is not None and len(self.keys) != self.n:
            raise _lib.MCXArgumentError("len(keys) must equal n_chains")
        self.D = np.broadcast_to(np.asarray(D, np.float64), (self.n,)).copy()
        self.capacity = int(series_capacity)
        d = _lib.BatchDesc()
        d.model, d.dim, d.n_chains = self.model, self.dim, self.n
        d.L_uniform = 0
        d.L = _ptr(self.L, C.c_int32)
        d.seed = self.seed
        d.chain_keys = _ptr(self.keys, C.c_uint32) if self.keys is not None else None
        d.D_uniform = 0.0
        d.D = _ptr(self.D, C.c_double)
        d.series_capacity = self.capacity
        self._h = C.c_void_p()
        call("mcx_batch_create", self.ctx.handle, C.byref(d), C.byref(self._h))
        self.nsites = (self.L.astype(np.int64) ** self.dim)
        # launch shapes an autotuner measured for this (model, dim, L) (results
        # are the same under any shape; only the speed changes)
        self.tuned_shapes = {}
        if tuned:
            from . import autotune
            for alg, shape in autotune.lookup(self.model, self.dim, self.L).items():
                self.set_launch_shape(alg, **shape)
                self.tuned_shapes[alg] = dict(shape)

    # -- lifetime ------------------------------------------------------------
    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            call("mcx_batch_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return self.n

    # -- temperatures ------------------------------------------------------------
    def set_betas(self, betas):
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float64), (self.n,)))
        call("mcx_batch_set_betas", self._h, _ptr(b, C.c_double))
        self.betas = b.copy()

    def table_layout(self, alg) -> int:
        n = C.c_int32()
        call("mcx_batch_table_layout", self._h, alg_code(alg), C.byref(n))
        return n.value

    def set_tables(self, alg, p):
        p = np.ascontiguousarray(p, np.float64)
        per = self.table_layout(alg)
        if p.size != per * self.n:
            raise _lib.MCXArgumentError(f"expected {per * self.n} probabilities, got {p.size}")
        call("mcx_batch_set_tables", self._h, alg_code(alg), _ptr(p, C.c_double), per)

    # -- spins -------------------------------------------------------------------
    def init(self, mode: Union[str, int] = "random"):
        m = _INITS[mode.lower()] if isinstance(mode, str) else int(mode)
        call("mcx_batch_init_spins", self._h, m)

    def set_spins(self, chain: int, spins):
        s = np.ascontiguousarray(spins, np.int8).ravel()
        call("mcx_batch_set_spins", self._h, int(chain), _ptr(s, C.c_int8), s.size)

    def get_spins(self, chain: int) -> np.ndarray:
        if not 0 <= chain < self.n:
            raise _lib.MCXBoundsError(f"chain {chain} not in [0, {self.n})")
        s = np.empty(int(self.nsites[chain]), np.int8)
        call("mcx_batch_get_spins", self._h, int(chain), _ptr(s, C.c_int8), s.size)
        return s

    # -- running -----------------------------------------------------------------
    def run(self, alg, nsweeps: int, record: bool = False, plain: bool = False) -> float:
        """Run nsweeps checkerboard sweeps of every chain; returns device ms."""
        flags = (_lib.RUN_RECORD if record else 0) | (_lib.RUN_PLAIN if plain else 0)
        call("mcx_batch_run", self._h, alg_code(alg), int(nsweeps), flags)
        return self.last_run_ms()

    def set_launch_shape(self, alg=None, threads: int = 0, chains_per_group: int = 0,
                         lanes_per_word: int = 0):
        """Launch shape of the batched kernels for `alg` (None: every algorithm):
        workgroup size, chains per workgroup, Blume-Capel lanes per word (0 =
        built-in).  Results are bit-identical under every shape."""
        s = _lib.LaunchShape(int(threads), int(chains_per_group), int(lanes_per_word), 0)
        call("mcx_batch_set_launch_shape", self._h, -1 if alg is None else alg_code(alg),
             C.byref(s))

    def launch_shape(self, alg) -> dict:
        s = _lib.LaunchShape()
        call("mcx_batch_get_launch_shape", self._h, alg_code(alg), C.byref(s))
        return {"threads": s.threads, "chains_per_group": s.chains_per_group,
                "lanes_per_word": s.lanes_per_word}

    def last_run_ms(self) -> float:
        ms = C.c_double()
        call("mcx_batch_last_run_ms", self._h, C.byref(ms))
        return ms.value

    # -- observables ---------------------------------------------------------------
    def observables(self):
        E = np.empty(self.n, np.int32)
        M = np.empty(self.n, np.int32)
        Q = np.empty(self.n, np.int32)
        call("mcx_batch_get_observables", self._h, _ptr(E, C.c_int32), _ptr(M, C.c_int32),
             _ptr(Q, C.c_int32))
        return E, M, Q

    def energies(self) -> np.ndarray:
        """Energy per chain: E_pair (+ D * Q for Blume-Capel), FP64."""
        E, M, Q = self.observables()
        if self.model == _lib.ISING:
            return E.astype(np.float64)
        return E.astype(np.float64) + self.D * Q.astype(np.float64)

    def series_length(self) -> int:
        n = C.c_int64()
        call("mcx_batch_series_length", self._h, C.byref(n))
        return n.value

    def series(self, chain: int, first: int = 0, count: Optional[int] = None):
        if count is None:
            count = self.series_length() - first
        E = np.empty(count, np.int32)
        M = np.empty(count, np.int32)
        Q = np.empty(count, np.int32) if self.model == _lib.BLUME_CAPEL else None
        call("mcx_batch_get_series", self._h, int(chain), int(first), int(count),
             _ptr(E, C.c_int32), _ptr(M, C.c_int32), _ptr(Q, C.c_int32) if Q is not None else None)
        return E, M, Q

    def reset_series(self):
        call("mcx_batch_reset_series", self._h)

    def moments(self) -> np.ndarray:
        out = np.empty((self.n, _lib.N_MOMENTS), np.float64)
        call("mcx_batch_get_moments", self._h, _ptr(out, C.c_double))
        return out

    def counters(self):
        st = np.empty(self.n, np.int64)
        ac = np.empty(self.n, np.int64)
        call("mcx_batch_get_counters", self._h, _ptr(st, C.c_int64), _ptr(ac, C.c_int64))
        return st, ac

    def reset_counters(self):
        call("mcx_batch_reset_counters", self._h)

    def sweeps(self) -> np.ndarray:
        s = np.empty(self.n, np.uint64)
        call("mcx_batch_get_sweeps", self._h, _ptr(s, C.c_uint64))
        return s

    def set_sweeps(self, sweeps):
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sweeps, np.uint64), (self.n,)))
        call("mcx_batch_set_sweeps", self._h, _ptr(s, C.c_uint64))

    def set_streams(self, seed: int, keys=None):
        """Rebind the chains' Philox streams (the alg.rng of the reference)."""
        k = None if keys is None else np.ascontiguousarray(np.broadcast_to(
            np.asarray(keys, np.uint32), (self.n,)))
        call("mcx_batch_set_streams", self._h, int(seed), _ptr(k, C.c_uint32) if k is not None else None)
        self.seed = int(seed)
        self.keys = k

    def tau_int(self, observable: str = "energy", first: int = 0, count: Optional[int] = None,
                min_points: int = 2, max_lag: Optional[int] = None, c: float = 5.0) -> np.ndarray:
        """integrated_autocorrelation_times(traces; min_points, max_lag, c)
        (autocorrelations.jl:84-110) of every slot's recorded series over sweeps
        [first, first + count), computed on the device.  observable: "energy"
        (E_pair + D*Q), "magnetization", "abs_magnetization" or "q"."""
        try:
            obs = _OBS[observable.lower()]
        except KeyError:
            raise _lib.MCXArgumentError(f"unknown observable {observable!r}") from None
        if count is None:
            count = self.series_length() - first
        out = np.empty(self.n, np.float64)
        call("mcx_batch_tau_int", self._h, obs, int(first), int(count), int(min_points),
             -1 if max_lag is None else int(max_lag), float(c), _ptr(out, C.c_double))
        return out

    # -- snapshots (checkpoint / resume) -------------------------------------------
    def save_state(self, series: bool = True) -> np.ndarray:
        """The batch's device state (chain records, counters, moments, tables,
        spins and, with series=True, the recorded series) as a uint8 array."""
        flags = _lib.STATE_SERIES if series else 0
        n = C.c_size_t()
        call("mcx_batch_state_size", self._h, flags, C.byref(n))
        buf = np.empty(n.value, np.uint8)
        call("mcx_batch_save_state", self._h, flags, buf.ctypes.data_as(C.c_void_p), buf.size)
        return buf

    def load_state(self, blob):
        """Restore a snapshot of a batch of the same model, dim and chain sizes."""
        b = np.ascontiguousarray(blob, np.uint8)
        call("mcx_batch_load_state", self._h, b.ctypes.data_as(C.c_void_p), b.size)
