"""SpinSystems mirror: device-resident Ising / Blume-Capel lattices.

A system here is one chain of a LatticeBatch on the GPU.  The reference's
per-attempt `spin_flip!` cannot cross an FFI at ns cost, so the drop-in call is
`sweep_(sys, alg, nsweeps)`: N*nsweeps `spin_flip!` attempts in checkerboard
order (colour 0 in increasing site index, then colour 1).
  Ising(dims; J, periodic, h)       SpinSystems/src/ising.jl:413-424
  IsingLatticeOptim(Lx, Ly)         SpinSystems/src/ising.jl:440-461
  BlumeCapel(dims; J, D, periodic)  SpinSystems/src/blume_capel.jl:500-511
  init!(sys, type; rng)             ising.jl:74-78, blume_capel.jl:105-109
  energy / magnetization            ising.jl:17-18, blume_capel.jl:18-19
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import MCXArgumentError, MCXAssertionError
from .algorithms import (HeatBath, ImportanceSampling, Metropolis, Philox, algorithm_beta,
                         algorithm_code)
from .batch import LatticeBatch


def _square_dims(dims) -> tuple:
    dims = list(dims)
    if len(dims) not in (2, 3) or any(d != dims[0] for d in dims):
        raise MCXArgumentError("the GPU path supports square (2D) and cubic (3D) lattices")
    return len(dims), int(dims[0])


class SpinSystemBatch:
    """n systems of one model on one GPU (the vector of systems a
    ParallelChains(ThreadsBackend(n), algs) loop drives)."""

    def __init__(self, model: str, dims, n: int = 1, J=1, D=0.0, periodic: bool = True, h=0,
                 L: Optional[Sequence[int]] = None, series_capacity: int = 0, ctx=None):
        if J != 1:
            raise MCXArgumentError("the GPU path implements J = 1")
        if not periodic:
            raise MCXArgumentError("the GPU path implements periodic boundaries")
        if np.any(np.asarray(h) != 0):
            raise MCXArgumentError("the GPU path implements h = 0")
        dim, L0 = _square_dims(dims)
        Ls = [L0] * n if L is None else list(L)
        self.model = model
        self.batch = LatticeBatch(model, dim, Ls, seed=0, D=D, series_capacity=series_capacity,
                                  ctx=ctx)
        self.dim = dim
        self._bound = None  # (seed, keys) currently bound
        self._tables = None  # betas currently uploaded
        self._pt = None  # ParallelTempering driving this batch, if any
        self._walkers = None  # GeneralizedWalkers bound to this batch, if any

    def __len__(self):
        return self.batch.n

    def __getitem__(self, k: int) -> "SystemView":
        if not 0 <= k < self.batch.n:
            raise IndexError(k)
        return SystemView(self, k)

    @property
    def D(self):
        return self.batch.D

    # stream / table binding -------------------------------------------------
    def bind_streams(self, rngs: Sequence[Philox]):
        seeds = {r.seed for r in rngs}
        if len(seeds) != 1:
            raise MCXArgumentError("all chains of one batch share the Philox seed; keys differ")
        seed = rngs[0].seed
        keys = np.array([r.key for r in rngs], np.uint32)
        if self._bound is None or self._bound[0] != seed or not np.array_equal(self._bound[1], keys):
            self.batch.set_streams(seed, keys)
            self._bound = (seed, keys)

    def bind(self, algs: Sequence):
        self.bind_streams([a.rng for a in algs])
        if self._pt is not None:
            return  # a replica-exchange ladder owns the acceptance tables
        betas = np.array([algorithm_beta(a) for a in algs], np.float64)
        if self._tables is None or not np.array_equal(self._tables, betas):
            self.batch.set_betas(betas)
            self._tables = betas

    def walkers(self, algs: Sequence):
        """Device multicanonical / Wang-Landau walkers bound to these algorithms."""
        from .muca import GeneralizedWalkers
        w = self._walkers
        if w is None or len(w.algs) != len(algs) or any(x is not y for x, y in zip(w.algs, algs)):
            if w is not None:
                w.close()
            w = self._walkers = GeneralizedWalkers(self, algs)
        return w

    def sweep_(self, algs, nsweeps: int = 1, record: bool = False, plain: bool = False) -> float:
        algs = list(algs) if isinstance(algs, (list, tuple)) else [algs] * self.batch.n
        if len(algs) != self.batch.n:
            raise MCXArgumentError("one algorithm per chain")
        from .muca import is_walker_alg
        if all(is_walker_alg(a) for a in algs):
            if plain:
                raise MCXArgumentError("multicanonical / Wang-Landau walkers have no plain path")
            return self.walkers(algs).run_(nsweeps, record=record)
        codes = {algorithm_code(a) for a in algs}
        if len(codes) != 1:
            raise MCXArgumentError("all chains of one sweep_ call use the same algorithm type")
        self.bind(algs)
        _, acc0 = self.batch.counters()
        ms = self.batch.run(codes.pop(), nsweeps, record=record, plain=plain)
        _, acc1 = self.batch.counters()
        for k, a in enumerate(algs):
            a.steps += int(self.batch.nsites[k]) * int(nsweeps)
            if hasattr(a, "accepted"):
                a.accepted += int(acc1[k] - acc0[k])
        return ms

    def init_(self, type: str = "random", rngs: Optional[Sequence[Philox]] = None):
        t = type.lstrip(":").lower()
        if t == "random":
            if rngs is None:
                raise MCXAssertionError("Random initialization requires rng")
            seeds = {r.seed for r in rngs}
            if len(seeds) != 1:
                raise MCXArgumentError("all chains of one batch share the Philox seed; keys differ")
            keys = np.array([r.key for r in rngs], np.uint32)
            self.batch.set_streams(rngs[0].seed, keys)
            self._bound = (rngs[0].seed, keys)
        elif t not in ("up", "down", "zero"):
            raise MCXArgumentError(f"Unknown initialization type: {type}")
        self.batch.init(t)

    def energies(self) -> np.ndarray:
        return self.batch.energies()

    def magnetizations(self) -> np.ndarray:
        return self.batch.observables()[1].astype(np.int64)


class SystemView:
    """One chain of a SpinSystemBatch, with the reference's per-system API."""

    def __init__(self, owner: SpinSystemBatch, k: int):
        self.owner, self.k = owner, k

    @property
    def spins(self) -> np.ndarray:
        return self.owner.batch.get_spins(self.k)

    @spins.setter
    def spins(self, v):
        self.owner.batch.set_spins(self.k, v)

    @property
    def N(self):
        return int(self.owner.batch.nsites[self.k])

    def energy(self, full: bool = False):
        E, M, Q = self.owner.batch.observables()
        if self.owner.batch.model == _lib.ISING:
            return int(E[self.k])
        return float(E[self.k]) + float(self.owner.batch.D[self.k]) * float(Q[self.k])

    def magnetization(self, full: bool = False):
        return int(self.owner.batch.observables()[1][self.k])


class _SingleSystem(SystemView):
    def __init__(self, model, dims, J=1, D=0.0, periodic=True, h=0, series_capacity=0):
        owner = SpinSystemBatch(model, dims, 1, J=J, D=D, periodic=periodic, h=h,
                                series_capacity=series_capacity)
        super().__init__(owner, 0)
        self.J, self.D = J, D


class Ising(_SingleSystem):
    """Ising(dims; J=1, periodic=true, h=0) on the GPU (spins start all up)."""

    def __init__(self, dims, J=1, periodic=True, h=0, series_capacity=0):
        super().__init__("ising", dims, J=J, periodic=periodic, h=h, series_capacity=series_capacity)


class IsingLatticeOptim(_SingleSystem):
    """IsingLatticeOptim(Lx, Ly): 2D periodic, J = 1, h = 0."""

    def __init__(self, Lx: int, Ly: int, series_capacity=0):
        if Lx != Ly:
            raise MCXArgumentError("the GPU path supports square lattices")
        super().__init__("ising", [Lx, Ly], series_capacity=series_capacity)
        self.Lx, self.Ly = Lx, Ly


class BlumeCapel(_SingleSystem):
    """BlumeCapel(dims; J=1, D=0, periodic=true, h=0) on the GPU."""

    def __init__(self, dims, J=1, D=0.0, periodic=True, h=0, series_capacity=0):
        super().__init__("blume_capel", dims, J=J, D=float(D), periodic=periodic, h=h,
                         series_capacity=series_capacity)


def init_(sys, type="random", rng: Optional[Philox] = None):
    """init!(sys, type; rng) for a system or a SpinSystemBatch (one rng per chain)."""
    if isinstance(sys, SpinSystemBatch):
        return sys.init_(type, rng)
    owner = sys.owner
    if owner.batch.n != 1:
        raise MCXArgumentError("init_ on one chain of a batch: call it on the batch")
    owner.init_(type, None if rng is None else [rng])
    return sys


def sweep_(sys, alg, nsweeps: int = 1, record: bool = False, plain: bool = False) -> float:
    """N*nsweeps spin_flip! attempts per chain, checkerboard order, on the GPU.
    Returns the device time in ms."""
    if isinstance(sys, SpinSystemBatch):
        return sys.sweep_(alg, nsweeps, record=record, plain=plain)
    owner = sys.owner
    if owner.batch.n != 1:
        raise MCXArgumentError("sweep_ on one chain of a batch: call it on the batch")
    return owner.sweep_([alg], nsweeps, record=record, plain=plain)


def spin_flip_(sys, alg):
    raise NotImplementedError(
        "per-attempt spin_flip! does not cross the FFI; use sweep_(sys, alg, nsweeps)")


def energy(sys, full: bool = False):
    if isinstance(sys, SpinSystemBatch):
        return sys.energies()
    return sys.energy(full=full)


def magnetization(sys, full: bool = False):
    if isinstance(sys, SpinSystemBatch):
        return sys.magnetizations()
    return sys.magnetization(full=full)
