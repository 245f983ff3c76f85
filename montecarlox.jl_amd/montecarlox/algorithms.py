"""Algorithms of MonteCarloX for the spin-lattice path (host-side objects).

An algorithm carries its random stream, its ensemble and its counters; the
GPU library applies it.  Counters are mirrored from the device after each
device call.
  ImportanceSampling / accept rules   src/algorithms/importance_sampling.jl:23-110
  Metropolis, Glauber                 src/algorithms/metropolis.jl:63-127
  HeatBath                            src/algorithms/heat_bath.jl:11-17
"""
from __future__ import annotations

from typing import Optional

from ._lib import MCXArgumentError
from .ensembles import BoltzmannEnsemble, MulticanonicalEnsemble, WangLandauEnsemble


class Philox:
    """Counter-based Philox4x32-10 stream (the device replacement for the
    reference's Xoshiro/MersenneTwister `rng`).  key = (seed mod 2^32, key);
    seed >> 32 enters counter word 3.  The stream position is the sweep counter
    of the chain it drives."""

    def __init__(self, seed: int = 0, key: int = 0):
        if seed < 0 or key < 0 or key >= 2 ** 32 or seed >= 2 ** 64:
            raise MCXArgumentError("Philox seed must be in [0, 2^64) and key in [0, 2^32)")
        self.seed = int(seed)
        self.key = int(key)

    def __eq__(self, o):
        return isinstance(o, Philox) and (self.seed, self.key) == (o.seed, o.key)

    def __repr__(self):
        return f"Philox(seed={self.seed}, key={self.key})"


class AbstractAlgorithm:
    pass


class AbstractImportanceSampling(AbstractAlgorithm):
    steps: int
    accepted: int


class ImportanceSampling(AbstractImportanceSampling):
    """Generic importance sampling with a tabulated ensemble (muca / WL)."""

    def __init__(self, rng: Philox, ensemble):
        self.rng = rng
        self.ensemble = ensemble
        self.steps = 0
        self.accepted = 0


class AbstractMetropolis(AbstractImportanceSampling):
    pass


class Metropolis(AbstractMetropolis):
    """Metropolis(rng; beta) / Metropolis(rng, ensemble) (metropolis.jl:79-96)."""

    def __init__(self, rng: Philox, ensemble=None, beta=None, **kw):
        b2 = kw.pop("β", None)
        if kw:
            raise TypeError(f"unexpected keywords {sorted(kw)}")
        if ensemble is None:
            if beta is None and b2 is None:
                raise MCXArgumentError("Metropolis needs an ensemble or beta")
            ensemble = BoltzmannEnsemble(beta=beta if beta is not None else b2)
        self.rng = rng
        self.ensemble = ensemble
        self.steps = 0
        self.accepted = 0

    @property
    def beta(self):
        return self.ensemble.beta


class Glauber(Metropolis):
    """Glauber(rng; beta) / Glauber(rng, ensemble): logistic acceptance
    p = logistic(logweight(ensemble, dE)) (metropolis.jl:108-127)."""


class AbstractHeatBath(AbstractAlgorithm):
    pass


class HeatBath(AbstractHeatBath):
    """HeatBath(rng; beta) (heat_bath.jl:11-17); counts steps only."""

    def __init__(self, rng: Philox, beta=None, **kw):
        b2 = kw.pop("β", None)
        if kw:
            raise TypeError(f"unexpected keywords {sorted(kw)}")
        if beta is None and b2 is None:
            raise MCXArgumentError("HeatBath needs beta")
        self.rng = rng
        self.beta = beta if beta is not None else b2
        self.steps = 0


def ensemble(alg):
    return alg.ensemble


def acceptance_rate(alg) -> float:
    """accepted / steps, 0.0 before the first step (importance_sampling.jl:95-96;
    for a replica-exchange object the total over edges, replica_exchange.jl:88-91)."""
    steps, acc = alg.steps, alg.accepted
    if hasattr(steps, "sum"):
        steps, acc = steps.sum(), acc.sum()
    return float(acc / steps) if steps > 0 else 0.0


def reset_(alg):
    """reset!(alg): zero the counters; for a multicanonical algorithm also the
    histogram (importance_sampling.jl:106-110, algorithms/multicanonical.jl:27-33)."""
    if hasattr(alg, "reset_"):
        return alg.reset_()
    dev = getattr(getattr(alg, "ensemble", None), "_dev", None)
    if dev is not None:
        dev[0].reset_(dev[1])
        return None
    alg.steps = 0
    if hasattr(alg, "accepted"):
        alg.accepted = 0
    ens = getattr(alg, "ensemble", None)
    if isinstance(ens, MulticanonicalEnsemble):
        ens.histogram.values[...] = 0
    return None


def Multicanonical(rng, bins=None, init: float = 0.0):
    """Multicanonical([rng,] bins; init) (algorithms/multicanonical.jl:9-25)."""
    if bins is None:
        rng, bins = Philox(0), rng
    ens = bins if isinstance(bins, MulticanonicalEnsemble) else MulticanonicalEnsemble(bins, init=init)
    return ImportanceSampling(rng, ens)


def WangLandau(rng, bins=None, init: float = 0.0, logf: float = 1.0):
    """WangLandau([rng,] bins; init, logf) (algorithms/wang_landau.jl:10-17)."""
    if bins is None:
        rng, bins = Philox(0), rng
    ens = bins if isinstance(bins, WangLandauEnsemble) else WangLandauEnsemble(bins, init=init, logf=logf)
    return ImportanceSampling(rng, ens)


def algorithm_code(alg) -> int:
    from . import _lib
    if isinstance(alg, HeatBath):
        return _lib.HEATBATH
    if isinstance(alg, Glauber):
        if not isinstance(alg.ensemble, BoltzmannEnsemble):
            raise MCXArgumentError("GPU Glauber sweeps need a BoltzmannEnsemble")
        return _lib.GLAUBER
    if isinstance(alg, Metropolis):
        if not isinstance(alg.ensemble, BoltzmannEnsemble):
            raise MCXArgumentError("GPU Metropolis sweeps need a BoltzmannEnsemble")
        return _lib.METROPOLIS
    raise MCXArgumentError(f"unsupported algorithm {type(alg).__name__}")


def algorithm_beta(alg) -> float:
    return alg.beta if isinstance(alg, HeatBath) else alg.ensemble.beta
