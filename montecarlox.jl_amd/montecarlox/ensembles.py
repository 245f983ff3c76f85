"""Ensembles (log-weight providers) of MonteCarloX, host-side objects.

They hold parameters and tabulated weights; the GPU library evaluates them.
  BoltzmannEnsemble       src/ensembles/boltzmann.jl:7-30
  BinnedObject            src/infrastructure/binned_object.jl:98-249
  MulticanonicalEnsemble  src/ensembles/multicanonical.jl:1-44
  WangLandauEnsemble      src/ensembles/wang_landau.jl:1-23
"""
from __future__ import annotations

import bisect
import math
from typing import Callable, Optional, Sequence, Tuple, Union

import numpy as np

from ._lib import MCXArgumentError, MCXAssertionError, MCXBoundsError


class AbstractEnsemble:
    def __eq__(self, other):
        if type(self) is not type(other):
            return NotImplemented
        return all(_eq(getattr(self, k), getattr(other, k)) for k in vars(self))


def _eq(a, b):
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return np.array_equal(a, b)
    return a == b


class BoltzmannEnsemble(AbstractEnsemble):
    """logweight(E) = -beta * E.  Exactly one of beta / β / T (boltzmann.jl:7-23)."""

    def __init__(self, beta=None, T=None, **kw):
        b2 = kw.pop("β", None)
        if kw:
            raise TypeError(f"unexpected keywords {sorted(kw)}")
        if beta is not None and b2 is not None:
            raise MCXArgumentError("Specify only one of `beta` or `β`")
        b = beta if beta is not None else b2
        if (b is None) == (T is None):
            raise MCXArgumentError("Specify exactly one of `beta`/`β` or `T`")
        self.beta = b if b is not None else 1.0 / T

    def logweight(self, E):
        if np.ndim(E) > 0:
            return -self.beta * float(np.sum(E))
        return -self.beta * E

    def __call__(self, E):
        return self.logweight(E)


# --------------------------------------------------------------------------
# Binned objects
# --------------------------------------------------------------------------

class DiscreteBinning:
    """start, step, num; integer values use div, floats round half to even
    (binned_object.jl:13-24)."""

    def __init__(self, start, step, num):
        self.start, self.step, self.num = start, step, int(num)

    @property
    def integer(self):
        return isinstance(self.start, (int, np.integer)) and isinstance(self.step, (int, np.integer))

    def index(self, x) -> int:
        if self.integer and isinstance(x, (int, np.integer)):
            d = int(x) - int(self.start)
            q = abs(d) // abs(int(self.step))
            return q if (d >= 0) == (self.step > 0) else -q  # truncating div
        return int(round((x - self.start) / self.step))  # Python round: half to even

    def centers(self):
        return [self.start + self.step * k for k in range(self.num)]

    def __eq__(self, o):
        return isinstance(o, DiscreteBinning) and (self.start, self.step, self.num) == (
            o.start, o.step, o.num)


class ContinuousBinning:
    def __init__(self, edges):
        self.edges = [float(e) for e in edges]
        self.cent = [(a + b) * 0.5 for a, b in zip(self.edges[:-1], self.edges[1:])]

    def index(self, x) -> int:
        return bisect.bisect_right(self.edges, x) - 1

    def centers(self):
        return list(self.cent)

    def __eq__(self, o):
        return isinstance(o, ContinuousBinning) and self.edges == o.edges


def _domain_bins(d, interpretation="auto"):
    vals = list(d)
    is_int = all(isinstance(v, (int, np.integer)) for v in vals)
    if interpretation == "auto":
        interpretation = "discrete" if is_int else "continuous"
    if interpretation == "discrete":
        if len(vals) < 2:
            raise MCXArgumentError("Cannot create bins from a single value.")
        steps = [b - a for a, b in zip(vals[:-1], vals[1:])]
        if any(s != steps[0] for s in steps):
            raise MCXArgumentError("Non-equidistant discrete bins not supported without ExplicitBinning.")
        return DiscreteBinning(vals[0], steps[0], len(vals))
    if interpretation == "continuous":
        if len(vals) < 2:
            raise MCXArgumentError("Continuous bin edges must contain at least two values.")
        return ContinuousBinning(vals)
    raise MCXArgumentError(f"Invalid interpretation={interpretation}. Use :auto, :discrete, or :continuous.")


class BinnedObject:
    """1D binned values on a domain (binned_object.jl:98-122); lw(x) / lw[x]
    look a value up by bin, out-of-range raises BoundsError."""

    def __init__(self, domain, init=0.0, interpretation="auto", dtype=np.float64):
        if isinstance(domain, (DiscreteBinning, ContinuousBinning)):
            self.bins = domain
        else:
            if isinstance(domain, (str, bytes)) or not hasattr(domain, "__iter__"):
                raise MCXArgumentError("Invalid domain type for BinnedObject")
            self.bins = _domain_bins(domain, interpretation)
        n = self.bins.num if isinstance(self.bins, DiscreteBinning) else len(self.bins.edges) - 1
        self.values = np.full(n, init, dtype=dtype)

    def _idx(self, x) -> int:
        i = self.bins.index(x)
        if not 0 <= i < len(self.values):
            raise MCXBoundsError(f"value {x} outside the binned domain")
        return i

    def __call__(self, x):
        return self.values[self._idx(x)]

    def __getitem__(self, x):
        return self.values[self._idx(x)]

    def __setitem__(self, x, v):
        self.values[self._idx(x)] = v

    def __len__(self):
        return len(self.values)

    def __eq__(self, o):
        return isinstance(o, BinnedObject) and self.bins == o.bins and np.array_equal(self.values, o.values)

    def size(self):
        return self.values.shape

    def zero(self):
        z = BinnedObject(self.bins)
        z.values = np.zeros_like(self.values)
        return z

    def copy(self):
        c = BinnedObject(self.bins)
        c.values = self.values.copy()
        return c


def get_centers(bo, dim: int = 1):
    return bo.bins.centers()


def get_values(bo):
    return bo.values


def get_edges(bo, dim: int = 1):
    b = bo.bins
    if isinstance(b, ContinuousBinning):
        return list(b.edges)
    return [b.start - b.step / 2 + b.step * k for k in range(b.num + 1)]


def _assert_same_domain(a: BinnedObject, b: BinnedObject):
    if type(a.bins) is not type(b.bins):
        raise MCXAssertionError("Bin types must match in each dimension.")
    if isinstance(a.bins, DiscreteBinning):
        if a.bins.start != b.bins.start:
            raise MCXAssertionError("Discrete bins must have the same start.")
        if a.bins.step != b.bins.step:
            raise MCXAssertionError("Discrete bins must have the same step.")
        if a.bins.num != b.bins.num:
            raise MCXAssertionError("Discrete bins must have the same num.")
    elif a.bins.edges != b.bins.edges:
        raise MCXAssertionError("Continuous bins must have the same edges.")


def set_(bo: BinnedObject, xrange, f: Optional[Callable] = None):
    """set!(bo, xrange, f) / set!(bo, f) (binned_object.jl:216-249)."""
    if f is None:
        f, xrange = xrange, None
    cs = get_centers(bo)
    n = len(cs)
    if xrange is None:
        xl, xr = float(cs[0]), float(cs[-1])
    else:
        vals = (xrange[0], xrange[1]) if isinstance(xrange, tuple) else (list(xrange)[0], list(xrange)[-1])
        xl, xr = float(min(vals)), float(max(vals))
    il = min(max(bisect.bisect_left(cs, xl), 0), n - 1)
    ir = min(max(bisect.bisect_right(cs, xr) - 1, 0), n - 1)
    if il > ir or cs[il] > xr or cs[ir] < xl:
        raise MCXArgumentError("selected range does not overlap any bin centers")
    for i in range(il, ir + 1):
        bo.values[i] = float(f(cs[i]))
    return None


# --------------------------------------------------------------------------
# Generalised ensembles
# --------------------------------------------------------------------------

class MulticanonicalEnsemble(AbstractEnsemble):
    """Tabulated log-weight + visit histogram (multicanonical.jl:1-30)."""

    def __init__(self, bins, init: float = 0.0, histogram: Optional[BinnedObject] = None,
                 record_visits: bool = True):
        lw = bins if isinstance(bins, BinnedObject) else BinnedObject(bins, float(init))
        h = histogram if histogram is not None else lw.zero()
        _assert_same_domain(lw, h)
        self.logweight = lw
        self.histogram = h
        self.record_visits = record_visits

    def __call__(self, x):
        return self.logweight(x)

    def update_(self, mode: str = "simple"):
        """update!(e; mode=:simple): lw -= log(h) where h > 0 (multicanonical.jl:32-44),
        evaluated on the GPU (on the bound walker, or on these arrays)."""
        if mode != "simple":
            raise MCXArgumentError(f"unsupported mode={mode}, currently only :simple")
        dev = getattr(self, "_dev", None)
        if dev is not None:
            dev[0].update_(dev[1])
        else:
            from .muca import muca_update_values
            muca_update_values(self.logweight.values, self.histogram.values)
        return None


class WangLandauEnsemble(AbstractEnsemble):
    """Tabulated log-weight + modification factor logf (wang_landau.jl:1-23)."""

    def __init__(self, bins, init: float = 0.0, logf: float = 1.0):
        self.logweight = bins if isinstance(bins, BinnedObject) else BinnedObject(bins, float(init))
        self.logf = float(logf)

    def __call__(self, x):
        return self.logweight(x)

    def update_(self, power: float = 0.5):
        self.logf *= power
        return None


class CustomEnsemble(AbstractEnsemble):
    """The two-part ensemble of docs/src/examples/spin_systems/muca_BlumeCapel.jl:43-68:
    a Boltzmann weight on the pair term and a multicanonical weight on sum s^2,
    logweight(H) = logweight(pair, H[0]) + logweight(spin2, H[1]) for
    H = (J * sum_pair_interactions, sum_spins2); visits go to spin2's histogram.
    On the GPU, Metropolis(rng, CustomEnsemble(...)) drives Blume-Capel walkers."""

    def __init__(self, pair: BoltzmannEnsemble, spin2: MulticanonicalEnsemble,
                 record_visits: bool = True):
        if not isinstance(pair, BoltzmannEnsemble) or not isinstance(spin2, MulticanonicalEnsemble):
            raise MCXArgumentError("CustomEnsemble(pair::BoltzmannEnsemble, spin2::MulticanonicalEnsemble)")
        self.pair = pair
        self.spin2 = spin2
        self.record_visits = bool(record_visits)

    def logweight(self, H):
        return self.pair.logweight(H[0]) + self.spin2.logweight(H[1])

    def __call__(self, H):
        return self.logweight(H)


def reset_histogram_(obj):
    """reset_histogram!(ens or alg) of muca_BlumeCapel.jl:63-69: zero spin2's
    histogram (on the device when the ensemble drives device walkers)."""
    ens = getattr(obj, "ensemble", obj)
    if isinstance(ens, CustomEnsemble):
        dev = getattr(ens, "_dev", None)
        if dev is not None:
            dev[0].reset_histogram_(dev[1])
        else:
            ens.spin2.histogram.values[...] = 0
        return None
    raise MCXArgumentError("reset_histogram! takes a CustomEnsemble or an algorithm using one")


def logweight(ens, x=None):
    if x is None:
        return ens.logweight if not isinstance(ens, BoltzmannEnsemble) else ens
    if isinstance(ens, BoltzmannEnsemble):
        return ens.logweight(x)
    return ens.logweight(x)


def update_(obj, *args, **kw):
    """update!(ensemble) / update!(replica_exchange, xs) dispatch."""
    return obj.update_(*args, **kw)
