"""Rolling checkpoint / restore of host objects and device state.

  init_checkpoint(file, state; kw...)   src/infrastructure/checkpointing.jl:62-79
  checkpoint!(ckpt; kw...)              checkpointing.jl:48-56 (atomic tmp + mv)
  restore_checkpoint(path)              checkpointing.jl:93-98
  finalize!(ckpt)                       checkpointing.jl:106-109
  ckpt.sys, ckpt.alg, ...               checkpointing.jl:33-39

The reference serializes the Julia objects with `Serialization`.  Here the
state is a dict of host objects (algorithms, ensembles, Philox streams, plain
values, numpy arrays) and device objects (SpinSystemBatch, LatticeBatch,
ParallelTempering, GeneralizedWalkers, the single-system views); a device
object is written as its constructor arguments plus the library's snapshot of
its device arrays (mcx_*_save_state), and restored by building a fresh object
and loading the snapshot into it.  Because the device streams are
counter-based (the stream position is the chain's sweep counter, part of the
snapshot), n sweeps, a checkpoint, a restore and m sweeps equal n + m sweeps.

File format: an uncompressed numpy .npz archive (written through a file
object, so the name keeps its suffix) holding `meta`, the JSON description of
the state as uint8, and one array entry per numpy array or device snapshot.
Nothing in it is pickled; it is read with allow_pickle=False, and only the
classes registered below are rebuilt.
"""
from __future__ import annotations

import json
import os
import tempfile
from typing import Any, Dict, Optional

import numpy as np

from ._lib import MCXArgumentError
from .algorithms import Glauber, HeatBath, ImportanceSampling, Metropolis, Philox
from .batch import LatticeBatch
from .ensembles import (BinnedObject, BoltzmannEnsemble, ContinuousBinning, CustomEnsemble,
                        DiscreteBinning, MulticanonicalEnsemble, WangLandauEnsemble)

FORMAT = "montecarlox-checkpoint"
FORMAT_VERSION = 1


def _host_classes():
    from .muca import ParallelChains, ThreadsBackend
    return {c.__name__: c for c in (
        Philox, Metropolis, Glauber, HeatBath, ImportanceSampling, BoltzmannEnsemble,
        MulticanonicalEnsemble, WangLandauEnsemble, CustomEnsemble, BinnedObject, DiscreteBinning,
        ContinuousBinning, ParallelChains, ThreadsBackend)}


def _single_classes():
    from .spinsystems import BlumeCapel, Ising, IsingLatticeOptim, SystemView
    return {c.__name__: c for c in (SystemView, Ising, IsingLatticeOptim, BlumeCapel)}


# ---------------------------------------------------------------------------
# encoding
# ---------------------------------------------------------------------------

class _Encoder:
    def __init__(self):
        self.arrays: Dict[str, np.ndarray] = {}
        self.memo: Dict[int, int] = {}
        self.keep = []  # keeps encoded objects alive so their ids stay unique

    def array(self, a: np.ndarray) -> dict:
        if a.dtype.hasobject:
            raise MCXArgumentError("cannot checkpoint an object array")
        k = f"a{len(self.arrays)}"
        self.arrays[k] = np.ascontiguousarray(a)
        return {"__array__": k}

    def _new(self, obj) -> Optional[dict]:
        i = self.memo.get(id(obj))
        if i is not None:
            return {"__ref__": i}
        self.memo[id(obj)] = len(self.memo)
        self.keep.append(obj)
        return None

    def _id(self, obj) -> int:
        """The memo index of obj, registering it if it is new."""
        self._new(obj)
        return self.memo[id(obj)]

    def value(self, v) -> Any:
        if v is None or isinstance(v, (bool, str)):
            return v
        if isinstance(v, (int, float)) and not isinstance(v, np.generic):
            return v
        if isinstance(v, (np.ndarray, np.generic)):
            return self.array(np.asarray(v))
        if isinstance(v, list):
            return {"__list__": [self.value(x) for x in v]}
        if isinstance(v, tuple):
            return {"__tuple__": [self.value(x) for x in v]}
        if isinstance(v, dict):
            if not all(isinstance(k, str) for k in v):
                raise MCXArgumentError("checkpointed dicts need string keys")
            return {"__dict__": {k: self.value(x) for k, x in v.items()}}
        return self.obj(v)

    def obj(self, v) -> dict:
        from .muca import GeneralizedWalkers
        from .pt import ParallelTempering
        from .spinsystems import SpinSystemBatch, SystemView
        ref = self._new(v)
        if ref is not None:
            return ref
        i = self.memo[id(v)]
        name = type(v).__name__
        if isinstance(v, SpinSystemBatch):
            return self.systems(v, i)
        if isinstance(v, LatticeBatch):
            return {"__obj__": "LatticeBatch", "__id__": i, "batch": self.batch(v)}
        if isinstance(v, ParallelTempering):
            return {"__obj__": "ParallelTempering", "__id__": i, "systems": self.value(v.systems)}
        if isinstance(v, GeneralizedWalkers):
            return {"__obj__": "GeneralizedWalkers", "__id__": i, "systems": self.value(v.systems)}
        if isinstance(v, SystemView) and name in _single_classes():
            fields = {k: self.value(x) for k, x in vars(v).items()
                      if not k.startswith("_") and k not in ("owner", "k")}
            return {"__obj__": name, "__id__": i, "owner": self.value(v.owner), "k": v.k,
                    "fields": fields}
        if name in _host_classes() and type(v) is _host_classes()[name]:
            fields = {k: self.value(x) for k, x in vars(v).items() if not k.startswith("_")}
            return {"__obj__": name, "__id__": i, "fields": fields}
        raise MCXArgumentError(f"cannot checkpoint an object of type {type(v).__name__}")

    def batch(self, b: LatticeBatch) -> dict:
        rec = {"model": b.model, "dim": b.dim, "L": self.array(b.L), "seed": b.seed,
               "keys": None if b.keys is None else self.array(b.keys), "D": self.array(b.D),
               "capacity": b.capacity, "device": b.ctx.device,
               "betas": self.array(b.betas) if hasattr(b, "betas") else None,
               "state": self.array(b.save_state(series=True))}
        # a later reference to this batch object resolves to the same batch
        rec["batch_id"] = self._id(b)
        return rec

    def systems(self, s, i: int) -> dict:
        rec = {"__obj__": "SpinSystemBatch", "__id__": i, "model": s.model, "dim": s.dim,
               "bound": None if s._bound is None else
               [int(s._bound[0]), self.array(np.asarray(s._bound[1]))],
               "tables": None if s._tables is None else self.array(np.asarray(s._tables)),
               "pt": None, "walkers": None}
        rec["batch"] = self.batch(s.batch)
        if s._pt is not None:
            p = s._pt
            rec["pt"] = {"ladder": self.array(p.ladder), "n_sets": p.n_sets,
                         "algs": self.value(p.algs), "state": self.array(p.save_state()),
                         "order": p.series_order, "pt_id": self._id(p)}
        if s._walkers is not None:
            w = s._walkers
            rec["walkers"] = {"algs": self.value(w.algs), "state": self.array(w.save_state()),
                              "walkers_id": self._id(w)}
        return rec


# ---------------------------------------------------------------------------
# decoding
# ---------------------------------------------------------------------------

class _Decoder:
    def __init__(self, arrays):
        self.arrays = arrays
        self.memo: Dict[int, Any] = {}
        self.batch_loads, self.pt_loads, self.gen_loads = [], [], []

    def array(self, d) -> np.ndarray:
        return np.array(self.arrays[d["__array__"]])

    def value(self, v) -> Any:
        if not isinstance(v, dict):
            return v
        if "__array__" in v:
            a = self.array(v)
            return a[()] if a.ndim == 0 else a
        if "__list__" in v:
            return [self.value(x) for x in v["__list__"]]
        if "__tuple__" in v:
            return tuple(self.value(x) for x in v["__tuple__"])
        if "__dict__" in v:
            return {k: self.value(x) for k, x in v["__dict__"].items()}
        if "__ref__" in v:
            return self.memo[v["__ref__"]]
        if "__obj__" in v:
            return self.obj(v)
        raise MCXArgumentError("corrupt checkpoint: unknown record")

    def obj(self, v) -> Any:
        name, i = v["__obj__"], v["__id__"]
        if name == "SpinSystemBatch":
            return self.systems(v)
        if name == "LatticeBatch":
            out = self.memo[i] = self.batch(v["batch"])
            return out
        if name == "ParallelTempering":
            s = self.value(v["systems"])
            out = self.memo[i] = s._pt
            return out
        if name == "GeneralizedWalkers":
            s = self.value(v["systems"])
            out = self.memo[i] = s._walkers
            return out
        single = _single_classes()
        if name in single:
            o = single[name].__new__(single[name])
            self.memo[i] = o
            o.owner = self.value(v["owner"])
            o.k = int(v["k"])
            for k, x in v["fields"].items():
                setattr(o, k, self.value(x))
            return o
        host = _host_classes()
        if name in host:
            o = host[name].__new__(host[name])
            self.memo[i] = o
            for k, x in v["fields"].items():
                setattr(o, k, self.value(x))
            return o
        raise MCXArgumentError(f"corrupt checkpoint: unknown class {name!r}")

    def batch(self, r) -> LatticeBatch:
        from ._lib import Context, default_context
        if r["batch_id"] in self.memo:
            return self.memo[r["batch_id"]]
        ctx = default_context() if r["device"] == 0 else Context(r["device"])
        keys = None if r["keys"] is None else self.array(r["keys"])
        b = LatticeBatch(r["model"], r["dim"], self.array(r["L"]), seed=r["seed"], keys=keys,
                         D=self.array(r["D"]), series_capacity=r["capacity"], ctx=ctx)
        if r["betas"] is not None:
            b.betas = self.array(r["betas"])
        self.memo[r["batch_id"]] = b
        self.batch_loads.append((b, self.array(r["state"])))
        return b

    def systems(self, r):
        from .muca import GeneralizedWalkers
        from .pt import ParallelTempering
        from .spinsystems import SpinSystemBatch
        s = SpinSystemBatch.__new__(SpinSystemBatch)
        self.memo[r["__id__"]] = s
        s.model, s.dim = r["model"], r["dim"]
        s._bound = s._tables = s._pt = s._walkers = None
        s.batch = self.batch(r["batch"])
        if r["pt"] is not None:
            p = r["pt"]
            algs = self.value(p["algs"])
            pt = ParallelTempering(s, self.array(p["ladder"]), n_sets=int(p["n_sets"]), algs=algs)
            pt.series_order = p["order"]  # the batch snapshot carries the device side
            self.memo[p["pt_id"]] = pt
            self.pt_loads.append((pt, self.array(p["state"])))
        if r["walkers"] is not None:
            w = r["walkers"]
            gw = s._walkers = GeneralizedWalkers(s, self.value(w["algs"]))
            self.memo[w["walkers_id"]] = gw
            self.gen_loads.append((gw, self.array(w["state"])))
        s._bound = None if r["bound"] is None else (int(r["bound"][0]), self.array(r["bound"][1]))
        s._tables = None if r["tables"] is None else self.array(r["tables"])
        return s

    def finish(self):
        # constructors above reset chain records and tables; the snapshots go
        # in afterwards, lattices first, then the ladders and walkers on them
        for b, blob in self.batch_loads:
            b.load_state(blob)
        for p, blob in self.pt_loads:
            p.load_state(blob)
        for w, blob in self.gen_loads:
            w.load_state(blob)


# ---------------------------------------------------------------------------
# session API
# ---------------------------------------------------------------------------

def _write(path: str, state: Dict[str, Any]):
    enc = _Encoder()
    meta = {"format": FORMAT, "version": FORMAT_VERSION,
            "state": {k: enc.value(v) for k, v in state.items()}}
    arrays = dict(enc.arrays)
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        np.savez(f, **arrays)
    os.replace(tmp, path)


def _read(path: str) -> Dict[str, Any]:
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    meta = json.loads(arrays.pop("meta").tobytes().decode())
    if meta.get("format") != FORMAT or meta.get("version") != FORMAT_VERSION:
        raise MCXArgumentError(f"{path} is not a checkpoint of this library")
    dec = _Decoder(arrays)
    state = {k: dec.value(v) for k, v in meta["state"].items()}
    dec.finish()
    return state


class CheckpointSession:
    """A checkpoint file and the named objects written on every checkpoint_
    call (checkpointing.jl:33-39); fields read as attributes (ckpt.sys)."""

    def __init__(self, file: str, state: Optional[Dict[str, Any]] = None):
        object.__setattr__(self, "file", str(file))
        object.__setattr__(self, "_state", dict(state or {}))

    def __getattr__(self, name):
        try:
            return self.__dict__["_state"][name]
        except KeyError:
            raise AttributeError(name) from None

    def __setattr__(self, name, value):
        raise AttributeError("CheckpointSession fields are set through checkpoint_(ckpt, **kw)")

    def keys(self):
        return self._state.keys()


def checkpoint_(ckpt: CheckpointSession, **kw) -> str:
    """checkpoint!(ckpt; kw...): write the stored objects, keyword arguments
    overriding same-named fields, through a temporary file and a rename.
    As in the reference, the overrides go to the file only; the session keeps
    its stored objects."""
    state = dict(ckpt._state)
    state.update(kw)
    _write(ckpt.file, state)
    return ckpt.file


def init_checkpoint(file=None, state: Optional[Dict[str, Any]] = None, **kw) -> CheckpointSession:
    """init_checkpoint(file, state; kw...): create the parent directory, the
    session and the first checkpoint.  Without a file the checkpoint goes to
    ckpt.mcx in a fresh temporary directory; a dict as the first argument is
    the state."""
    if isinstance(file, dict) and state is None:
        file, state = None, file
    if file is None:
        file = os.path.join(tempfile.mkdtemp(), "ckpt.mcx")
    d = os.path.dirname(os.path.abspath(file))
    os.makedirs(d, exist_ok=True)
    ckpt = CheckpointSession(file, state)
    checkpoint_(ckpt, **kw)
    return ckpt


def restore_checkpoint(path: str) -> CheckpointSession:
    """restore_checkpoint(path): rebuild every object of the checkpoint (device
    objects on the GPU, with their snapshots loaded) as a new session."""
    return CheckpointSession(path, _read(path))


def finalize_(ckpt: CheckpointSession) -> None:
    """finalize!(ckpt): remove the checkpoint file."""
    try:
        os.remove(ckpt.file)
    except FileNotFoundError:
        pass
    return None
