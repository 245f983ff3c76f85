"""Autotuner for the launch shapes of the sweep kernels and the job server.

Every launch shape gives the same results bit for bit (the shape only lays the
chains over workgroups, and lanes over plane words); the tuner measures the
candidates on the device, checks that each one reproduces the built-in
shape's spins, series, moments and counters, and keeps the fastest.  The
chosen shapes live in ``tuned_shapes.json`` beside this file, keyed by
(model, dim, lattice sizes, algorithm); ``LatticeBatch`` applies the matching
entry when it is created, and ``Server`` takes the tuned stream count, quantum
and band depth unless the caller sets them.

  python -m montecarlox.autotune --out tuned_shapes.json [--quick]

runs the tuning for the benchmark workloads (a GPU is needed).
"""
from __future__ import annotations

import argparse
import json
import os
import time
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import _lib

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_shapes.json")
_MODEL_NAMES = {_lib.ISING: "ising", _lib.BLUME_CAPEL: "blume_capel"}
_ALG_NAMES = {_lib.METROPOLIS: "metropolis", _lib.HEATBATH: "heatbath", _lib.GLAUBER: "glauber"}
_table_cache = None


def _model_name(model) -> str:
    return model if isinstance(model, str) else _MODEL_NAMES[int(model)]


def key(model, dim: int, Ls, alg: str) -> str:
    """Table key: model/dim/sorted distinct lattice sizes/algorithm."""
    sizes = sorted({int(x) for x in np.atleast_1d(Ls)})
    return f"{_model_name(model)}/{int(dim)}/{','.join(map(str, sizes))}/{alg}"


def load_table(path: Optional[str] = None) -> dict:
    global _table_cache
    if path is None and _table_cache is not None:
        return _table_cache
    p = path or TABLE_PATH
    t = {"batched": {}, "server": {}}
    if os.path.exists(p):
        with open(p) as fh:
            t.update(json.load(fh))
    if path is None:
        _table_cache = t
    return t


def lookup(model, dim: int, Ls) -> Dict[str, dict]:
    """{algorithm: launch shape} of the tuned entries for a batch of these sizes."""
    tab = load_table()["batched"]
    out = {}
    for alg in _ALG_NAMES.values():
        e = tab.get(key(model, dim, Ls, alg))
        if e:
            out[alg] = {k: int(e[k]) for k in ("threads", "chains_per_group", "lanes_per_word")}
    return out


def server_settings() -> dict:
    """Tuned job-server settings (n_streams, quantum_ms, band_sweeps), or {}."""
    s = load_table()["server"]
    return {k: s[k] for k in ("n_streams", "quantum_ms", "band_sweeps") if k in s}


def candidates(model, dim: int, Ls) -> List[dict]:
    """Launch shapes worth measuring: workgroup sizes from one wave up to four
    per plane-word set, a few chains per workgroup, and for Blume-Capel the
    lanes per word."""
    Lmax = int(np.max(Ls))
    nw = (Lmax // 2 + 31) // 32
    words = (Lmax * Lmax if dim == 3 else Lmax) * nw
    if _model_name(model) == "ising" and Lmax < 64 and Lmax % 8 == 0:
        words = Lmax ** dim // 64  # flat planes: whole 32-site words
    base = min(256, max(64, (words + 63) // 64 * 64))
    threads = sorted({t for t in (64, 128, 256, 512, 1024) if base // 2 <= t <= 4 * base})
    out = []
    for t in threads:
        for cpg in (1, 2, 4):
            if _model_name(model) == "ising":
                out.append({"threads": t, "chains_per_group": cpg, "lanes_per_word": 0})
            else:
                for lpw in (1, 2, 4):
                    if words * lpw <= 2 * t:
                        out.append({"threads": t, "chains_per_group": cpg, "lanes_per_word": lpw})
    return out


def _state(b):
    E, M, Q = b.observables()
    st, ac = b.counters()
    spins = np.concatenate([b.get_spins(k) for k in range(0, b.n, max(1, b.n // 64))])
    return [spins, E, M, Q, st, ac, b.moments()]


def _same(x, y) -> bool:
    return all(np.array_equal(a, c) for a, c in zip(x, y))


def tune_batched(model: str, dim: int, Ls: Sequence[int], alg: str, sweeps: int = 200,
                 D: float = 1.0, seed: int = 7, reps: int = 3,
                 shapes: Optional[Iterable[dict]] = None, log=print) -> dict:
    """Measure the candidate shapes on a batch of lattices `Ls` (one chain per
    entry), temperatures uniform over T in [1.5, 3.5]; returns the table entry
    of the fastest shape that reproduces the built-in shape bit for bit."""
    from .batch import LatticeBatch
    Ls = np.asarray(Ls, np.int32)
    n = len(Ls)
    betas = 1.0 / (1.5 + 2.0 * np.arange(n) / max(1, n - 1))

    def make():
        b = LatticeBatch(model, dim, Ls, seed=seed, D=D, series_capacity=sweeps, tuned=False)
        b.set_betas(betas)
        b.init("random")
        return b

    def timed(b):
        b.run(alg, sweeps, record=True)  # the checked run (warms the kernel too)
        st = _state(b)
        ms = []
        for _ in range(reps):
            ms.append(b.run(alg, sweeps))
        return st, float(np.median(ms))

    ref = make()
    ref_state, ref_ms = timed(ref)
    ref.close()
    best = {"threads": 0, "chains_per_group": 0, "lanes_per_word": 0, "ms": ref_ms}
    tried = []
    for s in (shapes if shapes is not None else candidates(model, dim, Ls)):
        b = make()
        b.set_launch_shape(alg, **s)
        st, ms = timed(b)
        b.close()
        same = _same(st, ref_state)
        tried.append(dict(s, ms=round(ms, 4), identical=same))
        if log:
            log(f"  {key(model, dim, Ls, alg)} {s} {ms:.3f} ms {'' if same else 'DIFFERS'}")
        if not same:
            raise _lib.MCXError(f"launch shape {s} changed the results of {key(model, dim, Ls, alg)}")
        if ms < best["ms"]:
            best = dict(s, ms=ms)
    return {"threads": best["threads"], "chains_per_group": best["chains_per_group"],
            "lanes_per_word": best["lanes_per_word"], "ms": round(best["ms"], 4),
            "default_ms": round(ref_ms, 4), "chains": n, "sweeps": sweeps,
            "speedup": round(ref_ms / best["ms"], 4), "tried": tried}


def tune_server(n_jobs: int = 200, budget: int = 1 << 28, grid: Optional[List[dict]] = None,
                log=print) -> dict:
    """Replays the Poisson serving stream (200/s, seed 19) for each candidate
    (n_streams, quantum_ms, band_sweeps) and keeps the lowest p99 latency."""
    from .serving import Job, Server, poisson_workload
    jobs = poisson_workload(n_jobs)
    with Server(hbm_budget=budget, max_jobs=64) as w:  # warm every kernel class
        w.run([Job(m, L, 0.5, 2, "metropolis", D=0.5) for m in ("ising", "blume_capel")
               for L in (16, 64, 256, 512)])
    grid = grid or [{"n_streams": s, "quantum_ms": q, "band_sweeps": 0}
                    for s in (3, 4, 6) for q in (1.5, 2.0, 3.0)]
    best, tried = None, []
    for g in grid:
        srv = Server(hbm_budget=budget, max_jobs=1024, **g)
        t0 = time.perf_counter()
        for j in jobs:
            dt = t0 + j.arrival_s - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            srv.submit(j)
            srv.poll(1024)
        srv.drain()
        st = srv.stats()
        srv.close()
        r = dict(g, p50_ms=round(st["p50_ms"], 2), p99_ms=round(st["p99_ms"], 2))
        tried.append(r)
        if log:
            log(f"  server {r}")
        if best is None or r["p99_ms"] < best["p99_ms"]:
            best = r
    return dict(best, jobs=n_jobs, tried=tried)


# the benchmark workloads: (model, dim, lattice sizes, algorithm, D)
def _workloads(quick: bool):
    pt_L = np.repeat(np.random.Generator(np.random.PCG64(11)).choice(
        np.array([16, 24, 32, 48]), size=256), 64)
    n = 1024 if quick else 4096
    return [("ising", 2, [64] * n, "metropolis", 0.0),
            ("ising", 2, [32], "metropolis", 0.0),
            ("ising", 3, pt_L[: 4096 if quick else None], "metropolis", 0.0),
            ("blume_capel", 2, [64] * n, "metropolis", 1.0)]


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--out", default=TABLE_PATH)
    p.add_argument("--quick", action="store_true", help="fewer chains and no server tuning")
    p.add_argument("--sweeps", type=int, default=200)
    a = p.parse_args(argv)
    table = {"batched": {}, "server": {},
             "note": "written by python -m montecarlox.autotune on the device it describes"}
    for model, dim, Ls, alg, D in _workloads(a.quick):
        k = key(model, dim, Ls, alg)
        print(f"tuning {k} ({len(Ls)} chains)", flush=True)
        e = tune_batched(model, dim, Ls, alg, sweeps=a.sweeps if len(Ls) > 1 else 20 * a.sweeps,
                         D=D)
        table["batched"][k] = e
        print(f"  -> {e['threads']}/{e['chains_per_group']}/{e['lanes_per_word']}: "
              f"{e['ms']} ms (built-in {e['default_ms']} ms)", flush=True)
    if not a.quick:
        print("tuning the job server", flush=True)
        table["server"] = tune_server()
        print(f"  -> {table['server']}", flush=True)
    with open(a.out, "w") as fh:
        json.dump(table, fh, indent=1)
        fh.write("\n")
    print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
