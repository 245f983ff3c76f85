"""Continuous-batching job server (libmcx_hip.so's mcx_server_*).

No reference counterpart: the reference runs one chain per Julia thread to
completion (ParallelChains(ThreadsBackend(n), ...), src/infrastructure/parallel_chains.jl:99-105).
Here a scheduler thread inside the library keeps the lattices of many jobs
resident in a device pool bounded by an HBM budget, advances them together in
quanta and admits new jobs as soon as finished ones free their space.  A job's
result depends only on its own descriptor, so it is the same alone, in any mix
and as a one-chain LatticeBatch run with the same seed, key and random start.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import (BLUME_CAPEL, GLAUBER, HEATBATH, ISING, METROPOLIS, JobDesc, JobResult,
                   MCXArgumentError, ServerDesc, ServerStats, call)

class SLOReport(C.Structure):
    _fields_ = [("slo_ms", C.c_double), ("slo_attempts_per_ns", C.c_double),
                ("jobs", C.c_int64), ("met", C.c_int64), ("missed", C.c_int64),
                ("jobs_by_model", C.c_int64 * 2), ("met_by_model", C.c_int64 * 2),
                ("attainment", C.c_double), ("worst_lateness_ms", C.c_double),
                ("mean_deadline_frac", C.c_double), ("p50_ms", C.c_double),
                ("p90_ms", C.c_double), ("p99_ms", C.c_double), ("p999_ms", C.c_double),
                ("max_ms", C.c_double), ("mean_ms", C.c_double)]


_P = C.c_void_p
for _n, _a in {
    "mcx_server_create": [_P, C.POINTER(ServerDesc), C.POINTER(_P)],
    "mcx_server_destroy": [_P],
    "mcx_server_submit": [_P, C.POINTER(JobDesc), C.POINTER(C.c_uint64)],
    "mcx_server_poll": [_P, C.POINTER(JobResult), C.c_int32, C.POINTER(C.c_int32)],
    "mcx_server_drain": [_P, C.c_double],
    "mcx_server_stats": [_P, C.POINTER(ServerStats)],
    "mcx_server_set_slo": [_P, C.c_double, C.c_double],
    "mcx_server_slo_report": [_P, C.POINTER(SLOReport)],
}.items():
    _lib.declare(_n, _a)

_MODELS = {"ising": ISING, "blume_capel": BLUME_CAPEL, "blumecapel": BLUME_CAPEL}
_ALGS = {"metropolis": METROPOLIS, "heatbath": HEATBATH, "glauber": GLAUBER}


@dataclass
class Job:
    """One simulation job: nsweeps checkerboard sweeps of a random start."""
    model: str
    L: int
    beta: float
    nsweeps: int
    alg: str = "metropolis"
    dim: int = 2
    D: float = 0.0
    seed: int = 0
    key: int = 0
    tag: int = 0
    arrival_s: float = 0.0  # offset from the start of a workload (not sent to the device)

    def desc(self) -> JobDesc:
        m = _MODELS.get(self.model.lower())
        if m is None:
            raise MCXArgumentError(f"unknown model {self.model!r}")
        a = _ALGS.get(self.alg.lower())
        if a is None:
            raise MCXArgumentError(f"unknown algorithm {self.alg!r} (metropolis, heatbath or glauber)")
        return JobDesc(model=m, dim=self.dim, L=self.L, alg=a,
                       beta=float(self.beta), D=float(self.D), seed=self.seed & (2**64 - 1),
                       key=self.key & 0xFFFFFFFF, nsweeps=self.nsweeps, tag=self.tag)

    @property
    def attempts(self) -> int:
        return self.L ** self.dim * self.nsweeps


def _result_dict(r: JobResult) -> dict:
    return {"job_id": r.job_id, "tag": r.tag, "status": r.status, "L": r.L,
            "nsweeps": r.nsweeps, "accepted": r.accepted, "E_pair": r.E_pair, "M": r.M,
            "Q": r.Q, "moments": np.array(r.moments[:], dtype=np.float64),
            "latency_ms": r.latency_ms, "queue_ms": r.queue_ms, "device_ms": r.device_ms,
            "slo_met": r.slo_met}


class Server:
    """A job server on one device.

    hbm_budget: bytes for resident lattices (L^dim int8 each, 256-aligned);
    max_jobs: resident slots; quantum_ms: target device time of one quantum
    (0 = 2 ms); n_streams: streams the kernel classes run on (0 = 4);
    tile_sites: lattices at least this big run spread over many workgroups
    (0 = 2^15 Ising / 2^14 Blume-Capel sites, negative = never);
    band_sweeps: sweeps per band launch of those lattices (0 = built-in shapes,
    k > 0 = k, negative = per-colour launches without bands).  With tuned=True
    the autotuner's stream count, quantum and band depth replace the built-in
    ones that the caller leaves at 0; results do not depend on them."""

    def __init__(self, hbm_budget: int, max_jobs: int = 1024, quantum_ms: float = 0.0,
                 n_streams: int = 0, tile_sites: int = 0, band_sweeps: int = 0,
                 ctx: Optional[_lib.Context] = None, tuned: bool = True):
        self.ctx = ctx or _lib.default_context()
        self.settings = {}
        if tuned:
            from .autotune import server_settings
            self.settings = server_settings()
            n_streams = n_streams or int(self.settings.get("n_streams", 0))
            quantum_ms = quantum_ms or float(self.settings.get("quantum_ms", 0.0))
            band_sweeps = band_sweeps or int(self.settings.get("band_sweeps", 0))
        d = ServerDesc(hbm_budget=int(hbm_budget), max_jobs=int(max_jobs),
                       n_streams=int(n_streams), quantum_ms=float(quantum_ms),
                       tile_sites=int(tile_sites), band_sweeps=int(band_sweeps))
        self._h = _P()
        call("mcx_server_create", self.ctx.handle, C.byref(d), C.byref(self._h))

    def submit(self, job: Job) -> int:
        jid = C.c_uint64()
        d = job.desc()
        call("mcx_server_submit", self._h, C.byref(d), C.byref(jid))
        return jid.value

    def poll(self, max_results: int = 256) -> List[dict]:
        buf = (JobResult * max_results)()
        n = C.c_int32()
        call("mcx_server_poll", self._h, buf, max_results, C.byref(n))
        return [_result_dict(buf[i]) for i in range(n.value)]

    def drain(self, timeout_ms: float = -1.0) -> None:
        call("mcx_server_drain", self._h, float(timeout_ms))

    def run(self, jobs: List[Job], timeout_ms: float = -1.0) -> List[dict]:
        """Submit every job, wait for all and return their results in the order
        of `jobs` (results of jobs submitted earlier by other callers are
        consumed and dropped)."""
        ids = [self.submit(j) for j in jobs]
        self.drain(timeout_ms)
        got = {}
        while True:
            batch = self.poll(1024)
            if not batch:
                break
            got.update((r["job_id"], r) for r in batch)
        return [got[i] for i in ids]

    def stats(self) -> dict:
        s = ServerStats()
        call("mcx_server_stats", self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def set_slo(self, slo_ms: float, attempts_per_ns: float = 0.0) -> None:
        """Deadline of every job submitted from now on: slo_ms plus its attempts
        at attempts_per_ns (0: slo_ms alone)."""
        call("mcx_server_set_slo", self._h, float(slo_ms), float(attempts_per_ns))

    def slo_report(self) -> dict:
        """Deadlines met / missed (per model too), worst lateness, latency quantiles."""
        r = SLOReport()
        call("mcx_server_slo_report", self._h, C.byref(r))
        out = {f: getattr(r, f) for f, _ in SLOReport._fields_}
        for f in ("jobs_by_model", "met_by_model"):
            out[f] = {"ising": int(out[f][0]), "blume_capel": int(out[f][1])}
        return out

    def close(self):
        if self._h:
            call("mcx_server_destroy", self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_workload(n_jobs: int, rate_per_s: float = 200.0, seed: int = 19,
                     ising_fraction: float = 0.7, L_range=(16, 512), sweep_range=(1e3, 1e5),
                     T_range=(1.5, 3.5), D_range=(0.0, 1.9), alg: str = "metropolis",
                     sweep_scale: float = 1.0) -> List[Job]:
    """The serving stream: Poisson arrivals at rate_per_s; each job 2D Ising
    (ising_fraction) or Blume-Capel with D uniform in D_range; L log-uniform over
    L_range rounded to even; T uniform over T_range (beta = 1/T); sweeps
    log-uniform over sweep_range (times sweep_scale, at least 1).  Drawn with
    numpy's PCG64 from `seed`; job k gets Philox seed `seed` and chain key k."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    jobs = []
    lo, hi = L_range
    s_lo, s_hi = sweep_range
    for k in range(n_jobs):
        t += rng.exponential(1.0 / rate_per_s)
        ising = rng.random() < ising_fraction
        D = 0.0 if ising else float(rng.uniform(*D_range))
        L = lo * (hi / lo) ** rng.random()
        L = int(min(hi, max(lo, 2 * round(L / 2))))
        T = float(rng.uniform(*T_range))
        sw = s_lo * (s_hi / s_lo) ** rng.random()
        sw = max(1, int(round(sw * sweep_scale)))
        jobs.append(Job(model="ising" if ising else "blume_capel", L=L, beta=1.0 / T,
                        nsweeps=sw, alg=alg, dim=2, D=D, seed=seed, key=k, tag=k,
                        arrival_s=t))
    return jobs
