"""Benchmark: spin-flip attempts/ns on one MI355X.

Workload (the batched-chains configuration): 4096 independent 2D Ising chains,
L = 64 periodic, J = 1, h = 0, T_k = 1.5 + 2.0*k/4095 (beta_k = 1/T_k), spins
i.i.d. uniform from the Philox INIT stream with key (7, k), checkerboard
Metropolis, per-sweep int32 E/M series and FP64 moments recorded on the device.
One step = one mcx_batch_run of --sweeps-per-step sweeps over all 4096 chains
(default 1000, so the default 10 timed steps are the full 10^4-sweep run).

Printed on one JSON line: whole-job attempts/ns of the batched path, the
roofline of its kernel (2 algorithmic bytes per attempt: one int8 read + one
int8 write per site per sweep; the measured HBM traffic and VALU figures from
the committed rocprofv3 summaries beside it), the plain one-chain-per-launch
GPU path on a sample, the reference CPU path (random-site Metropolis with
Xoshiro256++, restated in C under oracle/) timed on the host cores on a sample,
and the further lines selected with --lines:
  c1       one 2D Ising chain, L=32, beta_c, seed 42: 10^4 + 10^5 sweeps on the
           plain path and on the batched path, and the reference CPU path on one
           thread over the same run
  pt       256 replica-exchange sets x 64 replicas of 3D Ising, L in {16..48}
  bc       4096 Blume-Capel chains, L=64, D=1
  walkers  1024 Blume-Capel multicanonical / Wang-Landau walkers, fixed weights
  c4       the Wang-Landau iteration of the same 1024 walkers (ln f from 1,
           flatness 0.8, weights shared per lattice size every 10^4 sweeps),
           --c4-updates updates
  serving  200 Poisson-arriving jobs replayed in real time against the server
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "montecarlox.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "spin-flip attempts/ns, batched vs one-chain-per-launch; p99 job latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_ATTEMPT = 2.0
# 256 CUs x 64 lanes x 2.4 GHz: 32-bit VALU lane-operations per second
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
ALL_LINES = ("c1", "pt", "bc", "walkers", "c4", "serving")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--sweeps-per-step", type=int, default=1000)
    p.add_argument("--chains", type=int, default=4096)
    p.add_argument("--L", type=int, default=64)
    p.add_argument("--lines", default=",".join(ALL_LINES),
                   help="comma list of further lines: " + ",".join(ALL_LINES) + " (or none)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    p.add_argument("--no-plain", action="store_true", help="skip the plain-path sample")
    p.add_argument("--no-headline", action="store_true",
                   help="skip the 4096-chain run (profiling other lines)")
    p.add_argument("--cpu-sweeps", type=int, default=1000)
    p.add_argument("--plain-chains", type=int, default=16)
    p.add_argument("--plain-sweeps", type=int, default=20)
    p.add_argument("--no-serving", action="store_true", help="skip the serving stream")
    p.add_argument("--no-kernels", action="store_true",
                   help="skip the per-family lines (pt, bc, walkers, c4, c1)")
    p.add_argument("--pt-rounds", type=int, default=2000,
                   help="exchange rounds of the 256-set replica-exchange line (10 sweeps each)")
    p.add_argument("--walker-sweeps", type=int, default=2000,
                   help="sweeps of the fixed-weight walker lines")
    p.add_argument("--c4-updates", type=int, default=1,
                   help="weight updates (10^4 sweeps each) of the Wang-Landau iteration line")
    p.add_argument("--c4-sweeps", type=int, default=10000, help="sweeps per weight update")
    p.add_argument("--serving-jobs", type=int, default=200,
                   help="jobs of the Poisson serving stream (200/s, seed 19)")
    p.add_argument("--serving-budget", type=int, default=1 << 28,
                   help="HBM budget of the job server in bytes")
    p.add_argument("--copy-probe", type=int, default=0,
                   help="also run a device copy of this many bytes (HBM counter scaling)")
    a = p.parse_args()
    lines = set() if a.lines in ("", "none") else set(a.lines.split(","))
    bad = lines - set(ALL_LINES)
    if bad:
        p.error(f"unknown lines {sorted(bad)}")
    if a.no_kernels:
        lines -= {"c1", "pt", "bc", "walkers", "c4"}
    if a.no_serving or a.serving_jobs <= 0:
        lines.discard("serving")
    a.line_set = lines
    return a


def _profile_value(name, key):
    """A `key: value` figure of a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1].split()[0])
    return None


def host_info():
    """CPU model, logical CPUs of the machine, CPUs this process may use, and the
    threads a CPU baseline uses (the process's CPU share, at most 16: the GPU
    machine allots 16 CPUs per GPU)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = nproc
    return {"cpu_model": model, "nproc": nproc, "cpus_usable": share,
            "threads": max(1, min(16, share))}


def stream_copy_gbs(torch, nbytes=1 << 30, reps=10):
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del x, y
    return 2.0 * nbytes * reps / dt / 1e9


def _roof(attempts, ms, bytes_per_attempt=BYTES_PER_ATTEMPT):
    ach = bytes_per_attempt * attempts / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_attempt": bytes_per_attempt}


def _cpu(oracle, host, sample, threads=None, **kw):
    th = host["threads"] if threads is None else threads
    t0 = time.perf_counter()
    n = oracle.refpath_run(**kw, nthreads=th)
    dt = time.perf_counter() - t0
    return {"value": round(n / (dt * 1e9), 4), "unit": "attempts/ns", "cores": th,
            "kind": "port", "sample": sample, "seconds": round(dt, 2),
            "cpu_model": host["cpu_model"], "nproc": host["nproc"]}


# ---------------------------------------------------------------------------
# lines
# ---------------------------------------------------------------------------

def headline(a, mcx, np, torch, out):
    n, L, S = a.chains, a.L, a.sweeps_per_step
    temps = 1.5 + 2.0 * np.arange(n) / max(n - 1, 1)
    betas = 1.0 / temps
    cap = S * (a.steps + a.warmup)
    b = mcx.LatticeBatch("ising", 2, L, n_chains=n, seed=7, series_capacity=cap)
    b.set_betas(betas)
    b.init("random")
    shape = b.tuned_shapes.get("metropolis", "built-in")
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        b.run("metropolis", S, record=True)
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dev_ms.append(b.run("metropolis", S, record=True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    attempts_per_step = float(n) * L * L * S
    value = attempts_per_step * a.steps / (wall * 1e9)
    kern_ms = float(np.mean(dev_ms))
    achieved = BYTES_PER_ATTEMPT * attempts_per_step / (kern_ms * 1e-3) / 1e9
    # sanity: the recorded series ends at the live observables
    E, M, _ = b.observables()
    Es, Ms, _ = b.series(n - 1, b.series_length() - 1, 1)
    assert Es[0] == E[n - 1] and Ms[0] == M[n - 1]
    _, acc = b.counters()
    b.close()
    prof = "k_ising_sweeps_2d_L64_4096chains"
    traffic = _profile_value(prof + "_traffic.txt", "traffic_bytes_per_launch")
    lane_ops = _profile_value(prof + "_valu.txt", "valu_lane_ops_per_attempt")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_ising_sweeps<2,0>", "kernel_ms": round(kern_ms, 3),
            "launch_shape": shape,
            "bytes_per_attempt": BYTES_PER_ATTEMPT,
            "algorithmic_bytes_per_launch": BYTES_PER_ATTEMPT * attempts_per_step,
            "note": "achieved = the algorithmic 2 B/attempt (an int8 site read and written "
                    "per sweep) over the launch time; the kernel keeps the lattice in LDS "
                    "for all the sweeps of a launch, so its HBM traffic is the small "
                    "'traffic' figure and frac can pass 1; the binding limit is 'valu'"}
    if traffic is not None and (n, L, S) == (4096, 64, 1000):
        roof["traffic_frac_of_algorithmic"] = round(traffic / roof["algorithmic_bytes_per_launch"], 5)
    if lane_ops is not None:
        # the binding limit: VALU lane-operations per attempt (committed counter
        # summary) x the live attempt rate, against the VALU peak
        rate = attempts_per_step / (kern_ms * 1e-3)
        roof["valu"] = {"lane_ops_per_attempt": lane_ops,
                        "achieved_lane_ops_per_s": round(lane_ops * rate, -9),
                        "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                        "frac": round(lane_ops * rate / VALU_PEAK_LANE_OPS, 4)}
    out.update({
        "value": round(value, 3), "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "config": {"workload": f"{n} independent 2D Ising chains, L={L} periodic, J=1, h=0, "
                               f"T uniform over [1.5, 3.5], Philox key (7, k), checkerboard "
                               f"Metropolis, per-sweep int32 E/M series + FP64 moments on device",
                   "chains": n, "L": L, "sweeps_per_step": S, "sweeps_timed": S * a.steps,
                   "acceptance": round(float(acc.sum()) / (attempts_per_step *
                                                          (a.steps + a.warmup)), 4)},
        "roofline": roof})
    bw = stream_copy_gbs(torch)
    roof["stream_copy_GBs"] = round(bw, 1)
    roof["frac_of_stream_copy"] = round(achieved / bw, 4)
    return betas, value


def plain_sample(a, mcx, np, torch, betas, value):
    pn, ps, L = a.plain_chains, a.plain_sweeps, a.L
    p = mcx.LatticeBatch("ising", 2, L, n_chains=pn, seed=7, series_capacity=ps + 2)
    p.set_betas(betas[:pn])
    p.init("random")
    p.run("metropolis", 2, record=True, plain=True)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.run("metropolis", ps, record=True, plain=True)
    torch.cuda.synchronize()
    pw = time.perf_counter() - t0
    pv = float(pn) * L * L * ps / (pw * 1e9)
    p.close()
    return {"value": round(pv, 4), "unit": "attempts/ns",
            "sample": f"{pn} chains x {ps} sweeps of the 4096-chain workload, one kernel per "
                      "colour per chain per sweep",
            "batched_over_plain": round(value / pv, 1),
            "roofline_frac": round(BYTES_PER_ATTEMPT * pv / HBM_PEAK_GBS, 6)}


def cpu_headline(a, oracle, host, betas):
    th = host["threads"]
    t0 = time.perf_counter()
    attempts, _ = oracle.refpath_ising2d(a.L, betas, 7, a.cpu_sweeps, th)
    cw = time.perf_counter() - t0
    return {"value": round(attempts / (cw * 1e9), 4), "unit": "attempts/ns", "cores": th,
            "kind": "port",
            "sample": f"all {len(betas)} chains x {a.cpu_sweeps} sweeps of random-site Metropolis "
                      "(C restatement of spin_flip!(IsingLatticeOptim, Metropolis) with "
                      "Xoshiro256++), one thread per chain group",
            "seconds": round(cw, 2), "cpu_model": host["cpu_model"], "nproc": host["nproc"],
            "cpus_usable": host["cpus_usable"]}


def c1_line(mcx, np, torch, oracle, host, cpu):
    """One chain, L=32, beta_c, seed 42 / key (42, 0), 10^4 + 10^5 sweeps."""
    L, beta, therm, meas = 32, 0.4406868, 10_000, 100_000
    att = float(L * L) * (therm + meas)
    out = {"workload": f"one 2D Ising chain, L={L}, beta={beta}, seed 42 / key (42, 0), "
                       f"{therm} + {meas} checkerboard Metropolis sweeps with per-sweep E/M series"}
    res = {}
    for plain in (True, False):
        b = mcx.LatticeBatch("ising", 2, L, n_chains=1, seed=42, keys=[0],
                             series_capacity=therm + meas)
        b.set_betas([beta])
        b.init("random")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run("metropolis", therm, record=True, plain=plain)
        b.run("metropolis", meas, record=True, plain=plain)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        E, _, _ = b.series(0, therm + meas - 1, 1)
        res[plain] = (wall, int(E[0]))
        b.close()
    assert res[True][1] == res[False][1]
    pv = att / (res[True][0] * 1e9)
    bv = att / (res[False][0] * 1e9)
    out.update({"value": round(pv, 4), "unit": "attempts/ns", "path": "plain (one kernel per "
                "colour per sweep)", "wall_s": round(res[True][0], 3),
                "roofline": _roof(att, res[True][0] * 1e3),
                "batched_one_chain": {"value": round(bv, 4), "wall_s": round(res[False][0], 3)}})
    if cpu:
        t0 = time.perf_counter()
        s = np.ones(L * L, np.int8)
        oracle.refpath_series_ising2d(L, beta, 42, therm + meas, s)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(att / (dt * 1e9), 4), "unit": "attempts/ns",
                               "cores": 1, "kind": "port",
                               "sample": "the whole run: random-site Metropolis on "
                                         "IsingLatticeOptim(32, 32) (C restatement, Xoshiro256++),"
                                         " one thread",
                               "seconds": round(dt, 2), "cpu_model": host["cpu_model"],
                               "nproc": host["nproc"]}
    return out


def pt_line(a, mcx, np, oracle, host, cpu):
    # 256 sets x 64 replicas of 3D Ising, L per set from {16, 24, 32, 48} (numpy
    # PCG64(11)), ladder set_betas(64, 0.20, 0.24, geometric), 10 sweeps per round
    from montecarlox.spinsystems import SpinSystemBatch
    R, S, rounds = 64, 256, a.pt_rounds
    set_L = np.random.Generator(np.random.PCG64(11)).choice(np.array([16, 24, 32, 48]), size=S)
    Ls = np.repeat(set_L, R)
    sys_ = SpinSystemBatch("ising", [16] * 3, n=len(Ls), L=Ls, series_capacity=rounds * 10)
    pt = mcx.ParallelTempering(sys_, mcx.set_betas(R, 0.20, 0.24, "geometric"), n_sets=S, seed=11)
    sys_.init_("random", [a_.rng for a_ in pt.algs])
    pt.run_(2, 10)
    # per-temperature E/M series on the device, the input of the adaptive
    # exchange interval (tau_int per ladder index)
    pt.set_series_order("temperature")
    ms = pt.run_(rounds, 10, record=True)
    att = float((Ls.astype(np.int64) ** 3).sum()) * rounds * 10
    acc = float(np.mean(mcx.acceptance_rates(pt)))
    t0 = time.perf_counter()
    taus = pt.tau_int(min_points=400, max_lag=200)
    tau_ms = (time.perf_counter() - t0) * 1e3
    sw = np.full((S, R), 10, np.int64)
    mcx.optimize_exchange_interval_(pt, None, sw, base_sweeps=10, min_sweeps=1, max_sweeps=400,
                                    min_points=400, max_lag=200)
    # the same 100-round block eagerly (one launch per round and step) and as one
    # hipGraph: the first graph call captures and instantiates it, later ones replay it
    gr = {}
    for name, graph in (("eager", False), ("graph_first", True), ("graph_replay", True),
                        ("graph_replay2", True)):
        t0 = time.perf_counter()
        dev = pt.run_(100, 10, graph=graph)
        gr[name] = {"wall_ms": round((time.perf_counter() - t0) * 1e3, 2),
                    "device_ms": round(dev, 2)}
    line = {"workload": f"{S} replica-exchange sets x {R} replicas, 3D Ising, L per set from "
                        "{16,24,32,48} (PCG64 seed 11: " +
                        ",".join(f"{L}:{int((set_L == L).sum())}" for L in (16, 24, 32, 48)) +
                        "), set_betas(64, 0.20, 0.24, geometric), checkerboard Metropolis, "
                        f"{rounds} rounds of 10 sweeps + one exchange",
            "value": round(att / (ms * 1e6), 3), "unit": "attempts/ns", "ms": round(ms, 2),
            "launch_shape": sys_.batch.tuned_shapes.get("metropolis", "built-in"),
            "exchange_acceptance": round(acc, 4), "roofline": _roof(att, ms),
            "tau_int": {"series": int(taus.size), "samples": rounds * 10, "max_lag": 200,
                        "ms": round(tau_ms, 2), "median": round(float(np.nanmedian(taus)), 3),
                        "max": round(float(np.nanmax(taus)), 3),
                        "retuned_sweeps_min_max": [int(sw.min()), int(sw.max())]},
            "hipgraph_vs_eager_100_rounds": gr}
    pt.close()
    sys_.batch.close()
    if cpu:
        line["cpu_baseline"] = _cpu(oracle, host,
                                    "one replica per set (256 chains, the same L mix) x 40 sweeps "
                                    "of random-site Metropolis on Ising([L,L,L])",
                                    model=0, dim=3, Ls=set_L, betas=0.22, seed=11, nsweeps=40)
    return line


def bc_line(mcx, np, oracle, host, cpu):
    n, L, sw = 4096, 64, 1000
    b = mcx.LatticeBatch("blume_capel", 2, L, n_chains=n, seed=5, D=1.0, series_capacity=2 * sw)
    b.set_betas(1.0 / (1.5 + 2.0 * np.arange(n) / (n - 1)))
    b.init("random")
    b.run("metropolis", sw, record=True)
    ms = b.run("metropolis", sw, record=True)
    att = float(n) * L * L * sw
    line = {"workload": f"{n} Blume-Capel chains, L={L}, D=1, T uniform [1.5, 3.5], checkerboard "
                        f"Metropolis, {sw} sweeps with per-sweep E/M/Q series + FP64 moments",
            "value": round(att / (ms * 1e6), 3), "unit": "attempts/ns", "ms": round(ms, 2),
            "launch_shape": b.tuned_shapes.get("metropolis", "built-in"),
            "roofline": _roof(att, ms)}
    b.close()
    # measured traffic and VALU work per launch / attempt (committed counter summary)
    prof = "k_bc_sweeps_2d_L64_4096chains_counters.txt"
    line["roofline"]["kernel"] = "k_bc_sweeps<2,0>"
    line["roofline"]["traffic"] = _profile_value(prof, "traffic_bytes_per_launch")
    lane_ops = _profile_value(prof, "valu_lane_ops_per_attempt")
    if lane_ops is not None:
        rate = att / (ms * 1e-3)
        line["roofline"]["valu"] = {"lane_ops_per_attempt": lane_ops,
                                    "achieved_lane_ops_per_s": round(lane_ops * rate, -9),
                                    "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                                    "frac": round(lane_ops * rate / VALU_PEAK_LANE_OPS, 4)}
    if cpu:
        line["cpu_baseline"] = _cpu(oracle, host, "256 chains L=64 x 300 sweeps of random-site "
                                    "Metropolis on BlumeCapel([L,L]; D=1)",
                                    model=1, dim=2, Ls=[64] * 256, betas=0.4, D=1.0, seed=5,
                                    nsweeps=300)
    return line


def _bc_walkers(mcx, np, kind, lw_of, logf=1.0):
    """1024 Blume-Capel walkers, 256 each at L = 16/32/48/64, D = 1, bins
    -N:1:3N (4N+1), Philox(3, k)."""
    from montecarlox.ensembles import DiscreteBinning
    from montecarlox.spinsystems import SpinSystemBatch
    Lw = np.repeat(np.array([16, 32, 48, 64]), 256)
    sys_ = SpinSystemBatch("blume_capel", [16, 16], n=len(Lw), L=Lw, D=1.0)
    algs = []
    for k, L in enumerate(Lw):
        N = int(L) * int(L)
        bo = mcx.BinnedObject(DiscreteBinning(-N, 1, 4 * N + 1), 0.0)
        bo.values[:] = lw_of(N)
        if kind == "multicanonical":
            algs.append(mcx.Multicanonical(mcx.Philox(3, k), bo))
        else:
            algs.append(mcx.WangLandau(mcx.Philox(3, k), bo, logf=logf))
    sys_.init_("random", [a_.rng for a_ in algs])
    return Lw, sys_, algs, mcx.GeneralizedWalkers(sys_, algs)


def walker_lines(a, mcx, np, oracle, host, cpu):
    """Fixed-weight walkers (lw = -0.5 E): the per-attempt cost of the walk."""
    out = {}
    for kind in ("multicanonical", "wang_landau"):
        sw = a.walker_sweeps
        Lw, sys_, algs, w = _bc_walkers(mcx, np, kind, lambda N: -0.5 * (np.arange(4 * N + 1) - N),
                                        logf=1e-3)
        w.run_(10)
        ms = w.run_(sw)
        att = float((Lw.astype(np.int64) ** 2).sum()) * sw
        line = {"workload": f"1024 Blume-Capel {kind.replace('_', '-')} walkers (256 each at "
                            "L=16/32/48/64), D=1, 4N+1 bins from -N, weights -0.5 E"
                            + (", ln f = 1e-3" if kind == "wang_landau" else "") +
                            f", one walker per workgroup in site order, {sw} sweeps",
                "value": round(att / (ms * 1e6), 3), "unit": "attempts/ns", "ms": round(ms, 2),
                "ns_per_attempt_L64": round(ms * 1e6 / (64 * 64 * sw), 2),
                "acceptance": round(sum(x.accepted for x in algs) /
                                    max(1, sum(x.steps for x in algs)), 4),
                "roofline": _roof(att, ms)}
        w.close()
        sys_.batch.close()
        if cpu:
            N = 64 * 64
            line["cpu_baseline"] = _cpu(
                oracle, host, f"64 walkers L=64 x 100 sweeps of random-site {kind.replace('_', '-')}"
                " Blume-Capel (D=1, 4N+1 bins)", model=1, dim=2, Ls=[64] * 64, betas=0.0, D=1.0,
                seed=3, nsweeps=100, lw=-0.5 * (np.arange(4 * N + 1) - N), bin_start=-N,
                bin_step=1, logf=1e-3 if kind == "wang_landau" else None)
        out[kind] = line
    return out


def c4_line(a, mcx, np, oracle, host, cpu):
    """The Wang-Landau iteration of the 1024 walkers: ln f from 1, updates of
    --c4-sweeps sweeps, per lattice size the walkers' mean log-weight shared
    after every update and ln f halved when the summed histogram is flat (0.8)
    over the reached bins, down to 1e-6."""
    Lw, sys_, algs, w = _bc_walkers(mcx, np, "wang_landau", lambda N: 0.0, logf=1.0)
    pr = w.iterate_(a.c4_sweeps, a.c4_updates, flatness=0.8, logf_final=1e-6, power=0.5)
    ms = pr["device_ms"]
    att = float((Lw.astype(np.int64) ** 2).sum()) * pr["sweeps"]
    flat = []
    for g in range(4):
        mn, mean, nb = w.flatness(g)
        flat.append({"L": int(16 * (g + 1)), "bins_reached": nb,
                     "root_min_over_mean": round(mn / mean, 3) if mean > 0 else None})
    logf = sorted({float(x.ensemble.logf) for x in algs})
    line = {"workload": "Wang-Landau iteration of 1024 Blume-Capel walkers (256 each at "
                        "L=16/32/48/64, seed 3), D=1, bins -N:1:3N, ln f from 1 halved to 1e-6 "
                        f"on flatness 0.8, weights shared per L every {a.c4_sweeps} sweeps; "
                        f"bounded to {a.c4_updates} update(s) (the full schedule needs >= 20)",
            "value": round(att / (ms * 1e6), 3), "unit": "attempts/ns", "ms": round(ms, 2),
            "updates": pr["updates"], "stages": pr["stages"], "sweeps": pr["sweeps"],
            "logf_now": logf, "flatness_per_L": flat, "roofline": _roof(att, ms)}
    w.close()
    sys_.batch.close()
    if cpu:
        N = 64 * 64
        line["cpu_baseline"] = _cpu(oracle, host, "64 Wang-Landau walkers L=64 x 100 sweeps of "
                                    "random-site Blume-Capel attempts (D=1, 4N+1 bins, ln f = 1)",
                                    model=1, dim=2, Ls=[64] * 64, betas=0.0, D=1.0, seed=3,
                                    nsweeps=100, lw=np.zeros(4 * N + 1), bin_start=-N, bin_step=1,
                                    logf=1.0)
    return line


def serving_stream(mcx, n_jobs, budget):
    """Replays the Poisson serving stream in real time against one job server:
    job k is submitted at its arrival time; latency runs from submission to the
    result being ready.  Returns the server's latency / throughput summary."""
    jobs = mcx.poisson_workload(n_jobs)
    # warm the kernels of every class once (not counted)
    with mcx.Server(hbm_budget=budget, max_jobs=64) as w:
        w.run([mcx.Job(m, L, 0.5, 2, alg, D=0.5) for m in ("ising", "blume_capel")
               for L in (16, 64, 256, 512) for alg in ("metropolis",)])
    srv = mcx.Server(hbm_budget=budget, max_jobs=1024)
    settings = dict(srv.settings) or "built-in"
    # latency SLO: every job within 50 ms plus its attempts at 20 attempts/ns
    slo_ms, slo_rate = 50.0, 20.0
    srv.set_slo(slo_ms, slo_rate)
    t0 = time.perf_counter()
    results = []
    for j in jobs:
        dt = t0 + j.arrival_s - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        srv.submit(j)
        results.extend(srv.poll(1024))
    srv.drain()
    results.extend(srv.poll(1024))
    wall = time.perf_counter() - t0
    st = srv.stats()
    slo = srv.slo_report()
    srv.close()
    assert len(results) == n_jobs and all(r["status"] == 0 for r in results)
    work = float(sum(j.attempts for j in jobs))
    return {
        "workload": f"{n_jobs} jobs, Poisson arrivals at 200/s (seed 19): 70% 2D Ising / 30% "
                    "Blume-Capel (D uniform [0, 1.9]), L log-uniform [16, 512] rounded to even, "
                    "T uniform [1.5, 3.5], sweeps log-uniform [1e3, 1e5], checkerboard "
                    "Metropolis from a random start, FP64 moments of every sweep",
        "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
        "max_ms": round(st["max_ms"], 2),
        "mean_queue_ms": round(float(sum(r["queue_ms"] for r in results)) / n_jobs, 3),
        "attempts": work,
        "attempts_per_ns": round(work / (wall * 1e9), 3),
        "busy_attempts_per_ns": round(st["busy_attempts_per_ns"], 3),
        "offered_attempts_per_ns": round(work / (jobs[-1].arrival_s * 1e9), 3),
        "wall_s": round(wall, 3), "quanta": st["quanta"], "hbm_budget": budget,
        "server_settings": settings,
        "slo": {"rule": f"latency <= {slo_ms:g} ms + attempts / ({slo_rate:g} attempts/ns)",
                "jobs": slo["jobs"], "met": slo["met"], "attainment": round(slo["attainment"], 4),
                "met_by_model": slo["met_by_model"], "jobs_by_model": slo["jobs_by_model"],
                "worst_lateness_ms": round(slo["worst_lateness_ms"], 2),
                "mean_latency_over_deadline": round(slo["mean_deadline_frac"], 4),
                "p90_ms": round(slo["p90_ms"], 2), "p999_ms": round(slo["p999_ms"], 2),
                "mean_ms": round(slo["mean_ms"], 2)},
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(chains are independent: N GPUs would each run their own share of chains)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import montecarlox as mcx

    cpu = not a.no_cpu
    oracle = None
    if cpu:
        import oracle
    host = host_info()
    out = {"metric": METRIC, "value": None, "unit": "attempts/ns", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8",
           "data": "synthetic"}
    if not a.no_headline:
        betas, value = headline(a, mcx, np, torch, out)
        if not a.no_plain:
            out["plain"] = plain_sample(a, mcx, np, torch, betas, value)
        if cpu:
            out["cpu_baseline"] = cpu_headline(a, oracle, host, betas)
    kern = {}
    if "c1" in a.line_set:
        out["c1"] = c1_line(mcx, np, torch, oracle, host, cpu)
    if "pt" in a.line_set:
        kern["replica_exchange"] = pt_line(a, mcx, np, oracle, host, cpu)
    if "bc" in a.line_set:
        kern["blume_capel"] = bc_line(mcx, np, oracle, host, cpu)
    if "walkers" in a.line_set:
        kern.update(walker_lines(a, mcx, np, oracle, host, cpu))
    if "c4" in a.line_set:
        kern["wang_landau_iteration"] = c4_line(a, mcx, np, oracle, host, cpu)
    if kern:
        out["kernels"] = kern
    if "serving" in a.line_set:
        out["serving"] = serving_stream(mcx, a.serving_jobs, a.serving_budget)
    if a.copy_probe:
        x = torch.empty(a.copy_probe, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = a.copy_probe
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
