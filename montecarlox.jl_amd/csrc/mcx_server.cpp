// mcx_server.cpp -- continuous-batching simulation server on one GPU.
//
// No counterpart in the reference, which runs one chain per thread to
// completion.  Jobs (model, dim, L, algorithm, beta, D, seed, key, sweeps)
// arrive at any time.  A scheduler thread keeps their lattices resident in a
// device pool bounded by an HBM budget and advances every resident job in
// quanta.  A quantum is one launch per kernel class (model, dim, algorithm,
// workgroup size; the multi-workgroup lattices one class per model, dim and
// algorithm), the classes spread over a few streams so that different kernels
// run side by side; within a launch every job runs the number of sweeps that
// fills its share of the quantum's target time, from the per-sweep time the
// kernel measured for it (device wall clock) in its previous quanta.  Finished
// jobs leave at the quantum boundary; their pool space and slot are freed and
// waiting jobs are admitted before the next quantum.
//
// Invariance: a job's Philox stream is keyed by its own (seed, key) and
// counted by its own sweep index, and the kernels continue the FP64 moment
// sums sweep by sweep from the stored ones, so nothing in a result depends on
// how its sweeps were split into quanta: spins, counters, last-sweep
// observables and moments are bit-identical alone, in any mix and as a
// one-chain mcx_batch run with the same seed, key and random start.
#include "mcx_api_impl.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

using namespace mcx;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

constexpr int64_t kAlign = 256;
constexpr int kMaxStreams = 8;
constexpr int kTileMaxSweeps = 512; // sweeps of a tiled job per quantum (per-sweep sum buffer)
constexpr int kMaxTileClasses = 16; // tiled classes (model, dim, algorithm) of one quantum
// Lanes per plane word of a tiled Blume-Capel launch: up to 8 (one Philox
// block each) while the launch stays near 16k lanes; more lanes only repeat
// the neighbour gathers once the chip is covered.
int tile_lanes(int max_words)
{
    int l = 4;
    while (l > 1 && (int64_t)max_words * l > 16384)
        l /= 2;
    return l;
}

// Band shape of a tiled class: B slabs per band, T sweeps per launch (a halo
// of 2T slabs on either side), lpw lanes per plane word and the workgroup
// size; T = 0: the band and its halo do not fit the LDS, and the class runs
// the per-colour kernels.
struct Band {
    int B, T, lpw, threads;
};

size_t band_lds(int model, int dim, int L, int B, int T)
{
    const int nw = (L / 2 + 31) / 32;
    const size_t per_slab = (size_t)(dim == 3 ? L : 1) * nw * 4 * (model == MCX_ISING ? 2 : 4);
    return per_slab * (size_t)(B + 4 * T);
}

Band band_shape(int model, int dim, int L, size_t lds_max, int band_sweeps)
{
    // static LDS of the band kernels (tables, reductions) and some slack
    const size_t budget = lds_max > 4096 ? lds_max - 4096 : 0;
    const int rw = (dim == 3 ? L : 1) * ((L / 2 + 31) / 32); // plane words per slab and colour
    // Ising: 32 rows (8 planes in 3D) and 8 (4) sweeps per launch, one lane per
    // word; Blume-Capel, whose word update costs about twice as much: 16 rows
    // (8 planes) and 6 (2) sweeps per launch, 2 lanes per word (two prefix
    // blocks each).  Chosen on the Poisson serving stream against other band
    // heights, depths and lane counts and against the per-colour path.
    Band b = model == MCX_ISING ? Band{dim == 3 ? 8 : 32, dim == 3 ? 4 : 8, 1, 0}
                                : Band{dim == 3 ? 8 : 16, dim == 3 ? 2 : 6, 2, 0};
    if (band_sweeps < 0) {
        b.T = 0; // asked for the per-colour kernels
        return b;
    }
    if (band_sweeps > 0)
        b.T = std::min(band_sweeps, kTileMaxSweeps);
    b.B = std::min(b.B, L);
    b.T = std::max(1, std::min(b.T, L / 4)); // a halo of 2T slabs: keep it within the lattice
    while (band_lds(model, dim, L, b.B, b.T) > budget && b.T > 1)
        --b.T;
    while (band_lds(model, dim, L, b.B, b.T) > budget && b.B > 1)
        b.B = std::max(1, b.B / 2);
    if (band_lds(model, dim, L, b.B, b.T) > budget) {
        b.T = 0;
        return b;
    }
    // one lane (group) per word of the first half-sweep, in whole waves
    const int64_t lanes = (int64_t)(b.B + 4 * b.T - 2) * rw * b.lpw;
    b.threads = (int)std::min<int64_t>(kBandThreads, (lanes + 63) / 64 * 64);
    return b;
}

int64_t align_up(int64_t n)
{
    return (n + kAlign - 1) / kAlign * kAlign;
}

struct Job {
    uint64_t id = 0;
    mcx_job_desc d{};
    int64_t remaining = 0; // sweeps still to run
    int64_t nsites = 0;
    int block = 0;         // threads per workgroup (LDS-resident path)
    int lpw = 1;           // lanes per plane word (Blume-Capel)
    bool tiled = false;    // multi-workgroup path (bit planes in the pool)
    int words = 0;         // plane words per colour (tiled path)
    double ticks_per_sweep = 0; // latest device measurement (0 = none yet)
    int64_t off = -1;      // byte offset of the lattice in the pool
    int64_t plane_off = -1; // byte offset of its bit planes (tiled path)
    int64_t obs_off = -1;   // byte offset of its per-sweep sums (tiled path)
    int64_t bytes = 0;
    int slot = -1;         // index into the per-job device arrays
    double t_submit = 0, t_admit = 0, device_ms = 0;
    double deadline_ms = -1; // latency target from submission (< 0: none)
};

// First-fit allocator over [0, size), with coalescing on release.
class PoolAlloc {
public:
    explicit PoolAlloc(int64_t size) { free_[0] = size; }
    int64_t alloc(int64_t n)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                const int64_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > n)
                    free_[off + n] = len - n;
                return off;
            }
        return -1;
    }
    void release(int64_t off, int64_t n)
    {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<int64_t, int64_t> free_; // offset -> length
};

struct ClassKey {
    int model, dim, alg, block, lpw;
    bool operator<(const ClassKey& o) const
    {
        return std::tie(model, dim, alg, block, lpw) <
               std::tie(o.model, o.dim, o.alg, o.block, o.lpw);
    }
    bool operator==(const ClassKey& o) const { return !(*this < o) && !(o < *this); }
};

using ShapeKey = std::tuple<int, int, int, int>; // model, dim, alg, L

// Launch class of an LDS-resident job: its kernel (model, dim, algorithm) and
// one of two workgroup sizes, 256 for lattices whose own shape needs at most
// 256 threads and 1024 above; Blume-Capel lanes per word are chosen per
// lattice in the kernel (lpw -1).  Few classes keep the quantum's launches on
// as many streams as there are classes, each with the whole quantum.
ClassKey class_of(const Job& j)
{
    return ClassKey{j.d.model, j.d.dim, j.d.alg, j.block <= 256 ? 256 : 1024,
                    j.d.model == MCX_ISING ? 1 : -1};
}

} // namespace

struct mcx_server {
    mcx_ctx* ctx = nullptr;
    int n_streams = 0;
    hipStream_t streams[kMaxStreams] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_in = nullptr, ev_done[kMaxStreams] = {};
    int64_t budget = 0;
    int max_slots = 0;
    double quantum_ms = 0;
    double ticks_per_ms = 0; // device wall clock
    int64_t tile_sites[2] = {0, 0}; // per model: lattices this big take the tiled path
    int band_sweeps = 0;            // mcx_server_desc::band_sweeps
    // latest measurement of one tiled sweep per tiled class (model, dim, alg)
    std::map<std::tuple<int, int, int>, double> tile_ms_per_sweep;
    // device arrays, one entry per slot
    int8_t* d_pool = nullptr;
    ChainRec* d_chains = nullptr;
    IsingTable* d_itab = nullptr;
    BCTable* d_btab = nullptr;
    double* d_mom = nullptr;             // [6 * slots]
    unsigned long long* d_acc = nullptr; // [slots]
    int32_t* d_last = nullptr;           // [3 * slots]: E, M, Q after the last sweep run
    int32_t* d_items = nullptr;          // [slots] work list of one quantum
    int64_t* d_nsw = nullptr;            // [slots] sweeps per work item
    int64_t* d_ticks = nullptr;          // [slots] wall-clock ticks per work item
    int32_t* d_titems = nullptr;         // [slots] tiled work list
    int64_t* d_tnsw = nullptr;           // [slots]
    int64_t* d_tplane = nullptr;         // [slots]
    int64_t* d_tobs = nullptr;           // [slots]
    hipEvent_t ev_t[kMaxTileClasses][2] = {}; // start / end of each tiled class's launches
    // worker-only state
    std::vector<Job> active;
    PoolAlloc* pool = nullptr;
    std::vector<int> free_slots;
    std::map<ShapeKey, double> learned; // last ticks/sweep seen per shape
    // shared state, guarded by mu
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Job> pending;
    std::deque<mcx_job_result> done;
    std::vector<double> latencies;
    // latency SLO: the rule for new jobs, and per finished job with a deadline
    // its model and latency / deadline
    double slo_ms = 0, slo_rate = 0;
    std::vector<std::tuple<int, double, double>> slo_done; // model, latency, deadline
    uint64_t next_id = 1;
    int64_t submitted = 0, completed = 0, n_active = 0, quanta = 0;
    double attempts = 0, busy_ms = 0, t_first = -1, t_last = 0;
    bool stop = false;
    std::string error;
    std::thread worker;
};

namespace {

int64_t job_sites(const mcx_job_desc& d)
{
    return (int64_t)d.L * d.L * (d.dim == 3 ? d.L : 1);
}

int job_words(const mcx_job_desc& d)
{
    const int nw = (d.L / 2 + 31) / 32;
    return (d.dim == 3 ? d.L * d.L : d.L) * nw;
}

bool job_tiled(const mcx_server* s, const mcx_job_desc& d)
{
    const int64_t t = s->tile_sites[d.model == MCX_ISING ? 0 : 1];
    // (flat Ising lattices, L < 64, always fit one workgroup's LDS)
    if (d.model == MCX_ISING && ising_flat(d.L))
        return false;
    return t > 0 && job_sites(d) >= t;
}

constexpr int64_t kTileObsBytes = (int64_t)kTileMaxSweeps * 4 * kTileObsReplicas * 4;

// Pool bytes of a job: the int8 lattice, plus on the tiled path its bit planes
// (2 planes for Ising, 4 for Blume-Capel, of 32-bit words) and its per-sweep
// sums.
int64_t job_bytes(const mcx_server* s, const mcx_job_desc& d)
{
    int64_t b = align_up(job_sites(d));
    if (job_tiled(s, d))
        b += align_up((int64_t)(d.model == MCX_ISING ? 2 : 4) * job_words(d) * 4) + kTileObsBytes;
    return b;
}

// Workgroup shape for one LDS-resident lattice: lanes per 32-site plane word
// (Blume-Capel up to 4, one prefix block each, while the workgroup stays
// within 1024 threads; Ising 1, its bit-serial draws serve all 32 sites at
// once) and threads = words * lanes, at most 1024 (a whole CU's worth of waves
// for a big lattice on its own), rounded up to a power of two so that few
// kernel classes exist.
void server_shape(int model, int dim, int L, int& block, int& lpw)
{
    const int64_t nw = (L / 2 + 31) / 32;
    const int64_t words =
        model == MCX_ISING ? ising_plane_words(dim, L) : (dim == 3 ? (int64_t)L * L : L) * nw;
    lpw = model == MCX_ISING ? 1 : 4;
    while (lpw > 1 && words * lpw > 1024)
        lpw /= 2;
    block = 64;
    while (block < words * lpw && block < 1024)
        block *= 2;
}

mcx_status validate(const mcx_server* s, const mcx_job_desc* d)
{
    if (d->model != MCX_ISING && d->model != MCX_BLUME_CAPEL)
        return fail(MCX_EINVAL, "mcx_server_submit: unknown model %d", d->model);
    if (d->dim != 2 && d->dim != 3)
        return fail(MCX_EINVAL, "mcx_server_submit: dim must be 2 or 3");
    if (d->L < 4 || (d->L & 1))
        return fail(MCX_EINVAL,
                    "mcx_server_submit: L = %d; the two-colour checkerboard needs an even L >= 4",
                    d->L);
    if (d->alg != MCX_METROPOLIS && d->alg != MCX_HEATBATH && d->alg != MCX_GLAUBER)
        return fail(MCX_EINVAL, "mcx_server_submit: alg must be MCX_METROPOLIS, MCX_HEATBATH or "
                                "MCX_GLAUBER");
    if (!std::isfinite(d->beta) || d->beta < 0 || !std::isfinite(d->D))
        return fail(MCX_EINVAL, "mcx_server_submit: beta must be finite and >= 0, D finite");
    if (d->nsweeps < 1)
        return fail(MCX_EINVAL, "mcx_server_submit: nsweeps must be >= 1");
    if (!job_tiled(s, *d) && lattice_lds_bytes(d->model, d->dim, d->L) > s->ctx->lds_max)
        return fail(MCX_EINVAL,
                    "mcx_server_submit: L = %d does not fit the device LDS and the server's "
                    "tile_sites keeps it off the multi-workgroup path",
                    d->L);
    if ((int64_t)job_words(*d) * 4 > INT32_MAX)
        return fail(MCX_EINVAL, "mcx_server_submit: L = %d is too large", d->L);
    if (job_bytes(s, *d) > s->budget)
        return fail(MCX_EINVAL,
                    "mcx_server_submit: a %lld-byte job exceeds the %lld-byte HBM budget",
                    (long long)job_bytes(s, *d), (long long)s->budget);
    return MCX_OK;
}

// Moves the waiting jobs that fit (arrival order, skipping those that do not
// fit yet) into `fresh`, with pool space and a slot.  Called with mu held.
void take_admissions(mcx_server* s, std::vector<Job>& fresh)
{
    for (auto it = s->pending.begin(); it != s->pending.end() && !s->free_slots.empty();) {
        const int64_t off = s->pool->alloc(it->bytes);
        if (off < 0) {
            ++it;
            continue;
        }
        Job j = *it;
        j.off = off;
        j.plane_off = j.tiled ? off + align_up(j.nsites) : -1;
        j.obs_off = j.tiled ? off + j.bytes - kTileObsBytes : -1;
        j.slot = s->free_slots.back();
        s->free_slots.pop_back();
        j.t_admit = now_ms();
        fresh.push_back(j);
        it = s->pending.erase(it);
    }
}

// Chain record, acceptance table, counters and random start of new jobs.
hipError_t setup_jobs(mcx_server* s, std::vector<Job>& fresh)
{
    hipStream_t st = s->streams[0];
    std::vector<ChainRec> rec(fresh.size());
    std::vector<IsingTable> it(fresh.size());
    std::vector<BCTable> bt(fresh.size());
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < fresh.size() && e == hipSuccess; ++i) {
        Job& j = fresh[i];
        auto lr = s->learned.find(ShapeKey{j.d.model, j.d.dim, j.d.alg, j.d.L});
        j.ticks_per_sweep = lr != s->learned.end() ? lr->second : 0.0;
        ChainRec& r = rec[i];
        std::memset(&r, 0, sizeof r);
        r.spin_off = j.off;
        r.sweep = 0;
        r.seed = j.d.seed;
        r.L = j.d.L;
        r.dim = j.d.dim;
        r.table = j.slot;
        r.key = j.d.key;
        r.D = j.d.model == MCX_ISING ? 0.0 : j.d.D;
        e = hipMemcpyAsync(s->d_chains + j.slot, &r, sizeof r, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && j.d.model == MCX_ISING) {
            fill_ising_table(it[i], j.d.dim, j.d.beta);
            e = hipMemcpyAsync(s->d_itab + j.slot, &it[i], sizeof it[i], hipMemcpyHostToDevice,
                               st);
        } else if (e == hipSuccess) {
            fill_bc_table(bt[i], j.d.dim, j.d.beta, j.d.D);
            e = hipMemcpyAsync(s->d_btab + j.slot, &bt[i], sizeof bt[i], hipMemcpyHostToDevice,
                               st);
        }
        if (e == hipSuccess)
            e = hipMemsetAsync(s->d_mom + 6 * (size_t)j.slot, 0, sizeof(double) * 6, st);
        if (e == hipSuccess)
            e = hipMemsetAsync(s->d_acc + j.slot, 0, sizeof(unsigned long long), st);
        if (e == hipSuccess)
            e = launch_init_spins(j.d.model, MCX_INIT_RANDOM, s->d_chains + j.slot, 1, j.nsites,
                                  s->d_pool, st);
        if (e == hipSuccess && j.tiled)
            e = hipMemsetAsync(s->d_pool + j.obs_off, 0, kTileObsBytes, st);
    }
    // tiled jobs: int8 lattice -> bit planes, per (model, dim)
    std::vector<int32_t> titems;
    std::vector<int64_t> tplane;
    for (int model = 0; model < 2 && e == hipSuccess; ++model)
        for (int dim = 2; dim <= 3 && e == hipSuccess; ++dim) {
            titems.clear();
            tplane.clear();
            int max_words = 0;
            for (const Job& j : fresh)
                if (j.tiled && j.d.model == model && j.d.dim == dim) {
                    titems.push_back(j.slot);
                    tplane.push_back(j.plane_off);
                    max_words = std::max(max_words, j.words);
                }
            if (titems.empty())
                continue;
            e = hipMemcpyAsync(s->d_titems, titems.data(), sizeof(int32_t) * titems.size(),
                               hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s->d_tplane, tplane.data(), sizeof(int64_t) * tplane.size(),
                                   hipMemcpyHostToDevice, st);
            TileArgs a{};
            a.items = s->d_titems;
            a.item_plane = s->d_tplane;
            a.n_items = (int32_t)titems.size();
            a.chains = s->d_chains;
            a.base = reinterpret_cast<uint8_t*>(s->d_pool);
            if (e == hipSuccess)
                e = model == MCX_ISING ? launch_ising_pack(dim, 0, a, max_words, st)
                                       : launch_bc_pack(dim, 0, a, max_words, st);
            // the staging vectors are refilled by the next (model, dim)
            if (e == hipSuccess)
                e = hipStreamSynchronize(st);
        }
    // the host staging vectors must outlive the copies
    const hipError_t e2 = hipStreamSynchronize(st);
    return e != hipSuccess ? e : e2;
}

// Sweeps a job runs in a quantum of `target` ticks.  Without a measurement
// yet, a conservative model (2 us + 1 ns per site and sweep) sets a short
// first quantum.
int64_t quantum_sweeps(const mcx_server* s, const Job& j, double target)
{
    const double per = j.ticks_per_sweep > 0
                           ? j.ticks_per_sweep
                           : s->ticks_per_ms * (2e-3 + 1e-6 * (double)j.nsites);
    const double n = std::floor(target / per);
    return n >= (double)j.remaining ? j.remaining : std::max<int64_t>(1, (int64_t)n);
}

using TileKey = std::tuple<int, int, int>; // model, dim, alg

// Device time of one tiled class in one quantum: ms for sweeps sweeps.
struct TileTime {
    TileKey key;
    float ms;
    int64_t sweeps;
};

// One quantum over every active job: ran[i] = sweeps run by active[i],
// ticks[i] = the wall-clock ticks its LDS-resident work item took (0 on the
// tiled path), ms = device time of the quantum, ttime = device time of each
// tiled class.
hipError_t run_quantum(mcx_server* s, std::vector<int64_t>& ran, std::vector<int64_t>& ticks,
                       float& ms, std::vector<TileTime>& ttime)
{
    const int n = (int)s->active.size();
    std::map<ClassKey, std::vector<int>> classes;
    std::map<TileKey, std::vector<int>> tclasses;
    for (int i = 0; i < n; ++i) {
        const Job& j = s->active[i];
        if (j.tiled)
            tclasses[TileKey{j.d.model, j.d.dim, j.d.alg}].push_back(i);
        else
            classes[class_of(j)].push_back(i);
    }
    // units of work: each tiled class, then the LDS-resident classes in
    // decreasing size.  The tiled classes hold the largest lattices, whose
    // latency is the tail of the job stream: with streams to spare each gets a
    // stream of its own (and so the whole quantum) and the LDS-resident classes
    // are dealt round-robin over the others; else every unit is dealt round-robin.
    std::vector<std::pair<ClassKey, std::vector<int>>> cl(classes.begin(), classes.end());
    std::stable_sort(cl.begin(), cl.end(),
                     [](const auto& a, const auto& b) { return a.second.size() > b.second.size(); });
    const int ntc = (int)tclasses.size();
    if (ntc > kMaxTileClasses)
        return hipErrorInvalidValue; // more (model, dim, algorithm) classes than exist
    const int n_units = (int)cl.size() + ntc;
    const int ns = std::max(1, std::min(s->n_streams, n_units));
    std::vector<int> ustream(n_units), per_stream(ns, 0);
    const bool own = ntc > 0 && ns > ntc && n_units > ntc;
    for (int u = 0; u < n_units; ++u)
        ustream[u] = !own ? u % ns : u < ntc ? u : ntc + (u - ntc) % (ns - ntc);
    for (int u = 0; u < n_units; ++u)
        ++per_stream[ustream[u]];
    auto unit_stream = [&](int u) { return ustream[u]; };
    const int unit0 = ntc; // unit of LDS class c is unit0 + c; of tiled class t, t

    ran.assign(n, 0);
    // LDS-resident work list: class by class, the longest items first
    std::vector<int32_t> items;
    std::vector<int64_t> nsw;
    std::vector<int> item_job;
    std::vector<std::pair<int, int>> span(cl.size());
    for (size_t c = 0; c < cl.size(); ++c) {
        const double target =
            s->quantum_ms * s->ticks_per_ms / per_stream[unit_stream(unit0 + (int)c)];
        auto& v = cl[c].second;
        for (int i : v)
            ran[i] = quantum_sweeps(s, s->active[i], target);
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
            return ran[a] * s->active[a].nsites > ran[b] * s->active[b].nsites;
        });
        span[c].first = (int)items.size();
        for (int i : v) {
            items.push_back(s->active[i].slot);
            nsw.push_back(ran[i]);
            item_job.push_back(i);
        }
        span[c].second = (int)v.size();
    }
    // tiled work list: every job of a tiled class runs the same number of
    // sweeps (or what it has left), sized from the class's last measured sweep
    std::vector<int32_t> titems;
    std::vector<int64_t> tnsw, tplane, tobs;
    std::vector<std::tuple<TileKey, int, int, int>> tspan; // key, first, count, max words
    std::vector<int> tspan_L;                              // largest L of each class
    std::vector<int64_t> tspan_sw;                         // sweeps of each class
    for (auto& kv : tclasses) {
        const int t = (int)tspan.size();
        const double target_ms = s->quantum_ms / per_stream[unit_stream(t)];
        const auto it = s->tile_ms_per_sweep.find(kv.first);
        const double per = it != s->tile_ms_per_sweep.end() ? it->second : 0.02;
        const int64_t nt = std::max<int64_t>(
            1, std::min<int64_t>(kTileMaxSweeps, (int64_t)std::floor(target_ms / per)));
        const int first = (int)titems.size();
        int max_words = 0, max_L = 0;
        int64_t sw = 0;
        for (int i : kv.second) {
            const Job& j = s->active[i];
            max_L = std::max(max_L, j.d.L);
            ran[i] = std::min(nt, j.remaining);
            sw = std::max(sw, ran[i]);
            titems.push_back(j.slot);
            tnsw.push_back(ran[i]);
            tplane.push_back(j.plane_off);
            tobs.push_back(j.obs_off);
            max_words = std::max(max_words, j.words);
        }
        tspan.emplace_back(kv.first, first, (int)kv.second.size(), max_words);
        tspan_L.push_back(max_L);
        tspan_sw.push_back(sw);
    }

    hipStream_t s0 = s->streams[0];
    const int nl = (int)items.size(), nt_items = (int)titems.size();
    hipError_t e = hipEventRecord(s->ev0, s0);
    if (e == hipSuccess && nl)
        e = hipMemcpyAsync(s->d_items, items.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, s0);
    if (e == hipSuccess && nl)
        e = hipMemcpyAsync(s->d_nsw, nsw.data(), sizeof(int64_t) * nl, hipMemcpyHostToDevice, s0);
    if (e == hipSuccess && nt_items)
        e = hipMemcpyAsync(s->d_titems, titems.data(), sizeof(int32_t) * nt_items,
                           hipMemcpyHostToDevice, s0);
    if (e == hipSuccess && nt_items)
        e = hipMemcpyAsync(s->d_tnsw, tnsw.data(), sizeof(int64_t) * nt_items,
                           hipMemcpyHostToDevice, s0);
    if (e == hipSuccess && nt_items)
        e = hipMemcpyAsync(s->d_tplane, tplane.data(), sizeof(int64_t) * nt_items,
                           hipMemcpyHostToDevice, s0);
    if (e == hipSuccess && nt_items)
        e = hipMemcpyAsync(s->d_tobs, tobs.data(), sizeof(int64_t) * nt_items,
                           hipMemcpyHostToDevice, s0);
    if (e == hipSuccess)
        e = hipEventRecord(s->ev_in, s0);
    for (int k = 1; k < ns && e == hipSuccess; ++k)
        e = hipStreamWaitEvent(s->streams[k], s->ev_in, 0);
    // tiled classes, each on the stream of its unit: band kernels (T sweeps per
    // launch) where the band and its halo fit the LDS, else the per-colour
    // kernels (one launch per colour and sweep); then the fold
    for (size_t t = 0; t < tspan.size() && e == hipSuccess; ++t) {
        const auto& ts = tspan[t];
        const TileKey& k = std::get<0>(ts);
        const int model = std::get<0>(k), dim = std::get<1>(k), alg = std::get<2>(k);
        const int64_t nsw_t = tspan_sw[t];
        hipStream_t st = s->streams[unit_stream((int)t)];
        e = hipEventRecord(s->ev_t[t][0], st);
        TileArgs a{};
        a.items = s->d_titems + std::get<1>(ts);
        a.item_nsw = s->d_tnsw + std::get<1>(ts);
        a.item_plane = s->d_tplane + std::get<1>(ts);
        a.n_items = std::get<2>(ts);
        a.chains = s->d_chains;
        a.itab = s->d_itab;
        a.btab = s->d_btab;
        a.base = reinterpret_cast<uint8_t*>(s->d_pool);
        a.item_obs = s->d_tobs + std::get<1>(ts);
        a.max_s = kTileMaxSweeps;
        a.accepted = s->d_acc;
        const Band b = band_shape(model, dim, tspan_L[t], s->ctx->lds_max, s->band_sweeps);
        if (b.T != 0) {
            a.band = b.B;
            a.lpw = b.lpw;
            const int max_bands = (tspan_L[t] + b.B - 1) / b.B;
            for (int64_t sw = 0; sw < nsw_t && e == hipSuccess; sw += b.T) {
                a.s = sw;
                a.T = (int32_t)std::min<int64_t>(b.T, nsw_t - sw);
                const size_t lds = band_lds(model, dim, tspan_L[t], b.B, a.T);
                e = model == MCX_ISING
                        ? launch_ising_band(dim, alg, a, max_bands, b.threads, lds, st)
                        : launch_bc_band(dim, alg, a, max_bands, b.threads, lds, st);
            }
        } else {
            a.lpw = tile_lanes(std::get<3>(ts));
            for (int64_t sw = 0; sw < nsw_t && e == hipSuccess; ++sw) {
                a.s = sw;
                for (int c = 0; c < 2 && e == hipSuccess; ++c) {
                    a.colour = c;
                    e = model == MCX_ISING
                            ? launch_ising_tiles(dim, alg, a, std::get<3>(ts), st)
                            : launch_bc_tiles(dim, alg, a, std::get<3>(ts), st);
                }
            }
        }
        if (e == hipSuccess)
            e = launch_tiles_fold(model, a, s->d_mom, s->d_last, s->d_last + s->max_slots,
                                  s->d_last + 2 * s->max_slots, st);
        if (e == hipSuccess)
            e = hipEventRecord(s->ev_t[t][1], st);
    }
    for (size_t c = 0; c < cl.size() && e == hipSuccess; ++c) {
        const ClassKey& k = cl[c].first;
        size_t lds = 0;
        int n_flat = 0, n_64 = 0;
        int64_t words = 0; // Ising: largest plane words per colour
        for (int i : cl[c].second) {
            lds = std::max(lds, lattice_lds_bytes(k.model, k.dim, s->active[i].d.L));
            if (k.model == MCX_ISING)
                words = std::max(words, ising_plane_words(k.dim, s->active[i].d.L));
            n_flat += k.model == MCX_ISING && ising_flat(s->active[i].d.L) ? 1 : 0;
            n_64 += s->active[i].d.L == 64 ? 1 : 0;
        }
        // Ising plane layout of the launch: row, flat, L = 64 only, or per item
        // when flat and row lattices mix
        const int nc = (int)cl[c].second.size();
        const int layout = n_flat == 0 ? (n_64 == nc ? 3 : 0) : (n_flat == nc ? 1 : 2);
        SweepArgs a{};
        a.items = s->d_items + span[c].first;
        a.n_items = span[c].second;
        a.record = 1;
        a.chains = s->d_chains;
        a.itab = s->d_itab;
        a.btab = s->d_btab;
        a.spins = s->d_pool;
        a.moments = s->d_mom;
        a.accepted = s->d_acc;
        a.last_E = s->d_last;
        a.last_M = s->d_last + s->max_slots;
        a.last_Q = s->d_last + 2 * s->max_slots;
        a.item_nsweeps = s->d_nsw + span[c].first;
        a.item_ticks = s->d_ticks + span[c].first;
        a.lanes_per_word = k.lpw;
        hipStream_t st = s->streams[unit_stream(unit0 + (int)c)];
        const int lay = ising_launch_layout(layout, a.n_items, k.block);
        e = k.model == MCX_ISING
                ? launch_ising_sweeps(k.dim, k.alg, lay, a, ising_launch_block(lay, k.block, words),
                                      a.n_items, lds, st)
                : launch_bc_sweeps(k.dim, k.alg, a, k.block, a.n_items, lds, st);
    }
    for (int k = 1; k < ns && e == hipSuccess; ++k) {
        e = hipEventRecord(s->ev_done[k], s->streams[k]);
        if (e == hipSuccess)
            e = hipStreamWaitEvent(s0, s->ev_done[k], 0);
    }
    std::vector<int64_t> tk(std::max(nl, 1));
    if (e == hipSuccess)
        e = hipEventRecord(s->ev1, s0);
    if (e == hipSuccess && nl)
        e = hipMemcpyAsync(tk.data(), s->d_ticks, sizeof(int64_t) * nl, hipMemcpyDeviceToHost, s0);
    const hipError_t e2 = hipStreamSynchronize(s0);
    if (e == hipSuccess)
        e = e2;
    ms = 0;
    ttime.clear();
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, s->ev0, s->ev1);
    for (size_t t = 0; t < tspan.size() && e == hipSuccess; ++t) {
        float tms = 0;
        e = hipEventElapsedTime(&tms, s->ev_t[t][0], s->ev_t[t][1]);
        ttime.push_back({std::get<0>(tspan[t]), tms, tspan_sw[t]});
    }
    ticks.assign(n, 0);
    for (int q = 0; q < nl; ++q)
        ticks[item_job[q]] = tk[q];
    return e;
}

void worker_main(mcx_server* s)
{
    const size_t n = (size_t)s->max_slots;
    std::vector<double> mom(6 * n);
    std::vector<unsigned long long> acc(n);
    std::vector<int32_t> last(3 * n);
    std::vector<int64_t> ran, ticks;
    std::vector<TileTime> ttime;
    auto die = [&](const char* what, hipError_t e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->error = std::string(what) + ": " + hipGetErrorString(e);
        s->cv_done.notify_all();
    };
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e != hipSuccess)
        return die("hipSetDevice", e);
    for (;;) {
        std::vector<Job> fresh;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->active.empty() || !s->pending.empty(); });
            if (s->stop)
                return;
            take_admissions(s, fresh);
            if (fresh.empty() && s->active.empty())
                continue; // unreachable: an idle pool always fits the first job
            if (s->t_first < 0)
                s->t_first = now_ms();
            s->n_active = (int64_t)(s->active.size() + fresh.size());
        }
        if (!fresh.empty()) {
            if ((e = setup_jobs(s, fresh)) != hipSuccess)
                return die("job set-up", e);
            s->active.insert(s->active.end(), fresh.begin(), fresh.end());
        }
        float ms = 0;
        if ((e = run_quantum(s, ran, ticks, ms, ttime)) != hipSuccess)
            return die("quantum", e);
        for (const TileTime& tt : ttime)
            if (tt.sweeps > 0 && tt.ms > 0)
                s->tile_ms_per_sweep[tt.key] = (double)tt.ms / (double)tt.sweeps;
        bool any_done = false;
        for (size_t i = 0; i < s->active.size(); ++i) {
            Job& j = s->active[i];
            j.remaining -= ran[i];
            j.device_ms += ms;
            if (ticks[i] > 0) {
                j.ticks_per_sweep = (double)ticks[i] / (double)ran[i];
                s->learned[ShapeKey{j.d.model, j.d.dim, j.d.alg, j.d.L}] = j.ticks_per_sweep;
            }
            any_done |= j.remaining == 0;
        }
        if (any_done) {
            e = hipMemcpy(mom.data(), s->d_mom, sizeof(double) * mom.size(), hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                e = hipMemcpy(acc.data(), s->d_acc, sizeof(unsigned long long) * n,
                              hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                e = hipMemcpy(last.data(), s->d_last, sizeof(int32_t) * last.size(),
                              hipMemcpyDeviceToHost);
            if (e != hipSuccess)
                return die("result read-back", e);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = now_ms();
        ++s->quanta;
        s->busy_ms += ms;
        for (size_t i = 0; i < s->active.size(); ++i)
            s->attempts += (double)ran[i] * (double)s->active[i].nsites;
        std::vector<Job> keep;
        for (const Job& j : s->active) {
            if (j.remaining > 0) {
                keep.push_back(j);
                continue;
            }
            mcx_job_result r;
            std::memset(&r, 0, sizeof r);
            r.job_id = j.id;
            r.tag = j.d.tag;
            r.status = MCX_OK;
            r.L = j.d.L;
            r.nsweeps = j.d.nsweeps;
            r.accepted = (int64_t)acc[j.slot];
            r.E_pair = last[j.slot];
            r.M = last[n + j.slot];
            r.Q = last[2 * n + j.slot];
            std::memcpy(r.moments, &mom[6 * (size_t)j.slot], sizeof(double) * 6);
            r.latency_ms = t - j.t_submit;
            r.queue_ms = j.t_admit - j.t_submit;
            r.device_ms = j.device_ms;
            r.slo_met = j.deadline_ms < 0 ? -1 : (r.latency_ms <= j.deadline_ms ? 1 : 0);
            if (j.deadline_ms >= 0)
                s->slo_done.emplace_back(j.d.model, r.latency_ms, j.deadline_ms);
            s->done.push_back(r);
            s->latencies.push_back(r.latency_ms);
            ++s->completed;
            s->t_last = t;
            s->pool->release(j.off, j.bytes);
            s->free_slots.push_back(j.slot);
        }
        s->active.swap(keep);
        s->n_active = (int64_t)s->active.size();
        if (any_done)
            s->cv_done.notify_all();
    }
}

void free_device(mcx_server* s)
{
    hipFree(s->d_pool);
    hipFree(s->d_chains);
    hipFree(s->d_itab);
    hipFree(s->d_btab);
    hipFree(s->d_mom);
    hipFree(s->d_acc);
    hipFree(s->d_last);
    hipFree(s->d_items);
    hipFree(s->d_nsw);
    hipFree(s->d_ticks);
    hipFree(s->d_titems);
    hipFree(s->d_tnsw);
    hipFree(s->d_tplane);
    hipFree(s->d_tobs);
    for (hipEvent_t ev : {s->ev0, s->ev1, s->ev_in})
        if (ev)
            hipEventDestroy(ev);
    for (auto& pr : s->ev_t)
        for (hipEvent_t ev : pr)
            if (ev)
                hipEventDestroy(ev);
    for (int k = 0; k < kMaxStreams; ++k) {
        if (s->ev_done[k])
            hipEventDestroy(s->ev_done[k]);
        if (s->streams[k])
            hipStreamDestroy(s->streams[k]);
    }
}

} // namespace

extern "C" mcx_status mcx_server_create(mcx_ctx* ctx, const mcx_server_desc* d, mcx_server** out)
{
    if (!ctx || !d || !out)
        return fail(MCX_EINVAL, "mcx_server_create: NULL argument");
    *out = nullptr;
    if (d->hbm_budget < kAlign || d->max_jobs < 1 || d->n_streams < 0 ||
        d->n_streams > kMaxStreams || !(d->quantum_ms >= 0) || !std::isfinite(d->quantum_ms))
        return fail(MCX_EINVAL,
                    "mcx_server_create: need hbm_budget >= %lld bytes, max_jobs >= 1, "
                    "0 <= n_streams <= %d and a finite quantum_ms >= 0",
                    (long long)kAlign, kMaxStreams);
    HIPCHK(hipSetDevice(ctx->device));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (khz <= 0)
        return fail(MCX_EHIP, "mcx_server_create: the device reports no wall-clock rate");
    auto* s = new mcx_server();
    s->ctx = ctx;
    s->budget = d->hbm_budget / kAlign * kAlign;
    s->max_slots = d->max_jobs;
    s->n_streams = d->n_streams > 0 ? d->n_streams : 4;
    s->quantum_ms = d->quantum_ms > 0 ? d->quantum_ms : 2.0;
    s->ticks_per_ms = (double)khz;
    // tiled path from 2^15 Ising / 2^14 Blume-Capel sites on, where one
    // workgroup per lattice gets slower per sweep than the tiled launches
    s->tile_sites[0] = d->tile_sites > 0 ? d->tile_sites : (d->tile_sites < 0 ? 0 : 1 << 15);
    s->tile_sites[1] = d->tile_sites > 0 ? d->tile_sites : (d->tile_sites < 0 ? 0 : 1 << 14);
    s->band_sweeps = d->band_sweeps;
    const size_t n = (size_t)s->max_slots;
    hipError_t e = hipSuccess;
    // One hardware queue per stream.  HIP spreads ordinary streams over a few
    // shared queues (GPU_MAX_HW_QUEUES), and kernels of two streams on one
    // queue run one after the other, so two classes of a quantum would not
    // overlap; a stream created with a CU mask always gets a queue of its own.
    // The mask enables every CU; where the runtime refuses one, an ordinary
    // non-blocking stream is used.
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
        ncu = 0;
    std::vector<uint32_t> cu_mask((size_t)(ncu + 31) / 32, 0u);
    for (int i = 0; i < ncu; ++i)
        cu_mask[(size_t)i / 32] |= 1u << (i % 32);
    for (int k = 0; k < s->n_streams && e == hipSuccess; ++k) {
        e = ncu > 0 ? hipExtStreamCreateWithCUMask(&s->streams[k], (uint32_t)cu_mask.size(),
                                                   cu_mask.data())
                    : hipErrorInvalidValue;
        if (e != hipSuccess) {
            (void)hipGetLastError();
            s->streams[k] = nullptr;
            e = hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking);
        }
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&s->ev_done[k], hipEventDisableTiming);
    }
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc(p, bytes);
    };
    alloc((void**)&s->d_pool, (size_t)s->budget);
    alloc((void**)&s->d_chains, sizeof(ChainRec) * n);
    alloc((void**)&s->d_itab, sizeof(IsingTable) * n);
    alloc((void**)&s->d_btab, sizeof(BCTable) * n);
    alloc((void**)&s->d_mom, sizeof(double) * 6 * n);
    alloc((void**)&s->d_acc, sizeof(unsigned long long) * n);
    alloc((void**)&s->d_last, sizeof(int32_t) * 3 * n);
    alloc((void**)&s->d_items, sizeof(int32_t) * n);
    alloc((void**)&s->d_nsw, sizeof(int64_t) * n);
    alloc((void**)&s->d_ticks, sizeof(int64_t) * n);
    alloc((void**)&s->d_titems, sizeof(int32_t) * n);
    alloc((void**)&s->d_tnsw, sizeof(int64_t) * n);
    alloc((void**)&s->d_tplane, sizeof(int64_t) * n);
    alloc((void**)&s->d_tobs, sizeof(int64_t) * n);
    if (e == hipSuccess)
        e = hipEventCreate(&s->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&s->ev1);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming);
    for (auto& pr : s->ev_t)
        for (hipEvent_t& ev : pr)
            if (e == hipSuccess)
                e = hipEventCreate(&ev);

    if (e == hipSuccess)
        e = hipMemsetAsync(s->d_last, 0, sizeof(int32_t) * 3 * n, s->streams[0]);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s->streams[0]);
    if (e != hipSuccess) {
        free_device(s);
        delete s;
        return fail(e == hipErrorOutOfMemory ? MCX_ENOMEM : MCX_EHIP, "mcx_server_create: %s",
                    hipGetErrorString(e));
    }
    s->pool = new PoolAlloc(s->budget);
    for (int i = s->max_slots - 1; i >= 0; --i)
        s->free_slots.push_back(i);
    ++ctx->n_users;
    s->worker = std::thread(worker_main, s);
    *out = s;
    return MCX_OK;
}

extern "C" mcx_status mcx_server_submit(mcx_server* s, const mcx_job_desc* d, uint64_t* job_id)
{
    if (!s || !d)
        return fail(MCX_EINVAL, "mcx_server_submit: NULL argument");
    mcx_status st = validate(s, d);
    if (st)
        return st;
    Job j;
    j.d = *d;
    j.remaining = d->nsweeps;
    j.nsites = job_sites(*d);
    j.tiled = job_tiled(s, *d);
    j.words = job_words(*d);
    j.bytes = job_bytes(s, *d);
    server_shape(d->model, d->dim, d->L, j.block, j.lpw);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->error.empty())
            return fail(MCX_EHIP, "mcx_server_submit: the server has failed (%s)",
                        s->error.c_str());
        j.id = s->next_id++;
        j.t_submit = now_ms();
        if (s->slo_ms > 0 || s->slo_rate > 0)
            j.deadline_ms = std::max(0.0, s->slo_ms) +
                            (s->slo_rate > 0 ? (double)j.nsites * (double)d->nsweeps /
                                                   (s->slo_rate * 1e6)
                                             : 0.0);
        ++s->submitted;
        s->pending.push_back(j);
    }
    s->cv.notify_all();
    if (job_id)
        *job_id = j.id;
    return MCX_OK;
}

extern "C" mcx_status mcx_server_poll(mcx_server* s, mcx_job_result* out, int32_t max, int32_t* n)
{
    if (!s || !n || max < 0 || (max > 0 && !out))
        return fail(MCX_EINVAL, "mcx_server_poll: NULL argument or negative max");
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    if (!s->error.empty())
        return fail(MCX_EHIP, "mcx_server_poll: the server has failed (%s)", s->error.c_str());
    return MCX_OK;
}

extern "C" mcx_status mcx_server_drain(mcx_server* s, double timeout_ms)
{
    if (!s)
        return fail(MCX_EINVAL, "mcx_server_drain: NULL handle");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto idle = [&] { return !s->error.empty() || s->completed == s->submitted; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, idle);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), idle))
        return fail(MCX_ESTATE, "mcx_server_drain: %lld of %lld jobs unfinished after %.0f ms",
                    (long long)(s->submitted - s->completed), (long long)s->submitted, timeout_ms);
    if (!s->error.empty())
        return fail(MCX_EHIP, "mcx_server_drain: the server has failed (%s)", s->error.c_str());
    return MCX_OK;
}

extern "C" mcx_status mcx_server_stats(mcx_server* s, mcx_server_stats_t* out)
{
    if (!s || !out)
        return fail(MCX_EINVAL, "mcx_server_stats: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->pending = (int64_t)s->pending.size();
    out->active = s->n_active;
    out->quanta = s->quanta;
    out->attempts = s->attempts;
    out->busy_ms = s->busy_ms;
    if (!s->latencies.empty()) {
        std::vector<double> v = s->latencies;
        std::sort(v.begin(), v.end());
        // linear interpolation between order statistics
        auto q = [&](double p) {
            const double x = p * (double)(v.size() - 1);
            const size_t i = (size_t)x;
            return i + 1 < v.size() ? v[i] + (x - (double)i) * (v[i + 1] - v[i]) : v[i];
        };
        out->p50_ms = q(0.50);
        out->p99_ms = q(0.99);
        out->max_ms = v.back();
    }
    out->wall_ms = s->t_first >= 0 && s->completed > 0 ? s->t_last - s->t_first : 0.0;
    out->attempts_per_ns = out->wall_ms > 0 ? s->attempts / (out->wall_ms * 1e6) : 0.0;
    out->busy_attempts_per_ns = s->busy_ms > 0 ? s->attempts / (s->busy_ms * 1e6) : 0.0;
    return MCX_OK;
}

extern "C" mcx_status mcx_server_set_slo(mcx_server* s, double slo_ms, double slo_attempts_per_ns)
{
    if (!s)
        return fail(MCX_EINVAL, "mcx_server_set_slo: NULL handle");
    if (!std::isfinite(slo_ms) || !std::isfinite(slo_attempts_per_ns))
        return fail(MCX_EINVAL, "mcx_server_set_slo: the SLO must be finite");
    std::lock_guard<std::mutex> lk(s->mu);
    s->slo_ms = slo_ms > 0 ? slo_ms : 0.0;
    s->slo_rate = slo_attempts_per_ns > 0 ? slo_attempts_per_ns : 0.0;
    return MCX_OK;
}

extern "C" mcx_status mcx_server_slo_report(mcx_server* s, mcx_slo_report* out)
{
    if (!s || !out)
        return fail(MCX_EINVAL, "mcx_server_slo_report: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->slo_ms = s->slo_ms;
    out->slo_attempts_per_ns = s->slo_rate;
    double frac = 0, worst = -1e300;
    for (const auto& e : s->slo_done) {
        const int m = std::get<0>(e);
        const double lat = std::get<1>(e), dl = std::get<2>(e);
        const bool met = lat <= dl;
        ++out->jobs;
        out->met += met ? 1 : 0;
        ++out->jobs_by_model[m];
        out->met_by_model[m] += met ? 1 : 0;
        worst = std::max(worst, lat - dl);
        frac += dl > 0 ? lat / dl : 0.0;
    }
    out->missed = out->jobs - out->met;
    out->attainment = out->jobs ? (double)out->met / (double)out->jobs : 1.0;
    out->worst_lateness_ms = out->jobs ? worst : 0.0;
    out->mean_deadline_frac = out->jobs ? frac / (double)out->jobs : 0.0;
    if (!s->latencies.empty()) {
        std::vector<double> v = s->latencies;
        std::sort(v.begin(), v.end());
        auto q = [&](double p) {
            const double x = p * (double)(v.size() - 1);
            const size_t i = (size_t)x;
            return i + 1 < v.size() ? v[i] + (x - (double)i) * (v[i + 1] - v[i]) : v[i];
        };
        out->p50_ms = q(0.50);
        out->p90_ms = q(0.90);
        out->p99_ms = q(0.99);
        out->p999_ms = q(0.999);
        out->max_ms = v.back();
        double sum = 0;
        for (double x : v)
            sum += x;
        out->mean_ms = sum / (double)v.size();
    }
    return MCX_OK;
}

extern "C" mcx_status mcx_server_destroy(mcx_server* s)
{
    if (!s)
        return MCX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    hipSetDevice(s->ctx->device);
    for (int k = 0; k < s->n_streams; ++k)
        hipStreamSynchronize(s->streams[k]);
    free_device(s);
    --s->ctx->n_users;
    delete s->pool;
    delete s;
    return MCX_OK;
}
