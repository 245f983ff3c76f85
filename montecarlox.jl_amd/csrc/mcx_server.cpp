// mcx_server.cpp -- continuous-batching simulation server on one GPU.
//
// No counterpart in the reference, which runs one chain per thread to
// completion.  Jobs (model, dim, L, algorithm, beta, D, seed, key, sweeps)
// arrive at any time.  A scheduler thread keeps their lattices resident in a
// device pool bounded by an HBM budget and advances every resident job in
// quanta: one launch per kernel class (model, dim, algorithm, block size), in
// which each job runs its own fixed number of sweeps per quantum, chosen so
// that every job's share of a quantum costs about the same.  Finished jobs
// leave at the quantum boundary; their pool space and slot are freed and
// waiting jobs are admitted before the next quantum.
//
// Invariance: a job's Philox stream is keyed by its own (seed, key) and
// counted by its own sweep index, and its quantum lengths depend only on the
// job, so its spins, counters and FP64 moments are the same alone, in any mix
// and as a one-chain mcx_batch run with the same seed, key and random start.
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

struct Job {
    uint64_t id = 0;
    mcx_job_desc d{};
    int64_t remaining = 0; // sweeps still to run
    int64_t nsites = 0;
    int64_t per_quantum = 0;
    int64_t off = -1;      // byte offset of the lattice in the pool
    int64_t bytes = 0;
    int slot = -1;         // index into the per-job device arrays
    double t_submit = 0, t_admit = 0, device_ms = 0;
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

} // namespace

struct mcx_server {
    mcx_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t budget = 0;
    int max_slots = 0;
    int64_t quantum_attempts = 0;
    int64_t min_sites = 0;
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
    // worker-only state
    std::vector<Job> active;
    PoolAlloc* pool = nullptr;
    std::vector<int> free_slots;
    // shared state, guarded by mu
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Job> pending;
    std::deque<mcx_job_result> done;
    std::vector<double> latencies;
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

int64_t job_bytes(const mcx_job_desc& d)
{
    return (job_sites(d) + kAlign - 1) / kAlign * kAlign;
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
    if (d->alg != MCX_METROPOLIS && d->alg != MCX_HEATBATH)
        return fail(MCX_EINVAL, "mcx_server_submit: alg must be MCX_METROPOLIS or MCX_HEATBATH");
    if (!std::isfinite(d->beta) || d->beta < 0 || !std::isfinite(d->D))
        return fail(MCX_EINVAL, "mcx_server_submit: beta must be finite and >= 0, D finite");
    if (d->nsweeps < 1)
        return fail(MCX_EINVAL, "mcx_server_submit: nsweeps must be >= 1");
    if (lattice_lds_bytes(d->model, d->dim, d->L) > s->ctx->lds_max)
        return fail(MCX_EINVAL, "mcx_server_submit: L = %d does not fit the device LDS", d->L);
    if (job_bytes(*d) > s->budget)
        return fail(MCX_EINVAL,
                    "mcx_server_submit: a %lld-byte lattice exceeds the %lld-byte HBM budget",
                    (long long)job_bytes(*d), (long long)s->budget);
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
        j.slot = s->free_slots.back();
        s->free_slots.pop_back();
        j.t_admit = now_ms();
        fresh.push_back(j);
        it = s->pending.erase(it);
    }
}

// Chain record, acceptance table, counters and random start of new jobs.
hipError_t setup_jobs(mcx_server* s, const std::vector<Job>& fresh)
{
    std::vector<ChainRec> rec(fresh.size());
    std::vector<IsingTable> it(fresh.size());
    std::vector<BCTable> bt(fresh.size());
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < fresh.size() && e == hipSuccess; ++i) {
        const Job& j = fresh[i];
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
        e = hipMemcpyAsync(s->d_chains + j.slot, &r, sizeof r, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess && j.d.model == MCX_ISING) {
            std::memset(&it[i], 0, sizeof it[i]);
            double pm[4], ph[8];
            ising_probabilities(j.d.beta, j.d.dim, pm, ph);
            fill_ising_metro(it[i], j.d.dim, pm);
            fill_ising_hb(it[i], j.d.dim, ph);
            e = hipMemcpyAsync(s->d_itab + j.slot, &it[i], sizeof it[i], hipMemcpyHostToDevice,
                               s->stream);
        } else if (e == hipSuccess) {
            fill_bc_table(bt[i], j.d.dim, j.d.beta, j.d.D);
            e = hipMemcpyAsync(s->d_btab + j.slot, &bt[i], sizeof bt[i], hipMemcpyHostToDevice,
                               s->stream);
        }
        if (e == hipSuccess)
            e = hipMemsetAsync(s->d_mom + 6 * (size_t)j.slot, 0, sizeof(double) * 6, s->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync(s->d_acc + j.slot, 0, sizeof(unsigned long long), s->stream);
        if (e == hipSuccess)
            e = launch_init_spins(j.d.model, MCX_INIT_RANDOM, s->d_chains + j.slot, 1, j.nsites,
                                  s->d_pool, s->stream);
    }
    // the host staging vectors must outlive the copies
    const hipError_t e2 = hipStreamSynchronize(s->stream);
    return e != hipSuccess ? e : e2;
}

struct ClassKey {
    int model, dim, alg, block;
    bool operator<(const ClassKey& o) const
    {
        return std::tie(model, dim, alg, block) < std::tie(o.model, o.dim, o.alg, o.block);
    }
    bool operator==(const ClassKey& o) const { return !(*this < o) && !(o < *this); }
};

// One quantum over every active job: ran[i] = sweeps run by active[i], ms =
// device time between the first and the last launch.
hipError_t run_quantum(mcx_server* s, std::vector<int64_t>& ran, float& ms)
{
    const int n = (int)s->active.size();
    std::vector<int> order(n);
    std::vector<ClassKey> key(n);
    ran.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        const Job& j = s->active[i];
        order[i] = i;
        key[i] = {j.d.model, j.d.dim, j.d.alg, lattice_block(j.d.dim, j.d.L)};
        ran[i] = std::min(j.per_quantum, j.remaining);
    }
    // class by class, the most expensive items first
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (!(key[a] == key[b]))
            return key[a] < key[b];
        return ran[a] * s->active[a].nsites > ran[b] * s->active[b].nsites;
    });
    std::vector<int32_t> items(n);
    std::vector<int64_t> nsw(n);
    for (int q = 0; q < n; ++q) {
        items[q] = s->active[order[q]].slot;
        nsw[q] = ran[order[q]];
    }
    hipError_t e = hipMemcpyAsync(s->d_items, items.data(), sizeof(int32_t) * n,
                                  hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_nsw, nsw.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                           s->stream);
    if (e == hipSuccess)
        e = hipEventRecord(s->ev0, s->stream);
    for (int q = 0; q < n && e == hipSuccess;) {
        const ClassKey k = key[order[q]];
        const int first = q;
        size_t lds = 0;
        for (; q < n && key[order[q]] == k; ++q)
            lds = std::max(lds, lattice_lds_bytes(k.model, k.dim, s->active[order[q]].d.L));
        SweepArgs a{};
        a.items = s->d_items + first;
        a.n_items = q - first;
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
        a.item_nsweeps = s->d_nsw + first;
        e = k.model == MCX_ISING
                ? launch_ising_sweeps(k.dim, k.alg, a, k.block, a.n_items, lds, s->stream)
                : launch_bc_sweeps(k.dim, k.alg, a, k.block, a.n_items, lds, s->stream);
    }
    if (e == hipSuccess)
        e = hipEventRecord(s->ev1, s->stream);
    const hipError_t e2 = hipStreamSynchronize(s->stream);
    if (e == hipSuccess)
        e = e2;
    ms = 0;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, s->ev0, s->ev1);
    return e;
}

void worker_main(mcx_server* s)
{
    const size_t n = (size_t)s->max_slots;
    std::vector<double> mom(6 * n);
    std::vector<unsigned long long> acc(n);
    std::vector<int32_t> last(3 * n);
    std::vector<int64_t> ran;
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
        if ((e = run_quantum(s, ran, ms)) != hipSuccess)
            return die("quantum", e);
        bool any_done = false;
        for (size_t i = 0; i < s->active.size(); ++i) {
            Job& j = s->active[i];
            j.remaining -= ran[i];
            j.device_ms += ms;
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
    if (s->ev0)
        hipEventDestroy(s->ev0);
    if (s->ev1)
        hipEventDestroy(s->ev1);
    if (s->stream)
        hipStreamDestroy(s->stream);
}

} // namespace

extern "C" mcx_status mcx_server_create(mcx_ctx* ctx, const mcx_server_desc* d, mcx_server** out)
{
    if (!ctx || !d || !out)
        return fail(MCX_EINVAL, "mcx_server_create: NULL argument");
    *out = nullptr;
    if (d->hbm_budget < kAlign || d->max_jobs < 1 || d->quantum_attempts < 0 || d->min_sites < 0)
        return fail(MCX_EINVAL,
                    "mcx_server_create: need hbm_budget >= %lld bytes, max_jobs >= 1 and "
                    "non-negative quantum_attempts and min_sites",
                    (long long)kAlign);
    HIPCHK(hipSetDevice(ctx->device));
    auto* s = new mcx_server();
    s->ctx = ctx;
    s->budget = d->hbm_budget / kAlign * kAlign;
    s->max_slots = d->max_jobs;
    s->quantum_attempts = d->quantum_attempts > 0 ? d->quantum_attempts : ((int64_t)1 << 24);
    s->min_sites = d->min_sites > 0 ? d->min_sites : 4096;
    const size_t n = (size_t)s->max_slots;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
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
    if (e == hipSuccess)
        e = hipEventCreate(&s->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&s->ev1);
    if (e == hipSuccess)
        e = hipMemsetAsync(s->d_last, 0, sizeof(int32_t) * 3 * n, s->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s->stream);
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
    j.bytes = job_bytes(*d);
    j.per_quantum = std::max<int64_t>(1, s->quantum_attempts / std::max(j.nsites, s->min_sites));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->error.empty())
            return fail(MCX_EHIP, "mcx_server_submit: the server has failed (%s)",
                        s->error.c_str());
        j.id = s->next_id++;
        j.t_submit = now_ms();
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
    hipStreamSynchronize(s->stream);
    free_device(s);
    --s->ctx->n_users;
    delete s->pool;
    delete s;
    return MCX_OK;
}
