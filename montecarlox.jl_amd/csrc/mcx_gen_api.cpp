// mcx_gen_api.cpp -- C ABI for multicanonical / Wang-Landau walkers.
#include "mcx_api_impl.h"
#include "mcx_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace mcx;

struct mcx_gen {
    mcx_batch* b = nullptr;
    int kind = MCX_MULTICANONICAL;
    int n_groups = 0;
    int64_t tot = 0; // weights / histogram entries over all walkers (bins padded to 32)
    std::vector<int32_t> group_of;
    std::vector<double> start, step;
    std::vector<int64_t> nbins, off, member_off;
    std::vector<int32_t> int_bins; // per group: integer-typed binning (Ising: truncating div)
    std::vector<int64_t> roff;     // per group: offset of its reached-bin mask
    std::vector<std::vector<int32_t>> members; // per group; members[g][0] is the root
    double* d_lw = nullptr;
    unsigned long long* d_hist = nullptr;
    int64_t* d_off = nullptr;
    double* d_logf = nullptr;
    unsigned long long* d_steps = nullptr;
    unsigned long long* d_acc = nullptr;
    int64_t* d_err = nullptr;
    int32_t* d_members = nullptr;
    uint8_t* d_reached = nullptr;         // per group: bins seen non-zero by a flatness test
    unsigned long long* d_part = nullptr; // per group: kGenStatBlocks * 3 partials
    double* d_stats = nullptr;            // per group: {min, mean, bins}
    int64_t rtot = 0;
    double* d_pbeta = nullptr; // kind MCX_MUCA_PAIR_SPIN2: beta of each walker's pair term
    std::vector<double> logf;
    std::vector<double> pbeta;
    // groups run side by side on their own streams (forked from / joined to
    // the context stream)
    static constexpr int kStreams = 4;
    hipStream_t gs[kStreams] = {};
    hipEvent_t fork = nullptr, join[kStreams] = {};
};

namespace {

void gen_free(mcx_gen* g)
{
    for (int k = 0; k < mcx_gen::kStreams; ++k) {
        if (g->gs[k])
            hipStreamDestroy(g->gs[k]);
        if (g->join[k])
            hipEventDestroy(g->join[k]);
    }
    if (g->fork)
        hipEventDestroy(g->fork);
    hipFree(g->d_lw);
    hipFree(g->d_hist);
    hipFree(g->d_off);
    hipFree(g->d_logf);
    hipFree(g->d_steps);
    hipFree(g->d_acc);
    hipFree(g->d_err);
    hipFree(g->d_members);
    hipFree(g->d_reached);
    hipFree(g->d_part);
    hipFree(g->d_stats);
    hipFree(g->d_pbeta);
}

mcx_status check_chain(mcx_gen* g, int32_t chain, const char* who)
{
    if (!g)
        return fail(MCX_EINVAL, "%s: NULL handle", who);
    if (chain < 0 || chain >= g->b->n)
        return fail(MCX_EBOUNDS, "%s: chain %d not in [0, %d)", who, chain, g->b->n);
    return MCX_OK;
}

mcx_status check_group(mcx_gen* g, int32_t group, const char* who)
{
    if (!g)
        return fail(MCX_EINVAL, "%s: NULL handle", who);
    if (group < 0 || group >= g->n_groups)
        return fail(MCX_EBOUNDS, "%s: group %d not in [0, %d)", who, group, g->n_groups);
    return MCX_OK;
}

// dynamic LDS of one walker: weights (a window of win bins between max_dbin
// padding slots on each side), optional histogram (whole table), spins
size_t gen_lds(int64_t win, int max_dbin, int64_t nbins, int64_t nsites, bool hist_in_lds)
{
    return 8 * ((size_t)win + 2 * (size_t)max_dbin) + (hist_in_lds ? 4 * (size_t)nbins : 0)
         + (size_t)((nsites + 15) / 16 * 16);
}

// Weight tables above this size live in HBM and the walker keeps a window of
// them in LDS, so that walkers of every lattice size fit side by side on a CU.
constexpr size_t kWindowAbove = 24 * 1024;
constexpr int64_t kMinWindow = 2048;

struct GenWindow {
    int64_t win;   // bins held in LDS (= nbins: the whole table)
    int max_dbin;  // largest bin change of one attempt
};

// The window must hold every bin 66 attempts can reach on either side of the
// current one (a chunk of 64 plus the walk's lookahead).
GenWindow gen_window(const mcx_gen* g, int k)
{
    const mcx_batch* b = g->b;
    double max_de;
    if (g->kind == MCX_MUCA_PAIR_SPIN2) {
        max_de = 1.0; // Q = sum s^2 changes by at most 1
    } else if (b->model == MCX_ISING) {
        max_de = 4.0 * b->dim; // |2 s sum_j s_j|
    } else {
        double dmax = 0;
        for (int c : g->members[k])
            dmax = std::max(dmax, std::fabs(b->rec[c].D));
        max_de = 4.0 * b->dim + dmax; // |ds| * |sum| + |D| |ds2|
    }
    GenWindow w;
    w.max_dbin = (int)std::ceil(max_de / std::fabs(g->step[k])) + 2;
    const int64_t nb = g->nbins[k];
    const int64_t need = 2 * 66 * (int64_t)w.max_dbin + 64;
    int64_t win = std::max(kMinWindow, (need + 63) / 64 * 64);
    if (8 * (size_t)nb <= kWindowAbove || win >= nb)
        win = nb;
    w.win = win;
    return w;
}

} // namespace

extern "C" mcx_status mcx_gen_create(mcx_batch* b, const mcx_gen_desc* d, mcx_gen** out)
{
    if (!b || !d || !out)
        return fail(MCX_EINVAL, "mcx_gen_create: NULL argument");
    *out = nullptr;
    if (d->kind != MCX_MULTICANONICAL && d->kind != MCX_WANGLANDAU && d->kind != MCX_MUCA_PAIR_SPIN2)
        return fail(MCX_EINVAL, "mcx_gen_create: kind must be MCX_MULTICANONICAL, MCX_WANGLANDAU or "
                                "MCX_MUCA_PAIR_SPIN2");
    if (d->n_groups < 1 || !d->bin_start || !d->bin_step || !d->n_bins)
        return fail(MCX_EINVAL, "mcx_gen_create: need n_groups >= 1 and per-group binnings");
    if (d->kind == MCX_MUCA_PAIR_SPIN2) {
        if (b->model != MCX_BLUME_CAPEL)
            return fail(MCX_EINVAL, "mcx_gen_create: MCX_MUCA_PAIR_SPIN2 walkers need a Blume-Capel "
                                    "batch (sum s^2 is constant for Ising)");
        if (!d->pair_beta)
            return fail(MCX_EINVAL, "mcx_gen_create: MCX_MUCA_PAIR_SPIN2 needs pair_beta");
        for (int i = 0; i < b->n; ++i)
            if (!std::isfinite(d->pair_beta[i]))
                return fail(MCX_EINVAL, "mcx_gen_create: pair_beta[%d] is not finite", i);
        for (int k = 0; k < d->n_groups; ++k)
            if (d->bin_start[k] != std::floor(d->bin_start[k]) || !(d->bin_step[k] >= 1) ||
                d->bin_step[k] != std::floor(d->bin_step[k]) || std::fabs(d->bin_start[k]) > 1e9)
                return fail(MCX_EINVAL, "mcx_gen_create: group %d: sum s^2 is an integer; start and "
                                        "step must be integral (step >= 1)", k);
    }
    auto* g = new mcx_gen();
    g->b = b;
    g->kind = d->kind;
    g->n_groups = d->n_groups;
    g->start.assign(d->bin_start, d->bin_start + d->n_groups);
    g->step.assign(d->bin_step, d->bin_step + d->n_groups);
    g->nbins.assign(d->n_bins, d->n_bins + d->n_groups);
    g->int_bins.assign(d->n_groups, b->model == MCX_ISING ? 1 : 0);
    if (d->bin_integer)
        for (int k = 0; k < d->n_groups; ++k)
            g->int_bins[k] = d->bin_integer[k] ? 1 : 0;
    g->group_of.resize(b->n);
    g->members.resize(d->n_groups);
    for (int i = 0; i < b->n; ++i) {
        const int grp = d->group_of_chain ? d->group_of_chain[i] : 0;
        if (grp < 0 || grp >= d->n_groups) {
            delete g;
            return fail(MCX_EINVAL, "mcx_gen_create: chain %d in group %d of %d", i, grp, d->n_groups);
        }
        g->group_of[i] = grp;
        g->members[grp].push_back(i);
    }
    for (int k = 0; k < d->n_groups; ++k) {
        if (g->members[k].empty()) {
            delete g;
            return fail(MCX_EINVAL, "mcx_gen_create: group %d has no walkers", k);
        }
        if (g->nbins[k] < 1 || !(g->step[k] != 0) || !std::isfinite(g->start[k]) ||
            !std::isfinite(g->step[k])) {
            delete g;
            return fail(MCX_EINVAL, "mcx_gen_create: group %d: invalid binning", k);
        }
        if (b->model == MCX_ISING && g->int_bins[k] &&
            (g->start[k] != std::floor(g->start[k]) || g->step[k] != std::floor(g->step[k]))) {
            delete g;
            return fail(MCX_EINVAL, "mcx_gen_create: group %d: an integer binning needs an "
                                    "integral start and step", k);
        }
        const int L0 = b->rec[g->members[k][0]].L;
        for (int c : g->members[k])
            if (b->rec[c].L != L0) {
                delete g;
                return fail(MCX_EINVAL, "mcx_gen_create: group %d mixes lattice sizes", k);
            }
        if (std::fabs(g->start[k]) > 1e15 || (double)g->nbins[k] > 2e9) {
            delete g;
            return fail(MCX_EINVAL, "mcx_gen_create: group %d: binning beyond 32-bit bin indices", k);
        }
    }
    for (int k = 0; k < d->n_groups; ++k) {
        const GenWindow w = gen_window(g, k);
        if (gen_lds(w.win, w.max_dbin, g->nbins[k], b->nsites[g->members[k][0]], false) + kGenStaticLds >
            b->ctx->lds_max) {
            delete g;
            return fail(MCX_EINVAL, "mcx_gen_create: group %d: the lattice (%lld sites) and a window "
                                    "of %lld bins do not fit the device LDS",
                        k, (long long)b->nsites[g->members[k][0]], (long long)w.win);
        }
    }
    g->off.resize(b->n);
    int64_t tot = 0;
    for (int i = 0; i < b->n; ++i) {
        g->off[i] = tot;
        tot += (g->nbins[g->group_of[i]] + 31) / 32 * 32;
    }
    g->tot = tot;
    for (int k = 0; k < d->n_groups; ++k) {
        g->roff.push_back(g->rtot);
        g->rtot += (g->nbins[k] + 255) / 256 * 256;
    }
    std::vector<int32_t> flat;
    for (int k = 0; k < g->n_groups; ++k) {
        g->member_off.push_back((int64_t)flat.size());
        flat.insert(flat.end(), g->members[k].begin(), g->members[k].end());
    }
    g->logf.assign(b->n, 1.0);
    g->pbeta.assign(b->n, 0.0);
    if (d->kind == MCX_MUCA_PAIR_SPIN2)
        g->pbeta.assign(d->pair_beta, d->pair_beta + b->n);
    HIPCHK(hipSetDevice(b->ctx->device));
    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc(p, bytes);
    };
    alloc((void**)&g->d_lw, sizeof(double) * tot);
    alloc((void**)&g->d_hist, sizeof(unsigned long long) * tot);
    alloc((void**)&g->d_off, sizeof(int64_t) * b->n);
    alloc((void**)&g->d_logf, sizeof(double) * b->n);
    alloc((void**)&g->d_steps, sizeof(unsigned long long) * b->n);
    alloc((void**)&g->d_acc, sizeof(unsigned long long) * b->n);
    alloc((void**)&g->d_err, sizeof(int64_t) * b->n);
    alloc((void**)&g->d_members, sizeof(int32_t) * b->n);
    alloc((void**)&g->d_reached, (size_t)g->rtot);
    alloc((void**)&g->d_part, sizeof(unsigned long long) * 3 * kGenStatBlocks * d->n_groups);
    alloc((void**)&g->d_stats, sizeof(double) * 3 * d->n_groups);
    alloc((void**)&g->d_pbeta, sizeof(double) * b->n);
    for (int k = 0; k < mcx_gen::kStreams && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&g->gs[k], hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&g->join[k], hipEventDisableTiming);
    }
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    hipStream_t st = b->ctx->stream;
    if (e == hipSuccess)
        e = hipMemsetAsync(g->d_lw, 0, sizeof(double) * tot, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(g->d_hist, 0, sizeof(unsigned long long) * tot, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(g->d_reached, 0, (size_t)g->rtot, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(g->d_steps, 0, sizeof(unsigned long long) * b->n, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(g->d_acc, 0, sizeof(unsigned long long) * b->n, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->d_off, g->off.data(), sizeof(int64_t) * b->n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->d_logf, g->logf.data(), sizeof(double) * b->n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->d_pbeta, g->pbeta.data(), sizeof(double) * b->n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->d_members, flat.data(), sizeof(int32_t) * b->n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        gen_free(g);
        delete g;
        return fail(e == hipErrorOutOfMemory ? MCX_ENOMEM : MCX_EHIP, "mcx_gen_create: %s",
                    hipGetErrorString(e));
    }
    ++b->n_gen;
    *out = g;
    return MCX_OK;
}

extern "C" mcx_status mcx_gen_destroy(mcx_gen* g)
{
    if (!g)
        return MCX_OK;
    hipSetDevice(g->b->ctx->device);
    hipStreamSynchronize(g->b->ctx->stream);
    --g->b->n_gen;
    gen_free(g);
    delete g;
    return MCX_OK;
}

extern "C" mcx_status mcx_gen_set_logweight(mcx_gen* g, int32_t chain, const double* lw, int64_t n)
{
    mcx_status s = check_chain(g, chain, "mcx_gen_set_logweight");
    if (s)
        return s;
    if (!lw || n != g->nbins[g->group_of[chain]])
        return fail(MCX_EINVAL, "mcx_gen_set_logweight: chain %d has %lld bins", chain,
                    (long long)g->nbins[g->group_of[chain]]);
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(hipMemcpyAsync(g->d_lw + g->off[chain], lw, sizeof(double) * n, hipMemcpyHostToDevice,
                          g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_set_logweight");
}

extern "C" mcx_status mcx_gen_get_logweight(mcx_gen* g, int32_t chain, double* lw, int64_t n)
{
    mcx_status s = check_chain(g, chain, "mcx_gen_get_logweight");
    if (s)
        return s;
    if (!lw || n != g->nbins[g->group_of[chain]])
        return fail(MCX_EINVAL, "mcx_gen_get_logweight: chain %d has %lld bins", chain,
                    (long long)g->nbins[g->group_of[chain]]);
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(hipMemcpyAsync(lw, g->d_lw + g->off[chain], sizeof(double) * n, hipMemcpyDeviceToHost,
                          g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_get_logweight");
}

extern "C" mcx_status mcx_gen_get_histogram(mcx_gen* g, int32_t chain, int64_t* h, int64_t n)
{
    mcx_status s = check_chain(g, chain, "mcx_gen_get_histogram");
    if (s)
        return s;
    if (!h || n != g->nbins[g->group_of[chain]])
        return fail(MCX_EINVAL, "mcx_gen_get_histogram: chain %d has %lld bins", chain,
                    (long long)g->nbins[g->group_of[chain]]);
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(hipMemcpyAsync(h, g->d_hist + g->off[chain], sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                          g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_get_histogram");
}

extern "C" mcx_status mcx_gen_set_logf(mcx_gen* g, int32_t chain, double logf)
{
    if (!g)
        return fail(MCX_EINVAL, "mcx_gen_set_logf: NULL handle");
    if (chain >= 0) {
        mcx_status s = check_chain(g, chain, "mcx_gen_set_logf");
        if (s)
            return s;
    }
    if (!std::isfinite(logf))
        return fail(MCX_EINVAL, "mcx_gen_set_logf: logf must be finite");
    for (int i = 0; i < g->b->n; ++i)
        if (chain < 0 || chain == i)
            g->logf[i] = logf;
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(hipMemcpyAsync(g->d_logf, g->logf.data(), sizeof(double) * g->b->n, hipMemcpyHostToDevice,
                          g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_set_logf");
}

extern "C" mcx_status mcx_gen_get_logf(mcx_gen* g, double* out)
{
    if (!g || !out)
        return fail(MCX_EINVAL, "mcx_gen_get_logf: NULL argument");
    std::memcpy(out, g->logf.data(), sizeof(double) * g->b->n);
    return MCX_OK;
}

namespace {

// nsweeps sweeps of the walkers of every group with active[k] (nullptr: all);
// adds the device time of the call to *ms when ms is set
mcx_status gen_run(mcx_gen* g, int64_t nsweeps, uint32_t flags, const std::vector<char>* active,
                   const char* who, double* ms_acc = nullptr)
{
    mcx_batch* b = g->b;
    const bool record = (flags & MCX_RUN_RECORD) != 0;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipMemsetAsync(g->d_err, 0xff, sizeof(int64_t) * b->n, st));
    HIPCHK(hipEventRecord(b->ctx->ev0, st));
    // groups by decreasing lattice size: the longest walks start first
    std::vector<int> order;
    for (int k = 0; k < g->n_groups; ++k)
        if (!active || (*active)[k])
            order.push_back(k);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return b->nsites[g->members[x][0]] > b->nsites[g->members[y][0]];
    });
    const int nstr = std::min((int)order.size(), (int)mcx_gen::kStreams);
    HIPCHK(hipEventRecord(g->fork, st));
    for (int q = 0; q < nstr; ++q)
        HIPCHK(hipStreamWaitEvent(g->gs[q], g->fork, 0));
    for (size_t oi = 0; oi < order.size(); ++oi) {
        const int k = order[oi];
        GenArgs a{};
        a.items = g->d_members + g->member_off[k];
        a.n_items = (int32_t)g->members[k].size();
        a.kind = g->kind;
        a.model = b->model;
        a.dim = b->dim;
        a.record = record ? 1 : 0;
        a.int_bins = g->int_bins[k];
        const int64_t ns = b->nsites[g->members[k][0]];
        const GenWindow gw = gen_window(g, k);
        a.win = (int32_t)gw.win;
        a.max_dbin = gw.max_dbin;
        // a windowed walker counts visits in HBM; a whole table may count in LDS
        a.hist_in_lds = gw.win == g->nbins[k] &&
                        gen_lds(gw.win, gw.max_dbin, g->nbins[k], ns, true) + kGenStaticLds <= b->ctx->lds_max ? 1 : 0;
        // the uint32 LDS counts of one launch hold at most nsweeps * N visits per bin
        if ((double)nsweeps * (double)ns >= 4294967296.0)
            a.hist_in_lds = 0;
        a.bin_start = g->start[k];
        a.bin_step = g->step[k];
        {
            // a power-of-two step divides exactly by a multiply (FP64) or a
            // shift (integer energies, which need an integral step)
            int e = 0;
            const double m = std::frexp(a.bin_step, &e);
            const bool pow2 = a.bin_step > 0 && m == 0.5;
            a.step_pow2 = pow2 && (b->model != MCX_ISING || e - 1 >= 0) ? 1 : 0;
            a.step_shift = pow2 && e - 1 >= 0 ? e - 1 : 0;
            a.inv_step = pow2 ? std::ldexp(1.0, 1 - e) : 0.0;
        }
        a.nbins = g->nbins[k];
        a.chains = b->d_chains;
        a.spins = b->d_spins;
        a.lw = g->d_lw;
        a.hist = g->d_hist;
        a.lw_off = g->d_off;
        a.logf = g->d_logf;
        a.steps = g->d_steps;
        a.accepted = g->d_acc;
        a.err_site = g->d_err;
        a.ser_E = b->d_E;
        a.ser_M = b->d_M;
        a.ser_Q = b->d_Q;
        a.nsweeps = nsweeps;
        a.rec_pos = b->rec_len;
        a.pair_beta = g->d_pbeta;
        HIPCHK(launch_gen_sweeps(a, gen_lds(gw.win, gw.max_dbin, g->nbins[k], ns, a.hist_in_lds), g->gs[oi % nstr]));
    }
    for (int q = 0; q < nstr; ++q) {
        HIPCHK(hipEventRecord(g->join[q], g->gs[q]));
        HIPCHK(hipStreamWaitEvent(st, g->join[q], 0));
    }
    HIPCHK(hipEventRecord(b->ctx->ev1, st));
    mcx_status s = sync_check(b, who);
    if (s)
        return s;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
    b->last_ms = ms;
    if (ms_acc)
        *ms_acc += ms;
    s = pull_chains(b);
    if (s)
        return s;
    if (record)
        b->rec_len += nsweeps;
    std::vector<int64_t> err(b->n);
    HIPCHK(hipMemcpy(err.data(), g->d_err, sizeof(int64_t) * b->n, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n; ++i)
        if (err[i] >= 0)
            return fail(MCX_EBOUNDS, "%s: chain %d: the energy of the attempt at site %lld "
                                     "lies outside the binned domain", who, i, (long long)err[i]);
    return MCX_OK;
}

} // namespace

extern "C" mcx_status mcx_gen_run(mcx_gen* g, int64_t nsweeps, uint32_t flags)
{
    if (!g)
        return fail(MCX_EINVAL, "mcx_gen_run: NULL handle");
    mcx_batch* b = g->b;
    if (nsweeps < 0)
        return fail(MCX_EINVAL, "mcx_gen_run: nsweeps must be >= 0");
    if (flags & ~MCX_RUN_RECORD)
        return fail(MCX_EINVAL, "mcx_gen_run: only MCX_RUN_RECORD is supported");
    if ((flags & MCX_RUN_RECORD) && b->rec_len + nsweeps > b->cap)
        return fail(MCX_EINVAL, "mcx_gen_run: recording %lld sweeps exceeds the series capacity",
                    (long long)nsweeps);
    return gen_run(g, nsweeps, flags, nullptr, "mcx_gen_run");
}

namespace {

mcx_status gen_reset(mcx_gen* g, int32_t chain, bool hist, bool counters, const char* who)
{
    if (!g)
        return fail(MCX_EINVAL, "%s: NULL handle", who);
    if (chain >= 0) {
        mcx_status s = check_chain(g, chain, who);
        if (s)
            return s;
    }
    mcx_batch* b = g->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    for (int i = 0; i < b->n; ++i) {
        if (chain >= 0 && chain != i)
            continue;
        const int64_t nb = g->nbins[g->group_of[i]];
        if (hist)
            HIPCHK(hipMemsetAsync(g->d_hist + g->off[i], 0, sizeof(unsigned long long) * nb, st));
        if (counters) {
            HIPCHK(hipMemsetAsync(g->d_steps + i, 0, sizeof(unsigned long long), st));
            HIPCHK(hipMemsetAsync(g->d_acc + i, 0, sizeof(unsigned long long), st));
        }
    }
    return sync_check(b, who);
}

} // namespace

extern "C" mcx_status mcx_gen_reset(mcx_gen* g, int32_t chain)
{
    // reset!(alg): the multicanonical algorithm clears its histogram too; a
    // Metropolis algorithm over a CustomEnsemble only its counters
    return gen_reset(g, chain, !g || g->kind != MCX_MUCA_PAIR_SPIN2, true, "mcx_gen_reset");
}

extern "C" mcx_status mcx_gen_reset_histogram(mcx_gen* g, int32_t chain)
{
    return gen_reset(g, chain, true, false, "mcx_gen_reset_histogram");
}

extern "C" mcx_status mcx_gen_set_pair_beta(mcx_gen* g, int32_t chain, double beta)
{
    if (!g)
        return fail(MCX_EINVAL, "mcx_gen_set_pair_beta: NULL handle");
    if (g->kind != MCX_MUCA_PAIR_SPIN2)
        return fail(MCX_ESTATE, "mcx_gen_set_pair_beta: only MCX_MUCA_PAIR_SPIN2 walkers have a "
                                "pair term");
    if (chain >= 0) {
        mcx_status s = check_chain(g, chain, "mcx_gen_set_pair_beta");
        if (s)
            return s;
    }
    if (!std::isfinite(beta))
        return fail(MCX_EINVAL, "mcx_gen_set_pair_beta: beta must be finite");
    for (int i = 0; i < g->b->n; ++i)
        if (chain < 0 || chain == i)
            g->pbeta[i] = beta;
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(hipMemcpyAsync(g->d_pbeta, g->pbeta.data(), sizeof(double) * g->b->n,
                          hipMemcpyHostToDevice, g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_set_pair_beta");
}

extern "C" mcx_status mcx_gen_update(mcx_gen* g, int32_t chain)
{
    if (!g)
        return fail(MCX_EINVAL, "mcx_gen_update: NULL handle");
    if (chain >= 0) {
        mcx_status s = check_chain(g, chain, "mcx_gen_update");
        if (s)
            return s;
    }
    mcx_batch* b = g->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    for (int k = 0; k < g->n_groups; ++k) {
        if (chain >= 0) {
            if (g->group_of[chain] != k)
                continue;
            const auto& m = g->members[k];
            const int pos = (int)(std::find(m.begin(), m.end(), chain) - m.begin());
            HIPCHK(launch_gen_update(g->d_lw, g->d_hist, g->d_off, g->d_members + g->member_off[k] + pos,
                                     1, g->nbins[k], st));
        } else {
            HIPCHK(launch_gen_update(g->d_lw, g->d_hist, g->d_off, g->d_members + g->member_off[k],
                                     (int32_t)g->members[k].size(), g->nbins[k], st));
        }
    }
    return sync_check(b, "mcx_gen_update");
}

extern "C" mcx_status mcx_gen_merge(mcx_gen* g, int32_t group)
{
    mcx_status s = check_group(g, group, "mcx_gen_merge");
    if (s)
        return s;
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(launch_gen_merge(g->d_hist, g->d_off, g->d_members + g->member_off[group],
                            (int32_t)g->members[group].size(), g->nbins[group], g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_merge");
}

extern "C" mcx_status mcx_gen_distribute(mcx_gen* g, int32_t group)
{
    mcx_status s = check_group(g, group, "mcx_gen_distribute");
    if (s)
        return s;
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(launch_gen_distribute(g->d_lw, g->d_off, g->d_members + g->member_off[group],
                                 (int32_t)g->members[group].size(), g->nbins[group],
                                 g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_distribute");
}

extern "C" mcx_status mcx_gen_average_logweight(mcx_gen* g, int32_t group)
{
    mcx_status s = check_group(g, group, "mcx_gen_average_logweight");
    if (s)
        return s;
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(launch_gen_average(g->d_lw, g->d_off, g->d_members + g->member_off[group],
                              (int32_t)g->members[group].size(), g->nbins[group], g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_average_logweight");
}

namespace {

// {min, mean, bins} of the summed histograms of n walkers of group k (the
// first n of its member list) over the group's reached bins, into
// d_stats[3k..3k+2] (device); the reached mask is updated
hipError_t group_stats_async(mcx_gen* g, int k, int n, hipStream_t st)
{
    return launch_gen_hist_stats(g->d_hist, g->d_off, g->d_members + g->member_off[k], n,
                                 g->nbins[k], g->d_reached + g->roff[k],
                                 g->d_part + 3 * kGenStatBlocks * k, g->d_stats + 3 * k, st);
}

} // namespace

extern "C" mcx_status mcx_gen_flatness(mcx_gen* g, int32_t group, double* out3)
{
    mcx_status s = check_group(g, group, "mcx_gen_flatness");
    if (s)
        return s;
    if (!out3)
        return fail(MCX_EINVAL, "mcx_gen_flatness: NULL output");
    HIPCHK(hipSetDevice(g->b->ctx->device));
    hipStream_t st = g->b->ctx->stream;
    HIPCHK(group_stats_async(g, group, 1, st)); // the root alone
    HIPCHK(hipMemcpyAsync(out3, g->d_stats + 3 * group, sizeof(double) * 3, hipMemcpyDeviceToHost,
                          st));
    return sync_check(g->b, "mcx_gen_flatness");
}

extern "C" mcx_status mcx_gen_iterate(mcx_gen* g, const mcx_gen_schedule* sc, mcx_gen_progress* out)
{
    const char* who = "mcx_gen_iterate";
    if (!g || !sc)
        return fail(MCX_EINVAL, "%s: NULL argument", who);
    if (sc->sweeps < 1 || sc->therm_sweeps < 0 || sc->max_updates < 1 || sc->flags != 0)
        return fail(MCX_EINVAL, "%s: need sweeps >= 1, therm_sweeps >= 0, max_updates >= 1 and "
                                "flags == 0", who);
    const bool wl = g->kind == MCX_WANGLANDAU;
    if (wl && (!(sc->flatness > 0 && sc->flatness <= 1) || !(sc->logf_final > 0) ||
               !(sc->power > 0 && sc->power < 1)))
        return fail(MCX_EINVAL, "%s: Wang-Landau needs 0 < flatness <= 1, logf_final > 0 and "
                                "0 < power < 1", who);
    mcx_gen_progress pr{};
    mcx_batch* b = g->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    // a Wang-Landau group runs while its root's ln f is not below logf_final
    std::vector<char> active(g->n_groups, 1);
    auto count_active = [&]() {
        int n = 0;
        for (int k = 0; k < g->n_groups; ++k) {
            if (wl && !(g->logf[g->members[k][0]] >= sc->logf_final))
                active[k] = 0;
            n += active[k] ? 1 : 0;
        }
        return n;
    };
    mcx_status s = MCX_OK;
    std::vector<double> stats(3 * (size_t)g->n_groups);
    while (pr.updates < sc->max_updates && count_active() > 0) {
        if (!wl) {
            // ParallelMulticanonical: thermalise, reset!, record, merge, update!, distribute
            if (sc->therm_sweeps > 0 &&
                (s = gen_run(g, sc->therm_sweeps, 0, &active, who, &pr.device_ms)) != MCX_OK)
                return s;
            for (int k = 0; k < g->n_groups; ++k)
                HIPCHK(launch_gen_zero_hist(g->d_hist, g->d_off, g->d_members + g->member_off[k],
                                            (int32_t)g->members[k].size(), g->nbins[k], st));
            HIPCHK(hipMemsetAsync(g->d_steps, 0, sizeof(unsigned long long) * b->n, st));
            HIPCHK(hipMemsetAsync(g->d_acc, 0, sizeof(unsigned long long) * b->n, st));
            if ((s = gen_run(g, sc->sweeps, 0, &active, who, &pr.device_ms)) != MCX_OK)
                return s;
            for (int k = 0; k < g->n_groups; ++k) {
                const int32_t* m = g->d_members + g->member_off[k];
                const int32_t n = (int32_t)g->members[k].size();
                HIPCHK(launch_gen_merge(g->d_hist, g->d_off, m, n, g->nbins[k], st));
                HIPCHK(launch_gen_update(g->d_lw, g->d_hist, g->d_off, m, 1, g->nbins[k], st));
                HIPCHK(launch_gen_distribute(g->d_lw, g->d_off, m, n, g->nbins[k], st));
            }
            if ((s = sync_check(b, who)) != MCX_OK)
                return s;
            pr.sweeps += sc->therm_sweeps + sc->sweeps;
            ++pr.updates;
            continue;
        }
        // Wang-Landau: sweep the active groups, share the weights, test flatness
        if ((s = gen_run(g, sc->sweeps, 0, &active, who, &pr.device_ms)) != MCX_OK)
            return s;
        pr.sweeps += sc->sweeps;
        ++pr.updates;
        for (int k = 0; k < g->n_groups; ++k) {
            if (!active[k])
                continue;
            const int32_t* m = g->d_members + g->member_off[k];
            const int32_t n = (int32_t)g->members[k].size();
            HIPCHK(launch_gen_average(g->d_lw, g->d_off, m, n, g->nbins[k], st));
            HIPCHK(launch_gen_distribute(g->d_lw, g->d_off, m, n, g->nbins[k], st));
            HIPCHK(group_stats_async(g, k, n, st));
        }
        HIPCHK(hipMemcpyAsync(stats.data(), g->d_stats, sizeof(double) * stats.size(),
                              hipMemcpyDeviceToHost, st));
        if ((s = sync_check(b, who)) != MCX_OK)
            return s;
        bool logf_changed = false;
        for (int k = 0; k < g->n_groups; ++k) {
            if (!active[k])
                continue;
            const double mn = stats[3 * k], mean = stats[3 * k + 1], cnt = stats[3 * k + 2];
            if (!(cnt > 0 && mn >= sc->flatness * mean))
                continue;
            const double lf = g->logf[g->members[k][0]] * sc->power;
            for (int c : g->members[k])
                g->logf[c] = lf;
            logf_changed = true;
            ++pr.stages;
            HIPCHK(launch_gen_zero_hist(g->d_hist, g->d_off, g->d_members + g->member_off[k],
                                        (int32_t)g->members[k].size(), g->nbins[k], st));
        }
        if (logf_changed)
            HIPCHK(hipMemcpyAsync(g->d_logf, g->logf.data(), sizeof(double) * b->n,
                                  hipMemcpyHostToDevice, st));
        if ((s = sync_check(b, who)) != MCX_OK)
            return s;
    }
    pr.active = count_active();
    if (out)
        *out = pr;
    return MCX_OK;
}

extern "C" mcx_status mcx_gen_get_counters(mcx_gen* g, int64_t* steps, int64_t* accepted)
{
    if (!g)
        return fail(MCX_EINVAL, "mcx_gen_get_counters: NULL handle");
    mcx_batch* b = g->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    if (steps)
        HIPCHK(hipMemcpyAsync(steps, g->d_steps, sizeof(int64_t) * b->n, hipMemcpyDeviceToHost, st));
    if (accepted)
        HIPCHK(hipMemcpyAsync(accepted, g->d_acc, sizeof(int64_t) * b->n, hipMemcpyDeviceToHost, st));
    return sync_check(b, "mcx_gen_get_counters");
}

extern "C" mcx_status mcx_gen_group_info(mcx_gen* g, int32_t group, int32_t* root, int32_t* count,
                                         int64_t* nbins)
{
    mcx_status s = check_group(g, group, "mcx_gen_group_info");
    if (s)
        return s;
    if (root)
        *root = g->members[group][0];
    if (count)
        *count = (int32_t)g->members[group].size();
    if (nbins)
        *nbins = g->nbins[group];
    return MCX_OK;
}

extern "C" mcx_status mcx_gen_set_histogram(mcx_gen* g, int32_t chain, const int64_t* h, int64_t n)
{
    mcx_status s = check_chain(g, chain, "mcx_gen_set_histogram");
    if (s)
        return s;
    if (!h || n != g->nbins[g->group_of[chain]])
        return fail(MCX_EINVAL, "mcx_gen_set_histogram: chain %d has %lld bins", chain,
                    (long long)g->nbins[g->group_of[chain]]);
    for (int64_t k = 0; k < n; ++k)
        if (h[k] < 0)
            return fail(MCX_EDOMAIN, "mcx_gen_set_histogram: bin %lld has a negative count",
                        (long long)k);
    HIPCHK(hipSetDevice(g->b->ctx->device));
    HIPCHK(hipMemcpyAsync(g->d_hist + g->off[chain], h, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                          g->b->ctx->stream));
    return sync_check(g->b, "mcx_gen_set_histogram");
}

extern "C" mcx_status mcx_muca_update_values(mcx_ctx* ctx, double* lw, const double* h, int64_t n)
{
    if (!ctx || !lw || !h || n < 1)
        return fail(MCX_EINVAL, "mcx_muca_update_values: NULL argument or n < 1");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    double *dl = nullptr, *dh = nullptr;
    HIPCHK(hipMalloc(&dl, sizeof(double) * n));
    hipError_t e = hipMalloc(&dh, sizeof(double) * n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dl, lw, sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dh, h, sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_muca_update_values(dl, dh, n, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(lw, dl, sizeof(double) * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    hipFree(dl);
    hipFree(dh);
    if (e != hipSuccess)
        return fail(MCX_EHIP, "mcx_muca_update_values: %s", hipGetErrorString(e));
    return MCX_OK;
}

// ---------------------------------------------------------------------------
// device-state snapshot (checkpoint / resume): log-weights, histograms, ln f
// and counters of every walker and the groups' reached-bin masks; restores
// into walkers of the same layout.
// ---------------------------------------------------------------------------

namespace {

// 2: the groups' reached-bin masks follow the counters
constexpr uint32_t kGenStateVersion = 2;

size_t gen_state_bytes(const mcx_gen* g)
{
    const size_t n = (size_t)g->b->n, t = (size_t)g->tot;
    return sizeof(StateHeader) + sizeof(int64_t) * n + sizeof(double) * t +
           sizeof(unsigned long long) * t + sizeof(double) * n + sizeof(unsigned long long) * 2 * n +
           (size_t)g->rtot;
}

} // namespace

extern "C" mcx_status mcx_gen_state_size(mcx_gen* g, size_t* bytes)
{
    if (!g || !bytes)
        return fail(MCX_EINVAL, "mcx_gen_state_size: NULL argument");
    *bytes = gen_state_bytes(g);
    return MCX_OK;
}

extern "C" mcx_status mcx_gen_save_state(mcx_gen* g, void* buf, size_t bytes)
{
    if (!g || !buf)
        return fail(MCX_EINVAL, "mcx_gen_save_state: NULL argument");
    if (bytes < gen_state_bytes(g))
        return fail(MCX_EINVAL, "mcx_gen_save_state: buffer of %zu bytes, %zu needed", bytes,
                    gen_state_bytes(g));
    mcx_batch* b = g->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    const size_t n = (size_t)b->n, t = (size_t)g->tot;
    StateHeader h = make_state_header(3, 0);
    h.version = kGenStateVersion;
    h.shape[0] = g->kind;
    h.shape[1] = b->n;
    h.shape[2] = g->n_groups;
    h.shape[3] = g->tot;
    BlobWriter w(buf, bytes);
    w.put(&h, sizeof h);
    w.put(g->off.data(), sizeof(int64_t) * n);
    hipStream_t st = b->ctx->stream;
    auto d2h = [&](const void* src, size_t len) -> hipError_t {
        return hipMemcpyAsync(w.reserve(len), src, len, hipMemcpyDeviceToHost, st);
    };
    HIPCHK(d2h(g->d_lw, sizeof(double) * t));
    HIPCHK(d2h(g->d_hist, sizeof(unsigned long long) * t));
    HIPCHK(d2h(g->d_logf, sizeof(double) * n));
    HIPCHK(d2h(g->d_steps, sizeof(unsigned long long) * n));
    HIPCHK(d2h(g->d_acc, sizeof(unsigned long long) * n));
    HIPCHK(d2h(g->d_reached, (size_t)g->rtot));
    return sync_check(b, "mcx_gen_save_state");
}

extern "C" mcx_status mcx_gen_load_state(mcx_gen* g, const void* buf, size_t bytes)
{
    if (!g || !buf)
        return fail(MCX_EINVAL, "mcx_gen_load_state: NULL argument");
    BlobReader r(buf, bytes);
    StateHeader h;
    if (!r.get(&h, sizeof h) || std::memcmp(h.magic, "MCXS", 4) != 0 ||
        h.version != kGenStateVersion || h.kind != 3)
        return fail(MCX_EINVAL, "mcx_gen_load_state: not a walker snapshot of this library");
    mcx_batch* b = g->b;
    if (h.shape[0] != g->kind || h.shape[1] != b->n || h.shape[2] != g->n_groups ||
        h.shape[3] != g->tot || bytes < gen_state_bytes(g))
        return fail(MCX_EINVAL, "mcx_gen_load_state: the snapshot is of another walker layout");
    const size_t n = (size_t)b->n, t = (size_t)g->tot;
    std::vector<int64_t> off(n);
    r.get(off.data(), sizeof(int64_t) * n);
    if (off != g->off)
        return fail(MCX_EINVAL, "mcx_gen_load_state: the snapshot's bin layout differs");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    auto h2d = [&](void* dst, size_t len) -> hipError_t {
        return hipMemcpyAsync(dst, r.take(len), len, hipMemcpyHostToDevice, st);
    };
    HIPCHK(h2d(g->d_lw, sizeof(double) * t));
    HIPCHK(h2d(g->d_hist, sizeof(unsigned long long) * t));
    const uint8_t* lf = r.take(sizeof(double) * n);
    std::memcpy(g->logf.data(), lf, sizeof(double) * n);
    HIPCHK(hipMemcpyAsync(g->d_logf, g->logf.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCHK(h2d(g->d_steps, sizeof(unsigned long long) * n));
    HIPCHK(h2d(g->d_acc, sizeof(unsigned long long) * n));
    HIPCHK(h2d(g->d_reached, (size_t)g->rtot));
    return sync_check(b, "mcx_gen_load_state");
}
