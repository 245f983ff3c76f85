// mcx_api.cpp -- C ABI of libmcx_hip.so: contexts and lattice batches.
//
// Host-side bookkeeping only; every spin update, reduction and random draw
// runs in the HIP kernels (mcx_ising.hip, mcx_bc.hip, mcx_misc.hip).
#include "mcx_api_impl.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>

using namespace mcx;

namespace mcx {

static thread_local std::string g_err;

mcx_status fail(mcx_status code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

mcx_status ok()
{
    return MCX_OK;
}

uint64_t threshold32(double p)
{
    if (!(p > 0))
        return 0;
    if (p >= 1)
        return (uint64_t)1 << 32;
    return (uint64_t)std::ceil(std::ldexp(p, 32));
}

// logistic(x), src/infrastructure/utils.jl:82-88
double logistic(double x)
{
    if (x >= 0)
        return 1.0 / (1.0 + std::exp(-x));
    const double ex = std::exp(x);
    return ex / (1.0 + ex);
}

void fill_ising_metro(IsingTable& t, int dim, const double* p)
{
    t.metro_always = t.metro_never = 0;
    for (int k = 0; k < 4; ++k)
        t.metro_thr[k] = 0;
    for (int k = 0; k < dim; ++k) {
        const uint64_t T = threshold32(p[k]);
        if (T >> 32)
            t.metro_always |= 1u << k;
        else if (T == 0)
            t.metro_never |= 1u << k;
        t.metro_thr[k] = (uint32_t)T;
    }
}

void fill_ising_hb(IsingTable& t, int dim, const double* p)
{
    t.hb_always = t.hb_never = 0;
    for (int k = 0; k < 8; ++k)
        t.hb_thr[k] = 0;
    for (int k = 0; k <= 2 * dim; ++k) {
        const uint64_t T = threshold32(p[k]);
        if (T >> 32)
            t.hb_always |= 1u << k;
        else if (T == 0)
            t.hb_never |= 1u << k;
        t.hb_thr[k] = (uint32_t)T;
    }
}

void fill_ising_gl(IsingTable& t, int dim, const double* p)
{
    t.gl_always = t.gl_never = 0;
    for (int k = 0; k < 8; ++k)
        t.gl_thr[k] = 0;
    for (int k = 0; k <= 2 * dim; ++k) {
        const uint64_t T = threshold32(p[k]);
        if (T >> 32)
            t.gl_always |= 1u << k;
        else if (T == 0)
            t.gl_never |= 1u << k;
        t.gl_thr[k] = (uint32_t)T;
    }
}

// Metropolis: class k (anti-aligned neighbours) has dE = 2(2 dim - 2k) and
// p = exp(logweight(BoltzmannEnsemble(beta), dE)) = exp(-beta * dE)
// (src/ensembles/boltzmann.jl:28).  Heat bath: class k (up neighbours) has
// local field h = 2k - 2 dim, dE = 2 s h and p_plus = logistic((beta * s) * dE)
// (SpinSystems/src/ising.jl:49), which equals logistic(beta * 2h) for s = +-1.
// Glauber: class k (anti-aligned neighbours, 0..2 dim) flips with
// p = logistic(logweight(BoltzmannEnsemble(beta), dE)) = logistic(-beta * dE)
// (src/algorithms/metropolis.jl:121-127).
void ising_probabilities(double beta, int dim, double* metro, double* hb, double* gl)
{
    for (int k = 0; k < dim; ++k) {
        const int dE = 2 * (2 * dim - 2 * k);
        metro[k] = std::exp(-beta * (double)dE);
    }
    for (int k = 0; k <= 2 * dim; ++k) {
        const int h = 2 * k - 2 * dim;
        hb[k] = logistic(beta * (double)(2 * h));
        const int dE = 2 * (2 * dim - 2 * k);
        gl[k] = logistic(-beta * (double)dE);
    }
}

void fill_ising_table(IsingTable& t, int dim, double beta)
{
    std::memset(&t, 0, sizeof t);
    double pm[4], ph[8], pg[8];
    ising_probabilities(beta, dim, pm, ph, pg);
    fill_ising_metro(t, dim, pm);
    fill_ising_hb(t, dim, ph);
    fill_ising_gl(t, dim, pg);
}

} // namespace mcx

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

extern "C" const char* mcx_last_error(void)
{
    return g_err.c_str();
}

extern "C" mcx_status mcx_version(int32_t* major, int32_t* minor)
{
    if (major)
        *major = MCX_VERSION_MAJOR;
    if (minor)
        *minor = MCX_VERSION_MINOR;
    return MCX_OK;
}

extern "C" mcx_status mcx_ctx_create(int32_t device, mcx_ctx** out)
{
    if (!out)
        return fail(MCX_EINVAL, "mcx_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(MCX_EINVAL, "mcx_ctx_create: device %d not in [0, %d)", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    auto* c = new mcx_ctx();
    c->device = device;
    c->n_cu = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlock;
    std::snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev1);
    // the allocation-free entry points' static device array on this device
    if (e == hipSuccess)
        e = tau_scratch_address(device, &c->tau_scratch);
    if (e != hipSuccess) {
        if (c->ev1)
            hipEventDestroy(c->ev1);
        if (c->ev0)
            hipEventDestroy(c->ev0);
        if (c->stream)
            hipStreamDestroy(c->stream);
        delete c;
        return fail(MCX_EHIP, "mcx_ctx_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return MCX_OK;
}

extern "C" mcx_status mcx_ctx_destroy(mcx_ctx* c)
{
    if (!c)
        return MCX_OK;
    if (c->n_users)
        return fail(MCX_ESTATE, "mcx_ctx_destroy: %d batch / server object(s) still use the context",
                    c->n_users);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipStreamDestroy(c->stream);
    delete c;
    return MCX_OK;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

namespace mcx {

size_t lattice_lds_bytes(int model, int dim, int L)
{
    const int half = L / 2;
    const int nw = (half + 31) / 32;
    const int64_t rows = dim == 3 ? (int64_t)L * L : L;
    // Ising: 192-word header (partials, 64-sweep E/M ring) + 2 planes;
    // Blume-Capel: static class tables + kBCLdsHeaderWords-word header
    // (partials, E/M/Q ring) + occupancy and up planes per colour
    if (model == MCX_ISING)
        return 768 + 2 * (size_t)ising_plane_words(dim, L) * 4;
    return (size_t)kBCStaticLds + 4 * (size_t)kBCLdsHeaderWords + 4 * (size_t)rows * (size_t)nw * 4;
}

int lattice_block(int model, int dim, int L)
{
    const int64_t words = model == MCX_ISING ? ising_plane_words(dim, L)
                                             : (dim == 3 ? (int64_t)L * L : L) * ((L / 2 + 31) / 32);
    int64_t b = (words + 63) / 64 * 64;
    if (b < 64)
        b = 64;
    if (b > 256)
        b = 256;
    return (int)b;
}

void batch_free(mcx_batch* b)
{
    hipFree(b->d_spins);
    hipFree(b->d_chains);
    hipFree(b->d_itab);
    hipFree(b->d_btab);
    hipFree(b->d_E);
    hipFree(b->d_M);
    hipFree(b->d_Q);
    hipFree(b->d_mom);
    hipFree(b->d_acc);
    hipFree(b->d_items);
    hipFree(b->d_obs);
    hipFree(b->d_last);
    hipFree(b->d_planes);
    hipFree(b->d_plane_off);
}

// Bit-plane storage of every chain (2 planes per colour word for Ising, 4 for
// Blume-Capel: the LDS layout of the batched kernels), allocated once.
mcx_status ensure_planes(mcx_batch* b)
{
    if (b->d_planes)
        return MCX_OK;
    std::vector<int64_t> off(b->n);
    int64_t total = 0;
    for (int k = 0; k < b->n; ++k) {
        const int64_t L = b->rec[k].L;
        const int64_t nw = (L / 2 + 31) / 32;
        const int64_t rows = b->dim == 3 ? L * L : L;
        off[k] = total;
        total += b->model == MCX_ISING ? 2 * ising_plane_words(b->dim, (int)L) : 4 * rows * nw;
    }
    HIPCHK(hipMalloc((void**)&b->d_plane_off, sizeof(int64_t) * (size_t)b->n));
    hipError_t e = hipMalloc((void**)&b->d_planes, sizeof(uint32_t) * (size_t)std::max<int64_t>(total, 1));
    if (e != hipSuccess) {
        hipFree(b->d_plane_off);
        b->d_plane_off = nullptr;
        b->d_planes = nullptr;
        return fail(MCX_EHIP, "plane storage (%lld words): %s", (long long)total,
                    hipGetErrorString(e));
    }
    HIPCHK(hipMemcpy(b->d_plane_off, off.data(), sizeof(int64_t) * (size_t)b->n,
                     hipMemcpyHostToDevice));
    return MCX_OK;
}

// The device copy of the chain records is authoritative (kernels advance the
// sweep counters, replica exchange rewrites the table indices): refresh the
// host mirror before changing a field and uploading it again.
mcx_status pull_chains(mcx_batch* b)
{
    HIPCHK(hipMemcpyAsync(b->rec.data(), b->d_chains, sizeof(ChainRec) * b->n,
                          hipMemcpyDeviceToHost, b->ctx->stream));
    return sync_check(b, "pull_chains");
}

// Acceptance tables from one beta per table index (the chain's own D).
mcx_status set_chain_tables(mcx_batch* b, const double* beta)
{
    if (b->model == MCX_ISING) {
        std::vector<IsingTable> t(b->n);
        for (int i = 0; i < b->n; ++i)
            fill_ising_table(t[i], b->dim, beta[i]);
        HIPCHK(hipMemcpyAsync(b->d_itab, t.data(), sizeof(IsingTable) * b->n,
                              hipMemcpyHostToDevice, b->ctx->stream));
    } else {
        std::vector<BCTable> t(b->n);
        for (int i = 0; i < b->n; ++i)
            fill_bc_table(t[i], b->dim, beta[i], b->rec[i].D);
        HIPCHK(hipMemcpyAsync(b->d_btab, t.data(), sizeof(BCTable) * b->n, hipMemcpyHostToDevice,
                              b->ctx->stream));
    }
    b->betas.assign(beta, beta + b->n);
    b->tables_ready = true;
    return sync_check(b, "set_chain_tables");
}

mcx_status upload_chains(mcx_batch* b)
{
    HIPCHK(hipMemcpyAsync(b->d_chains, b->rec.data(), sizeof(ChainRec) * b->n,
                          hipMemcpyHostToDevice, b->ctx->stream));
    return MCX_OK;
}

mcx_status sync_check(mcx_batch* b, const char* what)
{
    hipError_t e = hipStreamSynchronize(b->ctx->stream);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e != hipSuccess)
        return fail(MCX_EHIP, "%s: %s", what, hipGetErrorString(e));
    return MCX_OK;
}

} // namespace mcx

extern "C" mcx_status mcx_batch_create(mcx_ctx* ctx, const mcx_batch_desc* d, mcx_batch** out)
{
    if (!ctx || !d || !out)
        return fail(MCX_EINVAL, "mcx_batch_create: NULL argument");
    *out = nullptr;
    if (d->model != MCX_ISING && d->model != MCX_BLUME_CAPEL)
        return fail(MCX_EINVAL, "mcx_batch_create: unknown model %d", d->model);
    if (d->dim != 2 && d->dim != 3)
        return fail(MCX_EINVAL, "mcx_batch_create: dim must be 2 or 3 (got %d)", d->dim);
    if (d->n_chains < 1)
        return fail(MCX_EINVAL, "mcx_batch_create: n_chains must be >= 1 (got %d)", d->n_chains);
    if (d->series_capacity < 0)
        return fail(MCX_EINVAL, "mcx_batch_create: series_capacity must be >= 0");
    HIPCHK(hipSetDevice(ctx->device));

    auto* b = new mcx_batch();
    b->ctx = ctx;
    b->model = d->model;
    b->dim = d->dim;
    b->n = d->n_chains;
    b->seed = d->seed;
    b->cap = d->series_capacity;
    b->rec.resize(b->n);
    b->nsites.resize(b->n);
    b->steps.assign(b->n, 0);
    int64_t off = 0;
    for (int i = 0; i < b->n; ++i) {
        const int L = d->L ? d->L[i] : d->L_uniform;
        if (L < 4 || (L & 1)) {
            delete b;
            return fail(MCX_EINVAL,
                        "mcx_batch_create: chain %d: L = %d; the checkerboard needs an even L >= 4",
                        i, L);
        }
        const size_t lds = lattice_lds_bytes(d->model, d->dim, L);
        if (lds > ctx->lds_max) {
            delete b;
            return fail(MCX_EINVAL,
                        "mcx_batch_create: chain %d: L = %d needs %zu B of LDS (device limit %zu B)",
                        i, L, lds, ctx->lds_max);
        }
        const int64_t N = (int64_t)L * L * (d->dim == 3 ? L : 1);
        ChainRec& r = b->rec[i];
        r.spin_off = off;
        r.series_off = (int64_t)i * b->cap;
        r.sweep = 0;
        r.L = L;
        r.dim = d->dim;
        r.table = i;
        r.key = d->chain_keys ? d->chain_keys[i] : (uint32_t)i;
        r.D = d->D ? d->D[i] : d->D_uniform;
        r.seed = d->seed;
        if (d->model == MCX_ISING)
            r.D = 0.0;
        b->nsites[i] = N;
        b->max_sites = std::max(b->max_sites, N);
        off += (N + 255) / 256 * 256;
    }
    b->pool_bytes = off;

    // launch plan: chains grouped by workgroup size (and, for Ising, plane
    // layout), largest lattices first
    std::vector<int32_t> order(b->n);
    for (int i = 0; i < b->n; ++i)
        order[i] = i;
    auto flat_of = [&](int i) { return b->model == MCX_ISING ? ising_layout(b->rec[i].L) : 0; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        const int ba = lattice_block(b->model, b->dim, b->rec[a].L);
        const int bc = lattice_block(b->model, b->dim, b->rec[c].L);
        if (ba != bc)
            return ba > bc;
        if (flat_of(a) != flat_of(c))
            return flat_of(a) < flat_of(c);
        return b->nsites[a] > b->nsites[c];
    });
    for (int k = 0; k < b->n;) {
        const int blk = lattice_block(b->model, b->dim, b->rec[order[k]].L);
        const int fl = flat_of(order[k]);
        mcx_batch::Launch ln{k, 0, blk, 0, fl, 0};
        while (k < b->n && lattice_block(b->model, b->dim, b->rec[order[k]].L) == blk &&
               flat_of(order[k]) == fl) {
            ln.lds = std::max(ln.lds, lattice_lds_bytes(b->model, b->dim, b->rec[order[k]].L));
            if (b->model == MCX_ISING)
                ln.words = std::max(ln.words, ising_plane_words(b->dim, b->rec[order[k]].L));
            ++ln.count;
            ++k;
        }
        b->launches.push_back(ln);
    }

    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess && bytes)
            e = hipMalloc(p, bytes);
    };
    alloc((void**)&b->d_spins, (size_t)b->pool_bytes);
    alloc((void**)&b->d_chains, sizeof(ChainRec) * b->n);
    alloc((void**)&b->d_itab, sizeof(IsingTable) * b->n);
    alloc((void**)&b->d_btab, sizeof(BCTable) * b->n);
    alloc((void**)&b->d_mom, sizeof(double) * 6 * b->n);
    alloc((void**)&b->d_acc, sizeof(unsigned long long) * b->n);
    alloc((void**)&b->d_items, sizeof(int32_t) * b->n);
    alloc((void**)&b->d_obs, sizeof(int32_t) * 3 * b->n);
    alloc((void**)&b->d_last, sizeof(int32_t) * 3 * b->n);
    if (b->cap > 0) {
        const size_t ser = sizeof(int32_t) * (size_t)b->cap * (size_t)b->n;
        alloc((void**)&b->d_E, ser);
        alloc((void**)&b->d_M, ser);
        if (b->model == MCX_BLUME_CAPEL)
            alloc((void**)&b->d_Q, ser);
    }
    if (e != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(e == hipErrorOutOfMemory ? MCX_ENOMEM : MCX_EHIP,
                    "mcx_batch_create: device allocation failed: %s", hipGetErrorString(e));
    }
    hipStream_t st = ctx->stream;
    e = hipMemsetAsync(b->d_spins, 0x01, (size_t)b->pool_bytes, st); // all up: ones(Int8, N)
    if (e == hipSuccess)
        e = hipMemsetAsync(b->d_mom, 0, sizeof(double) * 6 * b->n, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(b->d_acc, 0, sizeof(unsigned long long) * b->n, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(b->d_itab, 0, sizeof(IsingTable) * b->n, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(b->d_btab, 0, sizeof(BCTable) * b->n, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_items, order.data(), sizeof(int32_t) * b->n,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_chains, b->rec.data(), sizeof(ChainRec) * b->n,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(MCX_EHIP, "mcx_batch_create: %s", hipGetErrorString(e));
    }
    ++ctx->n_users;
    *out = b;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_destroy(mcx_batch* b)
{
    if (!b)
        return MCX_OK;
    if (b->pt)
        return fail(MCX_ESTATE, "mcx_batch_destroy: destroy the batch's replica-exchange object first");
    if (b->n_gen)
        return fail(MCX_ESTATE, "mcx_batch_destroy: destroy the batch's %d walker object(s) first",
                    b->n_gen);
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    --b->ctx->n_users;
    batch_free(b);
    delete b;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_num_sites(mcx_batch* b, int32_t chain, int64_t* n)
{
    if (!b || !n)
        return fail(MCX_EINVAL, "mcx_batch_num_sites: NULL argument");
    if (chain < 0 || chain >= b->n)
        return fail(MCX_EBOUNDS, "mcx_batch_num_sites: chain %d not in [0, %d)", chain, b->n);
    *n = b->nsites[chain];
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_table_layout(mcx_batch* b, int32_t alg, int32_t* n)
{
    if (!b || !n)
        return fail(MCX_EINVAL, "mcx_batch_table_layout: NULL argument");
    if (b->model == MCX_ISING) {
        if (alg == MCX_METROPOLIS)
            *n = b->dim;
        else if (alg == MCX_HEATBATH || alg == MCX_GLAUBER)
            *n = 2 * b->dim + 1;
        else
            return fail(MCX_EINVAL, "mcx_batch_table_layout: alg %d has no table", alg);
    } else {
        if (alg == MCX_METROPOLIS || alg == MCX_GLAUBER)
            *n = 6 * (4 * b->dim + 1);
        else if (alg == MCX_HEATBATH)
            *n = 3 * (4 * b->dim + 1);
        else
            return fail(MCX_EINVAL, "mcx_batch_table_layout: alg %d has no table", alg);
    }
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_set_betas(mcx_batch* b, const double* beta)
{
    if (!b || !beta)
        return fail(MCX_EINVAL, "mcx_batch_set_betas: NULL argument");
    if (b->pt)
        return fail(MCX_ESTATE, "mcx_batch_set_betas: the batch's temperatures belong to its "
                                "replica-exchange object (destroy it first)");
    for (int i = 0; i < b->n; ++i)
        if (!(beta[i] >= 0) || !std::isfinite(beta[i]))
            return fail(MCX_EINVAL, "mcx_batch_set_betas: chain %d: beta = %g must be finite and >= 0",
                        i, beta[i]);
    HIPCHK(hipSetDevice(b->ctx->device));
    mcx_status s = pull_chains(b);
    if (s)
        return s;
    for (int i = 0; i < b->n; ++i)
        b->rec[i].table = i;
    s = upload_chains(b);
    if (s)
        return s;
    return set_chain_tables(b, beta);
}

extern "C" mcx_status mcx_batch_set_tables(mcx_batch* b, int32_t alg, const double* p, int32_t n)
{
    if (!b || !p)
        return fail(MCX_EINVAL, "mcx_batch_set_tables: NULL argument");
    if (b->pt)
        return fail(MCX_ESTATE, "mcx_batch_set_tables: the batch's temperatures belong to its "
                                "replica-exchange object (destroy it first)");
    int32_t want = 0;
    mcx_status st = mcx_batch_table_layout(b, alg, &want);
    if (st)
        return st;
    if (n != want)
        return fail(MCX_EINVAL, "mcx_batch_set_tables: expected %d probabilities per chain, got %d",
                    want, n);
    for (int64_t k = 0; k < (int64_t)n * b->n; ++k)
        if (!(p[k] >= 0) || !std::isfinite(p[k]))
            return fail(MCX_EDOMAIN, "mcx_batch_set_tables: entry %lld = %g is not a finite value >= 0",
                        (long long)k, p[k]);
    HIPCHK(hipSetDevice(b->ctx->device));
    if (b->model == MCX_ISING) {
        std::vector<IsingTable> t(b->n);
        HIPCHK(hipMemcpy(t.data(), b->d_itab, sizeof(IsingTable) * b->n, hipMemcpyDeviceToHost));
        for (int i = 0; i < b->n; ++i) {
            if (alg == MCX_METROPOLIS)
                fill_ising_metro(t[i], b->dim, p + (int64_t)i * n);
            else if (alg == MCX_GLAUBER)
                fill_ising_gl(t[i], b->dim, p + (int64_t)i * n);
            else
                fill_ising_hb(t[i], b->dim, p + (int64_t)i * n);
        }
        HIPCHK(hipMemcpyAsync(b->d_itab, t.data(), sizeof(IsingTable) * b->n,
                              hipMemcpyHostToDevice, b->ctx->stream));
    } else {
        std::vector<BCTable> t(b->n);
        HIPCHK(hipMemcpy(t.data(), b->d_btab, sizeof(BCTable) * b->n, hipMemcpyDeviceToHost));
        for (int i = 0; i < b->n; ++i) {
            if (alg == MCX_METROPOLIS)
                fill_bc_metro_from_p(t[i], b->dim, p + (int64_t)i * n);
            else if (alg == MCX_GLAUBER)
                fill_bc_gl_from_p(t[i], b->dim, p + (int64_t)i * n);
            else
                fill_bc_hb_from_p(t[i], b->dim, p + (int64_t)i * n);
        }
        HIPCHK(hipMemcpyAsync(b->d_btab, t.data(), sizeof(BCTable) * b->n, hipMemcpyHostToDevice,
                              b->ctx->stream));
    }
    b->tables_ready = true;
    return sync_check(b, "mcx_batch_set_tables");
}

extern "C" mcx_status mcx_batch_init_spins(mcx_batch* b, int32_t mode)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_init_spins: NULL batch");
    if (mode < MCX_INIT_UP || mode > MCX_INIT_RANDOM)
        return fail(MCX_EINVAL, "Unknown initialization type: %d", mode);
    if (mode == MCX_INIT_ZERO && b->model == MCX_ISING)
        return fail(MCX_EINVAL, "Unknown initialization type: zero (Ising spins are +-1)");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(launch_init_spins(b->model, mode, b->d_chains, b->n, b->max_sites, b->d_spins,
                             b->ctx->stream));
    return sync_check(b, "mcx_batch_init_spins");
}

extern "C" mcx_status mcx_batch_set_spins(mcx_batch* b, int32_t chain, const int8_t* s, size_t n)
{
    if (!b || !s)
        return fail(MCX_EINVAL, "mcx_batch_set_spins: NULL argument");
    if (chain < 0 || chain >= b->n)
        return fail(MCX_EBOUNDS, "mcx_batch_set_spins: chain %d not in [0, %d)", chain, b->n);
    if ((int64_t)n != b->nsites[chain])
        return fail(MCX_EINVAL, "mcx_batch_set_spins: chain %d has %lld sites, got %zu", chain,
                    (long long)b->nsites[chain], n);
    for (size_t i = 0; i < n; ++i) {
        const int v = s[i];
        const bool good = b->model == MCX_ISING ? (v == 1 || v == -1) : (v >= -1 && v <= 1);
        if (!good)
            return fail(MCX_EDOMAIN, "mcx_batch_set_spins: site %zu has spin %d", i, v);
    }
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpyAsync(b->d_spins + b->rec[chain].spin_off, s, n, hipMemcpyHostToDevice,
                          b->ctx->stream));
    return sync_check(b, "mcx_batch_set_spins");
}

extern "C" mcx_status mcx_batch_get_spins(mcx_batch* b, int32_t chain, int8_t* s, size_t n)
{
    if (!b || !s)
        return fail(MCX_EINVAL, "mcx_batch_get_spins: NULL argument");
    if (chain < 0 || chain >= b->n)
        return fail(MCX_EBOUNDS, "mcx_batch_get_spins: chain %d not in [0, %d)", chain, b->n);
    if ((int64_t)n != b->nsites[chain])
        return fail(MCX_EINVAL, "mcx_batch_get_spins: chain %d has %lld sites, got %zu", chain,
                    (long long)b->nsites[chain], n);
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpyAsync(s, b->d_spins + b->rec[chain].spin_off, n, hipMemcpyDeviceToHost,
                          b->ctx->stream));
    return sync_check(b, "mcx_batch_get_spins");
}

namespace mcx {

mcx_status run_sweeps_async(mcx_batch* b, int32_t alg, int64_t nsweeps, uint32_t flags,
                            bool want_last, int planes)
{
    const bool record = (flags & MCX_RUN_RECORD) != 0;
    hipStream_t st = b->ctx->stream;
    if (flags & MCX_RUN_PLAIN) {
        if (record)
            HIPCHK(launch_zero_series(b->d_chains, b->n, b->d_E, b->d_M, b->d_Q, b->rec_len,
                                      nsweeps, st));
        PlainArgs a{};
        a.chains = b->d_chains;
        a.itab = b->d_itab;
        a.btab = b->d_btab;
        a.spins = b->d_spins;
        a.ser_E = b->d_E;
        a.ser_M = b->d_M;
        a.ser_Q = b->d_Q;
        a.accepted = b->d_acc;
        for (int64_t s = 0; s < nsweeps; ++s) {
            a.rec_index = record ? b->rec_len + s : -1;
            for (int ch = 0; ch < b->n; ++ch) {
                const ChainRec& r = b->rec[ch];
                a.chain = ch;
                a.sweep = r.sweep + (uint64_t)s;
                a.spin_off = r.spin_off;
                a.L = r.L;
                const int n_col = (int)(b->nsites[ch] / 2);
                for (int c = 0; c < 2; ++c) {
                    a.colour = c;
                    if (b->model == MCX_ISING)
                        HIPCHK(launch_ising_plain(b->dim, alg, a, n_col, st));
                    else
                        HIPCHK(launch_bc_plain(b->dim, alg, a, n_col, st));
                }
            }
        }
        HIPCHK(launch_advance_sweeps(b->d_chains, b->n, nsweeps, st));
        if (record)
            HIPCHK(launch_series_moments(b->d_chains, b->n, b->d_E, b->d_M, b->d_Q, b->rec_len,
                                         nsweeps, b->d_mom, st));
        if (want_last)
            HIPCHK(launch_observables(b->d_chains, b->n, b->max_sites, b->d_spins, b->d_last,
                                      b->d_last + b->n, b->d_last + 2 * b->n, st));
    } else {
        SweepArgs a{};
        a.record = record ? 1 : 0;
        a.chains = b->d_chains;
        a.itab = b->d_itab;
        a.btab = b->d_btab;
        a.spins = b->d_spins;
        a.ser_E = b->d_E;
        a.ser_M = b->d_M;
        a.ser_Q = b->d_Q;
        a.moments = b->d_mom;
        a.accepted = b->d_acc;
        a.nsweeps = nsweeps;
        a.rec_pos = b->rec_len;
        a.by_table = b->by_table ? 1 : 0;
        a.cap = b->cap;
        if (planes && b->d_planes) {
            a.planes = b->d_planes;
            a.plane_off = b->d_plane_off;
            a.planes_in = planes & 1;
            a.planes_out = (planes >> 1) & 1;
        }
        if (want_last && nsweeps > 0) {
            a.last_E = b->d_last;
            a.last_M = b->d_last + b->n;
            a.last_Q = b->d_last + 2 * b->n;
        }
        // the tuned launch shape of this algorithm, if any (results do not depend on it)
        const mcx_launch_shape& sh = b->shape[alg];
        a.lanes_per_word = sh.lanes_per_word;
        for (const auto& ln : b->launches) {
            a.items = b->d_items + ln.first;
            a.n_items = ln.count;
            const int block = sh.threads > 0 ? sh.threads : ln.block;
            const int cpg = sh.chains_per_group > 1 ? sh.chains_per_group : 1;
            const int grid = (ln.count + cpg - 1) / cpg;
            if (b->model == MCX_ISING) {
                const int lay = ising_launch_layout(ln.layout, ln.count, block);
                HIPCHK(launch_ising_sweeps(b->dim, alg, lay, a,
                                           ising_launch_block(lay, block, ln.words), grid, ln.lds,
                                           st));
            } else {
                HIPCHK(launch_bc_sweeps(b->dim, alg, a, block, grid, ln.lds, st));
            }
        }
    }
    for (int ch = 0; ch < b->n; ++ch) {
        b->rec[ch].sweep += (uint64_t)nsweeps;
        b->steps[ch] += b->nsites[ch] * nsweeps;
    }
    if (record)
        b->rec_len += nsweeps;
    return MCX_OK;
}

mcx_status check_run(mcx_batch* b, int32_t alg, int64_t nsweeps, uint32_t flags, const char* who)
{
    if (!b)
        return fail(MCX_EINVAL, "%s: NULL batch", who);
    if (alg != MCX_METROPOLIS && alg != MCX_HEATBATH && alg != MCX_GLAUBER)
        return fail(MCX_EINVAL, "%s: alg must be MCX_METROPOLIS, MCX_HEATBATH or MCX_GLAUBER (got %d)",
                    who, alg);
    if (nsweeps < 0)
        return fail(MCX_EINVAL, "%s: nsweeps must be >= 0", who);
    if (flags & ~(MCX_RUN_RECORD | MCX_RUN_PLAIN))
        return fail(MCX_EINVAL, "%s: unknown flags 0x%x", who, flags);
    if (!b->tables_ready)
        return fail(MCX_ESTATE, "%s: no temperatures set (call mcx_batch_set_betas first)", who);
    if ((flags & MCX_RUN_RECORD) && b->rec_len + nsweeps > b->cap)
        return fail(MCX_EINVAL, "%s: recording %lld sweeps exceeds the series capacity (%lld of %lld used)",
                    who, (long long)nsweeps, (long long)b->rec_len, (long long)b->cap);
    if ((flags & MCX_RUN_RECORD) && (flags & MCX_RUN_PLAIN) && b->by_table)
        return fail(MCX_ESTATE, "%s: the plain path records in chain order; this batch records in "
                                "temperature order (mcx_pt_set_series_order)", who);
    return MCX_OK;
}

} // namespace mcx

extern "C" mcx_status mcx_batch_run(mcx_batch* b, int32_t alg, int64_t nsweeps, uint32_t flags)
{
    mcx_status st = check_run(b, alg, nsweeps, flags, "mcx_batch_run");
    if (st)
        return st;
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipEventRecord(b->ctx->ev0, b->ctx->stream));
    st = run_sweeps_async(b, alg, nsweeps, flags);
    if (st)
        return st;
    HIPCHK(hipEventRecord(b->ctx->ev1, b->ctx->stream));
    st = sync_check(b, "mcx_batch_run");
    if (st)
        return st;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
    b->last_ms = ms;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_set_launch_shape(mcx_batch* b, int32_t alg,
                                                 const mcx_launch_shape* s)
{
    if (!b || !s)
        return fail(MCX_EINVAL, "mcx_batch_set_launch_shape: NULL argument");
    if (alg >= 0 && alg != MCX_METROPOLIS && alg != MCX_HEATBATH && alg != MCX_GLAUBER)
        return fail(MCX_EINVAL, "mcx_batch_set_launch_shape: alg must be MCX_METROPOLIS, "
                                "MCX_HEATBATH, MCX_GLAUBER or < 0 (got %d)", alg);
    if (s->threads < 0 || s->threads > 1024 || s->threads % 64 != 0)
        return fail(MCX_EINVAL, "mcx_batch_set_launch_shape: threads must be 0 or a multiple of 64 "
                                "up to 1024 (got %d)", s->threads);
    if (s->chains_per_group < 0)
        return fail(MCX_EINVAL, "mcx_batch_set_launch_shape: chains_per_group must be >= 0");
    const int l = s->lanes_per_word;
    if (!(l == -1 || l == 0 || l == 1 || l == 2 || l == 4 || l == 8))
        return fail(MCX_EINVAL, "mcx_batch_set_launch_shape: lanes_per_word must be -1, 0, 1, 2, 4 "
                                "or 8 (got %d)", l);
    if (l != 0 && b->model == MCX_ISING)
        return fail(MCX_EINVAL, "mcx_batch_set_launch_shape: lanes_per_word is a Blume-Capel "
                                "setting");
    for (int k = 0; k < 5; ++k)
        if (alg < 0 ? (k == MCX_METROPOLIS || k == MCX_HEATBATH || k == MCX_GLAUBER) : k == alg)
            b->shape[k] = *s;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_get_launch_shape(mcx_batch* b, int32_t alg, mcx_launch_shape* s)
{
    if (!b || !s)
        return fail(MCX_EINVAL, "mcx_batch_get_launch_shape: NULL argument");
    if (alg != MCX_METROPOLIS && alg != MCX_HEATBATH && alg != MCX_GLAUBER)
        return fail(MCX_EINVAL, "mcx_batch_get_launch_shape: unknown alg %d", alg);
    *s = b->shape[alg];
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_last_run_ms(mcx_batch* b, double* ms)
{
    if (!b || !ms)
        return fail(MCX_EINVAL, "mcx_batch_last_run_ms: NULL argument");
    *ms = b->last_ms;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_get_observables(mcx_batch* b, int32_t* E, int32_t* M, int32_t* Q)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_get_observables: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    HIPCHK(launch_observables(b->d_chains, b->n, b->max_sites, b->d_spins, b->d_obs,
                              b->d_obs + b->n, b->d_obs + 2 * b->n, st));
    std::vector<int32_t> h(3 * (size_t)b->n);
    HIPCHK(hipMemcpyAsync(h.data(), b->d_obs, sizeof(int32_t) * 3 * b->n, hipMemcpyDeviceToHost, st));
    mcx_status s = sync_check(b, "mcx_batch_get_observables");
    if (s)
        return s;
    if (E)
        std::memcpy(E, h.data(), sizeof(int32_t) * b->n);
    if (M)
        std::memcpy(M, h.data() + b->n, sizeof(int32_t) * b->n);
    if (Q)
        std::memcpy(Q, h.data() + 2 * b->n, sizeof(int32_t) * b->n);
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_series_length(mcx_batch* b, int64_t* n)
{
    if (!b || !n)
        return fail(MCX_EINVAL, "mcx_batch_series_length: NULL argument");
    *n = b->rec_len;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_get_series(mcx_batch* b, int32_t chain, int64_t first, int64_t count,
                                           int32_t* E, int32_t* M, int32_t* Q)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_get_series: NULL batch");
    if (chain < 0 || chain >= b->n)
        return fail(MCX_EBOUNDS, "mcx_batch_get_series: chain %d not in [0, %d)", chain, b->n);
    if (first < 0 || count < 0 || first + count > b->rec_len)
        return fail(MCX_EBOUNDS, "mcx_batch_get_series: [%lld, %lld) not within the %lld recorded sweeps",
                    (long long)first, (long long)(first + count), (long long)b->rec_len);
    if (Q && !b->d_Q)
        return fail(MCX_EINVAL, "mcx_batch_get_series: Q is recorded for Blume-Capel batches only");
    if (count == 0)
        return MCX_OK;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const int64_t off = b->rec[chain].series_off + first;
    if (E)
        HIPCHK(hipMemcpyAsync(E, b->d_E + off, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st));
    if (M)
        HIPCHK(hipMemcpyAsync(M, b->d_M + off, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st));
    if (Q)
        HIPCHK(hipMemcpyAsync(Q, b->d_Q + off, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st));
    return sync_check(b, "mcx_batch_get_series");
}

extern "C" mcx_status mcx_batch_reset_series(mcx_batch* b)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_reset_series: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    b->rec_len = 0;
    HIPCHK(hipMemsetAsync(b->d_mom, 0, sizeof(double) * 6 * b->n, b->ctx->stream));
    return sync_check(b, "mcx_batch_reset_series");
}

extern "C" mcx_status mcx_batch_get_moments(mcx_batch* b, double* out)
{
    if (!b || !out)
        return fail(MCX_EINVAL, "mcx_batch_get_moments: NULL argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpyAsync(out, b->d_mom, sizeof(double) * 6 * b->n, hipMemcpyDeviceToHost,
                          b->ctx->stream));
    return sync_check(b, "mcx_batch_get_moments");
}

extern "C" mcx_status mcx_batch_get_counters(mcx_batch* b, int64_t* steps, int64_t* accepted)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_get_counters: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    if (accepted) {
        std::vector<unsigned long long> h(b->n);
        HIPCHK(hipMemcpyAsync(h.data(), b->d_acc, sizeof(unsigned long long) * b->n,
                              hipMemcpyDeviceToHost, b->ctx->stream));
        mcx_status s = sync_check(b, "mcx_batch_get_counters");
        if (s)
            return s;
        for (int i = 0; i < b->n; ++i)
            accepted[i] = (int64_t)h[i];
    }
    if (steps)
        std::memcpy(steps, b->steps.data(), sizeof(int64_t) * b->n);
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_reset_counters(mcx_batch* b)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_reset_counters: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    std::fill(b->steps.begin(), b->steps.end(), 0);
    HIPCHK(hipMemsetAsync(b->d_acc, 0, sizeof(unsigned long long) * b->n, b->ctx->stream));
    return sync_check(b, "mcx_batch_reset_counters");
}

extern "C" mcx_status mcx_batch_get_sweeps(mcx_batch* b, uint64_t* sweeps)
{
    if (!b || !sweeps)
        return fail(MCX_EINVAL, "mcx_batch_get_sweeps: NULL argument");
    for (int i = 0; i < b->n; ++i)
        sweeps[i] = b->rec[i].sweep;
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_set_sweeps(mcx_batch* b, const uint64_t* sweeps)
{
    if (!b || !sweeps)
        return fail(MCX_EINVAL, "mcx_batch_set_sweeps: NULL argument");
    for (int i = 0; i < b->n; ++i)
        if (sweeps[i] >> 48)
            return fail(MCX_EINVAL, "mcx_batch_set_sweeps: chain %d: sweep %llu exceeds 2^48", i,
                        (unsigned long long)sweeps[i]);
    HIPCHK(hipSetDevice(b->ctx->device));
    mcx_status s = pull_chains(b);
    if (s)
        return s;
    for (int i = 0; i < b->n; ++i)
        b->rec[i].sweep = sweeps[i];
    s = upload_chains(b);
    if (s)
        return s;
    return sync_check(b, "mcx_batch_set_sweeps");
}

extern "C" mcx_status mcx_batch_set_streams(mcx_batch* b, uint64_t seed, const uint32_t* keys)
{
    if (!b)
        return fail(MCX_EINVAL, "mcx_batch_set_streams: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    mcx_status s = pull_chains(b);
    if (s)
        return s;
    b->seed = seed;
    for (int i = 0; i < b->n; ++i) {
        b->rec[i].key = keys ? keys[i] : (uint32_t)i;
        b->rec[i].seed = seed;
    }
    s = upload_chains(b);
    if (s)
        return s;
    return sync_check(b, "mcx_batch_set_streams");
}

// ---------------------------------------------------------------------------
// integrated autocorrelation times of the recorded series
// (src/measurements/autocorrelations.jl:28-110)
// ---------------------------------------------------------------------------

namespace {

static_assert(mcx::kTauScratchSlots == MCX_TAU_NOALLOC_MAX_SLOTS,
              "the static tau array and the public limit must agree");

// Shared body of mcx_batch_tau_int / _noalloc: the taus are staged in d_out
// (n_chains doubles on the batch's device; nullptr = allocate one here).
mcx_status batch_tau_int(mcx_batch* b, int32_t observable, int64_t first, int64_t count,
                         int32_t min_points, int32_t max_lag, double c, double* taus,
                         double* d_out, const char* who)
{
    if (!b || !taus)
        return fail(MCX_EINVAL, "%s: NULL argument", who);
    if (observable < MCX_OBS_ENERGY || observable > MCX_OBS_Q)
        return fail(MCX_EINVAL, "%s: unknown observable %d", who, observable);
    if (observable == MCX_OBS_Q && b->model != MCX_BLUME_CAPEL)
        return fail(MCX_EINVAL, "%s: MCX_OBS_Q needs a Blume-Capel batch", who);
    if (min_points < 2)
        return fail(MCX_EINVAL, "min_points must be >= 2");
    if (!(c > 0))
        return fail(MCX_EINVAL, "c must be positive");
    if (first < 0 || count < 0 || first + count > b->rec_len)
        return fail(MCX_EBOUNDS, "%s: sweeps [%lld, %lld) outside the %lld recorded", who,
                    (long long)first, (long long)(first + count), (long long)b->rec_len);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const int64_t cap = count / 2;
    const int64_t lag_cap = max_lag < 0 ? cap : std::min<int64_t>(max_lag, cap);
    if (count < min_points || lag_cap < 1) {
        for (int i = 0; i < b->n; ++i)
            taus[i] = nan;
        return MCX_OK;
    }
    // the 128-bit lag sums hold n^3 * max|x|^2 (|E_pair| <= dim N, |M|, Q <= N)
    int64_t max_sites = 0;
    for (int i = 0; i < b->n; ++i)
        max_sites = std::max(max_sites, b->nsites[i]);
    double vmax = (double)max_sites * (observable == MCX_OBS_ENERGY ? (double)b->dim : 1.0);
    if (observable == MCX_OBS_ENERGY && b->model == MCX_BLUME_CAPEL)
        vmax = std::max(vmax, (double)max_sites);
    if (3.0 * std::log2((double)count) + 2.0 * std::log2(vmax + 1.0) > 124.0)
        return fail(MCX_EINVAL, "%s: %lld samples of lattices of %lld sites exceed "
                                "the exact 128-bit lag sums; use a shorter window", who,
                    (long long)count, (long long)max_sites);
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    double* d_tau = d_out;
    if (!d_tau)
        HIPCHK(hipMalloc((void**)&d_tau, sizeof(double) * b->n));
    TauArgs a{};
    a.X = observable == MCX_OBS_ENERGY ? b->d_E : observable == MCX_OBS_Q ? b->d_Q : b->d_M;
    a.Y = (observable == MCX_OBS_ENERGY && b->model == MCX_BLUME_CAPEL) ? b->d_Q : nullptr;
    a.chains = b->d_chains;
    a.cap = b->cap;
    a.first = first;
    a.n = count;
    a.absx = observable == MCX_OBS_ABS_M ? 1 : 0;
    a.lag_cap = (int32_t)lag_cap;
    a.c = c;
    a.n_traces = b->n;
    a.taus = d_tau;
    hipError_t e = launch_tau_int(a, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(taus, d_tau, sizeof(double) * b->n, hipMemcpyDeviceToHost, st);
    mcx_status s = e == hipSuccess ? sync_check(b, who)
                                   : fail(MCX_EHIP, "%s: %s", who, hipGetErrorString(e));
    if (!d_out)
        hipFree(d_tau);
    return s;
}

} // namespace

extern "C" mcx_status mcx_batch_tau_int(mcx_batch* b, int32_t observable, int64_t first,
                                        int64_t count, int32_t min_points, int32_t max_lag,
                                        double c, double* taus)
{
    return batch_tau_int(b, observable, first, count, min_points, max_lag, c, taus, nullptr,
                         "mcx_batch_tau_int");
}

extern "C" mcx_status mcx_batch_tau_int_noalloc(mcx_batch* b, int32_t observable, int64_t first,
                                                int64_t count, int32_t min_points,
                                                int32_t max_lag, double c, double* taus)
{
    const char* who = "mcx_batch_tau_int_noalloc";
    if (!b || !taus)
        return fail(MCX_EINVAL, "%s: NULL argument", who);
    if (b->n > MCX_TAU_NOALLOC_MAX_SLOTS)
        return fail(MCX_EINVAL, "%s: %d series exceed the static limit of %d; use "
                                "mcx_batch_tau_int", who, b->n, MCX_TAU_NOALLOC_MAX_SLOTS);
    if (!b->ctx->tau_scratch)
        return fail(MCX_ESTATE, "%s: the context has no static tau array", who);
    return batch_tau_int(b, observable, first, count, min_points, max_lag, c, taus,
                         b->ctx->tau_scratch, who);
}
