// mcx_state.cpp -- device-state snapshots of a lattice batch, the device side
// of checkpoint / resume (src/infrastructure/checkpointing.jl:48-111 keeps the
// whole Julia objects; here the batch's device arrays are written to and read
// from a caller buffer).  A snapshot restores into a batch of the same shape
// (model, dim, chain sizes); because the streams are counter-based, a run of
// n sweeps, a snapshot and m more sweeps equals a run of n + m sweeps.
#include "mcx_api_impl.h"

#include <cmath>

using namespace mcx;

namespace {

// 2: Blume-Capel tables hold integer heat-bath thresholds, and the Blume-Capel
// site uniforms come from 16-bit fields (a version-1 snapshot would resume on
// other random words)
constexpr uint32_t kBatchStateVersion = 2;

size_t batch_state_bytes(const mcx_batch* b, uint32_t flags)
{
    const size_t n = (size_t)b->n;
    size_t bytes = sizeof(StateHeader) + sizeof(ChainRec) * n + sizeof(int64_t) * n // steps
                   + sizeof(double) * n                                            // betas
                   + sizeof(unsigned long long) * n                                // accepted
                   + sizeof(double) * 6 * n                                        // moments
                   + sizeof(int32_t) * 3 * n                                       // last E/M/Q
                   + (b->model == MCX_ISING ? sizeof(IsingTable) : sizeof(BCTable)) * n
                   + (size_t)b->pool_bytes;
    if (flags & MCX_STATE_SERIES)
        bytes += sizeof(int32_t) * (size_t)b->rec_len * n * (b->d_Q ? 3 : 2);
    return bytes;
}

} // namespace

extern "C" mcx_status mcx_batch_state_size(mcx_batch* b, uint32_t flags, size_t* bytes)
{
    if (!b || !bytes)
        return fail(MCX_EINVAL, "mcx_batch_state_size: NULL argument");
    if (flags & ~MCX_STATE_SERIES)
        return fail(MCX_EINVAL, "mcx_batch_state_size: unknown flags 0x%x", flags);
    *bytes = batch_state_bytes(b, flags);
    return MCX_OK;
}

extern "C" mcx_status mcx_batch_save_state(mcx_batch* b, uint32_t flags, void* buf, size_t bytes)
{
    if (!b || !buf)
        return fail(MCX_EINVAL, "mcx_batch_save_state: NULL argument");
    if (flags & ~MCX_STATE_SERIES)
        return fail(MCX_EINVAL, "mcx_batch_save_state: unknown flags 0x%x", flags);
    const size_t need = batch_state_bytes(b, flags);
    if (bytes < need)
        return fail(MCX_EINVAL, "mcx_batch_save_state: buffer of %zu bytes, %zu needed", bytes, need);
    HIPCHK(hipSetDevice(b->ctx->device));
    mcx_status s = pull_chains(b);
    if (s)
        return s;
    const size_t n = (size_t)b->n;
    StateHeader h = make_state_header(1, flags);
    h.version = kBatchStateVersion;
    h.shape[0] = b->model;
    h.shape[1] = b->dim;
    h.shape[2] = b->n;
    h.shape[3] = b->pool_bytes;
    h.shape[4] = b->cap;
    h.shape[5] = b->rec_len;
    h.shape[6] = b->tables_ready ? 1 : 0;
    h.shape[7] = b->by_table ? 1 : 0;
    BlobWriter w(buf, bytes);
    std::vector<double> betas(n, 0.0);
    for (size_t i = 0; i < n && i < b->betas.size(); ++i)
        betas[i] = b->betas[i];
    w.put(&h, sizeof h);
    w.put(b->rec.data(), sizeof(ChainRec) * n);
    w.put(b->steps.data(), sizeof(int64_t) * n);
    w.put(betas.data(), sizeof(double) * n);
    hipStream_t st = b->ctx->stream;
    auto d2h = [&](const void* src, size_t len) -> hipError_t {
        return hipMemcpyAsync(w.reserve(len), src, len, hipMemcpyDeviceToHost, st);
    };
    HIPCHK(d2h(b->d_acc, sizeof(unsigned long long) * n));
    HIPCHK(d2h(b->d_mom, sizeof(double) * 6 * n));
    HIPCHK(d2h(b->d_last, sizeof(int32_t) * 3 * n));
    if (b->model == MCX_ISING)
        HIPCHK(d2h(b->d_itab, sizeof(IsingTable) * n));
    else
        HIPCHK(d2h(b->d_btab, sizeof(BCTable) * n));
    HIPCHK(d2h(b->d_spins, (size_t)b->pool_bytes));
    if ((flags & MCX_STATE_SERIES) && b->rec_len > 0) {
        int32_t* series[3] = {b->d_E, b->d_M, b->d_Q};
        for (int q = 0; q < 3; ++q) {
            if (!series[q])
                continue;
            for (size_t i = 0; i < n; ++i)
                HIPCHK(d2h(series[q] + b->rec[i].series_off, sizeof(int32_t) * (size_t)b->rec_len));
        }
    }
    return sync_check(b, "mcx_batch_save_state");
}

extern "C" mcx_status mcx_batch_load_state(mcx_batch* b, const void* buf, size_t bytes)
{
    if (!b || !buf)
        return fail(MCX_EINVAL, "mcx_batch_load_state: NULL argument");
    BlobReader r(buf, bytes);
    StateHeader h;
    if (!r.get(&h, sizeof h) || std::memcmp(h.magic, "MCXS", 4) != 0 || h.version != kBatchStateVersion || h.kind != 1)
        return fail(MCX_EINVAL, "mcx_batch_load_state: not a lattice-batch snapshot of this library");
    const bool series = (h.flags & MCX_STATE_SERIES) != 0;
    if (h.shape[0] != b->model || h.shape[1] != b->dim || h.shape[2] != b->n ||
        h.shape[3] != b->pool_bytes)
        return fail(MCX_EINVAL, "mcx_batch_load_state: the snapshot is of another batch shape "
                                "(model %lld, dim %lld, %lld chains)",
                    (long long)h.shape[0], (long long)h.shape[1], (long long)h.shape[2]);
    if (series && h.shape[4] != b->cap)
        return fail(MCX_EINVAL, "mcx_batch_load_state: the snapshot holds series of capacity %lld, "
                                "this batch has %lld",
                    (long long)h.shape[4], (long long)b->cap);
    if ((size_t)bytes < batch_state_bytes(b, 0))
        return fail(MCX_EINVAL, "mcx_batch_load_state: truncated snapshot");
    const size_t n = (size_t)b->n;
    std::vector<ChainRec> rec(n);
    r.get(rec.data(), sizeof(ChainRec) * n);
    for (size_t i = 0; i < n; ++i) {
        if (rec[i].L != b->rec[i].L || rec[i].dim != b->rec[i].dim ||
            rec[i].spin_off != b->rec[i].spin_off || rec[i].series_off != b->rec[i].series_off)
            return fail(MCX_EINVAL, "mcx_batch_load_state: chain %zu differs in size or layout", i);
        // the kernels index the tables (and, in temperature order, the series
        // slots) by the record's table, and the Philox counter holds 48 bits
        // of the sweep count
        if (rec[i].table < 0 || rec[i].table >= b->n || rec[i].sweep >= (1ull << 48) ||
            !std::isfinite(rec[i].D))
            return fail(MCX_EINVAL, "mcx_batch_load_state: chain %zu: corrupted record (table %d, "
                                    "sweep %llu)", i, rec[i].table,
                        (unsigned long long)rec[i].sweep);
    }
    std::vector<int64_t> steps(n);
    std::vector<double> betas(n);
    r.get(steps.data(), sizeof(int64_t) * n);
    r.get(betas.data(), sizeof(double) * n);
    const size_t tab = (b->model == MCX_ISING ? sizeof(IsingTable) : sizeof(BCTable)) * n;
    const size_t fixed = sizeof(unsigned long long) * n + sizeof(double) * 6 * n +
                         sizeof(int32_t) * 3 * n + tab + (size_t)b->pool_bytes;
    const int64_t rec_len = series ? h.shape[5] : 0;
    const size_t ser = series ? sizeof(int32_t) * (size_t)rec_len * n * (b->d_Q ? 3 : 2) : 0;
    if (r.left() < fixed + ser || rec_len < 0 || rec_len > b->cap)
        return fail(MCX_EINVAL, "mcx_batch_load_state: truncated or inconsistent snapshot");
    {
        // the spins must lie in the model's domain, as mcx_batch_set_spins requires
        BlobReader peek = r;
        peek.take(fixed - (size_t)b->pool_bytes);
        const int8_t* pool = reinterpret_cast<const int8_t*>(peek.take((size_t)b->pool_bytes));
        for (size_t i = 0; i < n; ++i) {
            const int8_t* s = pool + rec[i].spin_off;
            for (int64_t k = 0; k < b->nsites[i]; ++k)
                if (s[k] < -1 || s[k] > 1 || (b->model == MCX_ISING && s[k] == 0))
                    return fail(MCX_EINVAL, "mcx_batch_load_state: chain %zu: spin %lld = %d is "
                                            "outside the model's states", i, (long long)k, (int)s[k]);
        }
    }
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    auto h2d = [&](void* dst, size_t len) -> hipError_t {
        return hipMemcpyAsync(dst, r.take(len), len, hipMemcpyHostToDevice, st);
    };
    HIPCHK(h2d(b->d_acc, sizeof(unsigned long long) * n));
    HIPCHK(h2d(b->d_mom, sizeof(double) * 6 * n));
    HIPCHK(h2d(b->d_last, sizeof(int32_t) * 3 * n));
    HIPCHK(h2d(b->model == MCX_ISING ? (void*)b->d_itab : (void*)b->d_btab, tab));
    HIPCHK(h2d(b->d_spins, (size_t)b->pool_bytes));
    if (series && rec_len > 0) {
        int32_t* dst[3] = {b->d_E, b->d_M, b->d_Q};
        for (int q = 0; q < 3; ++q) {
            if (!dst[q])
                continue;
            for (size_t i = 0; i < n; ++i)
                HIPCHK(h2d(dst[q] + rec[i].series_off, sizeof(int32_t) * (size_t)rec_len));
        }
    }
    HIPCHK(hipMemcpyAsync(b->d_chains, rec.data(), sizeof(ChainRec) * n, hipMemcpyHostToDevice, st));
    mcx_status s = sync_check(b, "mcx_batch_load_state");
    if (s)
        return s;
    b->rec = rec;
    b->steps = steps;
    b->betas = betas;
    b->rec_len = rec_len;
    b->tables_ready = h.shape[6] != 0;
    b->by_table = h.shape[7] != 0;
    return MCX_OK;
}
