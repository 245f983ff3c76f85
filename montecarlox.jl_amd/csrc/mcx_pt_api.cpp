// mcx_pt_api.cpp -- C ABI for replica exchange (parallel tempering).
#include "mcx_api_impl.h"
#include "mcx_pt.h"

#include <cmath>
#include <cstring>

using namespace mcx;

struct mcx_pt {
    mcx_batch* b = nullptr;
    int n_sets = 0, R = 0;
    std::vector<double> ladder;
    int32_t *d_idx = nullptr, *d_pos = nullptr, *d_steps = nullptr, *d_acc = nullptr,
            *d_stage = nullptr;
    uint64_t* d_round = nullptr;
    double* d_ladder = nullptr;
    double* d_x = nullptr;
    // the instantiated graph of the last MCX_RUN_GRAPH run, and what its launch
    // arguments depend on: a run with the same key replays it as it is (the
    // kernels read the streams, ladders and tables from device memory); one that
    // differs only in the series position updates it in place
    struct GraphKey {
        int alg = -1, by_table = 0, keep_planes = 0;
        int64_t rounds = 0, sweeps = 0, rec_len = 0;
        uint32_t rflags = 0;
        mcx_launch_shape shape{};
        bool same_shape(const GraphKey& o) const
        {
            return alg == o.alg && by_table == o.by_table && keep_planes == o.keep_planes &&
                   rounds == o.rounds && sweeps == o.sweeps && rflags == o.rflags &&
                   std::memcmp(&shape, &o.shape, sizeof shape) == 0;
        }
    };
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey;
};

namespace {

void pt_free(mcx_pt* p)
{
    if (p->gexec)
        hipGraphExecDestroy(p->gexec);
    p->gexec = nullptr;
    hipFree(p->d_idx);
    hipFree(p->d_pos);
    hipFree(p->d_steps);
    hipFree(p->d_acc);
    hipFree(p->d_stage);
    hipFree(p->d_round);
    hipFree(p->d_ladder);
    hipFree(p->d_x);
}

PTArgs pt_args(mcx_pt* p)
{
    PTArgs a{};
    a.n_sets = p->n_sets;
    a.R = p->R;
    a.chains = p->b->d_chains;
    a.idx = p->d_idx;
    a.pos = p->d_pos;
    a.ladder = p->d_ladder;
    a.steps = p->d_steps;
    a.acc = p->d_acc;
    a.stage = p->d_stage;
    a.round = p->d_round;
    a.E = p->b->d_last;
    a.Q = p->b->d_last + 2 * p->b->n;
    a.x = nullptr;
    return a;
}

// identity ladder: replica r holds ladder index r + 1, its table is its own
mcx_status pt_reset_state(mcx_pt* p)
{
    mcx_batch* b = p->b;
    const int n = p->n_sets * p->R;
    std::vector<int32_t> idx(n), pos(n);
    for (int s = 0; s < p->n_sets; ++s)
        for (int r = 0; r < p->R; ++r) {
            idx[s * p->R + r] = r + 1;
            pos[s * p->R + r] = r;
        }
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipMemcpyAsync(p->d_idx, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->d_pos, pos.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(p->d_steps, 0, sizeof(int32_t) * p->n_sets * (p->R - 1), st));
    HIPCHK(hipMemsetAsync(p->d_acc, 0, sizeof(int32_t) * p->n_sets * (p->R - 1), st));
    HIPCHK(hipMemsetAsync(p->d_stage, 0, sizeof(int32_t) * p->n_sets, st));
    mcx_status s = pull_chains(b);
    if (s)
        return s;
    for (int i = 0; i < n; ++i)
        b->rec[i].table = i;
    s = upload_chains(b);
    if (s)
        return s;
    return sync_check(b, "mcx_pt_reset");
}

} // namespace

extern "C" mcx_status mcx_pt_create(mcx_batch* b, const mcx_pt_desc* d, mcx_pt** out)
{
    if (!b || !d || !out || !d->betas)
        return fail(MCX_EINVAL, "mcx_pt_create: NULL argument");
    *out = nullptr;
    if (d->replicas < 2)
        return fail(MCX_EINVAL, "need at least 2 algorithms for replica exchange");
    if (d->n_sets < 1 || (int64_t)d->n_sets * d->replicas != b->n)
        return fail(MCX_EINVAL, "mcx_pt_create: n_sets * replicas (%d * %d) must equal the batch's %d chains",
                    d->n_sets, d->replicas, b->n);
    if (b->pt)
        return fail(MCX_ESTATE, "mcx_pt_create: the batch already has a replica-exchange object");
    const int n = b->n;
    for (int i = 0; i < n; ++i)
        if (!(d->betas[i] >= 0) || !std::isfinite(d->betas[i]))
            return fail(MCX_EINVAL, "mcx_pt_create: ladder beta %d = %g must be finite and >= 0", i,
                        d->betas[i]);
    HIPCHK(hipSetDevice(b->ctx->device));
    auto* p = new mcx_pt();
    p->b = b;
    p->n_sets = d->n_sets;
    p->R = d->replicas;
    p->ladder.assign(d->betas, d->betas + n);
    hipError_t e = hipSuccess;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc(ptr, bytes);
    };
    alloc((void**)&p->d_idx, sizeof(int32_t) * n);
    alloc((void**)&p->d_pos, sizeof(int32_t) * n);
    alloc((void**)&p->d_steps, sizeof(int32_t) * p->n_sets * (p->R - 1));
    alloc((void**)&p->d_acc, sizeof(int32_t) * p->n_sets * (p->R - 1));
    alloc((void**)&p->d_stage, sizeof(int32_t) * p->n_sets);
    alloc((void**)&p->d_round, sizeof(uint64_t) * p->n_sets);
    alloc((void**)&p->d_ladder, sizeof(double) * n);
    alloc((void**)&p->d_x, sizeof(double) * n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_ladder, p->ladder.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                           b->ctx->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(p->d_round, 0, sizeof(uint64_t) * p->n_sets, b->ctx->stream);
    if (e != hipSuccess) {
        pt_free(p);
        delete p;
        return fail(e == hipErrorOutOfMemory ? MCX_ENOMEM : MCX_EHIP, "mcx_pt_create: %s",
                    hipGetErrorString(e));
    }
    // acceptance tables indexed by (set, ladder index)
    mcx_status s = set_chain_tables(b, p->ladder.data());
    if (s == MCX_OK)
        s = pt_reset_state(p);
    if (s) {
        pt_free(p);
        delete p;
        return s;
    }
    b->pt = p;
    *out = p;
    return MCX_OK;
}

extern "C" mcx_status mcx_pt_destroy(mcx_pt* p)
{
    if (!p)
        return MCX_OK;
    hipSetDevice(p->b->ctx->device);
    hipStreamSynchronize(p->b->ctx->stream);
    p->b->pt = nullptr;
    p->b->by_table = false;
    pt_free(p);
    delete p;
    return MCX_OK;
}

extern "C" mcx_status mcx_pt_reset(mcx_pt* p)
{
    if (!p)
        return fail(MCX_EINVAL, "mcx_pt_reset: NULL handle");
    HIPCHK(hipSetDevice(p->b->ctx->device));
    return pt_reset_state(p);
}

extern "C" mcx_status mcx_pt_exchange(mcx_pt* p, const double* x)
{
    if (!p)
        return fail(MCX_EINVAL, "mcx_pt_exchange: NULL handle");
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    PTArgs a = pt_args(p);
    if (x) {
        for (int i = 0; i < b->n; ++i)
            if (!std::isfinite(x[i]))
                return fail(MCX_EINVAL, "mcx_pt_exchange: energy %d is not finite", i);
        HIPCHK(hipMemcpyAsync(p->d_x, x, sizeof(double) * b->n, hipMemcpyHostToDevice, st));
        a.x = p->d_x;
    } else {
        HIPCHK(launch_observables(b->d_chains, b->n, b->max_sites, b->d_spins, b->d_last,
                                  b->d_last + b->n, b->d_last + 2 * b->n, st));
    }
    HIPCHK(launch_pt_exchange(a, st));
    return sync_check(b, "mcx_pt_exchange");
}

extern "C" mcx_status mcx_pt_run(mcx_pt* p, int32_t alg, int64_t n_rounds, int64_t sweeps_between,
                                 uint32_t flags)
{
    if (!p)
        return fail(MCX_EINVAL, "mcx_pt_run: NULL handle");
    mcx_batch* b = p->b;
    if (n_rounds < 0 || sweeps_between < 1)
        return fail(MCX_EINVAL, "mcx_pt_run: need n_rounds >= 0 and sweeps_between >= 1");
    if (flags & ~(MCX_RUN_RECORD | MCX_RUN_GRAPH))
        return fail(MCX_EINVAL, "mcx_pt_run: flags must be a combination of MCX_RUN_RECORD, MCX_RUN_GRAPH");
    mcx_status s = check_run(b, alg, n_rounds * sweeps_between, flags & MCX_RUN_RECORD, "mcx_pt_run");
    if (s)
        return s;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const bool graph = (flags & MCX_RUN_GRAPH) != 0;
    const uint32_t rflags = flags & MCX_RUN_RECORD;
    PTArgs a = pt_args(p);
    // between the rounds of this run the lattices stay bit planes in HBM: the
    // first round packs the int8 spins, the last one unpacks them (the exchange
    // reads the energies the sweeps leave, not the spins)
    const bool keep_planes = n_rounds > 1;
    if (keep_planes) {
        s = ensure_planes(b);
        if (s)
            return s;
    }
    auto planes_of = [&](int64_t k) {
        return keep_planes ? (k > 0 ? 1 : 0) | (k < n_rounds - 1 ? 2 : 0) : 0;
    };
    HIPCHK(hipEventRecord(b->ctx->ev0, st));
    if (graph && n_rounds > 0) {
        // one graph holding every round: sweep launch(es) then the exchange
        mcx_pt::GraphKey key;
        key.alg = alg;
        key.by_table = b->by_table ? 1 : 0;
        key.keep_planes = keep_planes ? 1 : 0;
        key.rounds = n_rounds;
        key.sweeps = sweeps_between;
        key.rec_len = rflags ? b->rec_len : 0;
        key.rflags = rflags;
        key.shape = b->shape[alg];
        if (p->gexec && key.same_shape(p->gkey) && key.rec_len == p->gkey.rec_len) {
            // the same launch arguments as the instantiated graph: replay it, and
            // advance the host mirror of the stream positions as the launches would
            for (int64_t k = 0; k < n_rounds; ++k)
                for (int ch = 0; ch < b->n; ++ch) {
                    b->rec[ch].sweep += (uint64_t)sweeps_between;
                    b->steps[ch] += b->nsites[ch] * sweeps_between;
                }
        } else {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            mcx_status rs = MCX_OK;
            for (int64_t k = 0; k < n_rounds && rs == MCX_OK; ++k) {
                rs = run_sweeps_async(b, alg, sweeps_between, rflags, true, planes_of(k));
                if (rs == MCX_OK && launch_pt_exchange(a, st) != hipSuccess)
                    rs = fail(MCX_EHIP, "mcx_pt_run: exchange launch failed during capture");
            }
            hipError_t ce = hipStreamEndCapture(st, &g);
            if (rs != MCX_OK || ce != hipSuccess) {
                if (g)
                    hipGraphDestroy(g);
                return rs != MCX_OK ? rs
                                    : fail(MCX_EHIP, "mcx_pt_run: graph capture failed: %s",
                                           hipGetErrorString(ce));
            }
            hipError_t ie = hipSuccess;
            if (p->gexec && key.same_shape(p->gkey)) {
                // same topology, other series position: update the instantiated graph
                hipGraphNode_t err_node = nullptr;
                hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
                if (hipGraphExecUpdate(p->gexec, g, &err_node, &res) != hipSuccess ||
                    res != hipGraphExecUpdateSuccess) {
                    (void)hipGetLastError();
                    hipGraphExecDestroy(p->gexec);
                    p->gexec = nullptr;
                }
            } else if (p->gexec) {
                hipGraphExecDestroy(p->gexec);
                p->gexec = nullptr;
            }
            if (!p->gexec)
                ie = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ie != hipSuccess) {
                p->gexec = nullptr;
                return fail(MCX_EHIP, "mcx_pt_run: graph instantiation failed: %s",
                            hipGetErrorString(ie));
            }
            p->gkey = key;
        }
        hipError_t le = hipGraphLaunch(p->gexec, st);
        if (le != hipSuccess)
            return fail(MCX_EHIP, "mcx_pt_run: graph launch failed: %s", hipGetErrorString(le));
    } else {
        for (int64_t k = 0; k < n_rounds; ++k) {
            s = run_sweeps_async(b, alg, sweeps_between, rflags, true, planes_of(k));
            if (s)
                return s;
            HIPCHK(launch_pt_exchange(a, st));
        }
    }
    HIPCHK(hipEventRecord(b->ctx->ev1, st));
    s = sync_check(b, "mcx_pt_run");
    if (s)
        return s;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
    b->last_ms = ms;
    return MCX_OK;
}

extern "C" mcx_status mcx_pt_get_ladder(mcx_pt* p, int32_t* indices, int32_t* stages)
{
    if (!p)
        return fail(MCX_EINVAL, "mcx_pt_get_ladder: NULL handle");
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    if (indices)
        HIPCHK(hipMemcpyAsync(indices, p->d_idx, sizeof(int32_t) * b->n, hipMemcpyDeviceToHost, st));
    if (stages)
        HIPCHK(hipMemcpyAsync(stages, p->d_stage, sizeof(int32_t) * p->n_sets, hipMemcpyDeviceToHost, st));
    return sync_check(b, "mcx_pt_get_ladder");
}

extern "C" mcx_status mcx_pt_get_counts(mcx_pt* p, int32_t* steps, int32_t* accepted)
{
    if (!p)
        return fail(MCX_EINVAL, "mcx_pt_get_counts: NULL handle");
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const size_t n = sizeof(int32_t) * p->n_sets * (p->R - 1);
    if (steps)
        HIPCHK(hipMemcpyAsync(steps, p->d_steps, n, hipMemcpyDeviceToHost, st));
    if (accepted)
        HIPCHK(hipMemcpyAsync(accepted, p->d_acc, n, hipMemcpyDeviceToHost, st));
    return sync_check(b, "mcx_pt_get_counts");
}

extern "C" mcx_status mcx_pt_get_rounds(mcx_pt* p, uint64_t* rounds)
{
    if (!p || !rounds)
        return fail(MCX_EINVAL, "mcx_pt_get_rounds: NULL argument");
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpyAsync(rounds, p->d_round, sizeof(uint64_t) * p->n_sets, hipMemcpyDeviceToHost,
                          b->ctx->stream));
    return sync_check(b, "mcx_pt_get_rounds");
}

extern "C" mcx_status mcx_pt_set_series_order(mcx_pt* p, int32_t order)
{
    if (!p)
        return fail(MCX_EINVAL, "mcx_pt_set_series_order: NULL handle");
    if (order != MCX_ORDER_REPLICA && order != MCX_ORDER_TEMPERATURE)
        return fail(MCX_EINVAL, "mcx_pt_set_series_order: order must be MCX_ORDER_REPLICA or "
                                "MCX_ORDER_TEMPERATURE (got %d)", order);
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    // a series (and its moments) is in one order throughout
    mcx_status s = mcx_batch_reset_series(b);
    if (s)
        return s;
    b->by_table = order == MCX_ORDER_TEMPERATURE;
    return MCX_OK;
}

// ---------------------------------------------------------------------------
// device-state snapshot (checkpoint / resume): ladder positions, counters,
// stages and round counters; restores into an object of the same shape.
// ---------------------------------------------------------------------------

namespace {

size_t pt_state_bytes(const mcx_pt* p)
{
    const size_t n = (size_t)p->n_sets * p->R, e = (size_t)p->n_sets * (p->R - 1);
    return sizeof(StateHeader) + sizeof(double) * n + sizeof(int32_t) * 2 * n +
           sizeof(int32_t) * 2 * e + sizeof(int32_t) * p->n_sets + sizeof(uint64_t) * p->n_sets;
}

} // namespace

extern "C" mcx_status mcx_pt_state_size(mcx_pt* p, size_t* bytes)
{
    if (!p || !bytes)
        return fail(MCX_EINVAL, "mcx_pt_state_size: NULL argument");
    *bytes = pt_state_bytes(p);
    return MCX_OK;
}

extern "C" mcx_status mcx_pt_save_state(mcx_pt* p, void* buf, size_t bytes)
{
    if (!p || !buf)
        return fail(MCX_EINVAL, "mcx_pt_save_state: NULL argument");
    if (bytes < pt_state_bytes(p))
        return fail(MCX_EINVAL, "mcx_pt_save_state: buffer of %zu bytes, %zu needed", bytes,
                    pt_state_bytes(p));
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    const size_t n = (size_t)p->n_sets * p->R, e = (size_t)p->n_sets * (p->R - 1);
    StateHeader h = make_state_header(2, 0);
    h.shape[0] = p->n_sets;
    h.shape[1] = p->R;
    BlobWriter w(buf, bytes);
    w.put(&h, sizeof h);
    w.put(p->ladder.data(), sizeof(double) * n);
    hipStream_t st = b->ctx->stream;
    auto d2h = [&](const void* src, size_t len) -> hipError_t {
        return hipMemcpyAsync(w.reserve(len), src, len, hipMemcpyDeviceToHost, st);
    };
    HIPCHK(d2h(p->d_idx, sizeof(int32_t) * n));
    HIPCHK(d2h(p->d_pos, sizeof(int32_t) * n));
    HIPCHK(d2h(p->d_steps, sizeof(int32_t) * e));
    HIPCHK(d2h(p->d_acc, sizeof(int32_t) * e));
    HIPCHK(d2h(p->d_stage, sizeof(int32_t) * p->n_sets));
    HIPCHK(d2h(p->d_round, sizeof(uint64_t) * p->n_sets));
    return sync_check(b, "mcx_pt_save_state");
}

extern "C" mcx_status mcx_pt_load_state(mcx_pt* p, const void* buf, size_t bytes)
{
    if (!p || !buf)
        return fail(MCX_EINVAL, "mcx_pt_load_state: NULL argument");
    BlobReader r(buf, bytes);
    StateHeader h;
    if (!r.get(&h, sizeof h) || std::memcmp(h.magic, "MCXS", 4) != 0 || h.version != 1 || h.kind != 2)
        return fail(MCX_EINVAL, "mcx_pt_load_state: not a replica-exchange snapshot of this library");
    if (h.shape[0] != p->n_sets || h.shape[1] != p->R || bytes < pt_state_bytes(p))
        return fail(MCX_EINVAL, "mcx_pt_load_state: the snapshot is of %lld sets x %lld replicas",
                    (long long)h.shape[0], (long long)h.shape[1]);
    const size_t n = (size_t)p->n_sets * p->R, e = (size_t)p->n_sets * (p->R - 1);
    std::vector<double> ladder(n);
    r.get(ladder.data(), sizeof(double) * n);
    for (size_t i = 0; i < n; ++i)
        if (ladder[i] != p->ladder[i])
            return fail(MCX_EINVAL, "mcx_pt_load_state: the snapshot's ladder differs at %zu", i);
    {
        // ladder positions: per set, idx a permutation of 1..R and pos its
        // inverse (the exchange kernel indexes replicas and chains by them);
        // stages 0 or 1
        BlobReader peek = r;
        std::vector<int32_t> idx(n), pos(n), stage((size_t)p->n_sets);
        peek.get(idx.data(), sizeof(int32_t) * n);
        peek.get(pos.data(), sizeof(int32_t) * n);
        peek.take(sizeof(int32_t) * 2 * e);
        peek.get(stage.data(), sizeof(int32_t) * (size_t)p->n_sets);
        for (int s = 0; s < p->n_sets; ++s) {
            const int32_t* ix = idx.data() + (size_t)s * p->R;
            const int32_t* ps = pos.data() + (size_t)s * p->R;
            for (int k = 0; k < p->R; ++k)
                if (ps[k] < 0 || ps[k] >= p->R || ix[ps[k]] != k + 1)
                    return fail(MCX_EINVAL, "mcx_pt_load_state: set %d: the ladder positions are "
                                            "not a permutation and its inverse", s);
            if (stage[(size_t)s] != 0 && stage[(size_t)s] != 1)
                return fail(MCX_EINVAL, "mcx_pt_load_state: set %d: stage %d", s, stage[(size_t)s]);
        }
    }
    mcx_batch* b = p->b;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    auto h2d = [&](void* dst, size_t len) -> hipError_t {
        return hipMemcpyAsync(dst, r.take(len), len, hipMemcpyHostToDevice, st);
    };
    HIPCHK(h2d(p->d_idx, sizeof(int32_t) * n));
    HIPCHK(h2d(p->d_pos, sizeof(int32_t) * n));
    HIPCHK(h2d(p->d_steps, sizeof(int32_t) * e));
    HIPCHK(h2d(p->d_acc, sizeof(int32_t) * e));
    HIPCHK(h2d(p->d_stage, sizeof(int32_t) * p->n_sets));
    HIPCHK(h2d(p->d_round, sizeof(uint64_t) * p->n_sets));
    return sync_check(b, "mcx_pt_load_state");
}
