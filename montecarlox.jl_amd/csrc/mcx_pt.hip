// mcx_pt.hip -- replica-exchange (parallel tempering) rounds on the device.
//
// One workgroup per PT set runs update!(rx::ReplicaExchange{ThreadsBackend}, xs)
// (src/algorithms/replica_exchange.jl:158-178): pairs (pair_id, pair_id + 1) of
// ladder indices, starting at 1 on even stages and 2 on odd ones -- every
// other position, so a stage's pairs are disjoint and each thread decides its
// own pairs, with the same words and the same result as in order; the shared
// uniform comes from the stream of the replica holding the lower ladder index
// (rand(algorithm(rx, ri).rng), :168); exchange_log_ratio (:110-113) and
// _accept_exchange (:115) decide; on acceptance the ensembles (here: the
// acceptance-table index of each chain) and the ladder indices swap.
// Configurations never move.
#include "mcx_device.h"
#include "mcx_pt.h"

namespace mcx {

__global__ void k_pt_exchange(PTArgs a)
{
    const int set = blockIdx.x;
    const int R = a.R;
    int32_t* idx = a.idx + (int64_t)set * R;
    int32_t* pos = a.pos + (int64_t)set * R;
    const double* lb = a.ladder + (int64_t)set * R;
    int32_t* steps = a.steps + (int64_t)set * (R - 1);
    int32_t* acc = a.acc + (int64_t)set * (R - 1);
    const int stage = a.stage[set];
    const uint64_t round = a.round[set];
    const int first = (stage % 2 == 0) ? 1 : 2;
    for (int pid = first + 2 * (int)threadIdx.x; pid <= R - 1; pid += 2 * (int)blockDim.x) {
        const int ri = pos[pid - 1];
        const int rj = pos[pid];
        const int ci = set * R + ri;
        const int cj = set * R + rj;
        steps[pid - 1] += 1;
        const double xi = a.x ? a.x[ci] : (double)a.E[ci] + a.chains[ci].D * (double)a.Q[ci];
        const double xj = a.x ? a.x[cj] : (double)a.E[cj] + a.chains[cj].D * (double)a.Q[cj];
        const double bi = lb[pid - 1];
        const double bj = lb[pid];
        const double lr = ((-bi * xj) - (-bi * xi)) + ((-bj * xi) - (-bj * xj));
        bool ok = lr > 0;
        if (!ok) {
            const uint64_t seed = a.chains[ci].seed;
            const uint4 w = philox4x32_10((uint32_t)round, (uint32_t)(round >> 32), P_PT << 24,
                                          (uint32_t)(seed >> 32), (uint32_t)seed,
                                          a.chains[ci].key);
            ok = (double)w.x * 0x1.0p-32 < exp(lr);
        }
        if (ok) {
            acc[pid - 1] += 1;
            const int32_t t = idx[ri];
            idx[ri] = idx[rj];
            idx[rj] = t;
            pos[pid - 1] = rj;
            pos[pid] = ri;
            a.chains[ci].table = set * R + idx[ri] - 1;
            a.chains[cj].table = set * R + idx[rj] - 1;
        }
    }
    // every thread has read the set's stage and round
    __syncthreads();
    if (threadIdx.x == 0) {
        a.stage[set] = 1 - stage;
        a.round[set] = round + 1;
    }
}

hipError_t launch_pt_exchange(const PTArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(k_pt_exchange, dim3(a.n_sets), dim3(64), 0, st, a);
    return hipGetLastError();
}

} // namespace mcx
