// mcx_misc.hip -- small kernels around the sweeps: initial states, full
// observables, series moments, sweep-counter advance, series clearing.
#include "mcx_device.h"
#include "../../include/mcx_hip.h"

namespace mcx {

__device__ __forceinline__ void site_coords(int64_t i, int L, int dim, int& x, int& y, int& z)
{
    x = (int)(i % L);
    y = (int)((i / L) % L);
    z = dim == 3 ? (int)(i / ((int64_t)L * L)) : 0;
}

// init!(sys, type; rng): SpinSystems/src/ising.jl:60-78, blume_capel.jl:87-109.
// :random draws site word (P_INIT) of each site at the chain's sweep counter:
// Ising +1 iff its top bit is set; Blume-Capel state floor(3w / 2^32) - 1.
__global__ void k_init_spins(int model, int mode, ChainRec* chains, int32_t n_chains,
                             int8_t* spins)
{
    for (int ch = blockIdx.y; ch < n_chains; ch += gridDim.y) {
        const ChainRec cr = chains[ch];
        const int L = cr.L;
        const int64_t N = (int64_t)L * L * (cr.dim == 3 ? L : 1);
        int8_t* sp = spins + cr.spin_off;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
             i += (int64_t)gridDim.x * blockDim.x) {
            int8_t v;
            if (mode == 0)
                v = 1;
            else if (mode == 1)
                v = -1;
            else if (mode == 2)
                v = 0;
            else {
                int x, y, z;
                site_coords(i, L, cr.dim, x, y, z);
                const int c = (x + y + z) & 1;
                const int64_t r = (int64_t)y + (int64_t)L * z;
                const int j = x >> 1;
                const int nw = ((L >> 1) + 31) >> 5;
                const uint32_t g = (uint32_t)((r * 2 + c) * nw + (j >> 5));
                const uint32_t w = site_word(cr.seed, cr.key, g, j & 31, cr.sweep, P_INIT);
                if (model == 0)
                    v = (w >> 31) ? 1 : -1;
                else
                    v = (int8_t)((int)(((uint64_t)w * 3u) >> 32) - 1);
            }
            sp[i] = v;
        }
    }
}

// Full recompute (ising.jl:474-482,500-504; blume_capel.jl:185-225):
// E_pair = -sum over bonds s_i s_j, M = sum s_i, Q = sum s_i^2.
__global__ void __launch_bounds__(256) k_observables(ChainRec* chains, const int8_t* spins,
                                                     int32_t* E, int32_t* M, int32_t* Q)
{
    __shared__ long long red[3][4];
    const int ch = blockIdx.x;
    const ChainRec cr = chains[ch];
    const int L = cr.L;
    const int dim = cr.dim;
    const int64_t N = (int64_t)L * L * (dim == 3 ? L : 1);
    const int8_t* sp = spins + cr.spin_off;
    long long pair = 0, m = 0, q = 0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        int x, y, z;
        site_coords(i, L, dim, x, y, z);
        const int s = sp[i];
        int nb = sp[i - x + (x + 1 == L ? 0 : x + 1)];
        const int64_t zb = dim == 3 ? (int64_t)z * L * L : 0;
        nb += sp[zb + (int64_t)(y + 1 == L ? 0 : y + 1) * L + x];
        if (dim == 3)
            nb += sp[(int64_t)(z + 1 == L ? 0 : z + 1) * L * L + (int64_t)y * L + x];
        pair += s * nb;
        m += s;
        q += s * s;
    }
    pair = wave_sum(pair);
    m = wave_sum(m);
    q = wave_sum(q);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = pair;
        red[1][wave] = m;
        red[2][wave] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tp = 0, tm = 0, tq = 0;
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); ++k) {
            tp += red[0][k];
            tm += red[1][k];
            tq += red[2][k];
        }
        if (E)
            E[ch] = (int32_t)(-tp);
        if (M)
            M[ch] = (int32_t)tm;
        if (Q)
            Q[ch] = (int32_t)tq;
    }
}

// Moments of series entries [first, first+count) in sweep order (the order the
// batched kernels accumulate in); energy = E_pair (+ D * Q for Blume-Capel).
__global__ void k_series_moments(ChainRec* chains, int32_t n_chains, const int32_t* E,
                                 const int32_t* M, const int32_t* Q, int64_t first, int64_t count,
                                 double* moments)
{
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= n_chains)
        return;
    const ChainRec cr = chains[ch];
    // running sums continue from the stored ones (split-independent, as in the sweep kernels)
    double* mo = moments + (int64_t)ch * 6;
    double mE = mo[1], mE2 = mo[2], mAM = mo[3], mM2 = mo[4], mM4 = mo[5];
    for (int64_t k = 0; k < count; ++k) {
        const int64_t at = cr.series_off + first + k;
        const double e = Q ? (double)E[at] + cr.D * (double)Q[at] : (double)E[at];
        const double m = (double)M[at];
        mE += e;
        mE2 += e * e;
        mAM += fabs(m);
        mM2 += m * m;
        mM4 += (m * m) * (m * m);
    }
    mo[0] += (double)count;
    mo[1] = mE;
    mo[2] = mE2;
    mo[3] = mAM;
    mo[4] = mM2;
    mo[5] = mM4;
}

__global__ void k_advance_sweeps(ChainRec* chains, int32_t n_chains, int64_t nsweeps)
{
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch < n_chains)
        chains[ch].sweep += (uint64_t)nsweeps;
}

__global__ void k_zero_series(ChainRec* chains, int32_t n_chains, int32_t* E, int32_t* M,
                              int32_t* Q, int64_t first, int64_t count)
{
    for (int ch = blockIdx.y; ch < n_chains; ch += gridDim.y) {
        const int64_t off = chains[ch].series_off + first;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
             k += (int64_t)gridDim.x * blockDim.x) {
            E[off + k] = 0;
            M[off + k] = 0;
            if (Q)
                Q[off + k] = 0;
        }
    }
}

hipError_t launch_init_spins(int model, int mode, ChainRec* chains, int32_t n_chains,
                             int64_t max_sites, int8_t* spins, hipStream_t st)
{
    const int block = 256;
    const int64_t gx = (max_sites + block - 1) / block;
    dim3 grid((unsigned)(gx > 4096 ? 4096 : gx), (unsigned)(n_chains > 65535 ? 65535 : n_chains));
    hipLaunchKernelGGL(k_init_spins, grid, dim3(block), 0, st, model, mode, chains, n_chains,
                       spins);
    return hipGetLastError();
}

hipError_t launch_observables(ChainRec* chains, int32_t n_chains, int64_t, const int8_t* spins,
                              int32_t* E, int32_t* M, int32_t* Q, hipStream_t st)
{
    hipLaunchKernelGGL(k_observables, dim3(n_chains), dim3(256), 0, st, chains, spins, E, M, Q);
    return hipGetLastError();
}

// One workgroup per work item of a tile series: per-sweep sums -> E/M/Q, FP64
// moments continued sweep by sweep (the same operations, in the same order, as
// the batched sweep kernels), last-sweep observables, sweep counter.
__global__ void __launch_bounds__(kFoldThreads) k_tiles_fold(TileArgs a, int model,
                                                               double* moments, int32_t* last_E,
                                                               int32_t* last_M, int32_t* last_Q)
{
    // one workgroup per item: the threads sum (and clear) the replicas of
    // their sweeps, then one thread continues the FP64 moments in sweep order,
    // as the sweep kernels do, so that they do not depend on the path
    __shared__ double e_s[kFoldMaxSweeps];
    __shared__ int32_t m_s[kFoldMaxSweeps];
    __shared__ int32_t last[3];
    __shared__ unsigned long long fl[kFoldThreads / 64];
    const int item = blockIdx.x;
    const int tid = threadIdx.x;
    const int ch = a.items[item];
    const int64_t nsw = a.item_nsw[item] < a.max_s ? a.item_nsw[item] : a.max_s;
    const ChainRec cr = a.chains[ch];
    const int64_t N = (int64_t)cr.L * cr.L * (cr.dim == 3 ? cr.L : 1);
    int32_t* obs = reinterpret_cast<int32_t*>(a.base + a.item_obs[item]);
    unsigned long long flips = 0;
    for (int64_t s = tid; s < nsw; s += kFoldThreads) {
        int32_t f[4];
        for (int k = 0; k < 4; ++k) {
            int4* o = reinterpret_cast<int4*>(obs + (s * 4 + k) * kTileObsReplicas);
            int32_t v = 0;
#pragma unroll
            for (int r = 0; r < kTileObsReplicas / 4; ++r) {
                const int4 q = o[r];
                v += q.x + q.y + q.z + q.w;
                o[r] = make_int4(0, 0, 0, 0);
            }
            f[k] = v;
        }
        flips += (unsigned long long)(uint32_t)f[3];
        int32_t E, M, Q;
        double e;
        if (model == MCX_ISING) {
            E = (int32_t)(2 * (int64_t)(uint32_t)f[1] - (int64_t)cr.dim * N);
            M = (int32_t)(2 * (int64_t)(uint32_t)f[0] - N);
            Q = (int32_t)N;
            e = (double)E;
        } else {
            E = -f[0];
            M = f[1];
            Q = f[2];
            e = (double)E + cr.D * (double)Q;
        }
        e_s[s] = e;
        m_s[s] = M;
        if (s == nsw - 1) {
            last[0] = E;
            last[1] = M;
            last[2] = Q;
        }
    }
    // flips: wave sums, then the waves' partials
    unsigned long long w = flips;
    for (int off = 32; off > 0; off >>= 1)
        w += __shfl_xor(w, off, 64);
    if ((tid & 63) == 0)
        fl[tid >> 6] = w;
    __syncthreads();
    if (tid != 0 || nsw <= 0)
        return;
    double* mo = moments + (int64_t)ch * 6;
    double mE = mo[1], mE2 = mo[2], mAM = mo[3], mM2 = mo[4], mM4 = mo[5];
    for (int64_t s = 0; s < nsw; ++s) {
        const double e = e_s[s];
        const double m = (double)m_s[s];
        mE += e;
        mE2 += e * e;
        mAM += fabs(m);
        mM2 += m * m;
        mM4 += (m * m) * (m * m);
    }
    unsigned long long tot = 0;
    for (int k = 0; k < kFoldThreads / 64; ++k)
        tot += fl[k];
    a.accepted[ch] += tot;
    last_E[ch] = last[0];
    last_M[ch] = last[1];
    last_Q[ch] = last[2];
    mo[0] += (double)nsw;
    mo[1] = mE;
    mo[2] = mE2;
    mo[3] = mAM;
    mo[4] = mM2;
    mo[5] = mM4;
    a.chains[ch].sweep = cr.sweep + (uint64_t)nsw;
}

hipError_t launch_tiles_fold(int model, const TileArgs& a, double* moments, int32_t* last_E,
                             int32_t* last_M, int32_t* last_Q, hipStream_t st)
{
    if (a.max_s > kFoldMaxSweeps)
        return hipErrorInvalidValue; // the per-sweep staging holds kFoldMaxSweeps sweeps
    if (a.n_items > 0)
        hipLaunchKernelGGL(k_tiles_fold, dim3(a.n_items), dim3(kFoldThreads), 0, st, a, model,
                           moments, last_E, last_M, last_Q);
    return hipGetLastError();
}

hipError_t launch_series_moments(ChainRec* chains, int32_t n_chains, const int32_t* E,
                                 const int32_t* M, const int32_t* Q, int64_t first,
                                 int64_t count, double* moments, hipStream_t st)
{
    hipLaunchKernelGGL(k_series_moments, dim3((n_chains + 63) / 64), dim3(64), 0, st, chains,
                       n_chains, E, M, Q, first, count, moments);
    return hipGetLastError();
}

hipError_t launch_advance_sweeps(ChainRec* chains, int32_t n_chains, int64_t nsweeps,
                                 hipStream_t st)
{
    hipLaunchKernelGGL(k_advance_sweeps, dim3((n_chains + 255) / 256), dim3(256), 0, st, chains,
                       n_chains, nsweeps);
    return hipGetLastError();
}

hipError_t launch_zero_series(ChainRec* chains, int32_t n_chains, int32_t* E, int32_t* M,
                              int32_t* Q, int64_t first, int64_t count, hipStream_t st)
{
    const int64_t gx = (count + 255) / 256;
    dim3 grid((unsigned)(gx > 1024 ? 1024 : gx), (unsigned)(n_chains > 65535 ? 65535 : n_chains));
    hipLaunchKernelGGL(k_zero_series, grid, dim3(256), 0, st, chains, n_chains, E, M, Q, first,
                       count);
    return hipGetLastError();
}

} // namespace mcx
