// mcx_gen.hip -- multicanonical and Wang-Landau walkers on gfx950.
//
// Per attempt (SpinSystems/src/ising.jl:25-33, blume_capel.jl:41-50 with
// accept!(alg, x_new, x_old), src/algorithms/importance_sampling.jl:69-85 and
// the Wang-Landau accept!, src/algorithms/wang_landau.jl:29-37):
//   lr = lw(E_old + dE) - lw(E_old); accept iff lr > 0 or u < exp(lr);
//   multicanonical: hist[visited] += 1;  Wang-Landau: lw[visited] -= logf
//   (and hist[visited] += 1, kept for the flatness check).
// The acceptance depends on the walker's running energy, so a walker is a
// sequential chain: sites are attempted in checkerboard order (colour 0 in
// increasing site index, then colour 1).
//
// One wavefront per walker.  Sites of one colour never neighbour each other,
// so for 64 consecutive sites of the colour the lanes compute dE, the proposal
// and the uniform in parallel and stage them in LDS; the wave then walks the
// 64 attempts in order.  The walker's FP64 log-weights live in LDS; the weight
// of the current bin is kept in a register, so an attempt costs one LDS load.
// Histograms are uint32 in LDS when they fit, otherwise int64 atomics in HBM.
#include "mcx_device.h"
#include "mcx_gen.h"

namespace mcx {

__device__ __forceinline__ int gen_propose(int s_old, int alt)
{
    return s_old == -1 ? (alt ? 0 : 1) : (s_old == 0 ? (alt ? -1 : 1) : (alt ? -1 : 0));
}

template <int MODEL, int DIM>
__global__ void __launch_bounds__(64) k_gen_sweeps(GenArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    AttemptRec* rec = reinterpret_cast<AttemptRec*>(lds_raw);        // 64 records
    double* lw = reinterpret_cast<double*>(lds_raw + 64 * sizeof(AttemptRec));
    uint32_t* hl = reinterpret_cast<uint32_t*>(lw + a.nbins);
    int8_t* sp = reinterpret_cast<int8_t*>(hl + (a.hist_in_lds ? a.nbins : 0));
    const int lane = threadIdx.x;
    const int64_t nb = a.nbins;
    const bool wl = a.kind == 3;

    for (int item = blockIdx.x; item < a.n_items; item += gridDim.x) {
        const int wk = a.items[item];
        const ChainRec cr = a.chains[wk];
        const int L = cr.L;
        const int half = L >> 1;
        const int nw = (half + 31) >> 5;
        const int rows = DIM == 3 ? L * L : L;
        const int64_t N = (int64_t)rows * L;
        double* lw_g = a.lw + a.lw_off[wk];
        unsigned long long* hist_g = a.hist + a.lw_off[wk];
        int8_t* sp_g = a.spins + cr.spin_off;
        const double logf = a.logf[wk];

        for (int64_t b = lane; b < nb; b += 64) {
            lw[b] = lw_g[b];
            if (a.hist_in_lds)
                hl[b] = 0;
        }
        for (int64_t i = lane; i < N; i += 64)
            sp[i] = sp_g[i];
        __syncthreads();

        // full recompute of the cached sums (_recompute_cached!)
        long long pair = 0, mm = 0, qq = 0;
        for (int64_t i = lane; i < N; i += 64) {
            const int x = (int)(i % L);
            const int y = (int)((i / L) % L);
            const int z = DIM == 3 ? (int)(i / ((int64_t)L * L)) : 0;
            const int s = sp[i];
            int nbs = sp[i - x + (x + 1 == L ? 0 : x + 1)];
            const int64_t zb = DIM == 3 ? (int64_t)z * L * L : 0;
            nbs += sp[zb + (int64_t)(y + 1 == L ? 0 : y + 1) * L + x];
            if (DIM == 3)
                nbs += sp[(int64_t)(z + 1 == L ? 0 : z + 1) * L * L + (int64_t)y * L + x];
            pair += s * nbs;
            mm += s;
            qq += s * s;
        }
        pair = wave_sum(pair);
        mm = wave_sum(mm);
        qq = wave_sum(qq);
        const double D = cr.D;

        auto bin_of = [&](long long Ei, double Ef, int64_t& b) -> bool {
            if (MODEL == 0) {
                const long long d = Ei - (long long)a.bin_start;
                const long long bb = d / (long long)a.bin_step; // truncating div
                b = bb;
                return bb >= 0 && bb < nb;
            }
            const double v = rint((Ef - a.bin_start) / a.bin_step);
            if (!(v >= 0.0 && v < (double)nb))
                return false;
            b = (int64_t)v;
            return true;
        };

        int64_t err = -1;
        int64_t b_cur = -1;  // bin whose weight lw_cur holds (none yet)
        double lw_cur = 0.0;
        unsigned long long steps = 0, nacc = 0;
        int64_t s = 0;
        for (; s < a.nsweeps && err < 0; ++s) {
            const uint64_t sweep = cr.sweep + (uint64_t)s;
            for (int c = 0; c < 2 && err < 0; ++c) {
                const int n_col = rows * half;
                for (int base = 0; base < n_col && err < 0; base += 64) {
                    const int o = base + lane;
                    if (o < n_col) {
                        const int r = o / half;
                        const int j = o - r * half;
                        const int y = DIM == 3 ? r % L : r;
                        const int z = DIM == 3 ? r / L : 0;
                        const int x = 2 * j + ((c + y + z) & 1);
                        const int64_t rowb = (int64_t)L * r;
                        const int64_t zb = DIM == 3 ? (int64_t)z * L * L : 0;
                        int nbs = sp[rowb + (x == 0 ? L - 1 : x - 1)] + sp[rowb + (x == L - 1 ? 0 : x + 1)]
                                + sp[zb + (int64_t)(y == 0 ? L - 1 : y - 1) * L + x]
                                + sp[zb + (int64_t)(y == L - 1 ? 0 : y + 1) * L + x];
                        if (DIM == 3)
                            nbs += sp[(int64_t)(z == 0 ? L - 1 : z - 1) * L * L + (int64_t)y * L + x]
                                 + sp[(int64_t)(z == L - 1 ? 0 : z + 1) * L * L + (int64_t)y * L + x];
                        const int si = sp[rowb + x];
                        const uint32_t g = (uint32_t)((r * 2 + c) * nw + (j >> 5));
                        AttemptRec t;
                        t.w = site_word(cr.seed, cr.key, g, j & 31, sweep, P_GEN);
                        t.site = (int32_t)(rowb + x);
                        if (MODEL == 0) {
                            t.dE = 2 * si * nbs;
                            t.dpair = -2 * si * nbs;
                            t.dm = (int16_t)(-2 * si);
                            t.dq = 0;
                            t.snew = (int8_t)-si;
                            t.dEf = 0.0;
                        } else {
                            const int sn = gen_propose(si, (int)(t.w & 1u));
                            const int ds = sn - si;
                            const int ds2 = sn * sn - si * si;
                            t.dEf = (-((double)ds * (1.0 * (double)nbs)) - 0.0) + D * (double)ds2;
                            t.dE = 0;
                            t.dpair = ds * nbs;
                            t.dm = (int16_t)ds;
                            t.dq = (int16_t)ds2;
                            t.snew = (int8_t)sn;
                        }
                        rec[lane] = t;
                    }
                    __syncthreads();
                    const int cnt = n_col - base < 64 ? n_col - base : 64;
                    for (int k = 0; k < cnt; ++k) {
                        const AttemptRec t = rec[k];
                        // E_old from the cached sums, as _cached_energy evaluates it
                        const double Eo = -(double)pair + D * (double)qq;
                        int64_t b_new, b_old;
                        const bool okn = bin_of(-pair + t.dE, Eo + t.dEf, b_new);
                        const bool oko = bin_of(-pair, Eo, b_old);
                        if (!okn || !oko) {
                            err = t.site;
                            break;
                        }
                        const double lw_old = b_old == b_cur ? lw_cur : lw[b_old];
                        const double lw_new = b_new == b_old ? lw_old : lw[b_new];
                        const double lr = lw_new - lw_old;
                        const double u = MODEL == 0 ? (double)t.w * 0x1.0p-32
                                                    : (double)(t.w >> 1) * 0x1.0p-31;
                        const bool ok = lr > 0 || u < exp(lr);
                        ++steps;
                        nacc += ok ? 1u : 0u;
                        const int64_t vis = ok ? b_new : b_old;
                        if (lane == 0) {
                            if (a.hist_in_lds)
                                hl[vis] += 1u;
                            else
                                atomicAdd(&hist_g[vis], 1ull);
                        }
                        double lw_vis = ok ? lw_new : lw_old;
                        if (wl) {
                            lw_vis = lw_vis - logf;
                            if (lane == 0)
                                lw[vis] = lw_vis;
                        }
                        if (ok) {
                            if (lane == 0)
                                sp[t.site] = t.snew;
                            pair += t.dpair;
                            mm += t.dm;
                            qq += t.dq;
                        }
                        b_cur = vis;
                        lw_cur = lw_vis;
                    }
                    __syncthreads();
                }
            }
            if (err < 0 && a.record && lane == 0) {
                const int64_t at = cr.series_off + a.rec_pos + s;
                a.ser_E[at] = (int32_t)(-pair);
                a.ser_M[at] = (int32_t)mm;
                if (a.ser_Q)
                    a.ser_Q[at] = (int32_t)qq;
            }
        }
        __syncthreads();
        for (int64_t i = lane; i < N; i += 64)
            sp_g[i] = sp[i];
        for (int64_t b = lane; b < nb; b += 64) {
            if (wl)
                lw_g[b] = lw[b];
            if (a.hist_in_lds && hl[b])
                hist_g[b] += hl[b];
        }
        if (lane == 0) {
            a.steps[wk] += steps;
            a.accepted[wk] += nacc;
            a.err_site[wk] = err;
            // a failed sweep is not completed: the counter stays at its start
            a.chains[wk].sweep = cr.sweep + (uint64_t)(err < 0 ? s : s - 1);
        }
        __syncthreads();
    }
}

hipError_t launch_gen_sweeps(const GenArgs& a, size_t lds_bytes, hipStream_t st)
{
    const dim3 grid(a.n_items), block(64);
    if (a.model == 0 && a.dim == 2)
        hipLaunchKernelGGL((k_gen_sweeps<0, 2>), grid, block, lds_bytes, st, a);
    else if (a.model == 0)
        hipLaunchKernelGGL((k_gen_sweeps<0, 3>), grid, block, lds_bytes, st, a);
    else if (a.dim == 2)
        hipLaunchKernelGGL((k_gen_sweeps<1, 2>), grid, block, lds_bytes, st, a);
    else
        hipLaunchKernelGGL((k_gen_sweeps<1, 3>), grid, block, lds_bytes, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// weight recursions and group reductions (one workgroup per walker / group)
// ---------------------------------------------------------------------------

// update!(::MulticanonicalEnsemble; mode=:simple): lw -= log(h) where h > 0
// (src/ensembles/multicanonical.jl:32-44)
__global__ void k_gen_update(double* lw, const unsigned long long* hist, const int64_t* lw_off,
                             const int32_t* walkers, int64_t nbins)
{
    const int w = walkers[blockIdx.y];
    double* l = lw + lw_off[w];
    const unsigned long long* h = hist + lw_off[w];
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long hv = h[b];
        const double logh = hv > 0 ? log((double)hv) : 0.0;
        l[b] -= logh;
    }
}

// The same recursion on caller-held arrays (the reference's Float64 histogram).
__global__ void k_muca_update_values(double* lw, const double* h, int64_t n)
{
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < n;
         b += (int64_t)gridDim.x * blockDim.x) {
        const double hv = h[b];
        const double logh = hv > 0 ? log(hv) : 0.0;
        lw[b] -= logh;
    }
}

hipError_t launch_muca_update_values(double* lw, const double* h, int64_t n, hipStream_t st)
{
    const int64_t gx = (n + 255) / 256;
    hipLaunchKernelGGL(k_muca_update_values, dim3((unsigned)(gx > 1024 ? 1024 : (gx < 1 ? 1 : gx))),
                       dim3(256), 0, st, lw, h, n);
    return hipGetLastError();
}

// merge_histograms!: the root (walkers[0]) accumulates every other walker's
// histogram; the others are unchanged (parallel_multicanonical.jl:38-52)
__global__ void k_gen_merge(unsigned long long* hist, const int64_t* lw_off,
                            const int32_t* walkers, int32_t n_walkers, int64_t nbins)
{
    unsigned long long* root = hist + lw_off[walkers[0]];
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long acc = root[b];
        for (int k = 1; k < n_walkers; ++k)
            acc += hist[lw_off[walkers[k]] + b];
        root[b] = acc;
    }
}

// distribute_logweight!: copy the root's log-weights to every other walker
// (parallel_multicanonical.jl:59-73)
__global__ void k_gen_distribute(double* lw, const int64_t* lw_off, const int32_t* walkers,
                                 int32_t n_walkers, int64_t nbins)
{
    const double* root = lw + lw_off[walkers[0]];
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x) {
        const double v = root[b];
        for (int k = 1; k < n_walkers; ++k)
            lw[lw_off[walkers[k]] + b] = v;
    }
}

// Parallel Wang-Landau merge: the root takes the walkers' mean log-weight,
// summed in walker order.
__global__ void k_gen_average(double* lw, const int64_t* lw_off, const int32_t* walkers,
                              int32_t n_walkers, int64_t nbins)
{
    double* root = lw + lw_off[walkers[0]];
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int k = 0; k < n_walkers; ++k)
            acc += lw[lw_off[walkers[k]] + b];
        root[b] = acc / (double)n_walkers;
    }
}

__global__ void k_gen_zero_hist(unsigned long long* hist, const int64_t* lw_off,
                                const int32_t* walkers, int64_t nbins)
{
    unsigned long long* h = hist + lw_off[walkers[blockIdx.y]];
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x)
        h[b] = 0;
}

// Flatness of the root histogram over the bins the walk has reached (lw
// differs from its start or h > 0): out = {min h, mean h, bins counted}.
__global__ void k_gen_flatness(const unsigned long long* hist, const double* lw,
                               const int64_t* lw_off, int32_t root, int64_t nbins, double* out)
{
    __shared__ double smin[4], ssum[4], scnt[4];
    const unsigned long long* h = hist + lw_off[root];
    const double* l = lw + lw_off[root];
    double mn = 1e300, sm = 0, ct = 0;
    for (int64_t b = threadIdx.x; b < nbins; b += blockDim.x) {
        if (h[b] > 0 || l[b] != 0.0) {
            const double v = (double)h[b];
            mn = v < mn ? v : mn;
            sm += v;
            ct += 1;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(mn, off, 64);
        mn = o < mn ? o : mn;
        sm += __shfl_xor(sm, off, 64);
        ct += __shfl_xor(ct, off, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smin[wv] = mn;
        ssum[wv] = sm;
        scnt[wv] = ct;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            mn = smin[k] < mn ? smin[k] : mn;
            sm += ssum[k];
            ct += scnt[k];
        }
        out[0] = ct > 0 ? mn : 0.0;
        out[1] = ct > 0 ? sm / ct : 0.0;
        out[2] = ct;
    }
}

static dim3 bins_grid(int64_t nbins, int y)
{
    const int64_t gx = (nbins + 255) / 256;
    return dim3((unsigned)(gx > 1024 ? 1024 : gx), (unsigned)y);
}

hipError_t launch_gen_update(double* lw, const unsigned long long* hist, const int64_t* lw_off,
                             const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_update, bins_grid(nbins, n_walkers), dim3(256), 0, st, lw, hist,
                       lw_off, walkers, nbins);
    return hipGetLastError();
}

hipError_t launch_gen_merge(unsigned long long* hist, const int64_t* lw_off,
                            const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_merge, bins_grid(nbins, 1), dim3(256), 0, st, hist, lw_off, walkers,
                       n_walkers, nbins);
    return hipGetLastError();
}

hipError_t launch_gen_distribute(double* lw, const int64_t* lw_off, const int32_t* walkers,
                                 int32_t n_walkers, int64_t nbins, hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_distribute, bins_grid(nbins, 1), dim3(256), 0, st, lw, lw_off,
                       walkers, n_walkers, nbins);
    return hipGetLastError();
}

hipError_t launch_gen_average(double* lw, const int64_t* lw_off, const int32_t* walkers,
                              int32_t n_walkers, int64_t nbins, hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_average, bins_grid(nbins, 1), dim3(256), 0, st, lw, lw_off, walkers,
                       n_walkers, nbins);
    return hipGetLastError();
}

hipError_t launch_gen_zero_hist(unsigned long long* hist, const int64_t* lw_off,
                                const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_zero_hist, bins_grid(nbins, n_walkers), dim3(256), 0, st, hist,
                       lw_off, walkers, nbins);
    return hipGetLastError();
}

hipError_t launch_gen_flatness(const unsigned long long* hist, const double* lw,
                               const int64_t* lw_off, int32_t root, int64_t nbins, double* out,
                               hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_flatness, dim3(1), dim3(256), 0, st, hist, lw, lw_off, root, nbins,
                       out);
    return hipGetLastError();
}

} // namespace mcx
