// mcx_tau.hip -- integrated autocorrelation times of the recorded series, on
// the device (src/measurements/autocorrelations.jl:28-110).
//
// tau = 1/2 + sum_{t>=1} C(t), C(t) the normalised autocorrelation at lag t,
// the sum stopped at the first t with C(t) <= 0 or t > c * tau (:52-72).  The
// centred lag products are evaluated exactly: for integer series x, y of
// length n with sums Sx, Sy,
//   N_xy(t) = n^2 sum_{i<n-t} (x_i - Sx/n)(y_{i+t} - Sy/n)
//           = n^2 P_xy(t) - n Sy H_x(t) - n Sx T_y(t) + (n - t) Sx Sy,
//   P_xy(t) = sum_{i<n-t} x_i y_{i+t},  H_x(t) = sum_{i<n-t} x_i,
//   T_y(t) = sum_{i>=t} y_i,
// an integer held in 128 bits.  For the Blume-Capel energy E = E_pair + D Q,
// N_E = N_pp + D (N_pq + N_qp) + D^2 N_qq in FP64.  Then
// C(t) = (N(t) / (n - t)) / (N(0) / n), the reference's cov(t) / C0 with both
// scaled by n^2.  No result depends on a summation order.
//
// One workgroup (4 waves) per series; lane l of every wave owns lag t0 + l of
// the current block of 64 lags, wave w sums over a quarter of i.  After each
// block, thread 0 walks the 64 values of C(t) in order and decides whether the
// window closes; the next block is computed only if it does not.
#include "mcx_internal.h"

#include <mutex>

namespace mcx {

namespace {

constexpr int kTauThreads = 256;
constexpr int kTauWaves = kTauThreads / 64;

__device__ __forceinline__ int64_t tau_val(const int32_t* s, int64_t i, bool absx)
{
    const int32_t v = s[i];
    return absx ? (int64_t)(v < 0 ? -v : v) : (int64_t)v;
}

// n^2 P - n Sy H - n Sx T + (n - t) Sx Sy
__device__ __forceinline__ __int128 tau_centred(int64_t n, int64_t t, __int128 P, int64_t Sx,
                                                int64_t Sy, int64_t H, int64_t T)
{
    const __int128 nn = (__int128)n;
    return nn * nn * P - nn * (__int128)Sy * (__int128)H - nn * (__int128)Sx * (__int128)T +
           (__int128)(n - t) * (__int128)Sx * (__int128)Sy;
}

} // namespace

__global__ __launch_bounds__(kTauThreads) void k_tau_int(TauArgs a)
{
    __shared__ __int128 red[kTauWaves][4][64];
    __shared__ long long red64[kTauWaves][4];
    __shared__ double cvals[64];
    __shared__ double n0_shared;
    __shared__ int go;

    const int j = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t n = a.n;
    const bool absx = a.absx != 0;
    const int32_t* x = a.X + (int64_t)j * a.cap + a.first;
    const double D = a.chains[j].D;
    const int32_t* y = (a.Y && D != 0.0) ? a.Y + (int64_t)j * a.cap + a.first : nullptr;

    // ---- sums and lag-0 products ----
    long long sx = 0, sy = 0;
    __int128 pxx = 0, pxy = 0, pyy = 0;
    for (int64_t i = tid; i < n; i += kTauThreads) {
        const int64_t xi = tau_val(x, i, absx);
        sx += xi;
        pxx += (__int128)(xi * xi);
        if (y) {
            const int64_t yi = y[i];
            sy += yi;
            pxy += (__int128)(xi * yi);
            pyy += (__int128)(yi * yi);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o);
        sy += __shfl_xor(sy, o);
    }
    red[wave][0][lane] = pxx;
    red[wave][1][lane] = pxy;
    red[wave][2][lane] = pyy;
    if (lane == 0) {
        red64[wave][0] = sx;
        red64[wave][1] = sy;
    }
    __syncthreads();
    // every thread forms the totals (same order everywhere)
    long long Sx = 0, Sy = 0;
    __int128 Pxx0 = 0, Pxy0 = 0, Pyy0 = 0;
    for (int w = 0; w < kTauWaves; ++w) {
        Sx += red64[w][0];
        Sy += red64[w][1];
    }
    if (tid == 0) {
        for (int w = 0; w < kTauWaves; ++w)
            for (int l = 0; l < 64; ++l) {
                Pxx0 += red[w][0][l];
                Pxy0 += red[w][1][l];
                Pyy0 += red[w][2][l];
            }
    }
    double N0 = 0; // N(0), thread 0 only
    if (tid == 0) {
        if (y) {
            const __int128 n_xx = tau_centred(n, 0, Pxx0, Sx, Sx, Sx, Sx);
            const __int128 n_xy = tau_centred(n, 0, Pxy0, Sx, Sy, Sx, Sy);
            const __int128 n_yy = tau_centred(n, 0, Pyy0, Sy, Sy, Sy, Sy);
            N0 = ((double)n_xx + D * (double)(n_xy + n_xy)) + (D * D) * (double)n_yy;
        } else {
            N0 = (double)tau_centred(n, 0, Pxx0, Sx, Sx, Sx, Sx);
        }
        go = N0 > 0.0 ? 1 : 0;
        n0_shared = N0;
    }
    __syncthreads();
    const double C0 = n0_shared / (double)n;
    double tau = 0.5;

    // heads / tails at the start of the block: sum_{i<t0} and sum_{i>=n-t0}
    // (t0 = 1 at first), carried by thread 0
    long long hx0 = 0, hy0 = 0, tx0 = 0, ty0 = 0;
    if (tid == 0) {
        hx0 = tau_val(x, 0, absx);
        tx0 = tau_val(x, n - 1, absx);
        if (y) {
            hy0 = y[0];
            ty0 = y[n - 1];
        }
    }
    for (int64_t t0 = 1; t0 <= a.lag_cap; t0 += 64) {
        __syncthreads();
        if (!go)
            break;
        const int64_t t = t0 + lane;
        const bool live = t <= a.lag_cap;
        // P(t) over i in [0, n - t): wave w takes a contiguous quarter of [0, n - t0)
        __int128 q_xx = 0, q_xy = 0, q_yx = 0, q_yy = 0;
        const int64_t span = n - t0;
        const int64_t chunk = (span + kTauWaves - 1) / kTauWaves;
        const int64_t ib = wave * chunk;
        const int64_t ie = ib + chunk < span ? ib + chunk : span;
        const int64_t lim = live ? n - t : 0; // i < n - t
        for (int64_t i = ib; i < ie; ++i) {
            if (i < lim) {
                const int64_t xi = tau_val(x, i, absx);
                const int64_t xt = tau_val(x, i + t, absx);
                q_xx += (__int128)(xi * xt);
                if (y) {
                    const int64_t yi = y[i];
                    const int64_t yt = y[i + t];
                    q_xy += (__int128)(xi * yt);
                    q_yx += (__int128)(yi * xt);
                    q_yy += (__int128)(yi * yt);
                }
            }
        }
        red[wave][0][lane] = q_xx;
        red[wave][1][lane] = q_xy;
        red[wave][2][lane] = q_yx;
        red[wave][3][lane] = q_yy;
        __syncthreads();
        if (wave == 0) {
            __int128 P_xx = 0, P_xy = 0, P_yx = 0, P_yy = 0;
            for (int w = 0; w < kTauWaves; ++w) {
                P_xx += red[w][0][lane];
                P_xy += red[w][1][lane];
                P_yx += red[w][2][lane];
                P_yy += red[w][3][lane];
            }
            // heads sum_{i<t} and tails sum_{i>=n-t} of this lane's lag, from the
            // block-start values of thread 0
            long long hx = __shfl(hx0, 0), hy = __shfl(hy0, 0);
            long long tx = __shfl(tx0, 0), ty = __shfl(ty0, 0);
            for (int64_t k = t0; k < t && live; ++k) {
                hx += tau_val(x, k, absx);
                tx += tau_val(x, n - 1 - k, absx);
                if (y) {
                    hy += y[k];
                    ty += y[n - 1 - k];
                }
            }
            double C = 0.0;
            if (live) {
                // H_x(t) = Sx - tail_x(t), T_y(t) = Sy - head_y(t)
                if (y) {
                    const __int128 n_xx = tau_centred(n, t, P_xx, Sx, Sx, Sx - tx, Sx - hx);
                    const __int128 n_xy = tau_centred(n, t, P_xy, Sx, Sy, Sx - tx, Sy - hy);
                    const __int128 n_yx = tau_centred(n, t, P_yx, Sy, Sx, Sy - ty, Sx - hx);
                    const __int128 n_yy = tau_centred(n, t, P_yy, Sy, Sy, Sy - ty, Sy - hy);
                    const double Nt = ((double)n_xx + D * (double)(n_xy + n_yx)) + (D * D) * (double)n_yy;
                    C = (Nt / (double)(n - t)) / C0;
                } else {
                    const double Nt = (double)tau_centred(n, t, P_xx, Sx, Sx, Sx - tx, Sx - hx);
                    C = (Nt / (double)(n - t)) / C0;
                }
            }
            cvals[lane] = C;
            // block-start heads / tails of the next block: lane 63's values plus one term
            if (lane == 63) {
                const int64_t tn = t0 + 64;
                if (tn <= a.lag_cap) {
                    hx += tau_val(x, t, absx);
                    tx += tau_val(x, n - 1 - t, absx);
                    if (y) {
                        hy += y[t];
                        ty += y[n - 1 - t];
                    }
                }
            }
            const long long nhx = __shfl(hx, 63), nhy = __shfl(hy, 63);
            const long long ntx = __shfl(tx, 63), nty = __shfl(ty, 63);
            if (lane == 0) {
                hx0 = nhx;
                hy0 = nhy;
                tx0 = ntx;
                ty0 = nty;
            }
        }
        __syncthreads();
        if (tid == 0) {
            int cont = 1;
            for (int l = 0; l < 64 && cont; ++l) {
                const int64_t tl = t0 + l;
                if (tl > a.lag_cap) {
                    cont = 0;
                    break;
                }
                const double Cl = cvals[l];
                if (Cl <= 0.0) {
                    cont = 0;
                    break;
                }
                const double tau_next = tau + Cl;
                if ((double)tl > a.c * tau_next) {
                    cont = 0;
                    break;
                }
                tau = tau_next;
            }
            go = cont;
        }
    }
    if (tid == 0)
        a.taus[j] = tau > 0.5 ? tau : 0.5;
}

hipError_t launch_tau_int(const TauArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(k_tau_int, dim3(a.n_traces), dim3(kTauThreads), 0, st, a);
    return hipGetLastError();
}

} // namespace mcx

// Output staging of the allocation-free tau entry point: one FP64 per slot,
// kTauScratchSlots at most (512 KiB on each GPU the library runs on, held
// until the process ends).  Every call on a GPU shares that GPU's copy, so
// calls that overlap on one GPU race.  Host code never passes `scratch`
// itself to a kernel or a copy: in host code the name is a stand-in, not a
// device address.  It uses the per-device address that tau_scratch_address()
// obtains once per device with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[mcx::kTauScratchSlots];

namespace mcx {

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
} // namespace

hipError_t tau_scratch_address(int device, double** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess)
            return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device)
            hipSetDevice(cur);
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

} // namespace mcx
