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
// Two wavefronts per walker.  Sites of one colour never neighbour each other,
// so for 64 consecutive sites of the colour the lanes of the staging wave
// (wave 1) compute dE, the proposal, the uniform and its log in parallel and
// store them in an LDS stage; the walking wave (wave 0) walks the 64 attempts
// in order from the stage.  The stage is double-buffered: while wave 0 walks
// chunk i, wave 1 stages chunk i + 1 of the same colour, which reads only
// spins of the other colour and the chunk's own sites, none of which the walk
// of chunk i changes.  The first chunk of a colour is staged once the previous
// colour's spins are written.  One workgroup barrier per chunk hands the
// buffers over.  The walker's FP64 log-weights live in LDS; the weight and
// bin of the current energy are kept in registers, so an attempt costs one LDS
// load and no exp unless lr falls within 1e-9 of log(u).  In the integer walk
// the chunk's spin writes, counters, running sums and visit counts are done
// by wave 1 from the recorded outcomes while wave 0 walks the next chunk; the
// other walks apply spins and visits with wave 0's lanes after the walk.  When
// the bins are linear in the energy the integer walk carries only the byte
// offset of the current bin's weight (one add per proposal, no domain test:
// NaN padding slots around the window stop a pass that leaves the table).
// Histograms are uint32 in LDS when they fit, otherwise int64 atomics in HBM.
//
// KIND 2 (Blume-Capel): the CustomEnsemble of docs/src/examples/spin_systems/
// muca_BlumeCapel.jl:43-89 -- a Boltzmann weight on the pair term and a
// multicanonical weight on Q = sum s_i^2:
//   logweight(H) = logweight(pair, H[1]) + logweight(spin2, H[2])
//               = (-beta) * P + lw(Q),   H = (J * sum_pair_interactions, sum_spins2),
// P = sum_<ij> s_i s_j (J = 1), evaluated for the old and the new state as
// accept!(alg, H_new, H_old) does; the visit goes to Q's bin.
#include "mcx_device.h"
#include "mcx_gen.h"

namespace mcx {

__device__ __forceinline__ int readlane_i32(int v, int k)
{
    return __builtin_amdgcn_readlane(v, k);
}

__device__ __forceinline__ double readlane_f64(double v, int k)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Orders the LDS and global accesses of the calling wavefront's lanes before
// it against those after it (what a workgroup barrier does, without waiting
// for the other wavefront of the workgroup).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ double uniform_f64(double v)
{
    return readlane_f64(v, 0);
}

__device__ __forceinline__ int gen_propose(int s_old, int alt)
{
    return s_old == -1 ? (alt ? 0 : 1) : (s_old == 0 ? (alt ? -1 : 1) : (alt ? -1 : 0));
}

template <int MODEL, int DIM, int KIND>
__global__ void __launch_bounds__(kGenThreads) k_gen_sweeps(GenArgs a)
{
    constexpr bool WL = KIND == 1; // Wang-Landau
    constexpr bool PS = KIND == 2; // Boltzmann pair term + multicanonical sum s^2
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // the weight window, with max_dbin padding slots on each side for the
    // linear walk's reads past the ends of the table
    double* lw = reinterpret_cast<double*>(lds_raw) + a.max_dbin;
    // two chunks' staged attempts (stage [i & 1] holds chunk i of the colour)
    // and the lane-0 walk's results.  Entry 64 of st_dE / st_luh / st_lul is the
    // integer walk's read-ahead past the last attempt; its value is never used.
    // The FP64 walk keeps log(u) in st_luh and the FP64 energy change in st_lul.
    __shared__ int st_dE[2][65], st_dp[2][64], st_mq[2][64], st_site[2][64], st_snew[2][64];
    // the walk's outcomes per attempt, and how many attempts it ran, per stage
    __shared__ int st_res[2][64], sh_nd[2];
    __shared__ double st_luh[2][65], st_lul[2][65], st_u[2][64];
    // the linear Wang-Landau walk's weights of the visited bins before their
    // decrements, per attempt of the pass (to undo a pass after a band attempt)
    __shared__ double st_old[64];
    // the walk's error site, one slot per barrier parity (wave 0 writes the slot
    // of the barrier it is about to enter; both waves read it after that barrier)
    __shared__ int sh_err[2];
    // the weights in LDS are the window [wbase, wbase + wsz) of the walker's
    // table (the whole table when it fits: wsz = nbins, wbase = 0)
    const int wsz = a.win;
    uint32_t* hl = reinterpret_cast<uint32_t*>(lw + wsz + a.max_dbin);
    int8_t* sp = reinterpret_cast<int8_t*>(hl + (a.hist_in_lds ? a.nbins : 0));
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6); // 0 walks, 1 stages
    const int64_t nb = a.nbins;
    constexpr bool wl = WL;
    // padding slots around the weight window: NaN, so that a linear walk's
    // proposal past an end of the table fails both of its compares
    for (int i = tid; i < a.max_dbin; i += kGenThreads) {
        lw[-1 - i] = __builtin_nan("");
        lw[wsz + i] = __builtin_nan("");
    }
    if (tid < 2) {
        st_dE[tid][64] = 0;
        st_luh[tid][64] = 0.0;
        st_lul[tid][64] = 0.0;
    }

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

        if (a.hist_in_lds)
            for (int64_t b = tid; b < nb; b += kGenThreads)
                hl[b] = 0;
        for (int64_t i = tid; i < N; i += kGenThreads)
            sp[i] = sp_g[i];
        if (tid < 2)
            sh_err[tid] = -1;
        __syncthreads();

        // full recompute of the cached sums (_recompute_cached!), by each wave
        // over the whole lattice
        long long pair0 = 0, mm0 = 0, qq0 = 0;
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
            pair0 += s * nbs;
            mm0 += s;
            qq0 += s * s;
        }
        // the running sums fit 32 bits (|pair| <= 3N, N < 2^31 / 3); after the
        // reduction they are wave-uniform and live in scalar registers
        int pair = __builtin_amdgcn_readfirstlane((int)wave_sum(pair0));
        int mm = __builtin_amdgcn_readfirstlane((int)wave_sum(mm0));
        int qq = __builtin_amdgcn_readfirstlane((int)wave_sum(qq0));
        const double D = cr.D;

        // bin index as _binindex evaluates it (binned_object.jl:18-24): integer
        // energies (Ising) on an integer binning truncating div, everything
        // else round-half-even of (E - start) / step in FP64.  A power-of-two
        // step divides exactly as a multiply / shift.
        const bool step_pow2 = a.step_pow2 != 0;
        // (the walk's state is wave-uniform: bins and decisions are moved to
        // scalar registers so that its branches are scalar branches)
        auto bin_int = [&](long long Ei, int& b) -> bool {
            const long long d = Ei - (long long)a.bin_start;
            long long bb;
            if (step_pow2)
                bb = d >= 0 ? (d >> a.step_shift) : -((-d) >> a.step_shift);
            else
                bb = d / (long long)a.bin_step; // truncating div
            const bool in = bb >= 0 && bb < nb;
            b = __builtin_amdgcn_readfirstlane(in ? (int)bb : 0);
            return __builtin_amdgcn_readfirstlane(in ? 1 : 0) != 0;
        };
        auto bin_fp = [&](double Ef, int& b) -> bool {
            const double q = step_pow2 ? (Ef - a.bin_start) * a.inv_step
                                       : (Ef - a.bin_start) / a.bin_step;
            const double v = rint(q);
            const bool in = v >= 0.0 && v < (double)nb;
            b = __builtin_amdgcn_readfirstlane(in ? (int)v : 0);
            return __builtin_amdgcn_readfirstlane(in ? 1 : 0) != 0;
        };

        // Integer-energy walk: Ising always; Blume-Capel when D and the bin start
        // are integers and the step is 1, so that every FP64 energy is an exact
        // integer and round((E - start) / 1) = E - start -- the same bins and
        // the same chain as the FP64 walk, from scalar integer arithmetic.
        const double Dr = rint(D);
        // (32-bit: energies, offsets from the bin start and D * Q stay below 2^29)
        const double lim = 536870912.0;
        const bool unit_bins = a.bin_step == 1.0 && rint(a.bin_start) == a.bin_start;
        const bool int_e = fabs(a.bin_start) < lim && 3.0 * (double)N < lim &&
                           (MODEL == 0 ? step_pow2 && (a.int_bins || unit_bins)
                                       : (Dr == D && fabs(D) * (double)N < lim && unit_bins));
        // Ising energies on an integer binning take the truncating div everywhere
        const bool div_bins = MODEL == 0 && a.int_bins;
        const long long Di = (long long)Dr;
        // the integer walk (lane 0 walks E; the staging wave adds up each chunk)
        const bool int_walk = int_e && !PS;
        const long long start_i = (long long)a.bin_start;
        const int start32 = (int)start_i;
        const int smask = step_pow2 ? (1 << a.step_shift) - 1 : 0;
        // Linear bins: bin(E + dE) = bin(E) + dE / step exactly -- unit bins,
        // or Ising energies (which change by multiples of 4) on a step of 1, 2
        // or 4 that divides E - start.  The staging then stores each attempt's
        // bin change and the walk carries only the bin.
        const int lin_sh = MODEL == 0 ? a.step_shift : 0;
        const unsigned nbu = (unsigned)nb;
        auto ibin = [&](long long E, int& b) -> bool {
            if (MODEL == 0)
                return bin_int(E, b);
            const long long d = E - start_i;
            const bool in = d >= 0 && d < nb;
            b = in ? (int)d : 0;
            return in;
        };

        int64_t err = -1;
        int b_cur = -1;  // bin whose weight lw_cur holds (none yet)
        double lw_cur = 0.0;
        unsigned long long steps = 0, nacc = 0;
        // energy of the cached sums (_cached_energy) and its bin; recomputed
        // only when a move is accepted
        double Eo = MODEL == 0 ? -(double)pair : -(double)pair + D * (double)qq;
        long long Ei = -(long long)pair + (MODEL == 0 ? 0 : Di * (long long)qq);
        const bool lin = int_walk && nb < (1 << 27) &&
                         (MODEL == 1 || (lin_sh <= 2 && ((Ei - start_i) & smask) == 0));
        int b_old = 0;
        bool old_ok = PS ? bin_int((long long)qq, b_old)
                         : (int_e ? ibin(Ei, b_old)
                                  : (div_bins ? bin_int(Ei, b_old) : bin_fp(Eo, b_old)));
        const double nbeta = PS ? -a.pair_beta[wk] : 0.0;
        // ---- weight window: a chunk of 64 attempts moves the walker by at most
        // 64 * max_dbin bins, so a window that holds `reach` bins on both sides of
        // the chunk's starting bin serves every weight the chunk reads ----
        const int reach = 66 * a.max_dbin;
        int wbase = 0;
        auto place_window = [&](int bcur) {
            int nbase = bcur - wsz / 2;
            nbase = nbase < 0 ? 0 : nbase;
            nbase = nbase > (int)nb - wsz ? (int)nb - wsz : nbase;
            return nbase;
        };
        if (wsz < nb && old_ok)
            wbase = place_window(b_old);
        for (int b = tid; b < wsz; b += kGenThreads)
            lw[b] = lw_g[wbase + b];
        __syncthreads();
        // bin of Q (integer binning, truncating div), lane-0 state
        auto qbin0 = [&](int Qv, int& b) -> bool {
            const int d = Qv - start32;
            int bb;
            if (step_pow2)
                bb = d >= 0 ? (d >> a.step_shift) : -((-d) >> a.step_shift);
            else
                bb = d / (int)a.bin_step;
            const bool in = bb >= 0 && bb < nb;
            b = in ? bb : 0;
            return in;
        };
        // ---- wave 0: walk chunk [base, base + 64) of the colour from stage `buf`
        // (the window of weights first re-centred if the walker nears its edge) ----
        auto walk_chunk = [&](int base, int buf, int n_col) {
            if (wsz < nb) {
                const int bcur = __builtin_amdgcn_readfirstlane(b_old);
                const bool okc = __builtin_amdgcn_readfirstlane(old_ok ? 1 : 0) != 0;
                if (okc && (bcur - wbase < reach || wbase + wsz - 1 - bcur < reach)) {
                    const int nbase = place_window(bcur);
                    if (nbase != wbase) {
                        wave_sync();
                        // the old window goes back whole before the new one
                        // is read: the two overlap
                        if (WL) {
                            for (int b = lane; b < wsz; b += 64)
                                lw_g[wbase + b] = lw[b];
                            wave_sync();
                        }
                        for (int b = lane; b < wsz; b += 64)
                            lw[b] = lw_g[nbase + b];
                        wbase = nbase;
                        wave_sync();
                    }
                }
            }
            const int cnt = n_col - base < 64 ? n_col - base : 64;
            // ---- walk the attempts in site order (wave-uniform state) ----
            int my_vis = -1, my_ok = 0, n_done = cnt;
            if (PS) {
                // lane 0 walks the state (P, Q); every attempt evaluates the
                // old and new weights as the reference does (the old one
                // is the last accepted new one: lw does not change during
                // a multicanonical walk)
                int err_v = -1;
                if (lane == 0) {
                    int kk = 0;
                    if (!old_ok) {
                        n_done = 0;
                        err_v = st_site[buf][0];
                        kk = cnt;
                    }
                    double w_cur = kk < cnt ? nbeta * (double)pair + lw[b_old - wbase] : 0.0;
                    for (int k = kk; k < cnt; ++k) {
                        const int Pn = pair + st_dp[buf][k];
                        const int Qn = qq + st_dE[buf][k];
                        int bn;
                        if (!qbin0(Qn, bn)) {
                            n_done = k;
                            err_v = st_site[buf][k];
                            break;
                        }
                        const double w_new = nbeta * (double)Pn + lw[bn - wbase];
                        const double lr = w_new - w_cur;
                        // lr > 0 || u < exp(lr), exp only inside the 1e-9 band
                        bool ok = lr > 0 || lr > st_luh[buf][k];
                        if (!ok && !(lr < st_lul[buf][k]))
                            ok = st_u[buf][k] < exp(lr);
                        ++steps;
                        nacc += ok ? 1u : 0u;
                        if (ok) {
                            const int mq = st_mq[buf][k];
                            pair = Pn;
                            qq = Qn;
                            mm += (mq & 0xffff) - 8;
                            b_old = bn;
                            w_cur = w_new;
                        }
                        st_res[buf][k] = b_old | (ok ? (int)0x80000000 : 0);
                    }
                }
                n_done = __builtin_amdgcn_readfirstlane(n_done);
                err_v = __builtin_amdgcn_readfirstlane(err_v);
                if (err_v >= 0)
                    err = err_v;
                wave_sync();
                if (lane < n_done) {
                    const int r = st_res[buf][lane];
                    my_vis = r & 0x7fffffff;
                    my_ok = (r >> 31) & 1;
                }
            } else if (int_e) {
                // Integer walk on lane 0 alone: its state stays in vector
                // registers and every step is a VALU or LDS operation (no
                // scalar round trips, no per-attempt branches).  The
                // walker always sits in the bin it last visited, so
                // lw_cur holds that bin's weight; attempt k+1's staged
                // values are loaded while k decides.  The weight of k+1's
                // proposed bin is written as a choice between the two
                // outcomes of k, which the compiler turns into one load from
                // the chosen address (loading both and choosing measured
                // slower: more instructions on the chain than LDS latency
                // saved).  An
                // attempt out of the binned domain, or one whose lr lies
                // in the 1e-9 band around log(u), stops the fast loop; it
                // is settled exactly (with exp) and the loop resumes.  The
                // loop records each attempt's bin and outcome; the counters
                // and the running sums are added up from those records by
                // all lanes after the chunk.
                int err_v = -1;
                if (lane == 0 && lin) {
                    // Linear bins: the state is the byte offset Ea = 8 * bin of
                    // the current bin's weight and an attempt proposes Ea plus
                    // its staged offset change, so that the next weight's LDS
                    // address is one add.  No domain test in the fast loop: a
                    // chunk's proposals stay inside the window except past the
                    // ends of the table, where they read the NaN padding slots
                    // (lr = NaN fails both compares and stops the loop; the
                    // exact step below raises the bounds error).
                    if (b_cur != b_old && old_ok) {
                        lw_cur = lw[b_old - wbase];
                        b_cur = b_old;
                    }
                    // weight at byte offset e: *(lwb + e)
                    const char* lwb = reinterpret_cast<const char*>(lw - wbase);
                    char* lwbw = reinterpret_cast<char*>(lw - wbase);
                    auto wt = [&](int e) { return *reinterpret_cast<const double*>(lwb + e); };
                    int Ea = 8 * b_old;
                    int kk = 0;
                    if (!old_ok) {
                        n_done = 0;
                        err_v = st_site[buf][0];
                        kk = cnt;
                    }
                    while (kk < cnt) {
                        int stop = cnt;
                        int dB_c = st_dE[buf][kk];
                        double luh_c = st_luh[buf][kk], lul_c = st_lul[buf][kk];
                        double lw_pre = wt(Ea + dB_c);
                        // (a walking wave issues about one instruction per 7
                        // cycles, so the loop's own counter and pointer updates
                        // count: unrolled by 16 they are ~1 of ~20 per attempt)
                        if (!WL) {
                            // Multicanonical: the weights do not change, so the
                            // pass need not stop at an attempt in the band: it
                            // notes that one occurred, and the first such
                            // attempt is found afterwards from the records (the
                            // walk's state before attempt k is the bin recorded
                            // for k-1, and lw_cur is always that bin's weight)
                            const int Ea0 = Ea;
                            bool band = false;
#pragma unroll 16
                            for (int k = kk; k < cnt; ++k) {
                                const int dB_n = st_dE[buf][k + 1];
                                const double luh_n = st_luh[buf][k + 1], lul_n = st_lul[buf][k + 1];
                                const int En = Ea + dB_c;
                                const double lr = lw_pre - lw_cur;
                                const bool acc_c = lr > luh_c;
                                band = band | !(acc_c | (lr < lul_c));
                                Ea = acc_c ? En : Ea;
                                lw_cur = acc_c ? lw_pre : lw_cur;
                                st_res[buf][k] = Ea | (acc_c ? (int)0x80000000 : 0);
                                lw_pre = wt(Ea + dB_n);
                                dB_c = dB_n;
                                luh_c = luh_n;
                                lul_c = lul_n;
                            }
                            if (!band)
                                break;
                            int Es = Ea0;
                            for (stop = kk; stop < cnt; ++stop) {
                                const double lr = wt(Es + st_dE[buf][stop]) - wt(Es);
                                if (!(lr > st_luh[buf][stop]) && !(lr < st_lul[buf][stop]))
                                    break;
                                Es = st_res[buf][stop] & 0x7fffffff;
                            }
                            Ea = Es;
                            lw_cur = wt(Es);
                        } else {
                            // Wang-Landau: the pass likewise runs every attempt,
                            // one in the band taken as a rejection, and keeps
                            // each visited bin's weight from before its decrement;
                            // after a band attempt the pass is undone (weights
                            // restored last first) and replayed exactly up to it
                            const int Ea0 = Ea;
                            const double lw0 = lw_cur;
                            bool band = false;
#pragma unroll 16
                            for (int k = kk; k < cnt; ++k) {
                                const int dB_n = st_dE[buf][k + 1];
                                const double luh_n = st_luh[buf][k + 1], lul_n = st_lul[buf][k + 1];
                                const int En = Ea + dB_c;
                                // the current bin's weight has just been lowered
                                // after k+1's was read
                                const double lw_new = En == Ea ? lw_cur : lw_pre;
                                const double lr = lw_new - lw_cur;
                                const bool acc_c = lr > luh_c;
                                band = band | !(acc_c | (lr < lul_c));
                                Ea = acc_c ? En : Ea;
                                lw_cur = acc_c ? lw_new : lw_cur;
                                st_res[buf][k] = Ea | (acc_c ? (int)0x80000000 : 0);
                                lw_pre = wt(Ea + dB_n);
                                st_old[k] = lw_cur;
                                lw_cur = lw_cur - logf;
                                *reinterpret_cast<double*>(lwbw + Ea) = lw_cur;
                                dB_c = dB_n;
                                luh_c = luh_n;
                                lul_c = lul_n;
                            }
                            if (!band)
                                break;
                            for (int k = cnt - 1; k >= kk; --k)
                                *reinterpret_cast<double*>(lwbw + (st_res[buf][k] & 0x7fffffff)) =
                                    st_old[k];
                            Ea = Ea0;
                            lw_cur = lw0;
                            for (stop = kk; stop < cnt; ++stop) {
                                const int En = Ea + st_dE[buf][stop];
                                const double lw_new = En == Ea ? lw_cur : wt(En);
                                const double lr = lw_new - lw_cur;
                                const bool acc_c = lr > st_luh[buf][stop];
                                if (!acc_c && !(lr < st_lul[buf][stop]))
                                    break;
                                if (acc_c) {
                                    Ea = En;
                                    lw_cur = lw_new;
                                }
                                st_res[buf][stop] = Ea | (acc_c ? (int)0x80000000 : 0);
                                lw_cur = lw_cur - logf;
                                *reinterpret_cast<double*>(lwbw + Ea) = lw_cur;
                            }
                        }
                        if (stop == cnt)
                            break;
                        // the attempt that stopped the fast loop, exactly
                        const int En = Ea + st_dE[buf][stop];
                        if (!((unsigned)En < 8u * nbu)) {
                            n_done = stop;
                            err_v = st_site[buf][stop];
                            break;
                        }
                        const double lw_new = En == Ea ? lw_cur : wt(En);
                        const double lr = lw_new - lw_cur;
                        const bool ok = lr > 0 || st_u[buf][stop] < exp(lr);
                        if (ok) {
                            Ea = En;
                            lw_cur = lw_new;
                        }
                        st_res[buf][stop] = Ea | (ok ? (int)0x80000000 : 0);
                        if (WL) {
                            lw_cur = lw_cur - logf;
                            *reinterpret_cast<double*>(lwbw + Ea) = lw_cur;
                        }
                        kk = stop + 1;
                    }
                    b_old = Ea >> 3;
                    b_cur = b_old;
                    Ei = start_i + ((long long)b_old << lin_sh);
                } else if (lane == 0) {
                    // bin of an energy (32-bit): Ising floor-free div by 2^s
                    // rounding toward zero as div does; Blume-Capel E - start
                    auto bin32 = [&](int E) -> int {
                        const int d = E - start32;
                        return MODEL == 0 ? (d + ((d >> 31) & smask)) >> a.step_shift : d;
                    };
                    auto binc = [&](int E) -> int { // window index, clamped into it
                        const int b = bin32(E) - wbase;
                        return max(0, min(b, wsz - 1));
                    };
                    int E32 = (int)Ei;
                    if (b_cur != b_old && old_ok) {
                        lw_cur = lw[b_old - wbase];
                        b_cur = b_old;
                    }
                    int kk = 0;
                    if (!old_ok) {
                        // the walker's current energy lies outside the binned
                        // domain: its first attempt fails
                        n_done = 0;
                        err_v = st_site[buf][0];
                        kk = cnt;
                    }
                    while (kk < cnt) {
                        int stop = cnt;
                        bool alive = true;
                        int dE_c = st_dE[buf][kk];
                        double luh_c = st_luh[buf][kk], lul_c = st_lul[buf][kk];
                        double lw_pre = lw[binc(E32 + dE_c)];
#pragma unroll 8
                        for (int k = kk; k < cnt; ++k) {
                            const int kn = k + 1;
                            const int dE_n = st_dE[buf][kn];
                            const double luh_n = st_luh[buf][kn], lul_n = st_lul[buf][kn];
                            const int En = E32 + dE_c;
                            const int bb = bin32(En);
                            const bool okn = (unsigned)bb < nbu;
                            const int b_new = okn ? bb : 0;
                            const double lw_a = lw[binc(En + dE_n)];
                            const double lw_r = lw[binc(E32 + dE_n)];
                            // a multicanonical walk never writes the weights, so
                            // the prefetched weight is exact even for the current
                            // bin; a Wang-Landau walk has just lowered that one
                            const double lw_new = !WL ? lw_pre : (b_new == b_old ? lw_cur : lw_pre);
                            const double lr = lw_new - lw_cur;
                            const bool acc_c = lr > luh_c;
                            const bool go = alive && okn && (acc_c || lr < lul_c);
                            const bool ok = go && acc_c;
                            stop = alive && !go ? k : stop;
                            alive = go;
                            E32 = ok ? En : E32;
                            b_old = ok ? b_new : b_old;
                            lw_cur = ok ? lw_new : lw_cur;
                            st_res[buf][k] = b_old | (ok ? (int)0x80000000 : 0);
                            if (WL) {
                                // unchanged value when the attempt did not run
                                lw_cur = go ? lw_cur - logf : lw_cur;
                                lw[b_old - wbase] = lw_cur;
                            }
                            lw_pre = ok ? lw_a : lw_r;
                            dE_c = dE_n;
                            luh_c = luh_n;
                            lul_c = lul_n;
                        }
                        b_cur = b_old;
                        if (stop == cnt)
                            break;
                        // the attempt that stopped the fast loop, exactly
                        const int En = E32 + st_dE[buf][stop];
                        const int bb = bin32(En);
                        if (!((unsigned)bb < nbu)) {
                            n_done = stop;
                            err_v = st_site[buf][stop];
                            break;
                        }
                        const double lw_new = bb == b_old ? lw_cur : lw[bb - wbase];
                        const double lr = lw_new - lw_cur;
                        const bool ok = lr > 0 || st_u[buf][stop] < exp(lr);
                        if (ok) {
                            E32 = En;
                            b_old = bb;
                            lw_cur = lw_new;
                        }
                        st_res[buf][stop] = b_old | (ok ? (int)0x80000000 : 0);
                        if (WL) {
                            lw_cur = lw_cur - logf;
                            lw[b_old - wbase] = lw_cur;
                        }
                        b_cur = b_old;
                        kk = stop + 1;
                    }
                    Ei = E32;
                }
                n_done = __builtin_amdgcn_readfirstlane(n_done);
                err_v = __builtin_amdgcn_readfirstlane(err_v);
                if (err_v >= 0)
                    err = err_v;
                // the chunk's spin writes, counters, running sums and visit
                // counts are done by the staging wave during the next walk
                // (my_ok stays 0 here: nothing is deferred below)
                if (lane == 0)
                    sh_nd[buf] = n_done;
            } else
            for (int k = 0; k < cnt; ++k) {
                int b_new;
                bool okn;
                if (div_bins)
                    okn = bin_int(-(long long)pair + __builtin_amdgcn_readfirstlane(st_dE[buf][k]), b_new);
                else
                    okn = bin_fp(Eo + uniform_f64(st_lul[buf][k]), b_new);
                if (!okn || !old_ok) {
                    err = st_site[buf][k];
                    n_done = k;
                    break;
                }
                const double lw_old = b_old == b_cur ? lw_cur : lw[b_old - wbase];
                const double lw_new = b_new == b_old ? lw_old : lw[b_new - wbase];
                const double lr = lw_new - lw_old;
                // lr > 0 || u < exp(lr); log(u) from the staging settles
                // all but a 1e-9-wide band, where exp decides exactly
                bool ok = __builtin_amdgcn_readfirstlane(lr > 0 ? 1 : 0) != 0;
                if (!ok) {
                    const double lu = uniform_f64(st_luh[buf][k]);
                    // 1 = certain accept, 0 = certain reject, 2 = evaluate exp
                    const int cls = lr > lu + 1e-9 && lu > -700.0 ? 1
                                  : (lr < lu - 1e-9 ? 0 : 2);
                    const int clsu = __builtin_amdgcn_readfirstlane(cls);
                    if (clsu == 2)
                        ok = __builtin_amdgcn_readfirstlane(
                                 uniform_f64(st_u[buf][k]) < exp(lr) ? 1 : 0) != 0;
                    else
                        ok = clsu == 1;
                }
                ++steps;
                nacc += ok ? 1u : 0u;
                const int vis = ok ? b_new : b_old;
                if (lane == k) {
                    my_vis = vis;
                    my_ok = ok ? 1 : 0;
                }
                double lw_vis = ok ? lw_new : lw_old;
                if (wl) {
                    lw_vis = lw_vis - logf;
                    if (lane == 0)
                        lw[vis - wbase] = lw_vis;
                }
                if (ok) {
                    pair += __builtin_amdgcn_readfirstlane(st_dp[buf][k]);
                    mm += __builtin_amdgcn_readfirstlane((st_mq[buf][k] & 0xffff) - 8);
                    qq += __builtin_amdgcn_readfirstlane((st_mq[buf][k] >> 16) - 8);
                    if (MODEL == 0) {
                        // (Ising energies are exact in FP64: the new bin is the
                        // bin of the new cached energy)
                        b_old = b_new;
                        Eo = -(double)pair;
                    } else {
                        Eo = -(double)pair + D * (double)qq;
                        old_ok = bin_fp(Eo, b_old);
                    }
                }
                b_cur = vis;
                lw_cur = lw_vis;
            }
            // ---- deferred, in parallel: spins and visit counts (order-free) ----
            if (lane < n_done) {
                if (my_ok)
                    sp[st_site[buf][lane]] = (int8_t)st_snew[buf][lane];
                if (!int_walk) {
                    if (a.hist_in_lds)
                        atomicAdd(&hl[my_vis], 1u);
                    else
                        atomicAdd(&hist_g[my_vis], 1ull);
                }
            }
        };
        // ---- staging wave, integer walk: the spin writes, counters, running
        // sums and visit counts of a walked chunk (stage `pb`), from its
        // recorded outcomes (every attempt before n_done ran; accepted ones
        // move).  The spins are written before this wave stages again, and
        // staging reads them only from the other colour or from sites of
        // later chunks. ----
        auto add_up = [&](int pb) {
            const int nd = __builtin_amdgcn_readfirstlane(sh_nd[pb]);
            int my_ok = 0, my_vis = 0;
            if (lane < nd) {
                // (the linear walk records weight byte offsets, 8 * bin)
                const int r = st_res[pb][lane];
                my_vis = (r & 0x7fffffff) >> (lin ? 3 : 0);
                my_ok = (r >> 31) & 1;
            }
            const int s_ok = wave_sum32(my_ok);
            pair += wave_sum32(my_ok ? st_dp[pb][lane] : 0);
            mm += wave_sum32(my_ok ? (st_mq[pb][lane] & 0xffff) - 8 : 0);
            qq += wave_sum32(my_ok ? (st_mq[pb][lane] >> 16) - 8 : 0);
            steps += (unsigned)nd;
            nacc += (unsigned)s_ok;
            if (lane < nd) {
                if (my_ok)
                    sp[st_site[pb][lane]] = (int8_t)st_snew[pb][lane];
                if (a.hist_in_lds)
                    atomicAdd(&hl[my_vis], 1u);
                else
                    atomicAdd(&hist_g[my_vis], 1ull);
            }
        };
        // ---- stage chunk [base, base + 64) of colour c into stage `buf`: lane k
        // prepares attempt base + k (the sites of one colour are independent
        // given the other colour) ----
        auto stage = [&](int c, int base, int buf, uint64_t sweep) {
            const int n_col = rows * half;
            const int o = base + lane;
            int t_site = 0, t_snew = 0, t_dE = 0, t_dpair = 0, t_dm = 0, t_dq = 0;
            double t_dEf = 0.0, t_u = 0.0, t_lu = 0.0;
            int t_mq = 8 | (8 << 16); // packed dm+8 | (dq+8) << 16
            double t_luh = INFINITY, t_lul = -INFINITY;
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
                const uint32_t w = site_word(cr.seed, cr.key, g, j & 31, sweep, P_GEN);
                t_site = (int)(rowb + x);
                if (MODEL == 0) {
                    t_dE = 2 * si * nbs;
                    t_dpair = -2 * si * nbs;
                    t_dm = -2 * si;
                    t_snew = -si;
                    t_dEf = (double)t_dE;
                    t_u = (double)w * 0x1.0p-32;
                } else {
                    const int sn = gen_propose(si, (int)(w & 1u));
                    const int ds = sn - si;
                    const int ds2 = sn * sn - si * si;
                    t_dEf = (-((double)ds * (1.0 * (double)nbs)) - 0.0) + D * (double)ds2;
                    t_dE = int_e ? -(ds * nbs) + (int)Di * ds2 : 0;
                    t_dpair = ds * nbs;
                    t_dm = ds;
                    t_dq = ds2;
                    t_snew = sn;
                    t_u = (double)(w >> 1) * 0x1.0p-31;
                }
                t_lu = log(t_u);
                // u < exp(lr) is certain for lr > log(u) + 1e-9 and false for
                // lr < log(u) - 1e-9 (log and exp err by ~1e-16 relative)
                t_luh = t_lu > -700.0 ? t_lu + 1e-9 : INFINITY;
                t_lul = t_lu - 1e-9;
                t_mq = (t_dm + 8) | ((t_dq + 8) << 16);
            }
            st_site[buf][lane] = t_site;
            st_snew[buf][lane] = t_snew;
            st_dp[buf][lane] = t_dpair;
            st_mq[buf][lane] = t_mq;
            st_u[buf][lane] = t_u;
            if (PS) {
                st_dE[buf][lane] = t_dq;
                st_luh[buf][lane] = t_luh;
                st_lul[buf][lane] = t_lul;
            } else if (int_e) {
                st_dE[buf][lane] = lin ? 8 * (t_dE >> lin_sh) : t_dE;
                // lr > 0 || lr > log(u) + 1e-9  <=>  lr > min(0, log(u) + 1e-9)
                st_luh[buf][lane] = t_luh < 0.0 ? t_luh : 0.0;
                st_lul[buf][lane] = t_lul;
            } else {
                st_dE[buf][lane] = t_dE;
                st_luh[buf][lane] = t_lu;
                st_lul[buf][lane] = t_dEf;
            }
        };

        // barriers entered so far in the walk (parity selects the sh_err slot)
        int nbar = 0;
        // staging wave, integer walk: the stage whose chunk was walked last and
        // is not yet added up (-1: none).  It is added up before that stage is
        // staged again, and at the end of each sweep before the series record.
        int pend = -1;
        int64_t s = 0;
        for (; s < a.nsweeps && err < 0; ++s) {
            const uint64_t sweep = cr.sweep + (uint64_t)s;
            for (int c = 0; c < 2 && err < 0; ++c) {
                const int n_col = rows * half;
                // the colour's first chunk, once the other colour's spins are written
                if (wid == 1) {
                    if (pend >= 0)
                        add_up(pend);
                    pend = -1;
                    stage(c, 0, 0, sweep);
                }
                __syncthreads();
                ++nbar;
                for (int base = 0, buf = 0; base < n_col && err < 0; base += 64, buf ^= 1) {
                    if (wid == 1) {
                        // the previous chunk's stage is added up before it is
                        // staged again
                        if (pend >= 0)
                            add_up(pend);
                        pend = int_walk ? buf : -1;
                        if (base + 64 < n_col)
                            stage(c, base + 64, buf ^ 1, sweep);
                    } else {
                        walk_chunk(base, buf, n_col);
                        if (err >= 0 && lane == 0)
                            sh_err[nbar & 1] = (int)err;
                    }
                    __syncthreads();
                    err = __builtin_amdgcn_readfirstlane(sh_err[nbar & 1]);
                    ++nbar;
                }
            }
            if (wid == 1 && pend >= 0) {
                add_up(pend);
                pend = -1;
            }
            // (the integer walk's running sums are the staging wave's)
            if (err < 0 && a.record && tid == (int_walk ? 64 : 0)) {
                const int64_t at = cr.series_off + a.rec_pos + s;
                a.ser_E[at] = (int32_t)(-pair);
                a.ser_M[at] = (int32_t)mm;
                if (a.ser_Q)
                    a.ser_Q[at] = (int32_t)qq;
            }
        }
        __syncthreads();
        for (int64_t i = tid; i < N; i += kGenThreads)
            sp_g[i] = sp[i];
        // (the window's position is wave 0's state)
        if (wl && wid == 0)
            for (int b = lane; b < wsz; b += 64)
                lw_g[wbase + b] = lw[b];
        if (a.hist_in_lds)
            for (int64_t b = tid; b < nb; b += kGenThreads)
                if (hl[b])
                    hist_g[b] += hl[b];
        if (tid == (int_walk ? 64 : 0)) {
            a.steps[wk] += steps;
            a.accepted[wk] += nacc;
            a.err_site[wk] = err;
            // a failed sweep is not completed: the counter stays at its start
            a.chains[wk].sweep = cr.sweep + (uint64_t)(err < 0 ? s : s - 1);
        }
        __syncthreads();
    }
}

// kind -> KIND: 2 multicanonical -> 0, 3 Wang-Landau -> 1, 5 pair + sum s^2 -> 2
template <int M, int D>
static void gen_launch(const GenArgs& a, size_t lds_bytes, hipStream_t st)
{
    const dim3 grid(a.n_items), block(kGenThreads);
    if (a.kind == 3)
        hipLaunchKernelGGL((k_gen_sweeps<M, D, 1>), grid, block, lds_bytes, st, a);
    else if constexpr (M == 1) {
        if (a.kind == 5)
            hipLaunchKernelGGL((k_gen_sweeps<1, D, 2>), grid, block, lds_bytes, st, a);
        else
            hipLaunchKernelGGL((k_gen_sweeps<M, D, 0>), grid, block, lds_bytes, st, a);
    } else
        hipLaunchKernelGGL((k_gen_sweeps<M, D, 0>), grid, block, lds_bytes, st, a);
}

hipError_t launch_gen_sweeps(const GenArgs& a, size_t lds_bytes, hipStream_t st)
{
    if (a.model == 0 && a.dim == 2)
        gen_launch<0, 2>(a, lds_bytes, st);
    else if (a.model == 0)
        gen_launch<0, 3>(a, lds_bytes, st);
    else if (a.dim == 2)
        gen_launch<1, 2>(a, lds_bytes, st);
    else
        gen_launch<1, 3>(a, lds_bytes, st);
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

// Histogram statistics of a walker set (the group's walkers, or its root
// alone): per bin the sum of their histograms in 64-bit integers; a bin counts
// once it has been seen non-zero by any evaluation (`reached`, so that a bin
// the walkers left stays in the flatness test), and over the counted bins each
// block writes {min, sum, count} partials.  Integer sums: the result does not
// depend on the order of the additions.
__global__ void k_gen_hist_stats(const unsigned long long* hist, const int64_t* lw_off,
                                 const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                                 uint8_t* reached, unsigned long long* part)
{
    __shared__ unsigned long long smin[4], ssum[4], scnt[4];
    unsigned long long mn = ~0ull, sm = 0, ct = 0;
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long h = 0;
        for (int k = 0; k < n_walkers; ++k)
            h += hist[lw_off[walkers[k]] + b];
        bool r = reached[b] != 0;
        if (h > 0 && !r) {
            reached[b] = 1;
            r = true;
        }
        if (r) {
            mn = h < mn ? h : mn;
            sm += h;
            ct += 1;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(mn, off, 64);
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
        part[3 * blockIdx.x] = mn;
        part[3 * blockIdx.x + 1] = sm;
        part[3 * blockIdx.x + 2] = ct;
    }
}

// out = {min, mean, bins} over the counted bins (0, 0, 0 when none)
__global__ void k_gen_hist_stats_final(const unsigned long long* part, int n_part, double* out)
{
    if (threadIdx.x != 0)
        return;
    unsigned long long mn = ~0ull, sm = 0, ct = 0;
    for (int k = 0; k < n_part; ++k) {
        mn = part[3 * k] < mn ? part[3 * k] : mn;
        sm += part[3 * k + 1];
        ct += part[3 * k + 2];
    }
    out[0] = ct > 0 ? (double)mn : 0.0;
    out[1] = ct > 0 ? (double)sm / (double)ct : 0.0;
    out[2] = (double)ct;
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

hipError_t launch_gen_hist_stats(const unsigned long long* hist, const int64_t* lw_off,
                                 const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                                 uint8_t* reached, unsigned long long* part, double* out,
                                 hipStream_t st)
{
    int64_t blocks = (nbins + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > kGenStatBlocks ? kGenStatBlocks : blocks);
    hipLaunchKernelGGL(k_gen_hist_stats, dim3((unsigned)blocks), dim3(256), 0, st, hist, lw_off,
                       walkers, n_walkers, nbins, reached, part);
    hipLaunchKernelGGL(k_gen_hist_stats_final, dim3(1), dim3(64), 0, st, part, (int)blocks, out);
    return hipGetLastError();
}

} // namespace mcx
