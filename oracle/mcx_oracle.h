/*
 * mcx_oracle.h -- CPU oracle for the MonteCarloX spin-lattice sampling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Everything here is a straight, sequential restatement of the reference's
 * per-site update rules (zierenberg/MonteCarloX.jl + SpinSystems) under the
 * arithmetic contract the HIP library implements (checkerboard site order and
 * counter-based Philox4x32-10 streams instead of a mutable Xoshiro state).
 * Each function cites the reference lines it follows.
 */
#ifndef MCX_ORACLE_H
#define MCX_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random-stream purposes (top byte of counter word 2). */
enum {
    ORC_P_METRO = 1,   /* Ising Metropolis acceptance, bit-sliced uniform   */
    ORC_P_HB = 2,      /* Ising heat-bath draw, bit-sliced uniform          */
    ORC_P_INIT = 3,    /* random initial state, site word                   */
    ORC_P_BC_METRO = 4,/* Blume-Capel Metropolis / Glauber: proposal + U31  */
    ORC_P_BC_HB = 5,   /* Blume-Capel heat bath: 32-bit U                   */
    ORC_P_GEN = 6,     /* multicanonical / Wang-Landau acceptance, site word */
    ORC_P_PT = 8       /* replica-exchange shared uniform                   */
};

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Site -> (group, bit) map of the checkerboard layout. */
void orc_site_group(int dim, int L, int64_t site, uint32_t* group, int* bit, int* colour);
void orc_ising_group(int dim, int L, int64_t site, uint32_t* group, int* bit, int* colour);

/* 32-bit uniforms of the contract. */
uint32_t orc_bitsliced_word(uint64_t seed, uint32_t chain_key, uint32_t group, int bit,
                            uint64_t sweep, int purpose);
uint32_t orc_site_word(uint64_t seed, uint32_t chain_key, uint32_t group, int bit,
                       uint64_t sweep, int purpose);
uint32_t orc_pt_word(uint64_t seed, uint32_t chain_key, uint64_t round);
/* Blume-Capel site uniform: the site's 16-bit field f (word (bit & 15) % 4,
 * half bit >> 4, of the block at counter (8*group + (bit & 15) / 4, sweep)),
 * u16 = f ^ 0x8000, and the tail t16 = low 16 bits of word 0 of the block at
 * (8*group + 5, sweep, chunk 1 + bit).  Returns U = u16 * 2^16 + t16; *alt
 * (if not NULL) = f & 1, the Metropolis proposal bit, whose uniform is
 * U31 = (u16 >> 1) * 2^16 + t16. */
uint32_t orc_bc_site_uniform(uint64_t seed, uint32_t chain_key, uint32_t group, int bit,
                             uint64_t sweep, int purpose, int* alt);

/* Helpers of the reference arithmetic. */
double orc_logistic(double x);
uint64_t orc_threshold32(double p);   /* ceil(p * 2^32), p in [0, 1] */

/* Observables: E = -J sum_<ij> s_i s_j (+ D sum s^2), M = sum s_i. */
int64_t orc_pair_sum(int dim, int L, const int8_t* spins);      /* sum_<ij> s_i s_j */
int64_t orc_magnetization(int dim, int L, const int8_t* spins);
int64_t orc_sum_sq(int dim, int L, const int8_t* spins);

/* Random initial states (init!(sys, :random)). model: 0 Ising, 1 Blume-Capel */
void orc_init_random(int model, int dim, int L, uint64_t seed, uint32_t chain_key,
                     uint64_t sweep, int8_t* spins);

/* Canonical sweeps on one chain.
 *   model 0 (Ising): alg 0 Metropolis, 1 heat bath, 2 Glauber.
 *   model 1 (Blume-Capel): alg 0 Metropolis, 1 heat bath, 2 Glauber; D = crystal field.
 * Glauber (src/algorithms/metropolis.jl:121-127) accepts iff u < logistic(-beta dE),
 * u from the Metropolis stream of the model.
 * Per sweep, after both colours: E_pair[s] = -sum s_i s_j, M[s] = sum s_i,
 * Q[s] = sum s_i^2 (any of the three may be NULL).
 * *accepted accumulates accepted moves (Metropolis) / changed sites (heat bath). */
void orc_canonical_sweeps(int model, int alg, int dim, int L, double beta, double D,
                          uint64_t seed, uint32_t chain_key, uint64_t sweep0, int64_t nsweeps,
                          int8_t* spins, int32_t* E_pair, int32_t* M, int32_t* Q,
                          int64_t* accepted);

/* Generalised-ensemble (multicanonical / Wang-Landau) sweeps on one walker.
 * Bins: DiscreteBinning(start, step, num).  Ising energies use integer div
 * binning, Blume-Capel energies (Float64 in the reference) use round-half-even.
 * wl = 0: multicanonical (hist[visited] += 1 per attempt)
 * wl = 1: Wang-Landau (lw[visited] -= logf per attempt; hist still counted).
 * Returns 0, or 2 (bounds error: *err_site set; state as at the failing attempt). */
int orc_generalized_sweeps(int model, int wl, int dim, int L, double D,
                           double bin_start, double bin_step, int64_t nbins, int int_bins,
                           double* lw, int64_t* hist, double logf,
                           uint64_t seed, uint32_t chain_key, uint64_t sweep0, int64_t nsweeps,
                           int8_t* spins, int64_t* steps, int64_t* accepted,
                           int32_t* E_pair, int32_t* M, int32_t* Q, int64_t* err_site);

/* One replica-exchange round on one set of n replicas (1-based ladder indices
 * in `indices`, as in ReplicaExchange.indices).  beta_of_ladder[k-1] is the
 * beta of ladder index k; x[r] is replica r's energy.  steps/accepted have
 * n-1 entries.  *stage flips; round is the set's exchange-round counter. */
void orc_pt_round(int n, const double* beta_of_ladder, const double* x, int32_t* indices,
                  int32_t* steps, int32_t* accepted, int32_t* stage, uint64_t seed,
                  const uint32_t* chain_keys, uint64_t round);

/* exchange_log_ratio + _accept_exchange for Boltzmann ensembles
 * (replica_exchange.jl:110-115); returns accepted, *lr = log ratio. */
int orc_exchange_pair(double beta_i, double beta_j, double x_i, double x_j, double u, double* lr);

/* Reference CPU path on one 2D chain recording the cached E / M after every
 * sweep of N random-site attempts (start: spins, modified in place). */
void orc_refpath_series_ising2d(int L, double beta, uint64_t seed, int64_t nsweeps, int8_t* spins,
                                int32_t* E_pair, int32_t* M, int64_t* accepted);

/* Multicanonical weight recursion: lw[i] -= log(h[i]) where h[i] > 0. */
void orc_muca_update(int64_t nbins, double* lw, const double* h);

/* Blume-Capel (J = 1, h = 0) under Metropolis(rng, CustomEnsemble(BoltzmannEnsemble(beta),
 * MulticanonicalEnsemble(bins over sum s_i^2))): the spin_flip! / accept! of
 * docs/src/examples/spin_systems/muca_BlumeCapel.jl:43-89, in checkerboard order
 * with the site words of ORC_P_GEN.  Integer binning of sum s^2 (truncating div).
 * Returns 0, or 2 with *err_site when sum s^2 leaves the binned domain. */
int orc_pair_spin2_sweeps(int dim, int L, double beta, double bin_start, double bin_step,
                          int64_t nbins, const double* lw, int64_t* hist, uint64_t seed,
                          uint32_t chain_key, uint64_t sweep0, int64_t nsweeps, int8_t* spins,
                          int64_t* steps, int64_t* accepted, int32_t* E_pair, int32_t* M,
                          int32_t* Q, int64_t* err_site);

/* ---- reference CPU path (timed as the CPU baseline) ----------------------
 * Random-site single-spin-flip Metropolis exactly as spin_flip!(IsingLatticeOptim,
 * Metropolis) does it: i = rand(UInt) % N, rand() < exp(-beta*dE), exp per
 * uphill move; Xoshiro256++ per chain seeded by SplitMix64(seed + chain).
 * Runs `nchains` independent L x L chains over `nthreads` POSIX threads.
 * Returns total attempts made. */
int64_t orc_refpath_ising2d(int L, int nchains, const double* betas, uint64_t seed,
                            int64_t nsweeps, int nthreads, int8_t* spins_out);

/* General reference CPU path (timed baseline only): nchains independent chains
 * of model (0 Ising, 1 Blume-Capel) in dim 2 or 3, per-chain L, beta and D
 * (D may be NULL), random-site Metropolis, or multicanonical weights lw over
 * DiscreteBinning(bin_start, bin_step, nbins) when muca == 1, or Wang-Landau
 * walkers starting from lw (each lowers its own copy by logf per visit) when
 * muca == 2.  Returns the attempts made. */
int64_t orc_refpath_run(int model, int dim, int muca, int nchains, const int32_t* L,
                        const double* betas, const double* D, const double* lw, int64_t nbins,
                        double bin_start, double bin_step, uint64_t seed, int64_t nsweeps,
                        int nthreads, double logf);

/* Integrated autocorrelation time (src/measurements/autocorrelations.jl:28-73)
 * of one integer series: x (|x| if absx), or the energy x + D*y when y is not
 * NULL; lag window 1..lag_cap (1 <= lag_cap <= n/2), window constant c.  The
 * centred lag sums are evaluated exactly in 128-bit integers,
 *   N(t) = n^2 sum_{i<n-t} (x_i - mean)(x_{i+t} - mean),
 * and C(t) = (N(t) / (n - t)) / (N(0) / n) in FP64. */
double orc_tau_int(const int32_t* x, const int32_t* y, int64_t n, double D, int absx,
                   int32_t lag_cap, double c);
/* The reference's own evaluation, for an FP64 series: centred samples,
 * C0 = dot(c, c) / n, cov(t) = dot(c[0:n-t], c[t:n]) / (n - t), sequential sums. */
double orc_tau_int_centred(const double* x, int64_t n, int32_t lag_cap, double c);

#ifdef __cplusplus
}
#endif
#endif
