/*
 * mcx_hip.h -- C ABI of the MI355X (gfx950) spin-lattice sampling library.
 *
 * The library runs the sampling hot path of MonteCarloX.jl + SpinSystems on
 * one GPU: checkerboard Metropolis / heat-bath sweeps over Ising (2D/3D) and
 * Blume-Capel lattices, replica exchange (parallel tempering) and
 * multicanonical / Wang-Landau walkers.  Every entry point is plain C: opaque
 * handles, plain pointers and sizes, no torch or HIP types.
 *
 * The reference has no FFI: its "plugin API" is Julia multiple dispatch on the
 * exported names.  A per-attempt call across an FFI cannot run at ns cost, so
 * each entry point below replaces a reference call at sweep / batch
 * granularity; the comment on each names the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns mcx_status (0 = MCX_OK).  On error,
 *     mcx_last_error() returns a thread-local message.
 *   - Status codes map to the reference's exception types:
 *       MCX_EINVAL -> ArgumentError   (e.g. replica_exchange.jl:54, binned_object.jl:50-53)
 *       MCX_EBOUNDS-> BoundsError     (bin out of range, test_multicanonical.jl:39-43)
 *       MCX_EDOMAIN-> AssertionError  (binned_object.jl:198-213)
 *       MCX_ENOMEM, MCX_EHIP -> ErrorException carrying mcx_last_error()
 *   - The library owns all device memory.  Host buffers are caller-owned and
 *     copied synchronously before the call returns.
 *   - A handle may be used from one host thread at a time; all device work of
 *     a context is ordered on that context's stream.
 *   - Spins are int8, sites numbered x-fastest: i = x + L*y (+ L*L*z), 0-based,
 *     periodic boundaries (IsingLatticeOptim layout, SpinSystems/src/ising.jl:444-456).
 */
#ifndef MCX_HIP_H
#define MCX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mcx_status;

#define MCX_OK 0
#define MCX_EINVAL 1
#define MCX_EBOUNDS 2
#define MCX_EDOMAIN 3
#define MCX_ENOMEM 4
#define MCX_EHIP 5
#define MCX_ESTATE 6

/* models */
#define MCX_ISING 0
#define MCX_BLUME_CAPEL 1

/* algorithms */
#define MCX_METROPOLIS 0
#define MCX_HEATBATH 1
#define MCX_MULTICANONICAL 2
#define MCX_WANGLANDAU 3
#define MCX_GLAUBER 4     /* Glauber(rng; beta): logistic acceptance, metropolis.jl:108-127 */
/* walker kind: Metropolis(rng, CustomEnsemble(BoltzmannEnsemble(beta), MulticanonicalEnsemble(bins
 * over sum s_i^2))) on Blume-Capel, docs/src/examples/spin_systems/muca_BlumeCapel.jl:43-89 */
#define MCX_MUCA_PAIR_SPIN2 5

/* initial states, init!(sys, type) */
#define MCX_INIT_UP 0
#define MCX_INIT_DOWN 1
#define MCX_INIT_ZERO 2
#define MCX_INIT_RANDOM 3

/* run flags */
#define MCX_RUN_RECORD 1u /* write per-sweep E/M(/Q) series and FP64 moments */
#define MCX_RUN_PLAIN 2u  /* one-chain-per-launch path: a kernel per colour per chain per sweep */
#define MCX_RUN_GRAPH 4u  /* mcx_pt_run: capture every round in one hipGraph and launch it once */

/* per-chain FP64 moments, in this order */
/* observables of the recorded series (mcx_batch_tau_int) */
#define MCX_OBS_ENERGY 0   /* E_pair + D*Q (Ising: E_pair) */
#define MCX_OBS_M 1        /* magnetization */
#define MCX_OBS_ABS_M 2    /* |magnetization| */
#define MCX_OBS_Q 3        /* sum s_i^2 (Blume-Capel) */

/* order of the recorded series under replica exchange (mcx_pt_set_series_order) */
#define MCX_ORDER_REPLICA 0     /* slot k = chain k */
#define MCX_ORDER_TEMPERATURE 1 /* slot set*R + i - 1 = whichever replica holds ladder index i */

#define MCX_MOM_COUNT 0
#define MCX_MOM_E 1
#define MCX_MOM_E2 2
#define MCX_MOM_ABSM 3
#define MCX_MOM_M2 4
#define MCX_MOM_M4 5
#define MCX_N_MOMENTS 6

typedef struct mcx_ctx mcx_ctx;
typedef struct mcx_batch mcx_batch;
typedef struct mcx_pt mcx_pt;
typedef struct mcx_gen mcx_gen;
typedef struct mcx_server mcx_server;

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */

/* No reference counterpart (device selection).  Creates a context bound to
 * HIP device `device` with its own stream. */
mcx_status mcx_ctx_create(int32_t device, mcx_ctx** out);
mcx_status mcx_ctx_destroy(mcx_ctx* ctx);
/* Thread-local text of the last error (never NULL). */
const char* mcx_last_error(void);
mcx_status mcx_version(int32_t* major, int32_t* minor);

/* ------------------------------------------------------------------------ */
/* Lattice batches: N independent chains                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t model;              /* MCX_ISING or MCX_BLUME_CAPEL                     */
    int32_t dim;                /* 2 or 3                                           */
    int32_t n_chains;           /* >= 1                                              */
    int32_t L_uniform;          /* side length when L == NULL                       */
    const int32_t* L;           /* per-chain even side length >= 4, or NULL         */
    uint64_t seed;              /* Philox key word 0 = low 32 bits, counter word 3 = high 32 */
    const uint32_t* chain_keys; /* per-chain Philox key word 1, or NULL = chain index */
    double D_uniform;           /* Blume-Capel crystal field when D == NULL         */
    const double* D;            /* per-chain crystal field, or NULL                  */
    int64_t series_capacity;    /* sweeps of per-sweep series kept per chain (>= 0) */
} mcx_batch_desc;

/* Replaces: one system per chain, IsingLatticeOptim(L, L) / Ising([L,L(,L)]; J=1)
 * (SpinSystems/src/ising.jl:413-424,440-461) or BlumeCapel([L,L]; J=1, D)
 * (SpinSystems/src/blume_capel.jl:500-511), plus one algorithm per chain as in
 * ParallelChains(ThreadsBackend(n), algs) (src/infrastructure/parallel_chains.jl:14-105).
 * Spins start all up (the constructors' ones(Int8, N)). */
mcx_status mcx_batch_create(mcx_ctx* ctx, const mcx_batch_desc* desc, mcx_batch** out);
mcx_status mcx_batch_destroy(mcx_batch* b);
mcx_status mcx_batch_num_sites(mcx_batch* b, int32_t chain, int64_t* n_sites);

/* Replaces: Metropolis(rng; beta) / HeatBath(rng; beta) per chain
 * (src/algorithms/metropolis.jl:95-96, heat_bath.jl:17): one beta >= 0 per
 * chain; the library tabulates exp(-beta*dE) and logistic(beta*s*dE) exactly
 * as logweight(BoltzmannEnsemble) (src/ensembles/boltzmann.jl:28) and
 * logistic (src/infrastructure/utils.jl:82-88) evaluate them. */
mcx_status mcx_batch_set_betas(mcx_batch* b, const double* beta_per_chain);

/* Host-built acceptance probabilities (so a caller can supply its own exp()
 * bits).  alg = MCX_METROPOLIS or MCX_HEATBATH; p has n_chains * n_per_chain
 * entries in the class order documented by mcx_batch_table_layout. */
mcx_status mcx_batch_set_tables(mcx_batch* b, int32_t alg, const double* p, int32_t n_per_chain);
/* Number of probabilities per chain for (alg) and the class convention:
 * Ising Metropolis: class k = number of anti-aligned neighbours, k < dim,
 *   p = exp(-beta * dE) with dE = 2*(2*dim - 2k)
 * Ising heat bath: class k = number of up neighbours, 0 <= k <= 2*dim,
 *   p = logistic(beta * 2*(2k - 2*dim))
 * Blume-Capel Metropolis (NS = 4*dim + 1 neighbour sums, 6*NS entries):
 *   class ((s_old+1)*2 + alt)*NS + (sum_nbr + 2*dim), alt = proposal bit,
 *   p = exp(-beta * dE) (p >= 1 accepts always)
 * Blume-Capel heat bath (3*NS entries): per sum_nbr (w1, w1 + w2, z) of
 *   SpinSystems/src/blume_capel.jl:66-76 */
mcx_status mcx_batch_table_layout(mcx_batch* b, int32_t alg, int32_t* n_per_chain);

/* Replaces init!(sys, :up/:down/:zero/:random; rng) (SpinSystems/src/ising.jl:74-78,
 * blume_capel.jl:105-109).  :random draws from the Philox INIT stream at each
 * chain's current sweep counter. */
mcx_status mcx_batch_init_spins(mcx_batch* b, int32_t mode);
/* Replaces direct access to sys.spins.  n must equal the chain's site count;
 * values must be +-1 (Ising) or -1/0/+1 (Blume-Capel), else MCX_EDOMAIN. */
mcx_status mcx_batch_set_spins(mcx_batch* b, int32_t chain, const int8_t* spins, size_t n);
mcx_status mcx_batch_get_spins(mcx_batch* b, int32_t chain, int8_t* spins, size_t n);

/* Replaces the user loop `for _ in 1:N*nsweeps; spin_flip!(sys, alg); end`
 * over every chain (SpinSystems/src/ising.jl:35-58, blume_capel.jl:52-85),
 * in checkerboard order.  alg = MCX_METROPOLIS or MCX_HEATBATH. */
mcx_status mcx_batch_run(mcx_batch* b, int32_t alg, int64_t nsweeps, uint32_t flags);
/* Device time (ms) of the last run / pt run, measured with HIP events on the
 * context stream around the kernels of that call. */
mcx_status mcx_batch_last_run_ms(mcx_batch* b, double* ms);

/* No reference counterpart (launch tuning).  The launch shape of the batched
 * sweep kernels for one algorithm (alg < 0: every algorithm), as an autotuner
 * picks it per (model, dim, L, algorithm); any shape gives the same results,
 * bit for bit -- it changes only how the chains are laid over workgroups:
 *   threads           workgroup size (a multiple of 64, <= 1024; 0 = built-in:
 *                     one lane per plane word, 64..256)
 *   chains_per_group  chains one workgroup runs one after another (>= 1; 0 = 1)
 *   lanes_per_word    Blume-Capel: lanes sharing one 32-site plane word
 *                     (1, 2, 4, 8; 0 = 1; -1 = the most, up to 4, that fit the
 *                     workgroup; a word draws 4 prefix blocks, so lanes past
 *                     the fourth idle); must be 0 for Ising */
typedef struct {
    int32_t threads;
    int32_t chains_per_group;
    int32_t lanes_per_word;
    int32_t pad;
} mcx_launch_shape;
mcx_status mcx_batch_set_launch_shape(mcx_batch* b, int32_t alg, const mcx_launch_shape* shape);
mcx_status mcx_batch_get_launch_shape(mcx_batch* b, int32_t alg, mcx_launch_shape* shape);

/* Replaces energy(sys) / magnetization(sys) (SpinSystems/src/ising.jl:17-18,
 * blume_capel.jl:18-19): per chain E_pair = -sum_<ij> s_i s_j, M = sum s_i,
 * Q = sum s_i^2 (any output may be NULL).  Energy = E_pair + D*Q. */
mcx_status mcx_batch_get_observables(mcx_batch* b, int32_t* E_pair, int32_t* M, int32_t* Q);

/* Replaces measure!(measurements, sys, i) of energy/magnetization after each
 * sweep (src/measurements/measurements.jl:192-200).  Entries [first, first+count)
 * of chain `chain`'s recorded series. */
mcx_status mcx_batch_series_length(mcx_batch* b, int64_t* n);
mcx_status mcx_batch_get_series(mcx_batch* b, int32_t chain, int64_t first, int64_t count,
                                int32_t* E_pair, int32_t* M, int32_t* Q);
/* Clears the series and the moments. */
mcx_status mcx_batch_reset_series(mcx_batch* b);
/* n_chains * MCX_N_MOMENTS doubles: count, sum E, sum E^2, sum |M|, sum M^2, sum M^4
 * over recorded sweeps, accumulated in FP64 on the device. */
mcx_status mcx_batch_get_moments(mcx_batch* b, double* out);

/* Replaces alg.steps / alg.accepted and reset!(alg)
 * (src/algorithms/importance_sampling.jl:80-110). */
mcx_status mcx_batch_get_counters(mcx_batch* b, int64_t* steps, int64_t* accepted);
mcx_status mcx_batch_reset_counters(mcx_batch* b);

/* Replaces the alg.rng field (Metropolis(rng; beta), src/algorithms/metropolis.jl:63-96):
 * rebinds the chains' Philox streams to key (seed & 0xffffffff, chain_keys[i])
 * with counter word 3 = seed >> 32 (chain_keys NULL = chain index).  The
 * stream position (sweep counter) is kept. */
mcx_status mcx_batch_set_streams(mcx_batch* b, uint64_t seed, const uint32_t* chain_keys);

/* Per-chain sweep counters (the Philox counter position); setting them is how
 * a restored checkpoint resumes bit-exactly. */
mcx_status mcx_batch_get_sweeps(mcx_batch* b, uint64_t* sweeps);
mcx_status mcx_batch_set_sweeps(mcx_batch* b, const uint64_t* sweeps);

/* ------------------------------------------------------------------------ */
/* Replica exchange / parallel tempering                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t n_sets;       /* independent PT ladders                               */
    int32_t replicas;     /* R >= 2 per set; replica r of set s is chain s*R + r   */
    const double* betas;  /* n_sets*R: beta of ladder index k of set s = betas[s*R + k-1] */
} mcx_pt_desc;

/* Replaces ReplicaExchange(ThreadsBackend(R), algs) / ParallelTempering(betas; seed, rng)
 * (src/algorithms/replica_exchange.jl:52-58, parallel_tempering.jl:24-53) for
 * n_sets independent ladders over the chains of a batch (n_sets*R == n_chains).
 * Replica r starts on ladder index r+1 (indices = collect(1:n)), stage 0. */
mcx_status mcx_pt_create(mcx_batch* b, const mcx_pt_desc* desc, mcx_pt** out);
mcx_status mcx_pt_destroy(mcx_pt* pt);
/* Replaces reset!(rx) (replica_exchange.jl:67-73). */
mcx_status mcx_pt_reset(mcx_pt* pt);
/* Replaces update!(rx::ReplicaExchange{ThreadsBackend}, xs) (replica_exchange.jl:158-178):
 * one exchange round per set; xs = per-chain energies, or NULL for the chains'
 * current energies (E_pair + D*Q) computed on the device. */
mcx_status mcx_pt_exchange(mcx_pt* pt, const double* xs);
/* The PT user loop (docs/src/examples/spin_systems/pt_Ising2D.jl:84-109) on the
 * device: n_rounds x (sweeps_between sweeps of every replica at its current
 * ladder beta, then one exchange round with the energies after the sweeps).
 * flags: MCX_RUN_RECORD, MCX_RUN_GRAPH. */
mcx_status mcx_pt_run(mcx_pt* pt, int32_t alg, int64_t n_rounds, int64_t sweeps_between,
                      uint32_t flags);
/* rx.indices (n_chains, 1-based ladder index held by each replica) and rx.stage
 * per set (replica_exchange.jl:13-19). */
mcx_status mcx_pt_get_ladder(mcx_pt* pt, int32_t* indices, int32_t* stages);
/* rx.steps / rx.accepted per edge: n_sets*(R-1) each (acceptance_rates,
 * replica_exchange.jl:75-77). */
mcx_status mcx_pt_get_counts(mcx_pt* pt, int32_t* steps, int32_t* accepted);
/* Exchange rounds done per set (position of the PT random stream). */
mcx_status mcx_pt_get_rounds(mcx_pt* pt, uint64_t* rounds);
/* Record series and moments per temperature instead of per replica: the
 * (index(pt), E) samples of pt_Ising2D_fullpower_mpi.jl:88-98 grouped by ladder
 * index (parallel_tempering.jl:53-70, _group_samples), kept on the device.
 * With MCX_ORDER_TEMPERATURE, series / moment slot set*R + i - 1 holds, in
 * sweep order, the observables of whichever replica of the set held ladder
 * index i at that sweep.  Changing the order clears the series and moments. */
mcx_status mcx_pt_set_series_order(mcx_pt* pt, int32_t order);

/* ------------------------------------------------------------------------ */
/* Multicanonical / Wang-Landau walkers                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t kind;                  /* MCX_MULTICANONICAL, MCX_WANGLANDAU or MCX_MUCA_PAIR_SPIN2 */
    int32_t n_groups;              /* walkers of one group share a binning (and lattice size) */
    const int32_t* group_of_chain; /* n_chains entries, or NULL = all in group 0; the first  */
                                   /* chain of a group is its root                            */
    const double* bin_start;       /* per group: DiscreteBinning(start, step, num)            */
    const double* bin_step;
    const int64_t* n_bins;
    const double* pair_beta;       /* MCX_MUCA_PAIR_SPIN2: n_chains betas of the Boltzmann    */
                                   /* pair term; NULL for the other kinds                     */
    const int32_t* bin_integer;    /* per group: 1 = an integer-typed DiscreteBinning, 0 = a  */
                                   /* Float64 one; NULL = integer for Ising, Float64 for      */
                                   /* Blume-Capel                                             */
} mcx_gen_desc;

/* Replaces Multicanonical(rng, bins) / WangLandau(rng, bins; logf) per chain
 * (src/algorithms/multicanonical.jl:9-25, algorithms/wang_landau.jl:10-17) and
 * ParallelMulticanonical(ThreadsBackend(n), algs) (parallel_multicanonical.jl:21-28)
 * over the chains of a batch.  Each walker owns FP64 log-weights (initially 0)
 * and an int64 histogram.  Bins follow _binindex (binned_object.jl:18-24):
 * Ising (Int) energies on an integer binning with truncating div, every other
 * case (Float64 binning, Blume-Capel's Float64 energies) with round-half-even
 * of (E - start) / step. */
mcx_status mcx_gen_create(mcx_batch* b, const mcx_gen_desc* desc, mcx_gen** out);
mcx_status mcx_gen_destroy(mcx_gen* g);
/* ensemble(alg).logweight.values / .histogram.values of one walker. */
mcx_status mcx_gen_set_logweight(mcx_gen* g, int32_t chain, const double* lw, int64_t n);
mcx_status mcx_gen_get_logweight(mcx_gen* g, int32_t chain, double* lw, int64_t n);
mcx_status mcx_gen_get_histogram(mcx_gen* g, int32_t chain, int64_t* hist, int64_t n);
/* WangLandauEnsemble.logf (chain < 0: every walker). */
mcx_status mcx_gen_set_logf(mcx_gen* g, int32_t chain, double logf);
mcx_status mcx_gen_get_logf(mcx_gen* g, double* logf);
/* N*nsweeps spin_flip!(sys, alg::AbstractImportanceSampling) attempts per walker
 * (SpinSystems/src/ising.jl:25-33, blume_capel.jl:41-50) in checkerboard order,
 * with accept!(alg, E_new, E_old) (importance_sampling.jl:69-78) or the
 * Wang-Landau accept! (algorithms/wang_landau.jl:29-37).  An energy outside the
 * binned domain stops that walker at the attempt (counters unchanged for it)
 * and the call returns MCX_EBOUNDS.  flags: MCX_RUN_RECORD. */
mcx_status mcx_gen_run(mcx_gen* g, int64_t nsweeps, uint32_t flags);
/* reset!(alg): histogram and steps/accepted to zero (algorithms/multicanonical.jl:27-33);
 * for MCX_MUCA_PAIR_SPIN2 walkers (Metropolis algorithms) steps/accepted only, as
 * reset!(::AbstractImportanceSampling) (importance_sampling.jl:104-110); chain < 0:
 * every walker. */
mcx_status mcx_gen_reset(mcx_gen* g, int32_t chain);
/* The histogram alone to zero: reset_histogram!(::CustomEnsemble) of
 * muca_BlumeCapel.jl:63-66, or of any walker kind; chain < 0: every walker. */
mcx_status mcx_gen_reset_histogram(mcx_gen* g, int32_t chain);
/* MCX_MUCA_PAIR_SPIN2: the beta of a walker's Boltzmann pair term (chain < 0: every walker). */
mcx_status mcx_gen_set_pair_beta(mcx_gen* g, int32_t chain, double beta);
/* update!(::MulticanonicalEnsemble): lw -= log(h) where h > 0
 * (src/ensembles/multicanonical.jl:32-44); chain < 0: every walker. */
mcx_status mcx_gen_update(mcx_gen* g, int32_t chain);
/* merge_histograms! / distribute_logweight! within one group
 * (src/algorithms/parallel_multicanonical.jl:38-73). */
mcx_status mcx_gen_merge(mcx_gen* g, int32_t group);
mcx_status mcx_gen_distribute(mcx_gen* g, int32_t group);
/* No reference counterpart (parallel Wang-Landau): root log-weight = mean over
 * the group's walkers. */
mcx_status mcx_gen_average_logweight(mcx_gen* g, int32_t group);
/* No reference counterpart (Wang-Landau flatness test): {min, mean, count} of
 * the group root's histogram over the group's reached bins.  A bin is reached
 * once any flatness evaluation of the group (this call or mcx_gen_iterate)
 * has seen it non-zero; the mark persists when histograms are reset, so a bin
 * the walkers stopped visiting counts as 0.  Bins never visited (forbidden
 * energies, the untouched edges of the domain) do not count. */
mcx_status mcx_gen_flatness(mcx_gen* g, int32_t group, double* min_mean_count);

/* The weight iteration as one call, per binning group over its walkers:
 *  - multicanonical walkers (MCX_MULTICANONICAL, MCX_MUCA_PAIR_SPIN2): each
 *    update is the loop body of muca_Ising2D.jl:80-91 / muca_BlumeCapel.jl:103-112
 *    across the group as ParallelMulticanonical does it
 *    (parallel_chains_threads.jl:180-201): therm_sweeps sweeps, reset!(alg)
 *    (histogram and counters), `sweeps` sweeps, merge_histograms! into the root
 *    (parallel_multicanonical.jl:38-47), update! of the root's weights
 *    (ensembles/multicanonical.jl:32-44), distribute_logweight! (:59-68);
 *  - Wang-Landau walkers (no reference counterpart for the schedule): each
 *    update runs `sweeps` sweeps of the group's active walkers (lw[x] -= ln f per
 *    attempt, algorithms/wang_landau.jl:29-37), then the root takes the mean of
 *    the group's log-weights and distributes it; the group's summed histogram
 *    is flat when min >= flatness * mean over its reached bins (as
 *    mcx_gen_flatness), and then ln f <- power * ln f for the group
 *    (update!(::WangLandauEnsemble; power), ensembles/wang_landau.jl:23) and
 *    its histograms are zeroed; a group whose ln f falls below logf_final stops.
 * The call returns after max_updates updates or when every Wang-Landau group
 * has stopped. */
typedef struct {
    int64_t sweeps;       /* sweeps of every walker per update (>= 1), e.g. 10^4         */
    int64_t therm_sweeps; /* multicanonical: sweeps before reset! in each update (>= 0) */
    int64_t max_updates;  /* >= 1                                                       */
    double flatness;      /* Wang-Landau: (0, 1], e.g. 0.8                              */
    double logf_final;    /* Wang-Landau: > 0, e.g. 1e-6                                */
    double power;         /* Wang-Landau: (0, 1), e.g. 0.5                              */
    uint32_t flags;       /* 0                                                          */
    int32_t pad;
} mcx_gen_schedule;

typedef struct {
    int64_t updates;      /* updates done by the call                                  */
    int64_t stages;       /* Wang-Landau: ln f reductions done by the call, all groups */
    int64_t sweeps;       /* sweeps run by the longest-running walker                  */
    int32_t active;       /* Wang-Landau: groups still above logf_final                */
    int32_t pad;
    double device_ms;     /* device time of the sweeps                                 */
} mcx_gen_progress;

mcx_status mcx_gen_iterate(mcx_gen* g, const mcx_gen_schedule* sched, mcx_gen_progress* out);
/* alg.steps / alg.accepted of every walker. */
mcx_status mcx_gen_get_counters(mcx_gen* g, int64_t* steps, int64_t* accepted);
mcx_status mcx_gen_group_info(mcx_gen* g, int32_t group, int32_t* root, int32_t* count,
                              int64_t* n_bins);
/* Replaces direct writes to ensemble(alg).histogram.values (counts >= 0). */
mcx_status mcx_gen_set_histogram(mcx_gen* g, int32_t chain, const int64_t* hist, int64_t n);
/* update!(::MulticanonicalEnsemble) on caller-held arrays (the reference keeps a
 * Float64 histogram): lw[i] -= log(h[i]) where h[i] > 0, evaluated on the
 * device (src/ensembles/multicanonical.jl:32-44). */
mcx_status mcx_muca_update_values(mcx_ctx* ctx, double* lw, const double* h, int64_t n);

/* ------------------------------------------------------------------------ */
/* Integrated autocorrelation times of the recorded series                   */
/* ------------------------------------------------------------------------ */
/* Replaces integrated_autocorrelation_times(traces; min_points, max_lag, c)
 * (src/measurements/autocorrelations.jl:84-110, per trace :28-73) for the
 * series of every slot over sweeps [first, first + count) of the recording,
 * without copying them off the device.  taus[n_chains]: NaN for every slot if
 * count < min_points or the lag window is empty; max_lag < 0 means no cap
 * (count / 2).  The lag sums are exact (integer series, 128-bit sums), so the
 * result is the reference's up to the rounding of its FP64 sums.  In
 * temperature order (mcx_pt_set_series_order) slot set*R + i - 1 is ladder
 * index i: the taus of optimize_exchange_interval! (parallel_tempering.jl:83-121). */
mcx_status mcx_batch_tau_int(mcx_batch* b, int32_t observable, int64_t first, int64_t count,
                             int32_t min_points, int32_t max_lag, double c, double* taus);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving or retuning loops,
 * harnesses that forbid hipMalloc): mcx_batch_tau_int with the same arguments
 * and results, staging the taus in a static device array of
 * MCX_TAU_NOALLOC_MAX_SLOTS doubles (512 KiB per GPU, held until the process
 * ends) instead of a per-call allocation; it calls no allocating routine.
 * Batches of more slots get MCX_EINVAL.  Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (from several host threads or
 * contexts) race; calls on one context are ordered by its stream. */
#define MCX_TAU_NOALLOC_MAX_SLOTS 65536
mcx_status mcx_batch_tau_int_noalloc(mcx_batch* b, int32_t observable, int64_t first,
                                     int64_t count, int32_t min_points, int32_t max_lag,
                                     double c, double* taus);

/* ------------------------------------------------------------------------ */
/* Device-state snapshots (checkpoint / resume)                             */
/* ------------------------------------------------------------------------ */
/* Replace the state that checkpoint! / restore_checkpoint serialize
 * (src/infrastructure/checkpointing.jl:48-111): the objects' device arrays are
 * written to / read from a caller buffer of _state_size bytes (a versioned
 * header + the arrays; the caller stores it).  A snapshot restores into an
 * object of the same shape only (MCX_EINVAL otherwise).  The Philox streams are
 * counter-based, so n sweeps, a snapshot, a restore and m more sweeps give the
 * run of n + m sweeps (the deterministic restart of test_checkpointing.jl:100-147).
 * Batch: chain records (sweep counters, streams, tables), spins, counters,
 * moments, last observables and, with MCX_STATE_SERIES, the recorded series
 * (restored without it, the series restarts empty). */
#define MCX_STATE_SERIES 1u
mcx_status mcx_batch_state_size(mcx_batch* b, uint32_t flags, size_t* bytes);
mcx_status mcx_batch_save_state(mcx_batch* b, uint32_t flags, void* buf, size_t bytes);
mcx_status mcx_batch_load_state(mcx_batch* b, const void* buf, size_t bytes);
/* Replica exchange: ladder positions, per-edge counters, stages, round counters. */
mcx_status mcx_pt_state_size(mcx_pt* pt, size_t* bytes);
mcx_status mcx_pt_save_state(mcx_pt* pt, void* buf, size_t bytes);
mcx_status mcx_pt_load_state(mcx_pt* pt, const void* buf, size_t bytes);
/* Walkers: log-weights, histograms, ln f, counters. */
mcx_status mcx_gen_state_size(mcx_gen* g, size_t* bytes);
mcx_status mcx_gen_save_state(mcx_gen* g, void* buf, size_t bytes);
mcx_status mcx_gen_load_state(mcx_gen* g, const void* buf, size_t bytes);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of independent simulation jobs              */
/* ------------------------------------------------------------------------ */
/* No reference counterpart: the reference runs one chain per thread to
 * completion (ParallelChains(ThreadsBackend(n), ...), parallel_chains.jl:99-105).
 * A server owns a device pool of hbm_budget bytes for int8 lattices and
 * max_jobs job slots.  A scheduler thread admits queued jobs as pool space and
 * slots free up (arrival order; a job that does not fit yet does not block
 * smaller ones behind it) and advances all resident jobs in quanta of about
 * quantum_ms: one launch per kernel class over n_streams streams, each job
 * running as many sweeps as its measured per-sweep time fits in its share
 * (big lattices: up to 256 sweeps of per-colour launches over many workgroups).
 * A job's result depends only on its descriptor: it equals a one-chain
 * mcx_batch with the same model, dim, L, beta, D, seed, chain key,
 * MCX_INIT_RANDOM start and nsweeps sweeps of mcx_batch_run(alg, RECORD), for
 * spins, counters, last-sweep observables and moments alike, bit for bit. */
typedef struct {
    int64_t hbm_budget; /* bytes for resident jobs (see below), 256-aligned          */
    int32_t max_jobs;   /* resident job slots                                        */
    int32_t n_streams;  /* streams the kernel classes are spread over (0 = 4, <= 8)  */
    double quantum_ms;  /* target device time of one quantum (0 = 2 ms)             */
    int64_t tile_sites; /* lattices with at least this many sites run spread over    */
                        /* many workgroups (bit planes in the pool, bands of rows or  */
                        /* planes with halos, several sweeps per launch);             */
                        /* 0 = 2^15 Ising / 2^14 Blume-Capel,                         */
                        /* < 0 = never.  Other lattices run one workgroup each, from  */
                        /* LDS.  A job takes N bytes, plus N/4 (Ising) or N/2         */
                        /* (Blume-Capel) bytes of planes on the tiled path.           */
    int32_t band_sweeps; /* sweeps per band launch on the tiled path: 0 = the built-in */
                         /* shapes (Ising 8 in 2D / 4 in 3D, Blume-Capel 4 / 2), k > 0 */
                         /* = k (fewer where L or the LDS is too small), < 0 = no bands: */
                         /* one launch per colour and sweep, planes in HBM             */
    int32_t pad_;
} mcx_server_desc;

typedef struct {
    int32_t model, dim, L, alg; /* L even >= 4; alg MCX_METROPOLIS, _HEATBATH, _GLAUBER */
    double beta, D;             /* D: Blume-Capel crystal field (ignored for Ising)     */
    uint64_t seed;              /* Philox key word 0 = low 32 bits, counter word 3 = high */
    uint32_t key;               /* Philox key word 1 (the chain key)                    */
    int64_t nsweeps;            /* >= 1                                                 */
    uint64_t tag;               /* returned unchanged in the result                     */
} mcx_job_desc;

typedef struct {
    uint64_t job_id, tag;
    int32_t status; /* MCX_OK */
    int32_t L;
    int64_t nsweeps, accepted;
    int32_t E_pair, M, Q;      /* observables after the last sweep */
    int32_t slo_met;           /* 1 met its deadline, 0 missed it, -1 no SLO (mcx_server_set_slo) */
    double moments[MCX_N_MOMENTS]; /* as mcx_batch_get_moments, over every sweep */
    double latency_ms;  /* submit -> result ready                        */
    double queue_ms;    /* submit -> admission                           */
    double device_ms;   /* device time of the quanta the job took part in */
} mcx_job_result;

typedef struct {
    int64_t submitted, completed, pending, active, quanta;
    double attempts;             /* site visits done                                 */
    double busy_ms;              /* device time of all quanta                        */
    double wall_ms;              /* first admission -> last completion               */
    double p50_ms, p99_ms, max_ms; /* job latency (submit -> result), interpolated   */
    double attempts_per_ns;      /* attempts / wall_ms                               */
    double busy_attempts_per_ns; /* attempts / busy_ms                               */
} mcx_server_stats_t;

mcx_status mcx_server_create(mcx_ctx* ctx, const mcx_server_desc* desc, mcx_server** out);
/* Destroys after the current quantum; jobs not finished are dropped. */
mcx_status mcx_server_destroy(mcx_server* s);
/* Queues a job (thread-safe); MCX_EINVAL for a job that can never fit. */
mcx_status mcx_server_submit(mcx_server* s, const mcx_job_desc* job, uint64_t* job_id);
/* Moves up to max finished results, oldest first, into out; *n = number moved. */
mcx_status mcx_server_poll(mcx_server* s, mcx_job_result* out, int32_t max, int32_t* n);
/* Waits until every submitted job has finished (timeout_ms < 0: no limit);
 * MCX_ESTATE on timeout. */
mcx_status mcx_server_drain(mcx_server* s, double timeout_ms);
mcx_status mcx_server_stats(mcx_server* s, mcx_server_stats_t* out);

/* Latency SLO (no reference counterpart).  Jobs submitted after the call get
 * the deadline slo_ms + attempts / (attempts_per_ns * 1e6) ms from submission
 * (attempts = L^dim * nsweeps; attempts_per_ns <= 0: slo_ms alone; slo_ms <= 0
 * and attempts_per_ns <= 0: no SLO).  The report covers the completed jobs
 * that carry a deadline: how many met it, the worst lateness, the latency
 * quantiles (all completed jobs) and the mean latency / deadline ratio. */
typedef struct {
    double slo_ms, slo_attempts_per_ns; /* the rule in force                      */
    int64_t jobs, met, missed;          /* completed jobs with a deadline         */
    int64_t jobs_by_model[2], met_by_model[2]; /* [MCX_ISING], [MCX_BLUME_CAPEL]  */
    double attainment;                  /* met / jobs (1 when jobs == 0)          */
    double worst_lateness_ms;           /* max(latency - deadline) over them      */
    double mean_deadline_frac;          /* mean of latency / deadline             */
    double p50_ms, p90_ms, p99_ms, p999_ms, max_ms, mean_ms; /* every completed job */
} mcx_slo_report;
mcx_status mcx_server_set_slo(mcx_server* s, double slo_ms, double slo_attempts_per_ns);
mcx_status mcx_server_slo_report(mcx_server* s, mcx_slo_report* out);

#ifdef __cplusplus
}
#endif
#endif
