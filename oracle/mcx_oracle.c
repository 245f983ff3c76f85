/*
 * mcx_oracle.c -- CPU oracle (test infrastructure only; see mcx_oracle.h).
 *
 * Sequential, deliberately plain restatement of the reference update rules:
 *   spin_flip!(::AbstractIsing, ::AbstractMetropolis)   SpinSystems/src/ising.jl:35-41
 *   spin_flip!(::AbstractIsing, ::AbstractHeatBath)     SpinSystems/src/ising.jl:43-58
 *   spin_flip!(::AbstractBlumeCapel, ::AbstractMetropolis)  SpinSystems/src/blume_capel.jl:52-59
 *   spin_flip!(::AbstractBlumeCapel, ::AbstractHeatBath)    SpinSystems/src/blume_capel.jl:61-85
 *   spin_flip!(::AbstractIsing/BlumeCapel, ::AbstractImportanceSampling)
 *                                   SpinSystems/src/ising.jl:25-33, blume_capel.jl:41-50
 *   accept!/_accept!                src/algorithms/importance_sampling.jl:69-85
 *   Wang-Landau accept!             src/algorithms/wang_landau.jl:29-37
 *   update!(::ReplicaExchange)      src/algorithms/replica_exchange.jl:110-178
 * The site is visited in checkerboard order (colour 0 in increasing site
 * index, then colour 1) instead of pick_site's random choice, and the uniform
 * comes from a Philox4x32-10 counter stream instead of alg.rng.
 */
#include "mcx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                             */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void stream_ctr(uint32_t ctr[4], uint32_t c0, uint64_t sweep, int chunk, int purpose,
                       uint64_t seed)
{
    ctr[0] = c0;
    ctr[1] = (uint32_t)sweep;
    ctr[2] = (uint32_t)((sweep >> 32) & 0xffffu) | ((uint32_t)chunk << 16) | ((uint32_t)purpose << 24);
    ctr[3] = (uint32_t)(seed >> 32);
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    int dim, L;
    int64_t N;
    int nw;   /* 32-site words per (row, colour) */
} geo_t;

static geo_t make_geo(int dim, int L)
{
    geo_t g;
    g.dim = dim;
    g.L = L;
    g.N = (int64_t)L * L * (dim == 3 ? L : 1);
    g.nw = (L / 2 + 31) / 32;
    return g;
}

static void coords(const geo_t* G, int64_t i, int* x, int* y, int* z)
{
    *x = (int)(i % G->L);
    *y = (int)((i / G->L) % G->L);
    *z = G->dim == 3 ? (int)(i / ((int64_t)G->L * G->L)) : 0;
}

static int64_t site_of(const geo_t* G, int x, int y, int z)
{
    return (int64_t)x + (int64_t)G->L * y + (int64_t)G->L * G->L * z;
}

static int colour_of(const geo_t* G, int64_t i)
{
    int x, y, z;
    coords(G, i, &x, &y, &z);
    return (x + y + z) & 1;
}

void orc_site_group(int dim, int L, int64_t site, uint32_t* group, int* bit, int* colour)
{
    geo_t G = make_geo(dim, L);
    int x, y, z;
    coords(&G, site, &x, &y, &z);
    int c = (x + y + z) & 1;
    int64_t r = (int64_t)y + (int64_t)L * z;
    int j = x >> 1;
    *group = (uint32_t)((r * 2 + c) * G.nw + (j >> 5));
    *bit = j & 31;
    *colour = c;
}

/* The bit-sliced sweep stream of an Ising site.  Lattices with L < 64 and
 * L % 8 == 0 lay each colour's sites out flat, in the order f = (y + L z) L/2
 * + j (x = 2j + ((c + y + z) & 1)), so that every 32-site word is full: word
 * w = f / 32, bit f % 32, group 2w + c.  Other lattices use the row words of
 * orc_site_group. */
void orc_ising_group(int dim, int L, int64_t site, uint32_t* group, int* bit, int* colour)
{
    if (L < 64 && (L & 7) == 0) {
        geo_t G = make_geo(dim, L);
        int x, y, z;
        coords(&G, site, &x, &y, &z);
        int c = (x + y + z) & 1;
        int64_t f = ((int64_t)y + (int64_t)L * z) * (L / 2) + (x >> 1);
        *group = (uint32_t)(2 * (f >> 5) + c);
        *bit = (int)(f & 31);
        *colour = c;
        return;
    }
    orc_site_group(dim, L, site, group, bit, colour);
}

/* sum of the 2*dim nearest-neighbour spins, periodic boundaries
 * (the neighbour set of IsingLatticeOptim, ising.jl:444-456, and of the
 * periodic Graphs.grid lattice used by Ising(dims)/BlumeCapel(dims)). */
static int nbr_sum(const geo_t* G, const int8_t* s, int64_t i)
{
    int x, y, z, L = G->L;
    coords(G, i, &x, &y, &z);
    int acc = s[site_of(G, (x + 1) % L, y, z)] + s[site_of(G, (x + L - 1) % L, y, z)]
            + s[site_of(G, x, (y + 1) % L, z)] + s[site_of(G, x, (y + L - 1) % L, z)];
    if (G->dim == 3)
        acc += s[site_of(G, x, y, (z + 1) % L)] + s[site_of(G, x, y, (z + L - 1) % L)];
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Uniform words                                                             */
/* ------------------------------------------------------------------------ */

uint32_t orc_bitsliced_word(uint64_t seed, uint32_t chain_key, uint32_t group, int bit,
                            uint64_t sweep, int purpose)
{
    uint32_t key[2] = {(uint32_t)seed, chain_key};
    uint32_t w = 0;
    for (int chunk = 0; chunk < 8; ++chunk) {
        uint32_t ctr[4], out[4];
        stream_ctr(ctr, group, sweep, chunk, purpose, seed);
        orc_philox4x32_10(ctr, key, out);
        for (int b = 0; b < 4; ++b)
            w |= ((out[b] >> bit) & 1u) << (31 - (4 * chunk + b));
    }
    return w;
}

uint32_t orc_site_word(uint64_t seed, uint32_t chain_key, uint32_t group, int bit,
                       uint64_t sweep, int purpose)
{
    uint32_t key[2] = {(uint32_t)seed, chain_key};
    uint32_t ctr[4], out[4];
    stream_ctr(ctr, 8u * group + (uint32_t)(bit >> 2), sweep, 0, purpose, seed);
    orc_philox4x32_10(ctr, key, out);
    return out[bit & 3];
}

uint32_t orc_bc_site_uniform(uint64_t seed, uint32_t chain_key, uint32_t group, int bit,
                             uint64_t sweep, int purpose, int* alt)
{
    uint32_t key[2] = {(uint32_t)seed, chain_key};
    uint32_t ctr[4], out[4];
    int k = bit & 15;
    stream_ctr(ctr, 8u * group + (uint32_t)(k >> 2), sweep, 0, purpose, seed);
    orc_philox4x32_10(ctr, key, out);
    uint32_t f = (bit >= 16 ? out[k & 3] >> 16 : out[k & 3]) & 0xffffu;
    stream_ctr(ctr, 8u * group + 5u, sweep, 1 + bit, purpose, seed);
    orc_philox4x32_10(ctr, key, out);
    uint32_t t16 = out[0] & 0xffffu;
    if (alt)
        *alt = (int)(f & 1u);
    return ((f ^ 0x8000u) << 16) | t16;
}

uint32_t orc_pt_word(uint64_t seed, uint32_t chain_key, uint64_t round)
{
    uint32_t key[2] = {(uint32_t)seed, chain_key};
    uint32_t ctr[4] = {(uint32_t)round, (uint32_t)(round >> 32), (uint32_t)ORC_P_PT << 24,
                       (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[0];
}

/* Cache of one group's eight Philox blocks, so that the oracle builds each
 * bit-sliced word from the same 32 planes without recomputing them per site. */
typedef struct {
    uint64_t seed, sweep;
    uint32_t key, group;
    int purpose, valid;
    uint32_t planes[32];
} plane_cache_t;

static uint32_t cached_bitsliced(plane_cache_t* pc, uint64_t seed, uint32_t chain_key,
                                 uint32_t group, int bit, uint64_t sweep, int purpose)
{
    if (!pc->valid || pc->seed != seed || pc->sweep != sweep || pc->key != chain_key
        || pc->group != group || pc->purpose != purpose) {
        uint32_t key[2] = {(uint32_t)seed, chain_key};
        for (int chunk = 0; chunk < 8; ++chunk) {
            uint32_t ctr[4];
            stream_ctr(ctr, group, sweep, chunk, purpose, seed);
            orc_philox4x32_10(ctr, key, &pc->planes[4 * chunk]);
        }
        pc->seed = seed; pc->sweep = sweep; pc->key = chain_key;
        pc->group = group; pc->purpose = purpose; pc->valid = 1;
    }
    uint32_t w = 0;
    for (int b = 0; b < 32; ++b)
        w |= ((pc->planes[b] >> bit) & 1u) << (31 - b);
    return w;
}

/* ------------------------------------------------------------------------ */
/* Reference arithmetic helpers                                              */
/* ------------------------------------------------------------------------ */

/* logistic(x), src/infrastructure/utils.jl:82-88 */
double orc_logistic(double x)
{
    if (x >= 0)
        return 1.0 / (1.0 + exp(-x));
    double ex = exp(x);
    return ex / (1.0 + ex);
}

uint64_t orc_threshold32(double p)
{
    if (!(p > 0))
        return 0;
    if (p >= 1)
        return (uint64_t)1 << 32;
    return (uint64_t)ceil(ldexp(p, 32));
}

/* ------------------------------------------------------------------------ */
/* Observables (ising.jl:474-482, blume_capel.jl:185-225)                    */
/* ------------------------------------------------------------------------ */

int64_t orc_pair_sum(int dim, int L, const int8_t* s)
{
    geo_t G = make_geo(dim, L);
    int64_t acc = 0;
    for (int64_t i = 0; i < G.N; ++i)
        acc += (int64_t)s[i] * nbr_sum(&G, s, i);
    return acc / 2;
}

int64_t orc_magnetization(int dim, int L, const int8_t* s)
{
    geo_t G = make_geo(dim, L);
    int64_t acc = 0;
    for (int64_t i = 0; i < G.N; ++i)
        acc += s[i];
    return acc;
}

int64_t orc_sum_sq(int dim, int L, const int8_t* s)
{
    geo_t G = make_geo(dim, L);
    int64_t acc = 0;
    for (int64_t i = 0; i < G.N; ++i)
        acc += (int64_t)s[i] * s[i];
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Initial states (ising.jl:60-78, blume_capel.jl:87-109)                    */
/* ------------------------------------------------------------------------ */

void orc_init_random(int model, int dim, int L, uint64_t seed, uint32_t chain_key,
                     uint64_t sweep, int8_t* spins)
{
    geo_t G = make_geo(dim, L);
    for (int64_t i = 0; i < G.N; ++i) {
        uint32_t g;
        int t, c;
        orc_site_group(dim, L, i, &g, &t, &c);
        uint32_t w = orc_site_word(seed, chain_key, g, t, sweep, ORC_P_INIT);
        if (model == 0)
            spins[i] = (w >> 31) ? 1 : -1;
        else
            spins[i] = (int8_t)((int)(((uint64_t)w * 3u) >> 32) - 1);
    }
}

/* _propose_state, blume_capel.jl:21-30 ("u" true selects the first option) */
static int8_t bc_propose(int8_t s_old, int u)
{
    if (s_old == -1)
        return u ? 0 : 1;
    if (s_old == 0)
        return u ? -1 : 1;
    return u ? -1 : 0;
}

static void record(const geo_t* G, const int8_t* s, int64_t k, int32_t* E_pair, int32_t* M,
                   int32_t* Q)
{
    if (E_pair)
        E_pair[k] = (int32_t)(-orc_pair_sum(G->dim, G->L, s));
    if (M)
        M[k] = (int32_t)orc_magnetization(G->dim, G->L, s);
    if (Q)
        Q[k] = (int32_t)orc_sum_sq(G->dim, G->L, s);
}

/* ------------------------------------------------------------------------ */
/* Canonical sweeps                                                          */
/* ------------------------------------------------------------------------ */

void orc_canonical_sweeps(int model, int alg, int dim, int L, double beta, double D,
                          uint64_t seed, uint32_t chain_key, uint64_t sweep0, int64_t nsweeps,
                          int8_t* spins, int32_t* E_pair, int32_t* M, int32_t* Q,
                          int64_t* accepted)
{
    geo_t G = make_geo(dim, L);
    plane_cache_t pc;
    memset(&pc, 0, sizeof pc);
    int64_t acc = 0;
    for (int64_t k = 0; k < nsweeps; ++k) {
        uint64_t sweep = sweep0 + (uint64_t)k;
        for (int colour = 0; colour < 2; ++colour) {
            for (int64_t i = 0; i < G.N; ++i) {
                if (colour_of(&G, i) != colour)
                    continue;
                uint32_t g;
                int t, c;
                if (model == 0)
                    orc_ising_group(dim, L, i, &g, &t, &c);
                else
                    orc_site_group(dim, L, i, &g, &t, &c);
                int8_t si = spins[i];
                int nb = nbr_sum(&G, spins, i);
                if (model == 0 && alg == 0) {
                    /* flip_changes: dpair = -2 s sum(s_j); delta_energy = -dpair */
                    int dE = 2 * si * nb;
                    double lr = -beta * (double)dE; /* logweight(BoltzmannEnsemble, dE) */
                    int ok = lr > 0;
                    if (!ok) {
                        uint32_t w = cached_bitsliced(&pc, seed, chain_key, g, t, sweep, ORC_P_METRO);
                        ok = ldexp((double)w, -32) < exp(lr);
                    }
                    if (ok) {
                        spins[i] = (int8_t)-si;
                        ++acc;
                    }
                } else if (model == 0 && alg == 2) {
                    /* accept!(::Glauber, dE), metropolis.jl:121-127: rand < logistic(lr),
                     * the uniform from the Metropolis stream */
                    int dE = 2 * si * nb;
                    double lr = -beta * (double)dE;
                    uint32_t w = cached_bitsliced(&pc, seed, chain_key, g, t, sweep, ORC_P_METRO);
                    if (ldexp((double)w, -32) < orc_logistic(lr)) {
                        spins[i] = (int8_t)-si;
                        ++acc;
                    }
                } else if (model == 0) {
                    int dE = 2 * si * nb;
                    double p_plus = orc_logistic(beta * (double)si * (double)dE);
                    uint32_t w = cached_bitsliced(&pc, seed, chain_key, g, t, sweep, ORC_P_HB);
                    int8_t s_new = ldexp((double)w, -32) < p_plus ? 1 : -1;
                    if (s_new != si) {
                        spins[i] = s_new;
                        ++acc;
                    }
                } else if (alg == 0 || alg == 2) {
                    /* Metropolis (alg 0) or Glauber (alg 2): same proposal, acceptance
                     * exp(lr) or logistic(lr) */
                    int alt;
                    uint32_t U = orc_bc_site_uniform(seed, chain_key, g, t, sweep, ORC_P_BC_METRO,
                                                     &alt);
                    uint32_t u31 = ((U >> 17) << 16) | (U & 0xffffu);
                    int8_t s_new = bc_propose(si, alt);
                    int ds = s_new - si;
                    int ds2 = s_new * s_new - si * si;
                    double coupling = 1.0 * (double)nb;          /* local_coupling, J = 1 */
                    double dpair = (double)ds * coupling;
                    double dE = (-dpair - 0.0) + D * (double)ds2; /* delta_energy, h = 0 */
                    double lr = -beta * dE;
                    double u = ldexp((double)u31, -31);
                    int ok = alg == 2 ? u < orc_logistic(lr) : (lr > 0 || u < exp(lr));
                    if (ok) {
                        spins[i] = s_new;
                        ++acc;
                    }
                } else {
                    uint32_t U = orc_bc_site_uniform(seed, chain_key, g, t, sweep, ORC_P_BC_HB,
                                                     NULL);
                    double coupling = 1.0 * (double)nb;
                    double h_i = 0.0;
                    double e1 = 1.0 * coupling - h_i * (-1.0) + D;
                    double e2 = 0.0;
                    double e3 = -1.0 * coupling - h_i * 1.0 + D;
                    double w1 = exp(-beta * e1);
                    double w2 = exp(-beta * e2);
                    double w3 = exp(-beta * e3);
                    double z = w1 + w2 + w3;
                    double r = ldexp((double)U, -32) * z;
                    int8_t s_new = r < w1 ? -1 : (r < (w1 + w2) ? 0 : 1);
                    if (s_new != si) {
                        spins[i] = s_new;
                        ++acc;
                    }
                }
            }
        }
        record(&G, spins, k, E_pair, M, Q);
    }
    if (accepted)
        *accepted += acc;
}

/* ------------------------------------------------------------------------ */
/* Generalised ensembles                                                     */
/* ------------------------------------------------------------------------ */

/* _binindex(::DiscreteBinning{<:Integer}, ::Integer): div(x - start, step) (+1),
 * binned_object.jl:22-24; the float method rounds half to even (:18-20). */
static int bin_int(int64_t x, int64_t start, int64_t step, int64_t num, int64_t* idx)
{
    int64_t b = (x - start) / step; /* C division truncates like Julia's div */
    *idx = b;
    return b >= 0 && b < num;
}

static int bin_float(double x, double start, double step, int64_t num, int64_t* idx)
{
    double v = nearbyint((x - start) / step);
    if (!(v >= 0 && v < (double)num))
        return 0;
    *idx = (int64_t)v;
    return 1;
}

int orc_generalized_sweeps(int model, int wl, int dim, int L, double D,
                           double bin_start, double bin_step, int64_t nbins, int int_bins,
                           double* lw, int64_t* hist, double logf,
                           uint64_t seed, uint32_t chain_key, uint64_t sweep0, int64_t nsweeps,
                           int8_t* spins, int64_t* steps, int64_t* accepted,
                           int32_t* E_pair, int32_t* M, int32_t* Q, int64_t* err_site)
{
    geo_t G = make_geo(dim, L);
    int64_t pair = orc_pair_sum(dim, L, spins);
    int64_t q = orc_sum_sq(dim, L, spins);
    for (int64_t k = 0; k < nsweeps; ++k) {
        uint64_t sweep = sweep0 + (uint64_t)k;
        for (int colour = 0; colour < 2; ++colour) {
            for (int64_t i = 0; i < G.N; ++i) {
                if (colour_of(&G, i) != colour)
                    continue;
                uint32_t g;
                int t, c;
                orc_site_group(dim, L, i, &g, &t, &c);
                uint32_t w = orc_site_word(seed, chain_key, g, t, sweep, ORC_P_GEN);
                int8_t si = spins[i];
                int nb = nbr_sum(&G, spins, i);
                int64_t b_new = -1, b_old = -1;
                int ok_new, ok_old;
                int8_t s_new;
                int64_t dpair_i = 0, dq = 0;
                double u;
                if (model == 0) {
                    /* ising.jl:25-33 with E = -sum_pair (J = 1, Int) */
                    int64_t E_old = -pair;
                    int64_t dE = 2 * (int64_t)si * nb;
                    int64_t E_new = E_old + dE;
                    /* _binindex(::DiscreteBinning{Int}, ::Int) divides; a Float64
                     * binning takes the generic rounding method */
                    if (int_bins) {
                        ok_new = bin_int(E_new, (int64_t)bin_start, (int64_t)bin_step, nbins, &b_new);
                        ok_old = bin_int(E_old, (int64_t)bin_start, (int64_t)bin_step, nbins, &b_old);
                    } else {
                        ok_new = bin_float((double)E_new, bin_start, bin_step, nbins, &b_new);
                        ok_old = bin_float((double)E_old, bin_start, bin_step, nbins, &b_old);
                    }
                    s_new = (int8_t)-si;
                    dpair_i = -2 * (int64_t)si * nb;
                    u = ldexp((double)w, -32);
                } else {
                    /* blume_capel.jl:41-50; cached energy -sum_pair + D sum s^2 */
                    s_new = bc_propose(si, (int)(w & 1u));
                    int ds = s_new - si;
                    int ds2 = s_new * s_new - si * si;
                    double dpair = (double)ds * (1.0 * (double)nb);
                    double dE = (-dpair - 0.0) + D * (double)ds2;
                    double E_old = -(double)pair + D * (double)q;
                    double E_new = E_old + dE;
                    ok_new = bin_float(E_new, bin_start, bin_step, nbins, &b_new);
                    ok_old = bin_float(E_old, bin_start, bin_step, nbins, &b_old);
                    dpair_i = (int64_t)ds * nb;
                    dq = ds2;
                    u = ldexp((double)(w >> 1), -31);
                }
                if (!ok_new || !ok_old) {
                    if (err_site)
                        *err_site = i;
                    return 2;
                }
                double lr = lw[b_new] - lw[b_old];
                ++*steps;
                int acc = lr > 0 || u < exp(lr);
                *accepted += acc;
                int64_t vis = acc ? b_new : b_old;
                hist[vis] += 1;
                if (wl)
                    lw[vis] -= logf;
                if (acc) {
                    spins[i] = s_new;
                    pair += dpair_i;
                    q += dq;
                }
            }
        }
        record(&G, spins, k, E_pair, M, Q);
    }
    return 0;
}

int orc_pair_spin2_sweeps(int dim, int L, double beta, double bin_start, double bin_step,
                          int64_t nbins, const double* lw, int64_t* hist, uint64_t seed,
                          uint32_t chain_key, uint64_t sweep0, int64_t nsweeps, int8_t* spins,
                          int64_t* steps, int64_t* accepted, int32_t* E_pair, int32_t* M,
                          int32_t* Q, int64_t* err_site)
{
    /* docs/src/examples/spin_systems/muca_BlumeCapel.jl:43-89: spin_flip! builds
     * H = (J * sum_pair_interactions, sum_spins2) for the old and the proposed
     * state (J = 1) and calls accept!(alg, H_new, H_old) of a Metropolis
     * algorithm whose CustomEnsemble has logweight(H) = logweight(pair, H[1]) +
     * logweight(spin2, H[2]) = -beta * H[1] + lw(H[2]); the visit goes to the
     * spin2 histogram (record_visit!, :59-62; importance_sampling.jl:69-78). */
    geo_t G = make_geo(dim, L);
    double sum_pair = (double)orc_pair_sum(dim, L, spins); /* sum_pair_interactions */
    int64_t sum_sq = orc_sum_sq(dim, L, spins);            /* sum_spins2 */
    for (int64_t k = 0; k < nsweeps; ++k) {
        uint64_t sweep = sweep0 + (uint64_t)k;
        for (int colour = 0; colour < 2; ++colour) {
            for (int64_t i = 0; i < G.N; ++i) {
                if (colour_of(&G, i) != colour)
                    continue;
                uint32_t g;
                int t, c;
                orc_site_group(dim, L, i, &g, &t, &c);
                uint32_t w = orc_site_word(seed, chain_key, g, t, sweep, ORC_P_GEN);
                int8_t si = spins[i];
                int nb = nbr_sum(&G, spins, i);
                int8_t s_new = bc_propose(si, (int)(w & 1u));
                int ds = s_new - si;
                int ds2 = s_new * s_new - si * si;
                double dpair = (double)ds * (1.0 * (double)nb); /* propose_changes */
                double H1_old = 1.0 * sum_pair;
                int64_t H2_old = sum_sq;
                double H1_new = H1_old + 1.0 * dpair;
                int64_t H2_new = H2_old + ds2;
                int64_t b_new, b_old;
                int ok_new = bin_int(H2_new, (int64_t)bin_start, (int64_t)bin_step, nbins, &b_new);
                int ok_old = bin_int(H2_old, (int64_t)bin_start, (int64_t)bin_step, nbins, &b_old);
                if (!ok_new || !ok_old) {
                    if (err_site)
                        *err_site = i;
                    return 2;
                }
                double lw_new = (-beta * H1_new) + lw[b_new];
                double lw_old = (-beta * H1_old) + lw[b_old];
                double lr = lw_new - lw_old;
                double u = ldexp((double)(w >> 1), -31);
                ++*steps;
                int acc = lr > 0 || u < exp(lr);
                *accepted += acc;
                hist[acc ? b_new : b_old] += 1;
                if (acc) {
                    spins[i] = s_new;
                    sum_pair += dpair;
                    sum_sq = H2_new;
                }
            }
        }
        record(&G, spins, k, E_pair, M, Q);
    }
    return 0;
}

void orc_muca_update(int64_t nbins, double* lw, const double* h)
{
    /* update!(::MulticanonicalEnsemble), src/ensembles/multicanonical.jl:32-44 */
    for (int64_t i = 0; i < nbins; ++i) {
        double logh = h[i] > 0 ? log(h[i]) : 0.0;
        lw[i] -= logh;
    }
}

/* ------------------------------------------------------------------------ */
/* Replica exchange (replica_exchange.jl:110-178)                            */
/* ------------------------------------------------------------------------ */

int orc_exchange_pair(double beta_i, double beta_j, double x_i, double x_j, double u, double* lr)
{
    double r = ((-beta_i * x_j) - (-beta_i * x_i)) + ((-beta_j * x_i) - (-beta_j * x_j));
    if (lr)
        *lr = r;
    return (r > 0) || (u < exp(r));
}

void orc_pt_round(int n, const double* beta_of_ladder, const double* x, int32_t* indices,
                  int32_t* steps, int32_t* accepted, int32_t* stage, uint64_t seed,
                  const uint32_t* chain_keys, uint64_t round)
{
    int first = (*stage % 2 == 0) ? 1 : 2;
    for (int pair_id = first; pair_id <= n - 1; pair_id += 2) {
        int ri = -1, rj = -1;
        for (int r = 0; r < n; ++r) {
            if (indices[r] == pair_id)
                ri = r;
            if (indices[r] == pair_id + 1)
                rj = r;
        }
        steps[pair_id - 1] += 1;
        double u = ldexp((double)orc_pt_word(seed, chain_keys[ri], round), -32);
        double bi = beta_of_ladder[indices[ri] - 1];
        double bj = beta_of_ladder[indices[rj] - 1];
        double xi = x[ri], xj = x[rj];
        double lr = ((-bi * xj) - (-bi * xi)) + ((-bj * xi) - (-bj * xj));
        int ok = (lr > 0) || (u < exp(lr));
        if (ok) {
            accepted[pair_id - 1] += 1;
            int32_t tmp = indices[ri];
            indices[ri] = indices[rj];
            indices[rj] = tmp;
        }
    }
    *stage = 1 - *stage;
}

/* ------------------------------------------------------------------------ */
/* Reference CPU path (random site, Xoshiro256++, exp per uphill move)       */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t s[4];
} xoshiro_t;

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t xoshiro_next(xoshiro_t* r)
{
    uint64_t* s = r->s;
    uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

typedef struct {
    int L, c0, c1;
    const double* betas;
    uint64_t seed;
    int64_t nsweeps;
    const int32_t (*nbr4)[4];
    int8_t* spins_out;
    int64_t attempts;
} ref_job_t;

static void* ref_worker(void* arg)
{
    ref_job_t* job = (ref_job_t*)arg;
    int L = job->L;
    int64_t N = (int64_t)L * L;
    int8_t* s = (int8_t*)malloc((size_t)N);
    for (int ch = job->c0; ch < job->c1; ++ch) {
        uint64_t sm = job->seed + (uint64_t)ch;
        xoshiro_t rng;
        for (int k = 0; k < 4; ++k)
            rng.s[k] = splitmix64(&sm);
        for (int64_t i = 0; i < N; ++i)
            s[i] = (xoshiro_next(&rng) >> 63) ? 1 : -1;
        double beta = job->betas[ch];
        int64_t steps = 0, accepted = 0, sum_pair = 0, sum_spins = 0;
        for (int64_t k = 0; k < job->nsweeps * N; ++k) {
            /* pick_site (abstractions.jl:19), flip_changes (ising.jl:484-489) */
            int64_t i = (int64_t)(xoshiro_next(&rng) % (uint64_t)N);
            const int32_t* nb = job->nbr4[i];
            int lpi = s[i] * (s[nb[0]] + s[nb[1]] + s[nb[2]] + s[nb[3]]);
            int dpair = -2 * lpi;
            int dspin = -2 * s[i];
            int dE = -dpair;
            /* accept! (metropolis.jl:14-17), _accept! (importance_sampling.jl:80-85) */
            double lr = -beta * (double)dE;
            ++steps;
            int ok = (lr > 0) || ((double)(xoshiro_next(&rng) >> 11) * 0x1.0p-53 < exp(lr));
            accepted += ok;
            if (ok) { /* modify! (ising.jl:493-498) */
                s[i] = (int8_t)-s[i];
                sum_pair += dpair;
                sum_spins += dspin;
            }
        }
        job->attempts += steps;
        (void)accepted;
        (void)sum_pair;
        (void)sum_spins;
        if (job->spins_out)
            memcpy(job->spins_out + (int64_t)ch * N, s, (size_t)N);
    }
    free(s);
    return NULL;
}

void orc_refpath_series_ising2d(int L, double beta, uint64_t seed, int64_t nsweeps, int8_t* s,
                                int32_t* E_pair, int32_t* M, int64_t* accepted)
{
    int64_t N = (int64_t)L * L;
    uint64_t sm = seed;
    xoshiro_t rng;
    for (int k = 0; k < 4; ++k)
        rng.s[k] = splitmix64(&sm);
    int64_t sum_pair = orc_pair_sum(2, L, s), sum_spins = orc_magnetization(2, L, s), acc = 0;
    for (int64_t sw = 0; sw < nsweeps; ++sw) {
        for (int64_t k = 0; k < N; ++k) {
            int64_t i = (int64_t)(xoshiro_next(&rng) % (uint64_t)N);
            int x = (int)(i % L), y = (int)(i / L);
            int nb = s[y * L + (x + L - 1) % L] + s[y * L + (x + 1) % L]
                   + s[((y + L - 1) % L) * L + x] + s[((y + 1) % L) * L + x];
            int dpair = -2 * s[i] * nb;
            double lr = -beta * (double)(-dpair);
            int ok = (lr > 0) || ((double)(xoshiro_next(&rng) >> 11) * 0x1.0p-53 < exp(lr));
            if (ok) {
                sum_spins += -2 * s[i];
                s[i] = (int8_t)-s[i];
                sum_pair += dpair;
                ++acc;
            }
        }
        E_pair[sw] = (int32_t)(-sum_pair);
        M[sw] = (int32_t)sum_spins;
    }
    if (accepted)
        *accepted += acc;
}

int64_t orc_refpath_ising2d(int L, int nchains, const double* betas, uint64_t seed,
                            int64_t nsweeps, int nthreads, int8_t* spins_out)
{
    int64_t N = (int64_t)L * L;
    int32_t(*nbr4)[4] = (int32_t(*)[4])malloc(sizeof(int32_t[4]) * (size_t)N);
    /* IsingLatticeOptim neighbour table (ising.jl:444-456), 0-based */
    for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
            int i = y * L + x;
            nbr4[i][0] = y * L + (x + L - 1) % L;
            nbr4[i][1] = y * L + (x + 1) % L;
            nbr4[i][2] = ((y + L - 1) % L) * L + x;
            nbr4[i][3] = ((y + 1) % L) * L + x;
        }
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > nchains)
        nthreads = nchains;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    ref_job_t* jobs = (ref_job_t*)calloc((size_t)nthreads, sizeof(ref_job_t));
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].L = L;
        jobs[t].c0 = (int)((int64_t)nchains * t / nthreads);
        jobs[t].c1 = (int)((int64_t)nchains * (t + 1) / nthreads);
        jobs[t].betas = betas;
        jobs[t].seed = seed;
        jobs[t].nsweeps = nsweeps;
        jobs[t].nbr4 = (const int32_t(*)[4])nbr4;
        jobs[t].spins_out = spins_out;
        pthread_create(&th[t], NULL, ref_worker, &jobs[t]);
    }
    int64_t total = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        total += jobs[t].attempts;
    }
    free(jobs);
    free(th);
    free(nbr4);
    return total;
}

/* ---- reference CPU path, general form -------------------------------------
 * One chain per thread slot, as ParallelChains(ThreadsBackend(n), algs) runs
 * the reference (parallel_chains.jl:99-105): random site i = rand(UInt) % N
 * (abstractions.jl:19) over a precomputed neighbour table (the Graphs.grid
 * adjacency of Ising([L,L(,L)]) / BlumeCapel, ising.jl:413-424,
 * blume_capel.jl:500-511), Xoshiro256++ per chain seeded by
 * SplitMix64(seed + chain), and
 *   Ising Metropolis       spin_flip! (ising.jl:35-41), accept! (metropolis.jl:14-17)
 *   Blume-Capel Metropolis spin_flip! (blume_capel.jl:52-59) with _propose_state
 *                          (:21-30, rand(rng, Bool)) and dE (:235-248)
 *   multicanonical         spin_flip!(sys, ::AbstractImportanceSampling)
 *                          (ising.jl:25-33, blume_capel.jl:41-50): lr = lw[bin(E_new)]
 *                          - lw[bin(E_old)], accept! (importance_sampling.jl:69-78),
 *                          then record_visit! (ensembles/multicanonical.jl:25-30)
 *                          into a per-chain histogram; E outside the bins counts
 *                          as a rejected move (the caller sizes the bins).
 * Used only as the timed CPU baseline. */
typedef struct {
    int model, dim, muca, c0, c1; /* muca: 0 canonical, 1 multicanonical, 2 Wang-Landau */
    double logf;
    const int32_t* L;
    const double* betas;
    const double* D;
    const double* lw;
    int64_t nbins;
    double bin_start, bin_step;
    uint64_t seed;
    int64_t nsweeps;
    int64_t attempts;
} refrun_t;

static void* refrun_worker(void* arg)
{
    refrun_t* job = (refrun_t*)arg;
    const int Z = 2 * job->dim;
    for (int ch = job->c0; ch < job->c1; ++ch) {
        const int L = job->L[ch];
        const int64_t N = job->dim == 3 ? (int64_t)L * L * L : (int64_t)L * L;
        int8_t* s = (int8_t*)malloc((size_t)N);
        int32_t* nbr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N * Z));
        double* hist = job->muca ? (double*)calloc((size_t)job->nbins, sizeof(double)) : NULL;
        /* Wang-Landau: the walker's own weights, lowered by logf at every visit
         * (algorithms/wang_landau.jl:29-37) */
        double* lww = NULL;
        if (job->muca == 2) {
            lww = (double*)malloc(sizeof(double) * (size_t)job->nbins);
            for (int64_t b = 0; b < job->nbins; ++b)
                lww[b] = job->lw[b];
        }
        const double* lw = lww ? lww : job->lw;
        for (int64_t i = 0; i < N; ++i) {
            const int x = (int)(i % L), y = (int)((i / L) % L), z = (int)(i / ((int64_t)L * L));
            int32_t* nb = nbr + i * Z;
            const int64_t pl = (int64_t)L * L * z;
            nb[0] = (int32_t)(pl + (int64_t)y * L + (x + L - 1) % L);
            nb[1] = (int32_t)(pl + (int64_t)y * L + (x + 1) % L);
            nb[2] = (int32_t)(pl + (int64_t)((y + L - 1) % L) * L + x);
            nb[3] = (int32_t)(pl + (int64_t)((y + 1) % L) * L + x);
            if (job->dim == 3) {
                nb[4] = (int32_t)((int64_t)L * L * ((z + L - 1) % L) + (int64_t)y * L + x);
                nb[5] = (int32_t)((int64_t)L * L * ((z + 1) % L) + (int64_t)y * L + x);
            }
        }
        uint64_t sm = job->seed + (uint64_t)ch;
        xoshiro_t rng;
        for (int k = 0; k < 4; ++k)
            rng.s[k] = splitmix64(&sm);
        for (int64_t i = 0; i < N; ++i) {
            const uint64_t r = xoshiro_next(&rng);
            s[i] = job->model == 0 ? ((r >> 63) ? 1 : -1) : (int8_t)((int)(r % 3) - 1);
        }
        const double beta = job->betas[ch], D = job->D ? job->D[ch] : 0.0;
        int64_t sum_pair = 0, sum_sq = 0;
        for (int64_t i = 0; i < N; ++i) {
            sum_sq += s[i] * s[i];
            for (int q = 1; q < Z; q += 2) /* each bond once, from its lower end */
                sum_pair += s[i] * s[nbr[i * Z + q]];
        }
        double E = -(double)sum_pair + D * (double)sum_sq;
        int64_t steps = 0, accepted = 0;
        for (int64_t k = 0; k < job->nsweeps * N; ++k) {
            const int64_t i = (int64_t)(xoshiro_next(&rng) % (uint64_t)N);
            const int32_t* nb = nbr + i * Z;
            int sum = 0;
            for (int q = 0; q < Z; ++q)
                sum += s[nb[q]];
            const int so = s[i];
            int sn;
            if (job->model == 0)
                sn = -so;
            else {
                const int alt = (int)(xoshiro_next(&rng) >> 63);
                sn = so == -1 ? (alt ? 0 : 1) : (so == 0 ? (alt ? -1 : 1) : (alt ? -1 : 0));
            }
            const double dE = -(double)((sn - so) * sum) + D * (double)(sn * sn - so * so);
            double lr;
            int64_t b_old = 0, b_new = 0;
            if (job->muca) {
                const int in_old = bin_float(E, job->bin_start, job->bin_step, job->nbins, &b_old);
                const int in_new = bin_float(E + dE, job->bin_start, job->bin_step, job->nbins, &b_new);
                lr = in_old && in_new ? lw[b_new] - lw[b_old] : -INFINITY;
            } else {
                lr = -beta * dE;
            }
            ++steps;
            const int ok = (lr > 0) || ((double)(xoshiro_next(&rng) >> 11) * 0x1.0p-53 < exp(lr));
            if (ok) {
                s[i] = (int8_t)sn;
                E += dE;
                ++accepted;
            }
            if (job->muca) {
                int64_t b;
                if (bin_float(E, job->bin_start, job->bin_step, job->nbins, &b)) {
                    hist[b] += 1.0;
                    if (lww)
                        lww[b] -= job->logf;
                }
            }
        }
        job->attempts += steps;
        (void)accepted;
        free(lww);
        free(hist);
        free(nbr);
        free(s);
    }
    return NULL;
}

int64_t orc_refpath_run(int model, int dim, int muca, int nchains, const int32_t* L,
                        const double* betas, const double* D, const double* lw, int64_t nbins,
                        double bin_start, double bin_step, uint64_t seed, int64_t nsweeps,
                        int nthreads, double logf)
{
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > nchains)
        nthreads = nchains;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    refrun_t* jobs = (refrun_t*)calloc((size_t)nthreads, sizeof(refrun_t));
    for (int t = 0; t < nthreads; ++t) {
        refrun_t* j = &jobs[t];
        j->model = model;
        j->dim = dim;
        j->muca = muca;
        j->logf = logf;
        j->c0 = (int)((int64_t)nchains * t / nthreads);
        j->c1 = (int)((int64_t)nchains * (t + 1) / nthreads);
        j->L = L;
        j->betas = betas;
        j->D = D;
        j->lw = lw;
        j->nbins = nbins;
        j->bin_start = bin_start;
        j->bin_step = bin_step;
        j->seed = seed;
        j->nsweeps = nsweeps;
        pthread_create(&th[t], NULL, refrun_worker, j);
    }
    int64_t total = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        total += jobs[t].attempts;
    }
    free(jobs);
    free(th);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Integrated autocorrelation time (src/measurements/autocorrelations.jl)    */
/* ------------------------------------------------------------------------ */

typedef __int128 orc_i128;

static int64_t orc_tau_at(const int32_t* x, int64_t i, int absx)
{
    int64_t v = x[i];
    return absx && v < 0 ? -v : v;
}

/* n^2 sum_{i<n-t} (x_i - Sx/n)(y_{i+t} - Sy/n), exactly */
static orc_i128 orc_tau_cross(const int32_t* x, int ax, const int32_t* y, int ay, int64_t n,
                              int64_t t)
{
    orc_i128 P = 0;
    int64_t Sx = 0, Sy = 0, Hx = 0, Ty = 0;
    for (int64_t i = 0; i < n; ++i) {
        Sx += orc_tau_at(x, i, ax);
        Sy += orc_tau_at(y, i, ay);
    }
    for (int64_t i = 0; i + t < n; ++i) {
        P += (orc_i128)(orc_tau_at(x, i, ax) * orc_tau_at(y, i + t, ay));
        Hx += orc_tau_at(x, i, ax);
        Ty += orc_tau_at(y, i + t, ay);
    }
    orc_i128 nn = n;
    return nn * nn * P - nn * (orc_i128)Sy * Hx - nn * (orc_i128)Sx * Ty +
           (orc_i128)(n - t) * (orc_i128)Sx * (orc_i128)Sy;
}

static double orc_tau_N(const int32_t* x, const int32_t* y, int64_t n, double D, int absx,
                        int64_t t)
{
    if (!y)
        return (double)orc_tau_cross(x, absx, x, absx, n, t);
    orc_i128 xx = orc_tau_cross(x, 0, x, 0, n, t);
    orc_i128 xy = orc_tau_cross(x, 0, y, 0, n, t);
    orc_i128 yx = orc_tau_cross(y, 0, x, 0, n, t);
    orc_i128 yy = orc_tau_cross(y, 0, y, 0, n, t);
    return ((double)xx + D * (double)(xy + yx)) + (D * D) * (double)yy;
}

double orc_tau_int(const int32_t* x, const int32_t* y, int64_t n, double D, int absx,
                   int32_t lag_cap, double c)
{
    /* integrated_autocorrelation_time, autocorrelations.jl:28-73: tau = 1/2 +
     * sum C(t), stopped at C(t) <= 0 or t > c * tau; 1/2 for zero variance */
    if (y && D == 0.0)
        y = NULL;
    double C0 = orc_tau_N(x, y, n, D, absx, 0) / (double)n;
    if (!(C0 > 0))
        return 0.5;
    double tau = 0.5;
    for (int64_t t = 1; t <= lag_cap; ++t) {
        double C = (orc_tau_N(x, y, n, D, absx, t) / (double)(n - t)) / C0;
        if (C <= 0)
            break;
        double tau_next = tau + C;
        if ((double)t > c * tau_next)
            break;
        tau = tau_next;
    }
    return tau > 0.5 ? tau : 0.5;
}

double orc_tau_int_centred(const double* x, int64_t n, int32_t lag_cap, double c)
{
    /* autocorrelations.jl:47-73 as written: mu = sum(x) / n, centred samples,
     * C0 = dot(c, c) / n, cov(t) = dot(c[1:n-t], c[1+t:n]) / (n - t) */
    double mu = 0;
    for (int64_t i = 0; i < n; ++i)
        mu += x[i];
    mu /= (double)n;
    double C0 = 0;
    for (int64_t i = 0; i < n; ++i)
        C0 += (x[i] - mu) * (x[i] - mu);
    C0 /= (double)n;
    if (!(C0 > 0))
        return 0.5;
    double tau = 0.5;
    for (int64_t t = 1; t <= lag_cap; ++t) {
        double cov = 0;
        for (int64_t i = 0; i + t < n; ++i)
            cov += (x[i] - mu) * (x[i + t] - mu);
        double C = (cov / (double)(n - t)) / C0;
        if (C <= 0)
            break;
        double tau_next = tau + C;
        if ((double)t > c * tau_next)
            break;
        tau = tau_next;
    }
    return tau > 0.5 ? tau : 0.5;
}
