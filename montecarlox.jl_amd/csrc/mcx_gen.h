// mcx_gen.h -- device state of multicanonical / Wang-Landau walkers.
#pragma once

#include "mcx_internal.h"

namespace mcx {

// threads of a walker's workgroup: wave 0 walks, wave 1 stages the next chunk
constexpr int kGenThreads = 128;

// static LDS of the walker kernel: two chunks' stages (ints dE[65], dp, mq,
// site, snew[64]; doubles luh[65], lul[65], u[64]), two chunks' walk results
// (64 ints and a count each), 64 pre-decrement weights, two error slots and
// 64 bytes for the compiler's alignment of the arrays
constexpr size_t kGenStaticLds =
    2 * (4 * (65 + 4 * 64) + 8 * (65 + 65 + 64)) + 2 * (4 * 64 + 4) + 8 * 64 + 8 + 64;

struct GenArgs {
    const int32_t* items;   // walkers of this launch (all from one binning group)
    int32_t n_items;
    int32_t kind;           // 2 = multicanonical, 3 = Wang-Landau, 5 = pair Boltzmann + muca on sum s^2
    int32_t model;          // 0 Ising (integer energy), 1 Blume-Capel (FP64 energy)
    int32_t dim;
    int32_t record;
    int32_t hist_in_lds;    // 1: uint32 histogram staged in LDS, else int64 atomics in HBM
    int32_t int_bins;       // Ising energies on an integer binning: truncating div, else the
                            // FP64 round-half-even of (E - start) / step (binned_object.jl:18-24)
    double bin_start, bin_step;
    double inv_step;        // 1 / bin_step when step_pow2
    int32_t step_pow2;      // bin_step is a power of two (multiply / shift instead of divide)
    int32_t step_shift;     // log2(bin_step) when step_pow2 and bin_step >= 1 is an integer
    int64_t nbins;
    ChainRec* chains;
    int8_t* spins;
    double* lw;             // per walker, at lw_off[w]
    unsigned long long* hist; // per walker, at lw_off[w]
    const int64_t* lw_off;
    const double* logf;     // per walker (Wang-Landau)
    unsigned long long* steps;
    unsigned long long* accepted;
    int64_t* err_site;      // per walker: -1, or the site whose attempt left the binned domain
    int32_t* ser_E;
    int32_t* ser_M;
    int32_t* ser_Q;
    int64_t nsweeps;
    int64_t rec_pos;
    const double* pair_beta; // per walker (kind 5): beta of the Boltzmann pair term
    int32_t win;            // weights held in LDS: a window of win bins (win = nbins: all)
    int32_t max_dbin;       // largest bin change of one attempt (window reach per chunk)
};

hipError_t launch_gen_sweeps(const GenArgs& a, size_t lds_bytes, hipStream_t st);
hipError_t launch_gen_update(double* lw, const unsigned long long* hist, const int64_t* lw_off,
                             const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                             hipStream_t st);
hipError_t launch_muca_update_values(double* lw, const double* h, int64_t n, hipStream_t st);
hipError_t launch_gen_merge(unsigned long long* hist, const int64_t* lw_off,
                            const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                            hipStream_t st);
hipError_t launch_gen_distribute(double* lw, const int64_t* lw_off, const int32_t* walkers,
                                 int32_t n_walkers, int64_t nbins, hipStream_t st);
hipError_t launch_gen_average(double* lw, const int64_t* lw_off, const int32_t* walkers,
                              int32_t n_walkers, int64_t nbins, hipStream_t st);
hipError_t launch_gen_zero_hist(unsigned long long* hist, const int64_t* lw_off,
                                const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                                hipStream_t st);
// Histogram statistics of a set of walkers of one group: per bin the sum of
// their histograms; `reached` (one byte per bin) is OR-ed with sum > 0, and
// over the reached bins out = {min sum, mean sum, bins} (FP64).  `part` holds
// kGenStatBlocks * 3 uint64 partials.
constexpr int kGenStatBlocks = 64;
hipError_t launch_gen_hist_stats(const unsigned long long* hist, const int64_t* lw_off,
                                 const int32_t* walkers, int32_t n_walkers, int64_t nbins,
                                 uint8_t* reached, unsigned long long* part, double* out,
                                 hipStream_t st);

} // namespace mcx
