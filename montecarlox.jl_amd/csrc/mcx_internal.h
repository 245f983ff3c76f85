// mcx_internal.h -- device-side records shared by the host library and the
// HIP kernels of libmcx_hip.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcx {

// Random-stream purposes: top byte of Philox counter word 2.
enum Purpose : uint32_t {
    P_METRO = 1,    // Ising Metropolis acceptance (bit-sliced uniform)
    P_HB = 2,       // Ising heat-bath draw (bit-sliced uniform)
    P_INIT = 3,     // random initial state (site word)
    P_BC_METRO = 4, // Blume-Capel Metropolis / Glauber: 16-bit site fields (proposal + prefix)
    P_BC_HB = 5,    // Blume-Capel heat bath: 16-bit site fields (prefix)
    P_GEN = 6,      // multicanonical / Wang-Landau acceptance (site word)
    P_PT = 8        // replica-exchange shared uniform
};

// One chain (lattice + its stream position) as the kernels see it.
struct ChainRec {
    int64_t spin_off;   // byte offset of the chain's int8 lattice in the spin pool
    int64_t series_off; // element offset of the chain's series (chain * capacity)
    uint64_t sweep;     // sweeps done so far: the Philox counter position
    uint64_t seed;      // Philox key word 0 = seed mod 2^32, counter word 3 = seed >> 32
    int32_t L;          // even side length >= 4
    int32_t dim;        // 2 or 3
    int32_t table;      // index into the acceptance tables
    uint32_t key;       // Philox key word 1
    double D;           // Blume-Capel crystal field
};

// Canonical Ising acceptance for one temperature.  thr = low 32 bits of
// ceil(p * 2^32); a uniform word w accepts iff w < thr.  Bit k of `always`
// marks classes with ceil(p*2^32) >= 2^32 (always accept / always up), bit k
// of `never` classes with p == 0.
struct IsingTable {
    uint32_t metro_thr[4]; // class = number of anti-aligned neighbours (< dim)
    uint32_t metro_always, metro_never;
    uint32_t hb_thr[8];    // class = number of up neighbours (0..2*dim)
    uint32_t hb_always, hb_never;
    uint32_t gl_thr[8];    // Glauber: class = number of anti-aligned neighbours (0..2*dim)
    uint32_t gl_always, gl_never;
    uint32_t pad[4];
};

// Blume-Capel acceptance for one (beta, D).  Metropolis: per (s_old, alt,
// sum_nbr) the threshold ceil(p * 2^31) on a site's 31-bit uniform, with
// index ((s_old+1)*2 + alt) * NS + (sum_nbr + 2*dim), NS = 4*dim + 1; bit k of
// metro_always marks p >= 1.  Heat bath: per sum_nbr the integer thresholds
// (W1, W12) on a site's 32-bit uniform U: U < W1 exactly when
// fl(fl(U * 2^-32) * z) < w1, U < W12 exactly when it is < fl(w1 + w2), for the
// FP64 (w1, w1 + w2, z) of SpinSystems/src/blume_capel.jl:66-76 (fl(x * z) is
// monotone in x, so each set of such U is a prefix of [0, 2^32)).
// Words of the Blume-Capel batched kernel's dynamic LDS ahead of its bit
// planes: partial sums (96), the E/M/Q ring (192) and five FP64 running sums
// (10, and 2 of padding); its class tables are a static LDS block of
// kBCStaticLds bytes.
constexpr int kBCLdsHeaderWords = 300;
constexpr int kBCStaticLds = 1088;

// Ising lattices with L < 64 and L % 8 == 0 keep each colour's sites flat in
// (z, y, x) order, 32 to a word with no partial words (the sweep stream's
// group is 2 * word + colour); other sizes keep one or more words per row.
// Such lattices always run in one workgroup (never on the tiled path).
__host__ __device__ constexpr bool ising_flat(int L)
{
    return L < 64 && (L & 7) == 0;
}

// Launch layout of an Ising lattice alone (see launch_ising_sweeps).
__host__ __device__ constexpr int ising_layout(int L)
{
    return ising_flat(L) ? 1 : (L == 64 ? 3 : 0);
}

// Layout argument of a launch of `items` lattices of layout `layout` (0..3)
// in workgroups of `block` threads: launches with fewer waves than the device
// has SIMDs (1024) and workgroups of at most 256 threads share each word among
// a quad of lanes (layout + 4), in workgroups of ising_launch_block threads: a
// quad per word of the launch's largest lattice (`words` per colour), at most
// 4 * block -- the same words per pass, a quarter of the dependent work per
// lane.
__host__ __device__ constexpr int ising_launch_layout(int layout, int64_t items, int block)
{
    return block <= 256 && items * ((block + 63) / 64) < 1024 ? layout + 4 : layout;
}

__host__ __device__ constexpr int ising_launch_block(int launch_layout, int block, int64_t words)
{
    return launch_layout < 4 ? block
         : (4 * words + 63) / 64 * 64 < 4 * block ? (int)((4 * words + 63) / 64 * 64) : 4 * block;
}

// Ising bit-plane words per colour: L^dim / 64 flat, else rows * words per row.
__host__ __device__ constexpr int64_t ising_plane_words(int dim, int L)
{
    return ising_flat(L) ? (int64_t)L * L * (dim == 3 ? L : 1) / 64
                         : (int64_t)(dim == 3 ? L * L : L) * ((L / 2 + 31) / 32);
}

struct BCTable {
    uint32_t metro_thr[80];
    uint32_t metro_always[3];
    uint32_t pad;
    uint64_t hb_w[13][2];   // (W1, W12), each in [0, 2^32]
    uint32_t gl_thr[80];    // Glauber: as metro_thr with p = logistic(-beta * dE)
    uint32_t gl_always[3];
    uint32_t gl_pad;
};

struct SweepArgs {
    const int32_t* items; // chains to process, in launch order
    int32_t n_items;
    int32_t record;       // accumulate moments (and write the series when ser_E is set)
    ChainRec* chains;
    const IsingTable* itab;
    const BCTable* btab;
    int8_t* spins;
    int32_t* ser_E;       // E_pair = -sum s_i s_j
    int32_t* ser_M;
    int32_t* ser_Q;       // sum s_i^2 (Blume-Capel)
    double* moments;      // n_chains * 6
    unsigned long long* accepted;
    int64_t nsweeps;
    int64_t rec_pos;      // series index of the first sweep of this call
    int32_t* last_E;      // if set: E_pair / M / Q after the last sweep, per chain
    int32_t* last_M;
    int32_t* last_Q;
    const int64_t* item_nsweeps; // if set: sweeps of item k (overrides nsweeps)
    int64_t* item_ticks;         // if set: wall-clock ticks item k took
    int32_t lanes_per_word;      // Blume-Capel: lanes sharing one plane word (1, 2, 4, 8; 0 = 1;
                                 // negative: per lattice, the most up to 8 within the workgroup)
    int32_t by_table;            // series / moments of chain k go to slot chains[k].table
    int64_t cap;                 // series capacity per slot (slot j starts at j * cap)
    int32_t glauber;             // Blume-Capel: Glauber thresholds instead of Metropolis ones
    uint32_t* planes;            // if set: bit-plane storage of the chains, at plane_off[chain]
    const int64_t* plane_off;
    int32_t planes_in;           // read the lattice from planes instead of packing the int8 spins
    int32_t planes_out;          // leave it in planes instead of unpacking to the int8 spins
};

struct PlainArgs {
    ChainRec* chains;
    const IsingTable* itab;
    const BCTable* btab;
    int8_t* spins;
    int32_t* ser_E;
    int32_t* ser_M;
    int32_t* ser_Q;
    unsigned long long* accepted;
    int64_t rec_index;  // series index of this sweep (< 0: not recording)
    uint64_t sweep;     // sweep counter value for this sweep
    int64_t spin_off;   // the chain's lattice in `spins` and its side (the chain
    int32_t L;          // record's spin_off and L, passed so that the spin loads
                        // need not wait for the record's)
    int32_t chain;
    int32_t colour;
    int32_t glauber;    // Blume-Capel: Glauber thresholds instead of Metropolis ones
};

constexpr int kTileObsReplicas = 16;
constexpr int kFoldThreads = 256;   // threads of the fold of one tiled item
constexpr int kFoldMaxSweeps = 512; // sweeps one fold stages (TileArgs::max_s at most)

// Multi-workgroup path for big lattices: the lattice lives in bit planes in
// device memory (layout of the LDS planes) and one launch updates one colour
// of one sweep for every work item, one thread per 32-site plane word.
struct TileArgs {
    const int32_t* items;      // chain of work item k
    const int64_t* item_nsw;   // sweeps of item k in this series of launches
    const int64_t* item_plane; // byte offset of item k's planes in `base`
    int32_t n_items;
    int32_t colour;
    int64_t s;                 // sweep index within the series (items with nsw <= s skip)
    ChainRec* chains;
    const IsingTable* itab;
    const BCTable* btab;
    uint8_t* base;
    const int64_t* item_obs;   // byte offset of item k's per-sweep sums in `base`:
                               // int32 [max_s][4 fields][kTileObsReplicas], fields
                               // Ising (up, anti-aligned bonds, -, flips),
                               // Blume-Capel (pair sum, M, Q, flips); each block adds
                               // into replica blockIdx.x % kTileObsReplicas
    int32_t max_s;
    int32_t lpw;               // Blume-Capel: lanes sharing one plane word (1, 2, 4, 8; 0 = 1)
    unsigned long long* accepted;
    int32_t glauber;           // Blume-Capel: Glauber thresholds instead of Metropolis ones
    int32_t band;              // band kernels: slabs (rows in 2D, z-planes in 3D) per band
    int32_t T;                 // band kernels: sweeps per launch (halo 2T slabs each side)
};

// Launchers (defined in the .hip translation units).
// layout: 0 row planes, 1 flat planes, 2 either per item (ising_flat(L)),
// 3 L = 64 (one row per word); + 4: a small launch, a quad of lanes per word
// (ising_launch_layout)
hipError_t launch_ising_sweeps(int dim, int alg, int layout, const SweepArgs& a, int block,
                               int grid, size_t lds_bytes, hipStream_t st);
hipError_t launch_ising_plain(int dim, int alg, const PlainArgs& a, int n_sites_colour,
                              hipStream_t st);
hipError_t launch_bc_sweeps(int dim, int alg, const SweepArgs& a, int block, int grid,
                            size_t lds_bytes, hipStream_t st);
hipError_t launch_bc_plain(int dim, int alg, const PlainArgs& a, int n_sites_colour,
                           hipStream_t st);
hipError_t launch_ising_tiles(int dim, int alg, const TileArgs& a, int max_words, hipStream_t st);
hipError_t launch_bc_tiles(int dim, int alg, const TileArgs& a, int max_words, hipStream_t st);
// int8 lattice <-> planes of every work item (unpack = 1: planes -> int8)
hipError_t launch_ising_pack(int dim, int unpack, const TileArgs& a, int max_words, hipStream_t st);
hipError_t launch_bc_pack(int dim, int unpack, const TileArgs& a, int max_words, hipStream_t st);
// Folds the per-sweep sums of a series into each item's FP64 moments (sweep by
// sweep, as the batched kernels do), last-sweep E/M/Q and sweep counter, and
// clears the sums.
constexpr int kBandThreads = 1024; // largest band workgroup (a multiple of 64 is launched)
hipError_t launch_ising_band(int dim, int alg, const TileArgs& a, int max_bands, int threads,
                             size_t lds, hipStream_t st);
hipError_t launch_bc_band(int dim, int alg, const TileArgs& a, int max_bands, int threads,
                          size_t lds, hipStream_t st);
hipError_t launch_tiles_fold(int model, const TileArgs& a, double* moments, int32_t* last_E,
                             int32_t* last_M, int32_t* last_Q, hipStream_t st);
// Integrated autocorrelation times of recorded series, one workgroup per slot.
struct TauArgs {
    const int32_t* X;        // first series (E_pair, M or Q); slot j at j * cap + first
    const int32_t* Y;        // Q for the Blume-Capel energy E_pair + D Q (nullptr: X alone)
    const ChainRec* chains;  // D of slot j
    int64_t cap, first, n;   // capacity per slot, window start, samples per series
    int32_t absx;            // use |x| (|M|)
    int32_t lag_cap;         // largest lag, 1 <= lag_cap <= n / 2
    double c;                // window constant
    int32_t n_traces;
    double* taus;            // [n_traces]
};
hipError_t launch_tau_int(const TauArgs& a, hipStream_t st);
// Allocation-free tau: a static device array of kTauScratchSlots doubles per
// GPU stages the taus; tau_scratch_address gives its address on `device`
// (hipGetSymbolAddress, done once per device and cached).
constexpr int kTauScratchSlots = 65536;
constexpr int kMaxDevices = 64;
hipError_t tau_scratch_address(int device, double** out);

hipError_t launch_init_spins(int model, int mode, ChainRec* chains, int32_t n_chains,
                             int64_t max_sites, int8_t* spins, hipStream_t st);
hipError_t launch_observables(ChainRec* chains, int32_t n_chains, int64_t max_sites,
                              const int8_t* spins, int32_t* E, int32_t* M, int32_t* Q,
                              hipStream_t st);
hipError_t launch_series_moments(ChainRec* chains, int32_t n_chains, const int32_t* E,
                                 const int32_t* M, const int32_t* Q, int64_t first,
                                 int64_t count, double* moments, hipStream_t st);
hipError_t launch_advance_sweeps(ChainRec* chains, int32_t n_chains, int64_t nsweeps,
                                 hipStream_t st);
hipError_t launch_zero_series(ChainRec* chains, int32_t n_chains, int32_t* E, int32_t* M,
                              int32_t* Q, int64_t first, int64_t count, hipStream_t st);

} // namespace mcx
