// mcx_api_impl.h -- host-side object layouts behind the opaque C-ABI handles.
#pragma once

#include "../../include/mcx_hip.h"
#include "mcx_internal.h"

#include <cstring>
#include <string>
#include <vector>

#define MCX_VERSION_MAJOR 0
#define MCX_VERSION_MINOR 1

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::mcx::fail(e_ == hipErrorOutOfMemory ? MCX_ENOMEM : MCX_EHIP,     \
                               "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

struct mcx_ctx {
    int device = 0;
    int n_cu = 0;
    size_t lds_max = 0;
    char arch[64] = {0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int n_users = 0; // batches and servers built on this context
    double* tau_scratch = nullptr; // this device's static tau staging array (no allocation)
};

struct mcx_batch {
    struct Launch {
        int first, count, block;
        size_t lds;
        int layout;    // Ising: the chains' plane layout (ising_layout)
        int64_t words; // Ising: largest plane words per colour of the chains
    };
    mcx_ctx* ctx = nullptr;
    int model = 0, dim = 2, n = 0;
    uint64_t seed = 0;
    std::vector<mcx::ChainRec> rec; // host mirror of the device chain records
    std::vector<int64_t> nsites;
    std::vector<int64_t> steps;
    std::vector<double> betas;
    std::vector<Launch> launches;
    int64_t cap = 0, rec_len = 0, pool_bytes = 0, max_sites = 0;
    bool tables_ready = false;
    bool by_table = false; // series / moments recorded per table slot (temperature order)
    double last_ms = 0;
    int8_t* d_spins = nullptr;
    mcx::ChainRec* d_chains = nullptr;
    mcx::IsingTable* d_itab = nullptr;
    mcx::BCTable* d_btab = nullptr;
    int32_t* d_E = nullptr;
    int32_t* d_M = nullptr;
    int32_t* d_Q = nullptr;
    double* d_mom = nullptr;
    unsigned long long* d_acc = nullptr;
    int32_t* d_items = nullptr;
    int32_t* d_obs = nullptr;
    int32_t* d_last = nullptr; // [3 * n] E_pair, M, Q after the last sweep of a call
    // bit planes of every chain in HBM, held between the rounds of one replica-
    // exchange run instead of the int8 lattice (allocated on first use)
    uint32_t* d_planes = nullptr;
    int64_t* d_plane_off = nullptr; // [n] word offset of chain k's planes
    mcx_pt* pt = nullptr;      // replica-exchange object driving the tables, if any
    mcx_launch_shape shape[5] = {}; // per algorithm (MCX_METROPOLIS .. MCX_GLAUBER): tuned launch
    int n_gen = 0;             // multicanonical / Wang-Landau objects built on this batch
};

namespace mcx {

// Device-state snapshots (mcx_*_save_state / _load_state): a fixed header and
// the object's arrays in a fixed order, written to / read from a caller buffer.
struct StateHeader {
    char magic[4];    // "MCXS"
    uint32_t version; // 1
    uint32_t kind;    // 1 batch, 2 replica exchange, 3 walkers
    uint32_t flags;
    int64_t shape[8]; // kind-specific shape fields, checked on load
};

class BlobWriter {
public:
    BlobWriter(void* buf, size_t bytes) : p_(static_cast<uint8_t*>(buf)), left_(bytes) {}
    bool put(const void* src, size_t n)
    {
        if (n > left_)
            return false;
        if (p_)
            std::memcpy(p_, src, n);
        p_ = p_ ? p_ + n : nullptr;
        left_ -= n;
        return true;
    }
    uint8_t* reserve(size_t n) // n bytes to be filled by the caller (e.g. a device copy)
    {
        if (n > left_)
            return nullptr;
        uint8_t* r = p_;
        p_ += n;
        left_ -= n;
        return r;
    }

private:
    uint8_t* p_;
    size_t left_;
};

class BlobReader {
public:
    BlobReader(const void* buf, size_t bytes) : p_(static_cast<const uint8_t*>(buf)), left_(bytes) {}
    const uint8_t* take(size_t n)
    {
        if (n > left_)
            return nullptr;
        const uint8_t* r = p_;
        p_ += n;
        left_ -= n;
        return r;
    }
    bool get(void* dst, size_t n)
    {
        const uint8_t* r = take(n);
        if (!r)
            return false;
        std::memcpy(dst, r, n);
        return true;
    }
    size_t left() const { return left_; }

private:
    const uint8_t* p_;
    size_t left_;
};

inline StateHeader make_state_header(uint32_t kind, uint32_t flags)
{
    StateHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, "MCXS", 4);
    h.version = 1;
    h.kind = kind;
    h.flags = flags;
    return h;
}

mcx_status fail(mcx_status code, const char* fmt, ...);
uint64_t threshold32(double p);
double logistic(double x);
void ising_probabilities(double beta, int dim, double* metro, double* hb, double* gl);
void fill_ising_metro(IsingTable& t, int dim, const double* p);
void fill_ising_hb(IsingTable& t, int dim, const double* p);
void fill_ising_gl(IsingTable& t, int dim, const double* p);
void fill_ising_table(IsingTable& t, int dim, double beta);
void fill_bc_table(BCTable& t, int dim, double beta, double D);
void bc_probabilities(int dim, double beta, double D, double* metro, double* hb, double* gl);
void fill_bc_metro_from_p(BCTable& t, int dim, const double* p);
void fill_bc_hb_from_p(BCTable& t, int dim, const double* p);
void fill_bc_gl_from_p(BCTable& t, int dim, const double* p);
size_t lattice_lds_bytes(int model, int dim, int L);
int lattice_block(int model, int dim, int L);
mcx_status upload_chains(mcx_batch* b);
mcx_status pull_chains(mcx_batch* b);
mcx_status set_chain_tables(mcx_batch* b, const double* beta_per_table);
mcx_status sync_check(mcx_batch* b, const char* what);
// planes: bit 0 = read the lattices from b->d_planes instead of the int8
// spins, bit 1 = leave them in b->d_planes instead of the int8 spins (the
// batched kernels only; ensure_planes first)
mcx_status run_sweeps_async(mcx_batch* b, int32_t alg, int64_t nsweeps, uint32_t flags,
                            bool want_last = false, int planes = 0);
mcx_status ensure_planes(mcx_batch* b);
mcx_status check_run(mcx_batch* b, int32_t alg, int64_t nsweeps, uint32_t flags, const char* who);

} // namespace mcx
