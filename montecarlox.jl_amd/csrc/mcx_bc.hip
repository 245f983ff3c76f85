// mcx_bc.hip -- checkerboard Metropolis / heat-bath sweeps for Blume-Capel
// lattices (spin-1, J = 1, crystal field D, h = 0; 2D square / 3D cubic,
// periodic) on gfx950.
//
// Per-site rule:
//   Metropolis (SpinSystems/src/blume_capel.jl:52-59, :21-30, :235-248):
//     s_new = _propose_state(s_old, b) (b true selects the first alternative),
//     dE = -(ds * sum_j s_j) + D * ds2, accept iff u < exp(-beta * dE)
//   heat bath (blume_capel.jl:61-85):
//     w1 = exp(-beta*e1), w2 = exp(-beta*0), w3 = exp(-beta*e3), z = w1+w2+w3,
//     r = u * z; s_new = r < w1 ? -1 : (r < w1 + w2 ? 0 : 1)
// Random words (purpose P_BC_METRO for Metropolis and Glauber, P_BC_HB for
// heat bath), for the 32-site word `group` of one sweep:
//   prefix blocks: Philox block i (i = 0..3) at counter (8*group + i, sweep);
//     its word m holds the 16-bit fields of sites k = 4i + m (low half) and
//     k + 16 (high half); a site's field f gives u16 = f ^ 0x8000;
//   tail: the site's 32-bit uniform continues with t16 = low 16 bits of word 0
//     of the block at counter (8*group + 5, sweep, chunk 1 + bit).
// Metropolis / Glauber: the proposal bit b = u16 & 1 and the 31-bit uniform
// U31 = (u16 >> 1) * 2^16 + t16 (accept iff U31 < ceil(p * 2^31), i.e.
// U31 / 2^31 < p); heat bath: U = u16 * 2^16 + t16 and r = (U / 2^32) * z.
// The tail matters only when the prefix equals the threshold's leading bits
// (probability 2^-15 or 2^-16 per comparison): the kernels draw it then only.
//
// Batched kernel: one workgroup per chain; the lattice lives in LDS as four bit
// planes (occupied, up) x (colour 0, colour 1) in the word layout of the Ising
// kernels; one thread updates the 32 sites of a word: the neighbour sum is
// formed bit-sliced for all 32 sites at once, an 8 x 8 bit-matrix transpose
// turns the planes into one class byte per site, and each pair of sites is
// decided by one packed 16-bit saturating subtract of its two prefixes against
// its two thresholds (looked up by class in LDS).
#include "mcx_api_impl.h"
#include "mcx_device.h"

#include <cmath>
#include <cstring>

namespace mcx {

// ---------------------------------------------------------------------------
// host: tables
// ---------------------------------------------------------------------------

static int8_t propose_host(int s_old, int alt)
{
    if (s_old == -1)
        return alt ? 0 : 1;
    if (s_old == 0)
        return alt ? -1 : 1;
    return alt ? -1 : 0;
}

// Probabilities in the layout of mcx_batch_table_layout:
//   metro[((s_old+1)*2 + alt)*NS + (sum + 2*dim)] = exp(-beta * dE)
//   hb[(sum + 2*dim)*3 + {0,1,2}] = (w1, w1 + w2, z)
//   gl[same index as metro] = logistic(-beta * dE)  (Glauber, metropolis.jl:121-127)
void bc_probabilities(int dim, double beta, double D, double* metro, double* hb, double* gl)
{
    const int NS = 4 * dim + 1;
    for (int so = -1; so <= 1; ++so)
        for (int alt = 0; alt < 2; ++alt)
            for (int sum = -2 * dim; sum <= 2 * dim; ++sum) {
                const int sn = propose_host(so, alt);
                const int ds = sn - so;
                const int ds2 = sn * sn - so * so;
                const double coupling = 1.0 * (double)sum; // local_coupling, J = 1
                const double dpair = (double)ds * coupling;
                const double dE = (-dpair - 0.0) + D * (double)ds2;
                metro[((so + 1) * 2 + alt) * NS + (sum + 2 * dim)] = std::exp(-beta * dE);
                gl[((so + 1) * 2 + alt) * NS + (sum + 2 * dim)] = logistic(-beta * dE);
            }
    for (int sum = -2 * dim; sum <= 2 * dim; ++sum) {
        const double coupling = 1.0 * (double)sum;
        const double h_i = 0.0;
        const double e1 = 1.0 * coupling - h_i * (-1.0) + D;
        const double e2 = 0.0;
        const double e3 = -1.0 * coupling - h_i * 1.0 + D;
        const double w1 = std::exp(-beta * e1);
        const double w2 = std::exp(-beta * e2);
        const double w3 = std::exp(-beta * e3);
        const double z = w1 + w2 + w3;
        double* o = hb + (sum + 2 * dim) * 3;
        o[0] = w1;
        o[1] = w1 + w2;
        o[2] = z;
    }
}

// 31-bit thresholds ceil(p * 2^31) on the uniform w >> 1; bit k of `always`
// marks p >= 1 (accept without comparing).
static void fill_bc_thr31(uint32_t* thr, uint32_t* always, int dim, const double* p)
{
    const int n = 6 * (4 * dim + 1);
    always[0] = always[1] = always[2] = 0;
    for (int k = 0; k < 80; ++k)
        thr[k] = 0;
    for (int k = 0; k < n; ++k) {
        if (p[k] >= 1) {
            always[k >> 5] |= 1u << (k & 31);
            thr[k] = 0x80000000u;
        } else {
            const double v = std::ceil(std::ldexp(p[k] > 0 ? p[k] : 0.0, 31));
            if (v >= 2147483648.0) {
                always[k >> 5] |= 1u << (k & 31);
                thr[k] = 0x80000000u;
            } else {
                thr[k] = (uint32_t)v;
            }
        }
    }
}

void fill_bc_metro_from_p(BCTable& t, int dim, const double* p)
{
    fill_bc_thr31(t.metro_thr, t.metro_always, dim, p);
}

void fill_bc_gl_from_p(BCTable& t, int dim, const double* p)
{
    fill_bc_thr31(t.gl_thr, t.gl_always, dim, p);
}

// The number of 32-bit uniforms U with fl(fl(U * 2^-32) * z) < w: the first U
// for which the product is not below w (2^32 if there is none).  For finite
// z >= 0 the product is monotone in U, so these U form a prefix of [0, 2^32);
// with z = inf every product is inf or NaN (U = 0), never below w, and the
// count is 0.  The same IEEE double operations as the per-site formula.
static uint64_t hb_threshold(double w, double z)
{
    uint64_t lo = 0, hi = (uint64_t)1 << 32;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        const double r = ((double)mid * 0x1.0p-32) * z;
        if (r < w)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

void fill_bc_hb_from_p(BCTable& t, int dim, const double* p)
{
    const int NS = 4 * dim + 1;
    for (int k = 0; k < 13; ++k)
        t.hb_w[k][0] = t.hb_w[k][1] = 0;
    for (int k = 0; k < NS; ++k) {
        t.hb_w[k][0] = hb_threshold(p[3 * k], p[3 * k + 2]);
        t.hb_w[k][1] = hb_threshold(p[3 * k + 1], p[3 * k + 2]);
    }
}

void fill_bc_table(BCTable& t, int dim, double beta, double D)
{
    std::memset(&t, 0, sizeof t);
    double metro[78], hb[39], gl[78];
    bc_probabilities(dim, beta, D, metro, hb, gl);
    fill_bc_metro_from_p(t, dim, metro);
    fill_bc_hb_from_p(t, dim, hb);
    fill_bc_gl_from_p(t, dim, gl);
}

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------

__device__ __forceinline__ int bc_propose(int s_old, int alt)
{
    return s_old == -1 ? (alt ? 0 : 1) : (s_old == 0 ? (alt ? -1 : 1) : (alt ? -1 : 0));
}

struct BCNbr {
    uint32_t occ[6], up[6]; // U, Dn, A, Bw, Zu, Zd
};

// Neighbour occupancy / up words of word q of plane row iA, given the plane
// rows of its y-1, y+1 (and z-1, z+1) neighbours in the other colour's planes.
// SMALL: every row index times nw fits 24 bits (LDS-resident planes).
template <int DIM, bool SMALL = false>
__device__ __forceinline__ void bc_nbrs_rows(const uint32_t* __restrict__ Oo,
                                             const uint32_t* __restrict__ So, int nw, int half,
                                             int iA0, int iU0, int iD0, int iZu0, int iZd0, int q,
                                             int p, int W, uint32_t valid, BCNbr& n)
{
    auto mul = [&](int a) {
        return SMALL ? (int)__umul24((uint32_t)a, (uint32_t)nw) : a * nw;
    };
    const int r = iA0;
    const int iU = mul(iU0) + q, iD = mul(iD0) + q, iA = mul(r) + q;
    n.occ[0] = Oo[iU];
    n.up[0] = So[iU];
    n.occ[1] = Oo[iD];
    n.up[1] = So[iD];
    n.occ[2] = Oo[iA];
    n.up[2] = So[iA];
    if (DIM == 3) {
        const int iZu = mul(iZu0) + q, iZd = mul(iZd0) + q;
        n.occ[4] = Oo[iZu];
        n.up[4] = So[iZu];
        n.occ[5] = Oo[iZd];
        n.up[5] = So[iZd];
    } else {
        n.occ[4] = n.up[4] = n.occ[5] = n.up[5] = 0;
    }
    const uint32_t A0 = n.occ[2], A1 = n.up[2];
    if (nw == 1) {
        // the W-bit row rotated right by 1 (p) or left by 1 = right by W - 1,
        // without a branch on the lane's row parity
        const uint32_t sr = p ? 1u : (uint32_t)(W - 1);
        const uint32_t sl = (uint32_t)W - sr;
        n.occ[3] = ((A0 >> sr) | (A0 << sl)) & valid;
        n.up[3] = ((A1 >> sr) | (A1 << sl)) & valid;
    } else if (p) {
        const int qn = q + 1 == nw ? 0 : q + 1;
        n.occ[3] = (A0 >> 1) | ((Oo[mul(r) + qn] & 1u) << (W - 1));
        n.up[3] = (A1 >> 1) | ((So[mul(r) + qn] & 1u) << (W - 1));
    } else {
        const int qp = q == 0 ? nw - 1 : q - 1;
        const int Wp = word_width(qp, half);
        n.occ[3] = ((A0 << 1) | ((Oo[mul(r) + qp] >> (Wp - 1)) & 1u)) & valid;
        n.up[3] = ((A1 << 1) | ((So[mul(r) + qp] >> (Wp - 1)) & 1u)) & valid;
    }
}

template <int DIM, bool SMALL = false>
__device__ __forceinline__ void bc_load_nbrs(const uint32_t* __restrict__ Oo,
                                             const uint32_t* __restrict__ So, int L, int nw,
                                             int half, int r, int q, int y, int z, int p, int W,
                                             uint32_t valid, BCNbr& n)
{
    const int base = DIM == 3 ? z * L : 0;
    const int yu = y == 0 ? L - 1 : y - 1;
    const int yd = y == L - 1 ? 0 : y + 1;
    const int zu = DIM == 3 ? (z == 0 ? L - 1 : z - 1) : 0;
    const int zd = DIM == 3 ? (z == L - 1 ? 0 : z + 1) : 0;
    bc_nbrs_rows<DIM, SMALL>(Oo, So, nw, half, r, base + yu, base + yd, zu * L + y, zd * L + y, q,
                             p, W, valid, n);
}

// Bit-sliced (sum_j s_j + 2*dim) for 32 sites: each neighbour adds 2 if up,
// 1 if empty, 0 if down.  Result bits v0..v3.
template <int DIM>
__device__ __forceinline__ void bc_shifted_sum(const BCNbr& n, uint32_t& v0, uint32_t& v1,
                                               uint32_t& v2, uint32_t& v3)
{
    uint32_t a0, a1, a2, b0, b1, b2;
    uint32_t a[6], e[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        a[k] = n.occ[k] & n.up[k];
        e[k] = ~n.occ[k];
    }
    if (DIM == 2) {
        count4(a[0], a[1], a[2], a[3], a0, a1, a2);
        count4(e[0], e[1], e[2], e[3], b0, b1, b2);
    } else {
        count6(a[0], a[1], a[2], a[3], a[4], a[5], a0, a1, a2);
        count6(e[0], e[1], e[2], e[3], e[4], e[5], b0, b1, b2);
    }
    // 2*A + B
    v0 = b0;
    v1 = a0 ^ b1;
    const uint32_t c1 = a0 & b1;
    v2 = a1 ^ b2 ^ c1;
    const uint32_t c2 = (a1 & b2) | (c1 & (a1 | b2));
    v3 = a2 ^ c2;
}

// int8 lattice -> bit planes [occ c0, occ c1, up c0, up c1] for plane word w.
template <int DIM>
__device__ __forceinline__ void bc_pack_word(const int8_t* __restrict__ sp,
                                             uint32_t* __restrict__ P, int w, int L, int nw,
                                             int half, int words)
{
    const int r = w / nw;
    const int q = w - r * nw;
    const int y = DIM == 3 ? r % L : r;
    const int z = DIM == 3 ? r / L : 0;
    const int par = (y + z) & 1;
    const int W = word_width(q, half);
    const int8_t* row = sp + (int64_t)r * L + 64 * q;
    uint32_t eo = 0, eu = 0, oo = 0, ou = 0;
    for (int j = 0; j < W; ++j) {
        const int se = row[2 * j], so = row[2 * j + 1];
        eo |= (uint32_t)(se != 0) << j;
        eu |= (uint32_t)(se > 0) << j;
        oo |= (uint32_t)(so != 0) << j;
        ou |= (uint32_t)(so > 0) << j;
    }
    P[par * words + w] = eo;
    P[(2 + par) * words + w] = eu;
    P[(par ^ 1) * words + w] = oo;
    P[(2 + (par ^ 1)) * words + w] = ou;
}

// Bit planes -> int8 lattice for plane word w.
template <int DIM>
__device__ __forceinline__ void bc_unpack_word(int8_t* __restrict__ sp,
                                               const uint32_t* __restrict__ P, int w, int L,
                                               int nw, int half, int words)
{
    const int r = w / nw;
    const int q = w - r * nw;
    const int y = DIM == 3 ? r % L : r;
    const int z = DIM == 3 ? r / L : 0;
    const int par = (y + z) & 1;
    const int W = word_width(q, half);
    int8_t* row = sp + (int64_t)r * L + 64 * q;
    const uint32_t eo = P[par * words + w], eu = P[(2 + par) * words + w];
    const uint32_t oo = P[(par ^ 1) * words + w], ou = P[(2 + (par ^ 1)) * words + w];
    for (int j = 0; j < W; ++j) {
        row[2 * j] = ((eo >> j) & 1u) ? (((eu >> j) & 1u) ? 1 : -1) : 0;
        row[2 * j + 1] = ((oo >> j) & 1u) ? (((ou >> j) & 1u) ? 1 : -1) : 0;
    }
}

// Class of a site, i = sidx | o << 4 | u << 5 | alt << 6 (sidx = sum of the
// neighbour spins + 2*DIM, o / u the site's occupancy and up bits, alt the
// proposal bit), and its threshold from the host layout
// k = ((s_old + 1) * 2 + alt) * NS + sidx.  Classes that no site can be in
// (sidx >= NS, or up without occupancy) get 0.
constexpr int kBCClasses = 128;

template <int DIM>
__device__ __forceinline__ uint32_t bc_class_thr(const uint32_t* thr, int i)
{
    constexpr int NS = 4 * DIM + 1;
    const int sidx = i & 15, o = (i >> 4) & 1, u = (i >> 5) & 1, alt = (i >> 6) & 1;
    if (sidx >= NS || (!o && u))
        return 0u;
    const int s_old = o ? (u ? 1 : -1) : 0;
    return thr[((s_old + 1) * 2 + alt) * NS + sidx];
}

// LDS tables of one chain for the word update.  A site's 16-bit field f read
// as int16 is u16 - 32768 (u16 = f ^ 0x8000), so that a saturating int16
// subtract of a stored (A - 32768) is negative exactly when u16 < A.
//   Metropolis / Glauber, by class: a16 = min(2 * (thr >> 16), 65535) - 32768:
//     u16 < 2 * (thr >> 16) exactly when the uniform's leading 15 bits
//     u16 >> 1 are below thr's (accept); a difference of 0 or 1 is a tie
//     (u16 >> 1 == thr >> 16, decided on the full uniform against thr).
//   Heat bath, by sidx: a32 = (A(W1) - 32768) | (A(W12) - 32768) << 16 with
//     A(W) = min(W >> 16, 65535); a difference of 0 is a tie (decided against
//     hbw = (W1, W12)).
struct BCTabs {
    int16_t a16[kBCClasses];
    uint32_t a32[16];
    uint32_t thr[kBCClasses];
    uint64_t hbw[16][2];
};
constexpr int kBCTabWords = (int)(sizeof(BCTabs) / 4);

template <int DIM, int ALG>
__device__ __forceinline__ void bc_fill_tabs(BCTabs& t, const BCTable* tb, int glauber, int tid,
                                             int nthr)
{
    if (ALG == 0) {
        const uint32_t* src = glauber ? tb->gl_thr : tb->metro_thr;
        for (int k = tid; k < kBCClasses; k += nthr) {
            const uint32_t thr = bc_class_thr<DIM>(src, k);
            const int v = 2 * (int)(thr >> 16) - 32768;
            t.a16[k] = (int16_t)(v > 32767 ? 32767 : v);
            t.thr[k] = thr;
        }
    } else {
        constexpr int NS = 4 * DIM + 1;
        for (int k = tid; k < 16; k += nthr) {
            const uint64_t w1 = k < NS ? tb->hb_w[k][0] : 0, w12 = k < NS ? tb->hb_w[k][1] : 0;
            const uint32_t a1 = (uint32_t)((w1 >> 16) < 65535 ? (w1 >> 16) : 65535) - 32768u;
            const uint32_t a2 = (uint32_t)((w12 >> 16) < 65535 ? (w12 >> 16) : 65535) - 32768u;
            t.a32[k] = (a1 & 0xffffu) | (a2 << 16);
            t.hbw[k][0] = w1;
            t.hbw[k][1] = w12;
        }
    }
}

// Rows p, p + S of an 8 x 8 bit-matrix stored by byte column: exchange the
// columns whose index has bit S set in row p with those where it is clear in
// row p + S (m = the columns with it clear).
template <int S>
__device__ __forceinline__ void bc_swap_rows(uint32_t& a, uint32_t& b, uint32_t m)
{
    const uint32_t na = (a & m) | ((b << S) & ~m);
    const uint32_t nb = ((a >> S) & m) | (b & ~m);
    a = na;
    b = nb;
}

// 8 x 8 bit-matrix transpose within each byte column of x[0..7]: afterwards
// bit p of byte b of x[r] is what bit r of byte b of x[p] was.
__device__ __forceinline__ void bc_transpose8(uint32_t* x)
{
#pragma unroll
    for (int p = 0; p < 4; ++p)
        bc_swap_rows<4>(x[p], x[p + 4], 0x0f0f0f0fu);
#pragma unroll
    for (int p = 0; p < 8; p += (p & 1) ? 3 : 1) // 0, 1, 4, 5
        bc_swap_rows<2>(x[p], x[p + 2], 0x33333333u);
#pragma unroll
    for (int p = 0; p < 8; p += 2)
        bc_swap_rows<1>(x[p], x[p + 1], 0x55555555u);
}

__device__ __forceinline__ uint32_t rotr32(uint32_t x, uint32_t s)
{
    return __builtin_amdgcn_alignbit(x, x, s);
}

// 16-bit field of site `bit` of word `group` (see the random-word layout above)
// and the 16 bits that continue its uniform.  w2: counter word 2 of chunk 0.
__device__ __forceinline__ uint32_t bc_site_field(uint32_t group, int bit, uint32_t sw_lo,
                                                  uint32_t w2, uint32_t c3, uint32_t k0,
                                                  uint32_t k1)
{
    const int k = bit & 15;
    const uint4 R = philox4x32_10(8u * group + (uint32_t)(k >> 2), sw_lo, w2, c3, k0, k1);
    const int m = k & 3;
    const uint32_t w = m == 0 ? R.x : (m == 1 ? R.y : (m == 2 ? R.z : R.w));
    return bit >= 16 ? w >> 16 : w & 0xffffu;
}

__device__ __forceinline__ uint32_t bc_site_tail(uint32_t group, int bit, uint32_t sw_lo,
                                                 uint32_t w2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1)
{
    return philox4x32_10(8u * group + 5u, sw_lo, w2 | ((uint32_t)(1 + bit) << 16), c3, k0, k1).x &
           0xffffu;
}

typedef short bc_s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short bc_u16x2 __attribute__((ext_vector_type(2)));

// One 32-site word of colour c: updates its occupancy / up bits in place from
// the bit-sliced shifted neighbour sums v0..v3 (the per-site rule and random
// words above) and returns the number of changed sites this lane decided.
// The class bytes of all 32 sites come from one 8 x 8 bit-matrix transpose of
// the planes (byte = 2 * class for the int16 Metropolis table, 4 * sidx for
// the heat-bath pairs); pair k of prefix block i (k = 4i + m) holds sites k and
// k + 16, whose two thresholds are subtracted from the two fields at once.
// The sites of a word are independent (one colour), so a word may be shared
// by lpw adjacent lanes (lpw = 1, 2, 4 or 8, lane part = lane % lpw): each
// takes the prefix blocks part, part + lpw, ... and the lanes then OR their bit
// changes together, so every lane of the group ends with the whole word.
// Metropolis collects the accept and proposal bits of the sites and moves all
// accepted sites of the word at once: the proposal (the lower other state if
// alt, else the upper) changes the occupancy where ~o | (u ^ alt) and the up
// bit where ~alt | u.  Shared by the LDS-resident batched kernel and the
// multi-workgroup tile and band kernels.
// SHARED: the word's lanes are lpw > 1 (the caller branches on it, uniformly).
template <int DIM, int ALG, bool SHARED>
__device__ __forceinline__ uint32_t bc_word(uint32_t& occ, uint32_t& up, uint32_t v0, uint32_t v1,
                                            uint32_t v2, uint32_t v3, int W, const BCTabs& T,
                                            uint32_t group, uint32_t sw_lo, uint32_t w2,
                                            uint32_t c3, uint32_t k0, uint32_t k1, int part,
                                            int lpw)
{
    if (!SHARED) {
        part = 0;
        lpw = 1;
    }
    const uint32_t valid = width_mask(W);
    // Rotating the planes right by 4 * part puts the sites of prefix block
    // part + i where those of block i are, so that every lane runs the loop
    // below over the same compile-time pair positions.
    const uint32_t rot = 4u * (uint32_t)part;
    auto rin = [&](uint32_t v) { return SHARED ? rotr32(v, rot) : v; };
    uint32_t x[8];
    if (ALG == 0) {
        // class byte: bit 0 clear, bits 1-4 sidx, 5 occupancy, 6 up, 7 proposal
        x[0] = 0;
        x[1] = rin(v0);
        x[2] = rin(v1);
        x[3] = rin(v2);
        x[4] = rin(v3);
        x[5] = rin(occ);
        x[6] = rin(up);
        x[7] = 0;
    } else {
        // byte = 4 * sidx
        x[0] = x[1] = 0;
        x[2] = rin(v0);
        x[3] = rin(v1);
        x[4] = rin(v2);
        x[5] = rin(v3);
        x[6] = x[7] = 0;
    }
    bc_transpose8(x);
    // byte b of x[r]: site 8b + r (rotated positions)
    uint32_t lt1 = 0, lt2 = 0, altm = 0, own = 0;
    bc_u16x2 tz = {0xffff, 0xffff};
    const int nblk = W > 16 ? 4 : (W + 3) >> 2; // prefix blocks holding a valid site
    const uint32_t gb = 8u * group + (uint32_t)part;
    // Blocks i and i + 2 together: pair k = 4i + m and pair k + 8 hold sites
    // k, k + 16 and k + 8, k + 24, so one v_perm packs the four sites' sign
    // (or proposal) bits into one word whose shift by k puts them in place.
#pragma unroll
    for (int ip = 0; ip < 2; ++ip) {
        uint32_t d1[2][4], d2[2][4], fw[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = ip + 2 * h;
#pragma unroll
            for (int m = 0; m < 4; ++m)
                d1[h][m] = d2[h][m] = fw[h][m] = 0; // no decision: sign clear
            if ((i & (lpw - 1)) != 0 || part + i >= nblk)
                continue;
            const uint4 R = philox4x32_10(gb + (uint32_t)i, sw_lo, w2, c3, k0, k1);
            const uint32_t Rw[4] = {R.x, R.y, R.z, R.w};
            own |= 0x000f000fu << (4 * i);
            // the block's 8 table loads first, then the 4 packed compares
            uint32_t tlo[4], thi[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int k = 4 * i + m;
                // class bytes of sites k, k + 16: bytes 0 and 2 of x[k] (k < 8)
                // or bytes 1 and 3 of x[k - 8]; Metropolis adds the proposal
                // bits (field bits 0 and 16) as bit 7 of each byte
                uint32_t lo, hi;
                if (k < 8) {
                    const uint32_t c = ALG == 0
                        ? __builtin_amdgcn_bitop3_b32(Rw[m] << 7, 0x00800080u, x[k], 0xEA)
                        : x[k];
                    lo = c & 0xffu;
                    hi = __builtin_amdgcn_ubfe(c, 16u, 8u);
                } else {
                    const uint32_t c = ALG == 0
                        ? __builtin_amdgcn_bitop3_b32(Rw[m] << 15, 0x80008000u, x[k - 8], 0xEA)
                        : x[k - 8];
                    lo = __builtin_amdgcn_ubfe(c, 8u, 8u);
                    hi = c >> 24;
                }
                if (ALG == 0) {
                    tlo[m] = *reinterpret_cast<const uint16_t*>(
                        reinterpret_cast<const char*>(T.a16) + lo);
                    thi[m] = *reinterpret_cast<const uint16_t*>(
                        reinterpret_cast<const char*>(T.a16) + hi);
                } else {
                    tlo[m] = *reinterpret_cast<const uint32_t*>(
                        reinterpret_cast<const char*>(T.a32) + lo);
                    thi[m] = *reinterpret_cast<const uint32_t*>(
                        reinterpret_cast<const char*>(T.a32) + hi);
                }
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t f = Rw[m];
                const bc_s16x2 fs = __builtin_bit_cast(bc_s16x2, f);
                fw[h][m] = f;
                if (ALG == 0) {
                    const bc_s16x2 th = __builtin_bit_cast(
                        bc_s16x2, __builtin_amdgcn_perm(thi[m], tlo[m], 0x05040100u));
                    const bc_s16x2 d = __builtin_elementwise_sub_sat(fs, th);
                    d1[h][m] = __builtin_bit_cast(uint32_t, d);
                    tz = __builtin_elementwise_min(tz, __builtin_bit_cast(bc_u16x2, d));
                } else {
                    const bc_s16x2 th1 = __builtin_bit_cast(
                        bc_s16x2, __builtin_amdgcn_perm(thi[m], tlo[m], 0x05040100u));
                    const bc_s16x2 th2 = __builtin_bit_cast(
                        bc_s16x2, __builtin_amdgcn_perm(thi[m], tlo[m], 0x07060302u));
                    const bc_s16x2 e1 = __builtin_elementwise_sub_sat(fs, th1);
                    const bc_s16x2 e2 = __builtin_elementwise_sub_sat(fs, th2);
                    d1[h][m] = __builtin_bit_cast(uint32_t, e1);
                    d2[h][m] = __builtin_bit_cast(uint32_t, e2);
                    tz = __builtin_elementwise_min(
                        tz, __builtin_elementwise_min(__builtin_bit_cast(bc_u16x2, e1),
                                                      __builtin_bit_cast(bc_u16x2, e2)));
                }
            }
        }
        // sign bits 15 / 31 of pairs k and k + 8 -> bits 7, 15, 23, 31 (sites
        // k, k + 8, k + 16, k + 24); proposal bits 0 / 16 -> bits 0, 8, 16, 24
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int k = 4 * ip + m;
            const uint32_t msk = 0x01010101u << k;
            lt1 |= (__builtin_amdgcn_perm(d1[1][m], d1[0][m], 0x07030501u) >> (7 - k)) & msk;
            if (ALG == 0)
                altm |= (__builtin_amdgcn_perm(fw[1][m], fw[0][m], 0x06020400u) << k) & msk;
            else
                lt2 |= (__builtin_amdgcn_perm(d2[1][m], d2[0][m], 0x07030501u) >> (7 - k)) & msk;
        }
    }
    const uint32_t tzw = __builtin_bit_cast(uint32_t, tz);
    const uint32_t tie_bits = ALG == 0 ? 0xfffefffeu : 0xffffffffu;
    if ((tzw & tie_bits & 0xffffu) == 0 || (tzw & tie_bits & 0xffff0000u) == 0) {
        // rare: a site whose prefix equals its threshold's leading bits; mark
        // those sites, then decide each on its full uniform
        uint32_t tmask = 0;
        const bc_u16x2 one = {1, 1};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if ((i & (lpw - 1)) != 0 || part + i >= nblk)
                continue;
            const uint4 R = philox4x32_10(gb + (uint32_t)i, sw_lo, w2, c3, k0, k1);
            const uint32_t Rw[4] = {R.x, R.y, R.z, R.w};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int k = 4 * i + m;
                const uint32_t f = Rw[m];
                uint32_t c = k < 8 ? x[k] : x[k - 8] >> 8;
                const bc_s16x2 fs = __builtin_bit_cast(bc_s16x2, f);
                uint32_t z; // bit 15 / 31: the low / high site ties
                if (ALG == 0) {
                    c |= (f << 7) & 0x00800080u;
                    bc_s16x2 th;
                    th.x = *reinterpret_cast<const int16_t*>(
                        reinterpret_cast<const char*>(T.a16) + (c & 0xffu));
                    th.y = *reinterpret_cast<const int16_t*>(
                        reinterpret_cast<const char*>(T.a16) + ((c >> 16) & 0xffu));
                    const bc_u16x2 e = __builtin_bit_cast(
                        bc_u16x2, __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(fs, th)) &
                                      0xfffefffeu);
                    // a zero half of e: (e - 1) has bit 15 set and e has not
                    z = __builtin_bit_cast(uint32_t, e - one) & ~__builtin_bit_cast(uint32_t, e);
                } else {
                    const uint32_t elo = *reinterpret_cast<const uint32_t*>(
                        reinterpret_cast<const char*>(T.a32) + (c & 0xffu));
                    const uint32_t ehi = *reinterpret_cast<const uint32_t*>(
                        reinterpret_cast<const char*>(T.a32) + ((c >> 16) & 0xffu));
                    const uint32_t t1 = __builtin_amdgcn_perm(ehi, elo, 0x05040100u);
                    const uint32_t t2 = __builtin_amdgcn_perm(ehi, elo, 0x07060302u);
                    const bc_u16x2 e1 = __builtin_bit_cast(bc_u16x2, f ^ t1);
                    const bc_u16x2 e2 = __builtin_bit_cast(bc_u16x2, f ^ t2);
                    z = (__builtin_bit_cast(uint32_t, e1 - one) & ~__builtin_bit_cast(uint32_t, e1)) |
                        (__builtin_bit_cast(uint32_t, e2 - one) & ~__builtin_bit_cast(uint32_t, e2));
                }
                tmask |= (z >> (15 - k)) & (0x00010001u << k);
            }
        }
        while (tmask) {
            const int kv = __builtin_ctz(tmask); // rotated position
            tmask &= tmask - 1u;
            const int j = (kv + 4 * part) & 31; // the site's bit
            if (j >= W)
                continue;
            const uint32_t f = bc_site_field(group, j, sw_lo, w2, c3, k0, k1);
            const uint32_t t16 = bc_site_tail(group, j, sw_lo, w2, c3, k0, k1);
            const int sidx = (int)(((v0 >> j) & 1u) | (((v1 >> j) & 1u) << 1) |
                                   (((v2 >> j) & 1u) << 2) | (((v3 >> j) & 1u) << 3));
            if (ALG == 0) {
                const int cls = sidx | (int)(((occ >> j) & 1u) << 4) |
                                (int)(((up >> j) & 1u) << 5) | (int)((f & 1u) << 6);
                const uint32_t u31 = (((f ^ 0x8000u) >> 1) << 16) | t16;
                lt1 = (lt1 & ~(1u << kv)) | ((u31 < T.thr[cls] ? 1u : 0u) << kv);
            } else {
                const uint64_t U = ((uint64_t)(f ^ 0x8000u) << 16) | t16;
                lt1 = (lt1 & ~(1u << kv)) | ((U < T.hbw[sidx][0] ? 1u : 0u) << kv);
                lt2 = (lt2 & ~(1u << kv)) | ((U < T.hbw[sidx][1] ? 1u : 0u) << kv);
            }
        }
    }
    // back to the sites' own positions
    const uint32_t back = (32u - rot) & 31u;
    auto rout = [&](uint32_t v) { return SHARED ? rotr32(v, back) : v; };
    own = rout(own) & valid;
    lt1 = rout(lt1);
    uint32_t xo, xu, n_changed;
    if (ALG == 0) {
        altm = rout(altm);
        const uint32_t okm = lt1 & own;
        xo = okm & (~occ | (up ^ altm));
        xu = okm & (~altm | up);
        n_changed = (uint32_t)__popc(okm);
    } else {
        lt2 = rout(lt2);
        // s_new = -1 if U < W1, else 0 if U < W12, else +1
        const uint32_t nz = lt1 | ~lt2; // s_new != 0
        const uint32_t nu = ~(lt1 | lt2); // s_new = +1
        xo = (nz ^ occ) & own;
        xu = (nu ^ up) & own;
        n_changed = (uint32_t)__popc(xo | xu);
    }
    if (SHARED) {
        for (int off = 1; off < lpw; off <<= 1) {
            xo |= __shfl_xor(xo, off, 64);
            xu |= __shfl_xor(xu, off, 64);
        }
    }
    occ ^= xo;
    up ^= xu;
    return n_changed;
}

// ONE: one lane per word (lanes_per_word 0 or 1), its own instantiation without
// the shared-word path's registers.
template <int DIM, int ALG, bool ONE>
__global__ void __launch_bounds__(1024) k_bc_sweeps(SweepArgs a)
{
    extern __shared__ uint32_t lds[];
    // the class tables at a fixed LDS address (their loads take it as an
    // immediate offset), the rest in the dynamic allocation
    __shared__ BCTabs tabs;
    static_assert(96 + 192 + 10 <= kBCLdsHeaderWords, "LDS header");
    static_assert(sizeof(BCTabs) == kBCStaticLds, "static LDS");
    int32_t* part = reinterpret_cast<int32_t*>(lds); // [2][16][3]
    int32_t* ring = reinterpret_cast<int32_t*>(lds + 96); // E[64], M[64], Q[64] of the last sweeps
    // thread 0's running FP64 sums (sum E, E^2, |M|, M^2, M^4), 8-byte aligned
    double* msum = reinterpret_cast<double*>(lds + 288);
    uint32_t* P = lds + kBCLdsHeaderWords;            // [occ c0, occ c1, up c0, up c1]
    const int tid = threadIdx.x;
    const int B = blockDim.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = (B + 63) >> 6;
    constexpr uint32_t purpose = ALG == 0 ? P_BC_METRO : P_BC_HB;

    for (int item = blockIdx.x; item < a.n_items; item += gridDim.x) {
        const int ch = a.items[item];
        const ChainRec cr = a.chains[ch];
        // output slot: the chain, or its ladder slot when recording in temperature order
        const int64_t oslot = a.by_table ? (int64_t)cr.table : (int64_t)ch;
        const int64_t sbase = a.by_table ? oslot * a.cap : cr.series_off;
        const long long t_item = a.item_ticks && tid == 0 ? wall_clock64() : 0;
        const uint32_t k0 = (uint32_t)cr.seed;
        const uint32_t c3 = (uint32_t)(cr.seed >> 32);
        const int L = cr.L;
        const int half = L >> 1;
        const int nw = (half + 31) >> 5;
        const int rows = DIM == 3 ? L * L : L;
        const int words = rows * nw;
        // lanes per word: as asked, or (negative) the most, up to 4 (one prefix
        // block each), that keep words x lanes within the workgroup; uniform
        // over the workgroup
        int lpw = a.lanes_per_word > 1 && !ONE ? a.lanes_per_word : 1;
        if (a.lanes_per_word < 0 && !ONE) {
            lpw = 4;
            while (lpw > 1 && words * lpw > B)
                lpw >>= 1;
        }
        const int lpart = tid & (lpw - 1);
        int8_t* __restrict__ sp = a.spins + cr.spin_off;
        const BCTable* tb = a.btab + cr.table;
        bc_fill_tabs<DIM, ALG>(tabs, tb, a.glauber, tid, B);

        // int8 lattice (or the planes a previous launch left) -> planes
        if (a.planes_in) {
            const uint32_t* src = a.planes + a.plane_off[ch];
            for (int w = tid; w < 4 * words; w += B)
                P[w] = src[w];
        } else {
            for (int w = tid; w < words; w += B)
                bc_pack_word<DIM>(sp, P, w, L, nw, half, words);
        }
        __syncthreads();

        unsigned long long acc = 0;
        // running sums continue from the stored ones, sweep by sweep, so the
        // FP64 moments do not depend on how a run is split into launches
        if (a.record && tid == 0) {
            const double* mo = a.moments + oslot * 6;
            for (int k = 0; k < 5; ++k)
                msum[k] = mo[1 + k];
        }
        const int64_t nsw = a.item_nsweeps ? a.item_nsweeps[item] : a.nsweeps;
        const bool series = a.record && a.ser_E != nullptr;
        for (int64_t s = 0; s < nsw; ++s) {
            const uint64_t sweep = cr.sweep + (uint64_t)s;
            const uint32_t sw_lo = (uint32_t)sweep;
            const uint32_t w2 = ctr_word2(sweep, 0, purpose);
            const bool obs = a.record || (a.last_E != nullptr && s == nsw - 1);
            int32_t pr = 0, mm = 0, qq = 0;
            for (int c = 0; c < 2; ++c) {
                uint32_t* __restrict__ Oc = P + c * words;
                uint32_t* __restrict__ Sc = P + (2 + c) * words;
                const uint32_t* __restrict__ Oo = P + (c ^ 1) * words;
                const uint32_t* __restrict__ So = P + (2 + (c ^ 1)) * words;
                for (int w = tid / lpw; w < words; w += B / lpw) {
                    // (LDS-resident: every index fits 24 bits)
                    const int r = nw == 1 ? w : w / nw;
                    const int q = w - (int)__umul24((uint32_t)r, (uint32_t)nw);
                    const int y = DIM == 3 ? r % L : r;
                    const int z = DIM == 3 ? r / L : 0;
                    const int p = (c + y + z) & 1;
                    const int W = word_width(q, half);
                    const uint32_t valid = width_mask(W);
                    uint32_t occ = Oc[w], up = Sc[w];
                    BCNbr n;
                    bc_load_nbrs<DIM, true>(Oo, So, L, nw, half, r, q, y, z, p, W, valid, n);
                    uint32_t v0, v1, v2, v3;
                    bc_shifted_sum<DIM>(n, v0, v1, v2, v3);
                    const uint32_t group = __umul24((uint32_t)(r * 2 + c), (uint32_t)nw) + (uint32_t)q;
                    acc += !ONE && lpw > 1 ? bc_word<DIM, ALG, true>(occ, up, v0, v1, v2, v3, W, tabs,
                                                             group, sw_lo, w2, c3, k0, cr.key,
                                                             lpart, lpw)
                                   : bc_word<DIM, ALG, false>(occ, up, v0, v1, v2, v3, W, tabs,
                                                              group, sw_lo, w2, c3, k0, cr.key,
                                                              0, 1);
                    if (lpart == 0) {
                        Oc[w] = occ;
                        Sc[w] = up;
                    }
                    if (c == 1 && obs && lpart == 0) {
                        const uint32_t dn = occ & ~up;
                        const uint32_t upo = occ & up;
#pragma unroll
                        for (int k = 0; k < 2 * DIM; ++k) {
                            const uint32_t both = occ & n.occ[k];
                            const uint32_t anti = both & (up ^ n.up[k]);
                            pr += __popc(both) - 2 * __popc(anti);
                        }
                        const uint32_t Ao = n.occ[2], Au = n.up[2];
                        mm += __popc(upo) - __popc(dn) + __popc(Ao & Au) - __popc(Ao & ~Au);
                        qq += __popc(occ) + __popc(Ao);
                    }
                }
                if (c == 1 && obs) {
                    pr = wave_sum32(pr);
                    mm = wave_sum32(mm);
                    qq = wave_sum32(qq);
                    // (one wave: its sums stay in registers for the tail below)
                    if (lane == 0 && nwaves > 1) {
                        part[(s & 1) * 48 + wave * 3] = pr;
                        part[(s & 1) * 48 + wave * 3 + 1] = mm;
                        part[(s & 1) * 48 + wave * 3 + 2] = qq;
                    }
                }
                __syncthreads();
            }
            if (obs && tid == 0) {
                int64_t tp = pr, tm = mm, tq = qq;
                if (nwaves > 1) {
                    tp = tm = tq = 0;
                    for (int k = 0; k < nwaves; ++k) {
                        tp += part[(s & 1) * 48 + k * 3];
                        tm += part[(s & 1) * 48 + k * 3 + 1];
                        tq += part[(s & 1) * 48 + k * 3 + 2];
                    }
                }
                ring[s & 63] = (int32_t)(-tp);
                ring[64 + (s & 63)] = (int32_t)tm;
                ring[128 + (s & 63)] = (int32_t)tq;
                if (a.record) {
                    const double e = (double)(int32_t)(-tp) + cr.D * (double)(int32_t)tq;
                    const double m = (double)tm;
                    msum[0] += e;
                    msum[1] += e * e;
                    msum[2] += fabs(m);
                    msum[3] += m * m;
                    msum[4] += (m * m) * (m * m);
                }
            }
            // the series leaves the ring every 64 sweeps, coalesced
            if (series && ((s & 63) == 63 || s == nsw - 1)) {
                __syncthreads();
                const int64_t s0 = s & ~(int64_t)63;
                const int64_t at0 = sbase + a.rec_pos + s0;
                for (int t = tid; t <= (int)(s - s0); t += B) {
                    a.ser_E[at0 + t] = ring[t];
                    a.ser_M[at0 + t] = ring[64 + t];
                    a.ser_Q[at0 + t] = ring[128 + t];
                }
            }
        }
        if (a.last_E && nsw > 0 && tid == 0) {
            a.last_E[ch] = ring[(nsw - 1) & 63];
            a.last_M[ch] = ring[64 + ((nsw - 1) & 63)];
            a.last_Q[ch] = ring[128 + ((nsw - 1) & 63)];
        }

        // planes -> int8 lattice (or left as planes for the next launch)
        if (a.planes_out) {
            uint32_t* dst = a.planes + a.plane_off[ch];
            for (int w = tid; w < 4 * words; w += B)
                dst[w] = P[w];
        } else {
            for (int w = tid; w < words; w += B)
                bc_unpack_word<DIM>(sp, P, w, L, nw, half, words);
        }
        acc = wave_sum(acc);
        if (lane == 0 && acc)
            atomicAdd(&a.accepted[ch], acc);
        if (tid == 0) {
            if (a.record) {
                double* mo = a.moments + oslot * 6;
                mo[0] += (double)nsw;
                for (int k = 0; k < 5; ++k)
                    mo[1 + k] = msum[k];
            }
            a.chains[ch].sweep = cr.sweep + (uint64_t)nsw;
        }
        __syncthreads();
        if (a.item_ticks && tid == 0)
            a.item_ticks[item] = wall_clock64() - t_item;
    }
}

// Plain path: one thread per site of one colour of one chain, int8 in HBM.
template <int DIM, int ALG>
__global__ void __launch_bounds__(256) k_bc_plain(PlainArgs a, int n_col)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const ChainRec cr = a.chains[a.chain];
    const int L = a.L;
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int c = a.colour;
    constexpr int NS = 4 * DIM + 1;
    int e_part = 0, m_part = 0, q_part = 0;
    uint32_t flips = 0;
    if (idx < n_col) {
        const int r = idx / half;
        const int j = idx - r * half;
        const int y = DIM == 3 ? r % L : r;
        const int z = DIM == 3 ? r / L : 0;
        const int p = (c + y + z) & 1;
        const int x = 2 * j + p;
        int8_t* sp = a.spins + a.spin_off;
        const int64_t rowb = (int64_t)L * r;
        const int xl = x == 0 ? L - 1 : x - 1;
        const int xr = x == L - 1 ? 0 : x + 1;
        const int yu = y == 0 ? L - 1 : y - 1;
        const int yd = y == L - 1 ? 0 : y + 1;
        const int64_t zb = DIM == 3 ? (int64_t)z * L * L : 0;
        int nb = sp[rowb + xl] + sp[rowb + xr] + sp[zb + (int64_t)yu * L + x]
               + sp[zb + (int64_t)yd * L + x];
        if (DIM == 3) {
            const int zu = z == 0 ? L - 1 : z - 1;
            const int zd = z == L - 1 ? 0 : z + 1;
            nb += sp[(int64_t)zu * L * L + (int64_t)y * L + x]
                + sp[(int64_t)zd * L * L + (int64_t)y * L + x];
        }
        const int s_old = sp[rowb + x];
        const BCTable& tb = a.btab[cr.table];
        const uint32_t group = (uint32_t)((r * 2 + c) * nw + (j >> 5));
        const uint32_t sw_lo = (uint32_t)a.sweep;
        const uint32_t w2 = ctr_word2(a.sweep, 0, ALG == 0 ? P_BC_METRO : P_BC_HB);
        const uint32_t c3 = (uint32_t)(cr.seed >> 32), k0 = (uint32_t)cr.seed;
        const uint32_t f = bc_site_field(group, j & 31, sw_lo, w2, c3, k0, cr.key);
        const uint32_t u16 = f ^ 0x8000u;
        const int sidx = nb + 2 * DIM;
        int s_new;
        if (ALG == 0) {
            // U31 = (u16 >> 1) * 2^16 + tail < thr; the tail only on a tie
            const int alt = (int)(f & 1u);
            const int k = ((s_old + 1) * 2 + alt) * NS + sidx;
            const uint32_t thr = a.glauber ? tb.gl_thr[k] : tb.metro_thr[k];
            const uint32_t pre = u16 >> 1, A = thr >> 16;
            bool ok = pre < A;
            if (pre == A)
                ok = ((pre << 16) | bc_site_tail(group, j & 31, sw_lo, w2, c3, k0, cr.key)) < thr;
            s_new = ok ? bc_propose(s_old, alt) : s_old;
        } else {
            // U = u16 * 2^16 + tail against W1, W12
            const uint64_t W1 = tb.hb_w[sidx][0], W12 = tb.hb_w[sidx][1];
            const uint64_t A1 = W1 >> 16, A12 = W12 >> 16;
            uint64_t U = (uint64_t)u16 << 16;
            if (u16 == A1 || u16 == A12)
                U |= bc_site_tail(group, j & 31, sw_lo, w2, c3, k0, cr.key);
            s_new = U < W1 ? -1 : (U < W12 ? 0 : 1);
        }
        if (s_new != s_old) {
            sp[rowb + x] = (int8_t)s_new;
            flips = 1;
        }
        if (c == 1 && a.rec_index >= 0) {
            const int partner = sp[rowb + (p ? xl : xr)];
            e_part = -s_new * nb;
            m_part = s_new + partner;
            q_part = s_new * s_new + partner * partner;
        }
    }
    flips = wave_sum32(flips);
    if ((threadIdx.x & 63) == 0 && flips)
        atomicAdd(&a.accepted[a.chain], (unsigned long long)flips);
    if (c == 1 && a.rec_index >= 0) {
        e_part = wave_sum32(e_part);
        m_part = wave_sum32(m_part);
        q_part = wave_sum32(q_part);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.ser_E[cr.series_off + a.rec_index], e_part);
            atomicAdd(&a.ser_M[cr.series_off + a.rec_index], m_part);
            atomicAdd(&a.ser_Q[cr.series_off + a.rec_index], q_part);
        }
    }
}

// ---------------------------------------------------------------------------
// Tile kernel: one colour of one sweep of every work item, one thread per plane
// word, the planes in device memory.  Grid (word blocks, items).  Word update
// and observables as in k_bc_sweeps.
// ---------------------------------------------------------------------------

template <int DIM, int ALG>
__global__ void __launch_bounds__(256) k_bc_tiles(TileArgs a)
{
    __shared__ BCTabs tabs;
    const int item = blockIdx.y;
    if (a.s >= a.item_nsw[item])
        return;
    constexpr uint32_t purpose = ALG == 0 ? P_BC_METRO : P_BC_HB;
    const int tid = threadIdx.x;
    const int ch = a.items[item];
    const ChainRec cr = a.chains[ch];
    const BCTable* tb = a.btab + cr.table;
    bc_fill_tabs<DIM, ALG>(tabs, tb, a.glauber, tid, blockDim.x);
    __syncthreads();
    const int L = cr.L;
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int rows = DIM == 3 ? L * L : L;
    const int words = rows * nw;
    const int c = a.colour;
    const int lpw = a.lpw > 1 ? a.lpw : 1;
    const int part = tid & (lpw - 1);
    const int w = (blockIdx.x * blockDim.x + tid) / lpw;
    uint32_t flips = 0;
    int32_t pr = 0, mm = 0, qq = 0;
    if (w < words) {
        uint32_t* P = reinterpret_cast<uint32_t*>(a.base + a.item_plane[item]);
        uint32_t* Oc = P + c * words;
        uint32_t* Sc = P + (2 + c) * words;
        const uint32_t* Oo = P + (c ^ 1) * words;
        const uint32_t* So = P + (2 + (c ^ 1)) * words;
        const uint64_t sweep = cr.sweep + (uint64_t)a.s;
        const int r = nw == 1 ? w : w / nw;
        const int q = w - r * nw;
        const int y = DIM == 3 ? r % L : r;
        const int z = DIM == 3 ? r / L : 0;
        const int p = (c + y + z) & 1;
        const int W = word_width(q, half);
        const uint32_t valid = width_mask(W);
        uint32_t occ = Oc[w], up = Sc[w];
        BCNbr n;
        bc_load_nbrs<DIM>(Oo, So, L, nw, half, r, q, y, z, p, W, valid, n);
        uint32_t v0, v1, v2, v3;
        bc_shifted_sum<DIM>(n, v0, v1, v2, v3);
        const uint32_t group = (uint32_t)((r * 2 + c) * nw + q);
        flips = bc_word<DIM, ALG, true>(occ, up, v0, v1, v2, v3, W, tabs, group,
                                        (uint32_t)sweep, ctr_word2(sweep, 0, purpose),
                                        (uint32_t)(cr.seed >> 32), (uint32_t)cr.seed, cr.key,
                                        part, lpw);
        if (part == 0) {
            Oc[w] = occ;
            Sc[w] = up;
        }
        if (c == 1 && part == 0) {
            const uint32_t dn = occ & ~up;
            const uint32_t upo = occ & up;
#pragma unroll
            for (int k = 0; k < 2 * DIM; ++k) {
                const uint32_t both = occ & n.occ[k];
                const uint32_t anti = both & (up ^ n.up[k]);
                pr += __popc(both) - 2 * __popc(anti);
            }
            const uint32_t Ao = n.occ[2], Au = n.up[2];
            mm += __popc(upo) - __popc(dn) + __popc(Ao & Au) - __popc(Ao & ~Au);
            qq += __popc(occ) + __popc(Ao);
        }
    }
    tile_block_add(a, item, pr, mm, qq, (int32_t)flips);
}

template <int DIM>
__global__ void __launch_bounds__(256) k_bc_pack(TileArgs a, int unpack)
{
    const int item = blockIdx.y;
    const ChainRec cr = a.chains[a.items[item]];
    const int L = cr.L;
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int words = (DIM == 3 ? L * L : L) * nw;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words)
        return;
    uint32_t* P = reinterpret_cast<uint32_t*>(a.base + a.item_plane[item]);
    int8_t* sp = reinterpret_cast<int8_t*>(a.base) + cr.spin_off;
    if (unpack)
        bc_unpack_word<DIM>(sp, P, w, L, nw, half, words);
    else
        bc_pack_word<DIM>(sp, P, w, L, nw, half, words);
}

// Multi-workgroup path with temporal blocking (the scheme of k_ising_band):
// one band of a.band slabs plus 2T halo slabs on either side in LDS, T sweeps
// per launch, half-sweep k updating the local slabs [k + 1, R - 2 - k], the
// band written back, per-sweep observables over the band.  a.lpw lanes per
// plane word, as in k_bc_tiles.  LDS: [occ c0, occ c1, up c0, up c1][R slabs][rows per slab][nw].
template <int DIM, int ALG>
__global__ void __launch_bounds__(kBandThreads) k_bc_band(TileArgs a)
{
    extern __shared__ uint32_t bq[];
    __shared__ BCTabs tabs;
    const int item = blockIdx.y;
    const int64_t nsw = a.item_nsw[item];
    if (a.s >= nsw)
        return;
    constexpr uint32_t purpose = ALG == 0 ? P_BC_METRO : P_BC_HB;
    const int tid = threadIdx.x;
    const int nthr = (int)blockDim.x;
    const int ch = a.items[item];
    const ChainRec cr = a.chains[ch];
    const int L = cr.L;
    const int B = a.band;
    const int s0 = blockIdx.x * B;
    if (s0 >= L)
        return;
    const BCTable* tb = a.btab + cr.table;
    bc_fill_tabs<DIM, ALG>(tabs, tb, a.glauber, tid, nthr);
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int spr = DIM == 3 ? L : 1;
    const int H = 2 * a.T;
    const int Bb = B < L - s0 ? B : L - s0;
    const int R = Bb + 2 * H;
    const int rw = spr * nw;
    const int nloc = R * rw;
    const int words = (DIM == 3 ? L * L : L) * nw;
    uint32_t* P = reinterpret_cast<uint32_t*>(a.base + a.item_plane[item]);
    // global slab of local slab 0, in [0, L): the halo may be longer than the lattice
    const int g0 = ((s0 - H) % L + L) % L;
    for (int i = tid; i < 4 * nloc; i += nthr) {
        const int pl = i / nloc;
        const int li = i - pl * nloc;
        const int ls = li / rw;
        bq[i] = P[pl * words + ((g0 + ls) % L) * rw + (li - ls * rw)];
    }
    __syncthreads();
    const uint32_t k0 = (uint32_t)cr.seed, c3 = (uint32_t)(cr.seed >> 32);
    const int lpw = a.lpw > 1 ? a.lpw : 1;
    const int part = tid & (lpw - 1);
    const int64_t left = nsw - a.s;
    const int Tn = left < (int64_t)a.T ? (int)left : a.T;
    for (int t = 0; t < Tn; ++t) {
        const uint64_t sweep = cr.sweep + (uint64_t)(a.s + t);
        const uint32_t w2 = ctr_word2(sweep, 0, purpose);
        uint32_t flips = 0;
        int32_t pr = 0, mm = 0, qq = 0;
        for (int c = 0; c < 2; ++c) {
            const int k = 2 * t + c;
            const int lo = k + 1, n_upd = (R - 2 - k - lo + 1) * rw;
            uint32_t* Oc = bq + c * nloc;
            uint32_t* Sc = bq + (2 + c) * nloc;
            const uint32_t* Oo = bq + (c ^ 1) * nloc;
            const uint32_t* So = bq + (2 + (c ^ 1)) * nloc;
            // the lanes of one word share i, so the group stays converged for its shuffles
            for (int i = tid / lpw; i < n_upd; i += nthr / lpw) {
                const int ls = lo + i / rw;
                const int rem = i - (ls - lo) * rw;
                const int yl = DIM == 3 ? rem / nw : 0;
                const int q = rem - yl * nw;
                const int gs = (g0 + ls) % L;
                const int y = DIM == 3 ? yl : gs;
                const int z = DIM == 3 ? gs : 0;
                const int r = DIM == 3 ? z * L + y : y;
                const int lr = ls * spr + yl;
                int iU, iD, iZu = 0, iZd = 0;
                if (DIM == 2) {
                    iU = lr - 1;
                    iD = lr + 1;
                } else {
                    iU = ls * L + (yl == 0 ? L - 1 : yl - 1);
                    iD = ls * L + (yl == L - 1 ? 0 : yl + 1);
                    iZu = lr - L;
                    iZd = lr + L;
                }
                const int p = (c + y + z) & 1;
                const int W = word_width(q, half);
                const uint32_t valid = width_mask(W);
                uint32_t occ = Oc[lr * nw + q], up = Sc[lr * nw + q];
                BCNbr n;
                bc_nbrs_rows<DIM>(Oo, So, nw, half, lr, iU, iD, iZu, iZd, q, p, W, valid, n);
                uint32_t v0, v1, v2, v3;
                bc_shifted_sum<DIM>(n, v0, v1, v2, v3);
                const uint32_t group = (uint32_t)((r * 2 + c) * nw + q);
                const uint32_t nchanged =
                    bc_word<DIM, ALG, true>(occ, up, v0, v1, v2, v3, W, tabs, group,
                                            (uint32_t)sweep, w2, c3, k0, cr.key, part, lpw);
                if (part == 0) {
                    Oc[lr * nw + q] = occ;
                    Sc[lr * nw + q] = up;
                }
                if (ls >= H && ls < H + Bb) {
                    flips += nchanged;
                    if (c == 1 && part == 0) {
                        const uint32_t dn = occ & ~up;
                        const uint32_t upo = occ & up;
#pragma unroll
                        for (int kk = 0; kk < 2 * DIM; ++kk) {
                            const uint32_t both = occ & n.occ[kk];
                            const uint32_t anti = both & (up ^ n.up[kk]);
                            pr += __popc(both) - 2 * __popc(anti);
                        }
                        const uint32_t Ao = n.occ[2], Au = n.up[2];
                        mm += __popc(upo) - __popc(dn) + __popc(Ao & Au) - __popc(Ao & ~Au);
                        qq += __popc(occ) + __popc(Ao);
                    }
                }
            }
            __syncthreads();
        }
        tile_block_add_s(a, item, a.s + t, pr, mm, qq, (int32_t)flips);
    }
    for (int i = tid; i < 4 * Bb * rw; i += nthr) {
        const int pl = i / (Bb * rw);
        const int li = i - pl * Bb * rw;
        const int b = li / rw;
        P[pl * words + ((s0 + b) % L) * rw + (li - b * rw)] = bq[pl * nloc + (H + b) * rw + (li - b * rw)];
    }
}

// Metropolis and Glauber share the kernels (ALG 0) and differ in the threshold
// table, chosen by the glauber flag; heat bath is ALG 1.
hipError_t launch_bc_tiles(int dim, int alg, const TileArgs& a0, int max_words, hipStream_t st)
{
    TileArgs a = a0;
    a.glauber = alg == MCX_GLAUBER ? 1 : 0;
    const bool hb = alg == MCX_HEATBATH;
    const int lpw = a.lpw > 1 ? a.lpw : 1;
    const dim3 grid(((int64_t)max_words * lpw + 255) / 256, a.n_items);
    if (dim == 2 && !hb)
        hipLaunchKernelGGL((k_bc_tiles<2, 0>), grid, dim3(256), 0, st, a);
    else if (dim == 2)
        hipLaunchKernelGGL((k_bc_tiles<2, 1>), grid, dim3(256), 0, st, a);
    else if (!hb)
        hipLaunchKernelGGL((k_bc_tiles<3, 0>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_bc_tiles<3, 1>), grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bc_band(int dim, int alg, const TileArgs& a0, int max_bands, int threads,
                          size_t lds, hipStream_t st)
{
    TileArgs a = a0;
    a.glauber = alg == MCX_GLAUBER ? 1 : 0;
    const bool hb = alg == MCX_HEATBATH;
    const dim3 grid(max_bands, a.n_items), bl(threads);
    if (dim == 2 && !hb)
        hipLaunchKernelGGL((k_bc_band<2, 0>), grid, bl, lds, st, a);
    else if (dim == 2)
        hipLaunchKernelGGL((k_bc_band<2, 1>), grid, bl, lds, st, a);
    else if (!hb)
        hipLaunchKernelGGL((k_bc_band<3, 0>), grid, bl, lds, st, a);
    else
        hipLaunchKernelGGL((k_bc_band<3, 1>), grid, bl, lds, st, a);
    return hipGetLastError();
}

hipError_t launch_bc_pack(int dim, int unpack, const TileArgs& a, int max_words, hipStream_t st)
{
    const dim3 grid((max_words + 255) / 256, a.n_items);
    if (dim == 2)
        hipLaunchKernelGGL((k_bc_pack<2>), grid, dim3(256), 0, st, a, unpack);
    else
        hipLaunchKernelGGL((k_bc_pack<3>), grid, dim3(256), 0, st, a, unpack);
    return hipGetLastError();
}

hipError_t launch_bc_sweeps(int dim, int alg, const SweepArgs& a0, int block, int grid,
                            size_t lds_bytes, hipStream_t st)
{
    SweepArgs a = a0;
    a.glauber = alg == MCX_GLAUBER ? 1 : 0;
    const bool hb = alg == MCX_HEATBATH;
    // lds_bytes is the whole footprint; the class tables are static
    lds_bytes = lds_bytes > (size_t)kBCStaticLds ? lds_bytes - (size_t)kBCStaticLds : 0;
    const bool one = a.lanes_per_word == 0 || a.lanes_per_word == 1;
#define MCX_BCS(D, A)                                                                        \
    do {                                                                                     \
        if (one)                                                                             \
            hipLaunchKernelGGL((k_bc_sweeps<D, A, true>), dim3(grid), dim3(block), lds_bytes, \
                               st, a);                                                       \
        else                                                                                 \
            hipLaunchKernelGGL((k_bc_sweeps<D, A, false>), dim3(grid), dim3(block), lds_bytes, \
                               st, a);                                                       \
    } while (0)
    if (dim == 2 && !hb)
        MCX_BCS(2, 0);
    else if (dim == 2)
        MCX_BCS(2, 1);
    else if (!hb)
        MCX_BCS(3, 0);
    else
        MCX_BCS(3, 1);
#undef MCX_BCS
    return hipGetLastError();
}

hipError_t launch_bc_plain(int dim, int alg, const PlainArgs& a0, int n_col, hipStream_t st)
{
    PlainArgs a = a0;
    a.glauber = alg == MCX_GLAUBER ? 1 : 0;
    const bool hb = alg == MCX_HEATBATH;
    const int block = 256;
    const int grid = (n_col + block - 1) / block;
    if (dim == 2 && !hb)
        hipLaunchKernelGGL((k_bc_plain<2, 0>), dim3(grid), dim3(block), 0, st, a, n_col);
    else if (dim == 2)
        hipLaunchKernelGGL((k_bc_plain<2, 1>), dim3(grid), dim3(block), 0, st, a, n_col);
    else if (!hb)
        hipLaunchKernelGGL((k_bc_plain<3, 0>), dim3(grid), dim3(block), 0, st, a, n_col);
    else
        hipLaunchKernelGGL((k_bc_plain<3, 1>), dim3(grid), dim3(block), 0, st, a, n_col);
    return hipGetLastError();
}

} // namespace mcx
