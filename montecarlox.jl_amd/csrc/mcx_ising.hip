// mcx_ising.hip -- checkerboard Metropolis / heat-bath sweeps for Ising
// lattices (2D square, 3D simple cubic, periodic, J = 1, h = 0) on gfx950.
//
// Per-site rule (SpinSystems/src/ising.jl:35-58, src/algorithms/metropolis.jl:14-17,
// src/algorithms/importance_sampling.jl:80-85):
//   Metropolis: dE = 2 s_i sum_j s_j; accept iff -beta*dE > 0 or u < exp(-beta*dE)
//   heat bath:  s_new = +1 iff u < logistic(beta * s_i * dE); flip iff s_new != s_i
//   Glauber:    flip iff u < logistic(-beta * dE) (src/algorithms/metropolis.jl:121-127),
//               u from the Metropolis stream
// (ALG template argument: 0 Metropolis, 1 heat bath, 2 Glauber)
// visited in checkerboard order (colour 0 then colour 1; one colour's sites
// have disjoint neighbourhoods, so the parallel update of a colour equals the
// sequential one).
//
// Batched kernel (k_ising_sweeps): one workgroup per chain, the lattice held in
// LDS as bit planes for all the sweeps of the call.  Plane word (c, r, q) holds
// bit j of the 32 sites x = 2*(32q + j) + p of colour c in row r = y + L*z, with
// p = (c + y + z) & 1, so one 32-bit VALU op updates 32 sites.  Each site's
// uniform is bit-sliced: bit b (MSB first) of its 32-bit word w is bit j of
// word (b & 3) of Philox block (group, sweep, chunk = b >> 2, purpose), group =
// (2r + c)*nw + q.  Lattices with L < 64 and L % 8 == 0 use a flat layout
// instead (ising_flat): each colour's sites in (z, y, x) order, 32 to a word,
// rows split across words, group 2*word + c -- no partial words, so a 3D L=48
// lattice takes 1728 words per colour rather than 2304 of 24 sites.
// u < p  <=>  w < ceil(p * 2^32) is decided bit-serially
// MSB first, and a wavefront draws the next Philox block only while one of its
// sites is still tied with its threshold -- usually three blocks for 32 sites.
//
// Plain kernel (k_ising_plain): the one-chain-per-launch path -- one launch per
// colour per chain per sweep, one thread per site, int8 spins in HBM.  It
// evaluates the same bit-sliced uniform and therefore gives identical results.
#include "mcx_device.h"
#include "../../include/mcx_hip.h"

namespace mcx {

struct NbrWords {
    uint32_t A;  // other colour, same row, same bits (x-neighbour on one side)
    uint32_t Bw; // other colour, same row, shifted (x-neighbour on the other side)
    uint32_t U, Dn, Zu, Zd;
};

// Neighbour words of word q of plane row iA, given the plane rows of its
// y-1, y+1 (and z-1, z+1) neighbours in the other colour's plane Po.
// SMALL: every row index times nw fits 24 bits (LDS-resident planes), so the
// products use the full-rate 24-bit multiply.
template <int DIM, bool SMALL = false>
__device__ __forceinline__ void nbr_words_rows(const uint32_t* __restrict__ Po, int nw, int half,
                                               int iA, int iU, int iD, int iZu, int iZd, int q,
                                               int p, int W, uint32_t valid, NbrWords& n)
{
    auto mul = [&](int a) {
        return SMALL ? (int)__umul24((uint32_t)a, (uint32_t)nw) : a * nw;
    };
    const int r = iA;
    n.A = Po[mul(r) + q];
    n.U = Po[mul(iU) + q];
    n.Dn = Po[mul(iD) + q];
    if (DIM == 3) {
        n.Zu = Po[mul(iZu) + q];
        n.Zd = Po[mul(iZd) + q];
    } else {
        n.Zu = 0;
        n.Zd = 0;
    }
    if (nw == 1) {
        // whole half-row in one word: rotate within W bits, right by 1 (p) or
        // left by 1 = right by W - 1, the direction as a shift amount
        const uint32_t sr = p ? 1u : (uint32_t)(W - 1);
        n.Bw = ((n.A >> sr) | (n.A << ((uint32_t)W - sr))) & valid;
    } else if (p) {
        // x = 2j+1: right neighbour is other-colour bit j+1
        const int qn = q + 1 == nw ? 0 : q + 1;
        n.Bw = (n.A >> 1) | ((Po[mul(r) + qn] & 1u) << (W - 1));
    } else {
        // x = 2j: left neighbour is other-colour bit j-1
        const int qp = q == 0 ? nw - 1 : q - 1;
        const int Wp = word_width(qp, half);
        n.Bw = ((n.A << 1) | ((Po[mul(r) + qp] >> (Wp - 1)) & 1u)) & valid;
    }
}

template <int DIM, bool SMALL = false>
__device__ __forceinline__ void load_nbr_words(const uint32_t* __restrict__ Po, int L, int nw,
                                               int half, int r, int q, int y, int z, int p,
                                               int W, uint32_t valid, NbrWords& n)
{
    const int base = DIM == 3 ? z * L : 0;
    const int yu = y == 0 ? L - 1 : y - 1;
    const int yd = y == L - 1 ? 0 : y + 1;
    const int zu = DIM == 3 ? (z == 0 ? L - 1 : z - 1) : 0;
    const int zd = DIM == 3 ? (z == L - 1 ? 0 : z + 1) : 0;
    nbr_words_rows<DIM, SMALL>(Po, nw, half, r, base + yu, base + yd, zu * L + y, zd * L + y, q,
                               p, W, valid, n);
}

// Bit-serial "w < thr[class]" for the 32 sites of one word.  eq: sites still
// undecided (tied so far); m[k]: sites of class k.  Returns the "less" mask.
template <int NC, bool DRAW3>
__device__ __forceinline__ uint32_t bitserial_less(uint32_t eq, const uint32_t (&m)[NC],
                                                   const uint32_t (&thr)[NC], uint32_t group,
                                                   uint32_t sw_lo, uint32_t w2, uint32_t c3,
                                                   uint32_t k0, uint32_t k1)
{
    uint32_t lt = 0;
    if (!DRAW3) {
        for (uint32_t chunk = 0; chunk < 8; ++chunk) {
            if (!__any(eq != 0u))
                break;
            // (the chunk's counter word is wave-uniform: its Philox products
            // with the uniform words go to the scalar unit)
            const uint32_t cw = __builtin_amdgcn_readfirstlane(w2 | (chunk << 16));
            const uint4 R = philox4x32_10(group, sw_lo, cw, c3, k0, k1);
            const uint32_t Rb[4] = {R.x, R.y, R.z, R.w};
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t sh = 31u - (4u * chunk + (uint32_t)b);
                uint32_t Tb = 0;
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    // all-ones where the class's threshold has bit sh (wave-
                    // uniform); kept opaque so that the planes combine with
                    // and / and-or ops rather than one select per class
                    uint32_t tm = 0u - ((thr[k] >> sh) & 1u);
                    asm volatile("" : "+s"(tm));
                    Tb = k == 0 ? (m[k] & tm) : __builtin_amdgcn_bitop3_b32(m[k], tm, Tb, 0xEA);
                }
                // less where tied, the threshold bit is 1 and the random bit
                // 0; still tied where the two bits agree
                lt |= __builtin_amdgcn_bitop3_b32(eq, Tb, Rb[b], 0x40);
                eq = __builtin_amdgcn_bitop3_b32(eq, Rb[b], Tb, 0x90);
            }
        }
        return lt;
    }
    // the four bit-planes of chunk `chunk` (block R)
    auto planes = [&](const uint4& R, uint32_t chunk) {
        const uint32_t Rb[4] = {R.x, R.y, R.z, R.w};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t sh = 31u - (4u * chunk + (uint32_t)b);
            uint32_t Tb = 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                uint32_t tm = 0u - ((thr[k] >> sh) & 1u);
                asm volatile("" : "+s"(tm));
                Tb = k == 0 ? (m[k] & tm) : __builtin_amdgcn_bitop3_b32(m[k], tm, Tb, 0xEA);
            }
            lt |= __builtin_amdgcn_bitop3_b32(eq, Tb, Rb[b], 0x40);
            eq = __builtin_amdgcn_bitop3_b32(eq, Rb[b], Tb, 0x90);
        }
    };
    auto block = [&](uint32_t chunk) {
        const uint32_t cw = __builtin_amdgcn_readfirstlane(w2 | (chunk << 16));
        return philox4x32_10(group, sw_lo, cw, c3, k0, k1);
    };
    // DRAW3: a wavefront's 2048 sites nearly always need three chunks (a site
    // stays tied past 8 bits with probability ~2^-8), so the first three
    // blocks are drawn together -- three independent Philox chains the
    // scheduler interleaves -- and only later chunks wait for the tie test.
    // It costs registers, which pays where a launch has few waves per SIMD to
    // hide latency with (one 64-lane wave per L=64 chain: +5.5 %) and not
    // where occupancy does that (the flat layout's small lattices, or the
    // many-workgroup tile and band kernels).
    uint32_t chunk = 0;
    if (__any(eq != 0u)) {
        const uint4 R0 = block(0), R1 = block(1), R2 = block(2);
        planes(R0, 0);
        planes(R1, 1);
        planes(R2, 2);
        chunk = 3;
    }
    for (; chunk < 8; ++chunk) {
        if (!__any(eq != 0u))
            break;
        planes(block(chunk), chunk);
    }
    return lt;
}

// int8 lattice row word -> bit planes (up = 1) for plane word w.
template <int DIM>
__device__ __forceinline__ void ising_pack_word(const int8_t* __restrict__ sp,
                                                uint32_t* __restrict__ P, int w, int L, int nw,
                                                int half, int words)
{
    const int r = w / nw;
    const int q = w - r * nw;
    const int y = DIM == 3 ? r % L : r;
    const int z = DIM == 3 ? r / L : 0;
    const int par = (y + z) & 1; // colour of the even-x sites of this row
    const int W = word_width(q, half);
    const int8_t* row = sp + (int64_t)r * L + 64 * q;
    uint32_t ev = 0, od = 0;
    if ((L & 3) == 0) {
        const uint32_t* row4 = reinterpret_cast<const uint32_t*>(row);
        for (int k = 0; k < (W >> 1); ++k) {
            const uint32_t v = ~row4[k];
            ev |= (((v >> 7) & 1u) << (2 * k)) | (((v >> 23) & 1u) << (2 * k + 1));
            od |= (((v >> 15) & 1u) << (2 * k)) | (((v >> 31) & 1u) << (2 * k + 1));
        }
    } else {
        for (int j = 0; j < W; ++j) {
            ev |= (uint32_t)(row[2 * j] > 0) << j;
            od |= (uint32_t)(row[2 * j + 1] > 0) << j;
        }
    }
    P[par * words + w] = ev;
    P[(par ^ 1) * words + w] = od;
}

// Bit planes -> int8 lattice for plane word w.
template <int DIM>
__device__ __forceinline__ void ising_unpack_word(int8_t* __restrict__ sp,
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
    const uint32_t ev = P[par * words + w];
    const uint32_t od = P[(par ^ 1) * words + w];
    if ((L & 3) == 0) {
        uint32_t* row4 = reinterpret_cast<uint32_t*>(row);
        for (int k = 0; k < (W >> 1); ++k) {
            const uint32_t e0 = (ev >> (2 * k)) & 1u, e1 = (ev >> (2 * k + 1)) & 1u;
            const uint32_t o0 = (od >> (2 * k)) & 1u, o1 = (od >> (2 * k + 1)) & 1u;
            row4[k] = (0xFFu - 0xFEu * e0) | ((0xFFu - 0xFEu * o0) << 8)
                    | ((0xFFu - 0xFEu * e1) << 16) | ((0xFFu - 0xFEu * o1) << 24);
        }
    } else {
        for (int j = 0; j < W; ++j) {
            row[2 * j] = ((ev >> j) & 1u) ? 1 : -1;
            row[2 * j + 1] = ((od >> j) & 1u) ? 1 : -1;
        }
    }
}

// ---- flat layout (ising_flat(L)): bit i of word w of colour c is the colour-c
// site with flat index f = 32w + i = (y + L z) * L/2 + j, x = 2j + ((c + y + z) & 1).
// A z-plane is pw = L * L/2 / 32 whole words, so y-neighbours are the words
// shifted by L/2 bits within the plane (cyclically), z-neighbours the words one
// plane away, and the x-neighbours the other colour's same bit and its bit one
// to the left or right within the row (cyclically within the row).

// int8 lattice -> flat word w (0 .. 2 fw) of both colours' planes.
template <int DIM>
__device__ __forceinline__ void ising_pack_flat(const int8_t* __restrict__ sp,
                                                uint32_t* __restrict__ P, int w, int L, int half,
                                                int fw)
{
    const int c = w >= fw ? 1 : 0;
    const int f = 32 * (w - c * fw);
    int r = f / half, j = f - r * half;
    int y = DIM == 3 ? r % L : r, z = DIM == 3 ? r / L : 0;
    uint32_t bits = 0;
    for (int i = 0; i < 32; ++i) {
        const int x = 2 * j + ((c + y + z) & 1);
        bits |= (uint32_t)(sp[(int64_t)r * L + x] > 0) << i;
        if (++j == half) {
            j = 0;
            ++r;
            if (++y == L) {
                y = 0;
                ++z;
            }
        }
    }
    P[w] = bits;
}

// flat word w (0 .. 2 fw) -> int8 lattice
template <int DIM>
__device__ __forceinline__ void ising_unpack_flat(int8_t* __restrict__ sp,
                                                  const uint32_t* __restrict__ P, int w, int L,
                                                  int half, int fw)
{
    const int c = w >= fw ? 1 : 0;
    const int f = 32 * (w - c * fw);
    int r = f / half, j = f - r * half;
    int y = DIM == 3 ? r % L : r, z = DIM == 3 ? r / L : 0;
    const uint32_t bits = P[w];
    for (int i = 0; i < 32; ++i) {
        const int x = 2 * j + ((c + y + z) & 1);
        sp[(int64_t)r * L + x] = ((bits >> i) & 1u) ? 1 : -1;
        if (++j == half) {
            j = 0;
            ++r;
            if (++y == L) {
                y = 0;
                ++z;
            }
        }
    }
}

// Per-lattice constants of the flat neighbour words.
struct FlatGeo {
    int half, pw, fw;  // bits per row, words per z-plane, words per colour
    uint32_t starts;   // row starts of a word whose bit 0 starts a row
    uint32_t odd;      // bits of that word in rows 1, 3, ... after bit 0's row
    uint32_t mdiv;     // n / half = (n * mdiv) >> 16 for the n < 2^11 used here
    uint32_t mplane;   // w / pw = (w * mplane) >> 20 for w < 2^14
};

__device__ __forceinline__ FlatGeo flat_geo(int dim, int L)
{
    FlatGeo g;
    g.half = L >> 1;
    g.pw = (L * g.half) >> 5;
    g.fw = dim == 3 ? g.pw * L : g.pw;
    g.starts = 0;
    g.odd = 0;
    for (int k = 0; k * g.half < 32; ++k) {
        g.starts |= 1u << (k * g.half);
        if (k & 1)
            g.odd |= (0xffffffffu << (k * g.half))
                   & ((k + 1) * g.half >= 32 ? 0xffffffffu : ((1u << ((k + 1) * g.half)) - 1u));
    }
    g.mdiv = (65536u + (uint32_t)g.half - 1u) / (uint32_t)g.half;
    g.mplane = ((1u << 20) + (uint32_t)g.pw - 1u) / (uint32_t)g.pw;
    return g;
}

// Neighbour words of flat word w of colour c (in the other colour's plane Po).
// ROW32: L = 64, every word exactly one row (the flat and row layouts coincide).
template <int DIM, bool ROW32 = false>
__device__ __forceinline__ void load_nbr_flat(const uint32_t* __restrict__ Po, const FlatGeo& g,
                                              int L, int c, int w, NbrWords& n)
{
    if (ROW32) {
        // word w is row w: y-neighbours the adjacent words of the plane, the
        // far x-neighbour the other colour's word rotated by one bit
        const int zi = DIM == 3 ? w >> 6 : 0;
        const int pb = zi << 6;
        const int wl = w - pb;
        const uint32_t a0 = Po[w];
        n.A = a0;
        n.U = Po[pb + ((wl - 1) & 63)];
        n.Dn = Po[pb + ((wl + 1) & 63)];
        if (DIM == 3) {
            n.Zu = Po[(w - 64) & (64 * 64 - 1)];
            n.Zd = Po[(w + 64) & (64 * 64 - 1)];
        } else {
            n.Zu = 0;
            n.Zd = 0;
        }
        const uint32_t sr = ((c + wl + zi) & 1) ? 1u : 31u;
        n.Bw = __builtin_amdgcn_alignbit(a0, a0, sr);
        return;
    }
    const int zi = DIM == 3 ? (int)(__umul24((uint32_t)w, g.mplane) >> 20) : 0;
    const int pb = DIM == 3 ? (int)__umul24((uint32_t)zi, (uint32_t)g.pw) : 0;
    const int wl = w - pb;
    const int wp = pb + (wl == 0 ? g.pw - 1 : wl - 1);
    const int wn = pb + (wl == g.pw - 1 ? 0 : wl + 1);
    const uint32_t a0 = Po[w], am = Po[wp], ap = Po[wn];
    if (DIM == 3) {
        n.Zu = Po[zi == 0 ? w + g.fw - g.pw : w - g.pw];
        n.Zd = Po[zi == L - 1 ? w - g.fw + g.pw : w + g.pw];
    } else {
        n.Zu = 0;
        n.Zd = 0;
    }
    const uint32_t h = (uint32_t)g.half;
    // rows in this word: bit 0 is bit o of row y0; the next row starts at i0
    const uint32_t f0 = 32u * (uint32_t)wl;
    const uint32_t y0 = __umul24(f0, g.mdiv) >> 16;
    const uint32_t o = f0 - y0 * h;
    const uint32_t i0 = o == 0 ? 0u : h - o;
    const uint32_t y1 = __umul24(f0 + 32u, g.mdiv) >> 16;
    const uint32_t S = g.starts << i0;                                      // row starts
    const uint32_t E = (S >> 1) | (f0 + 32u == y1 * h ? 0x80000000u : 0u); // row ends
    // bits whose x-neighbour on the far side is to the right (p = 1)
    const uint32_t X = (g.odd << i0) | ((1u << i0) - 1u);
    const uint32_t yref = y0 + (i0 != 0 ? 1u : 0u);
    const uint32_t Pm = X ^ (0u - (((uint32_t)c + yref + (uint32_t)zi) & 1u));
    n.A = a0;
    n.U = __builtin_amdgcn_alignbit(a0, am, 32u - h); // f - half
    n.Dn = __builtin_amdgcn_alignbit(ap, a0, h);      // f + half
    const uint32_t left = __builtin_amdgcn_bitop3_b32(
        S, __builtin_amdgcn_alignbit(ap, a0, h - 1u), __builtin_amdgcn_alignbit(a0, am, 31u), 0xCA);
    const uint32_t right = __builtin_amdgcn_bitop3_b32(
        E, __builtin_amdgcn_alignbit(a0, am, 33u - h), __builtin_amdgcn_alignbit(ap, a0, 1u), 0xCA);
    n.Bw = __builtin_amdgcn_bitop3_b32(Pm, right, left, 0xCA);
}

// bitserial_less for a word shared by the four lanes of a quad (lane part p =
// lane % 4): in each round lane p draws chunk base + p and reduces its four
// bit-planes, from the sites still tied before the round, to (l, e) = (less
// within the chunk, still tied after it); the quad then composes the four
// pairs in chunk order, (l, e) then (l', e') -> (l | e & l', e & e'), with
// two DPP steps (from lane p + 1, then from lane p + 2; the last lanes read
// lane 3, which the composition leaves unchanged since l' and e' of one chunk
// are disjoint), and lane 0's result -- chunks base .. base + 3 -- goes to the
// whole quad.  The same bits and the same MSB-first comparison as
// bitserial_less, in a quarter of the dependent Philox work per lane: for
// launches with too few waves to hide latency.
template <int CTRL>
__device__ __forceinline__ uint32_t quad_from(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}

template <int NC>
__device__ __forceinline__ uint32_t bitserial_less_quad(uint32_t eq, const uint32_t (&m)[NC],
                                                        const uint32_t (&thr)[NC], uint32_t group,
                                                        uint32_t sw_lo, uint32_t w2, uint32_t c3,
                                                        uint32_t k0, uint32_t k1, uint32_t part)
{
    uint32_t lt = 0;
#pragma unroll
    for (uint32_t base = 0; base < 8; base += 4) {
        if (!__any(eq != 0u))
            break;
        const uint32_t chunk = base + part;
        const uint4 R = philox4x32_10(group, sw_lo, w2 | (chunk << 16), c3, k0, k1);
        const uint32_t Rb[4] = {R.x, R.y, R.z, R.w};
        uint32_t l = 0, e = eq;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t sh = 31u - (4u * chunk + (uint32_t)b);
            uint32_t Tb = 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const uint32_t tm = 0u - ((thr[k] >> sh) & 1u);
                Tb = k == 0 ? (m[k] & tm) : __builtin_amdgcn_bitop3_b32(m[k], tm, Tb, 0xEA);
            }
            l |= __builtin_amdgcn_bitop3_b32(e, Tb, Rb[b], 0x40);
            e = __builtin_amdgcn_bitop3_b32(e, Rb[b], Tb, 0x90);
        }
        uint32_t ln = quad_from<0xF9>(l), en = quad_from<0xF9>(e); // quad_perm [1,2,3,3]
        l |= e & ln;
        e &= en;
        ln = quad_from<0xFE>(l); // quad_perm [2,3,3,3]
        en = quad_from<0xFE>(e);
        l |= e & ln;
        e &= en;
        lt |= quad_from<0x00>(l); // quad_perm [0,0,0,0]
        eq = quad_from<0x00>(e);
    }
    return lt;
}

// One 32-site word of colour c: the new spins of the word from its old spins
// s0 and its neighbour words (the per-site rule above, bit-sliced).  Shared by
// the LDS-resident batched kernel and the multi-workgroup tile kernel.
// QUAD: the word is shared by the four lanes of a quad (bitserial_less_quad,
// lane part `part`); every lane of the quad returns the word's new spins.
template <int DIM, int ALG, bool DRAW3 = false, bool QUAD = false>
__device__ __forceinline__ uint32_t ising_word(uint32_t s0, const NbrWords& n, uint32_t valid,
                                               const IsingTable& tb, uint32_t group,
                                               uint32_t sw_lo, uint32_t w2, uint32_t c3,
                                               uint32_t k0, uint32_t k1, uint32_t part = 0)
{
    constexpr int NMC = DIM;         // Metropolis uphill classes
    constexpr int NHC = 2 * DIM + 1; // heat-bath classes
    if (ALG == 0) {
        uint32_t b0, b1, b2;
        if (DIM == 2)
            count4(s0 ^ n.U, s0 ^ n.Dn, s0 ^ n.A, s0 ^ n.Bw, b0, b1, b2);
        else
            count6(s0 ^ n.U, s0 ^ n.Dn, s0 ^ n.A, s0 ^ n.Bw, s0 ^ n.Zu, s0 ^ n.Zd, b0, b1, b2);
        uint32_t m[NMC];
        m[0] = count_eq<0>(b0, b1, b2) & valid;
        m[1] = count_eq<1>(b0, b1, b2) & valid;
        if (DIM == 3)
            m[NMC - 1] = count_eq<2>(b0, b1, b2) & valid;
        const uint32_t freem = DIM == 2 ? (b1 | b2) : (b2 | (b1 & b0));
        uint32_t forced = 0, need = 0;
        uint32_t thr[NMC];
#pragma unroll
        for (int k = 0; k < NMC; ++k) {
            thr[k] = tb.metro_thr[k];
            if ((tb.metro_always >> k) & 1u)
                forced |= m[k];
            else if (!((tb.metro_never >> k) & 1u))
                need |= m[k];
        }
        const uint32_t lt =
            QUAD ? bitserial_less_quad<NMC>(need, m, thr, group, sw_lo, w2, c3, k0, k1, part)
                 : bitserial_less<NMC, DRAW3>(need, m, thr, group, sw_lo, w2, c3, k0, k1);
        return s0 ^ ((freem | forced | lt) & valid);
    } else if (ALG == 2) {
        // Glauber: every class k (anti-aligned neighbours, 0..2 dim) has a threshold
        uint32_t b0, b1, b2;
        if (DIM == 2)
            count4(s0 ^ n.U, s0 ^ n.Dn, s0 ^ n.A, s0 ^ n.Bw, b0, b1, b2);
        else
            count6(s0 ^ n.U, s0 ^ n.Dn, s0 ^ n.A, s0 ^ n.Bw, s0 ^ n.Zu, s0 ^ n.Zd, b0, b1, b2);
        uint32_t m[NHC];
        m[0] = count_eq<0>(b0, b1, b2) & valid;
        m[1] = count_eq<1>(b0, b1, b2) & valid;
        m[2] = count_eq<2>(b0, b1, b2) & valid;
        m[3] = count_eq<3>(b0, b1, b2) & valid;
        m[4] = count_eq<4>(b0, b1, b2) & valid;
        if (DIM == 3) {
            m[NHC - 2] = count_eq<5>(b0, b1, b2) & valid;
            m[NHC - 1] = count_eq<6>(b0, b1, b2) & valid;
        }
        uint32_t forced = 0, need = 0;
        uint32_t thr[NHC];
#pragma unroll
        for (int k = 0; k < NHC; ++k) {
            thr[k] = tb.gl_thr[k];
            if ((tb.gl_always >> k) & 1u)
                forced |= m[k];
            else if (!((tb.gl_never >> k) & 1u))
                need |= m[k];
        }
        const uint32_t lt =
            QUAD ? bitserial_less_quad<NHC>(need, m, thr, group, sw_lo, w2, c3, k0, k1, part)
                 : bitserial_less<NHC, DRAW3>(need, m, thr, group, sw_lo, w2, c3, k0, k1);
        return s0 ^ ((forced | lt) & valid);
    } else {
        uint32_t b0, b1, b2;
        if (DIM == 2)
            count4(n.U, n.Dn, n.A, n.Bw, b0, b1, b2);
        else
            count6(n.U, n.Dn, n.A, n.Bw, n.Zu, n.Zd, b0, b1, b2);
        uint32_t m[NHC];
        m[0] = count_eq<0>(b0, b1, b2) & valid;
        m[1] = count_eq<1>(b0, b1, b2) & valid;
        m[2] = count_eq<2>(b0, b1, b2) & valid;
        m[3] = count_eq<3>(b0, b1, b2) & valid;
        m[4] = count_eq<4>(b0, b1, b2) & valid;
        if (DIM == 3) {
            m[NHC - 2] = count_eq<5>(b0, b1, b2) & valid;
            m[NHC - 1] = count_eq<6>(b0, b1, b2) & valid;
        }
        uint32_t up = 0, need = 0;
        uint32_t thr[NHC];
#pragma unroll
        for (int k = 0; k < NHC; ++k) {
            thr[k] = tb.hb_thr[k];
            if ((tb.hb_always >> k) & 1u)
                up |= m[k];
            else if (!((tb.hb_never >> k) & 1u))
                need |= m[k];
        }
        const uint32_t lt =
            QUAD ? bitserial_less_quad<NHC>(need, m, thr, group, sw_lo, w2, c3, k0, k1, part)
                 : bitserial_less<NHC, DRAW3>(need, m, thr, group, sw_lo, w2, c3, k0, k1);
        return (up | lt) & valid;
    }
}

// LAYOUT & 3 = 0: no item of the launch has the flat layout (ising_flat(L));
// 1: every item has it; 2: either, per item (the job server's mixed classes);
// 3: every item has L = 64, one row per word, run by the flat word loop with
// the row neighbours (fewer registers than the general row loop).  LAYOUT >= 4:
// the same for launches too small to fill the SIMDs, each word shared by a
// quad of lanes (bitserial_less_quad), in workgroups of up to four times the
// threads.  The word loops are separate instantiations: a loop for both
// layouts costs the row loop registers and ~3 % of its rate.
template <int DIM, int ALG, int LAYOUT>
__global__ void __launch_bounds__(1024) k_ising_sweeps(SweepArgs a)
{
    constexpr int BASE = LAYOUT & 3;
    constexpr bool QUAD = LAYOUT >= 4;
    extern __shared__ uint32_t lds[];
    uint32_t* part = lds;                                // [2 parity][16 waves][2]
    int32_t* ring = reinterpret_cast<int32_t*>(lds + 64); // E[64], M[64] of the last sweeps
    uint32_t* P = lds + 192;                             // planes [2 colours][rows][nw]
    const int tid = threadIdx.x;
    const int B = blockDim.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = (B + 63) >> 6;
    constexpr uint32_t purpose = ALG == 1 ? P_HB : P_METRO;
    constexpr int QPW = QUAD ? 4 : 1; // lanes per word
    const uint32_t qpart = (uint32_t)(lane & (QPW - 1));

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
        const bool flat = BASE == 2 ? ising_flat(L) : (BASE == 1 || BASE == 3);
        const int words = (int)ising_plane_words(DIM, L); // per colour
        const int64_t N = (int64_t)rows * L;
        // (once per item: its divisions and row-mask loop are scalar work that
        // a one-wave chain would otherwise repeat every half-sweep)
        const FlatGeo g = flat_geo(DIM, flat ? L : 8);
        int8_t* __restrict__ sp = a.spins + cr.spin_off;
        const IsingTable tb = a.itab[cr.table];

        // ---- int8 lattice (or the planes a previous launch left) -> bit planes (up = 1) ----
        if (a.planes_in) {
            const uint32_t* src = a.planes + a.plane_off[ch];
            for (int w = tid; w < 2 * words; w += B)
                P[w] = src[w];
        } else if (flat && BASE != 3) {
            for (int w = tid; w < 2 * words; w += B)
                ising_pack_flat<DIM>(sp, P, w, L, half, words);
        } else {
            for (int w = tid; w < words; w += B)
                ising_pack_word<DIM>(sp, P, w, L, nw, half, words);
        }
        __syncthreads();

        unsigned long long acc = 0;
        // running sums continue from the stored ones, sweep by sweep, so the
        // FP64 moments do not depend on how a run is split into launches
        double mE = 0, mE2 = 0, mAM = 0, mM2 = 0, mM4 = 0;
        if (a.record && tid == 0) {
            const double* mo = a.moments + oslot * 6;
            mE = mo[1];
            mE2 = mo[2];
            mAM = mo[3];
            mM2 = mo[4];
            mM4 = mo[5];
        }

        const int64_t nsw = a.item_nsweeps ? a.item_nsweeps[item] : a.nsweeps;
        const bool series = a.record && a.ser_E != nullptr;
        for (int64_t s = 0; s < nsw; ++s) {
            const uint64_t sweep = cr.sweep + (uint64_t)s;
            const uint32_t sw_lo = (uint32_t)sweep;
            const uint32_t w2 = ctr_word2(sweep, 0, purpose);
            // observables: every recorded sweep, and the last one when asked for
            const bool obs = a.record || (a.last_E != nullptr && s == nsw - 1);
            uint32_t nu = 0, na = 0;
            for (int c = 0; c < 2; ++c) {
                uint32_t* __restrict__ Pc = P + c * words;
                const uint32_t* __restrict__ Po = P + (c ^ 1) * words;
                // (wave-uniform trip count: the chunk loop inside sees uniform
                // control flow, so its per-chunk counter words stay scalar)
                if (flat) {
                    // QUAD: a quad of lanes per word (QPW = 4), part 0 writes it
                    for (int wb = (tid - lane) / QPW; wb < words; wb += B / QPW) {
                        const int w = wb + lane / QPW;
                        if (w >= words)
                            continue;
                        const uint32_t s0 = Pc[w];
                        NbrWords n;
                        load_nbr_flat<DIM, BASE == 3>(Po, g, L, c, w, n);
                        const uint32_t snew = ising_word<DIM, ALG, BASE != 1 && !QUAD, QUAD>(
                            s0, n, 0xffffffffu, tb, 2u * (uint32_t)w + (uint32_t)c, sw_lo, w2, c3,
                            k0, cr.key, qpart);
                        if (QPW > 1 && qpart != 0)
                            continue;
                        acc += (unsigned long long)__popc(snew ^ s0);
                        Pc[w] = snew;
                        if (c == 1 && obs) {
                            na += __popc(snew ^ n.U) + __popc(snew ^ n.Dn) + __popc(snew ^ n.A)
                                + __popc(snew ^ n.Bw);
                            if (DIM == 3)
                                na += __popc(snew ^ n.Zu) + __popc(snew ^ n.Zd);
                            nu += __popc(snew) + __popc(n.A);
                        }
                    }
                } else
                for (int wb = (tid - lane) / QPW; wb < words; wb += B / QPW) {
                    const int w = wb + lane / QPW;
                    if (w >= words)
                        continue;
                    const int r = nw == 1 ? w : w / nw;
                    const int q = w - (int)__umul24((uint32_t)r, (uint32_t)nw);
                    const int y = DIM == 3 ? r % L : r;
                    const int z = DIM == 3 ? r / L : 0;
                    const int p = (c + y + z) & 1;
                    const int W = word_width(q, half);
                    const uint32_t valid = width_mask(W);
                    const uint32_t s0 = Pc[w];
                    NbrWords n;
                    load_nbr_words<DIM, true>(Po, L, nw, half, r, q, y, z, p, W, valid, n);
                    const uint32_t group =
                        __umul24((uint32_t)(r * 2 + c), (uint32_t)nw) + (uint32_t)q;
                    const uint32_t snew = ising_word<DIM, ALG, !QUAD, QUAD>(
                        s0, n, valid, tb, group, sw_lo, w2, c3, k0, cr.key, qpart);
                    if (QUAD && qpart != 0)
                        continue;
                    const uint32_t flip = snew ^ s0;
                    acc += (unsigned long long)__popc(flip);
                    Pc[w] = snew;
                    if (c == 1 && obs) {
                        // every bond joins colour 0 and colour 1: count each from its colour-1 end
                        na += __popc(snew ^ n.U) + __popc(snew ^ n.Dn) + __popc(snew ^ n.A)
                            + __popc(snew ^ n.Bw);
                        if (DIM == 3)
                            na += __popc(snew ^ n.Zu) + __popc(snew ^ n.Zd);
                        nu += __popc(snew) + __popc(n.A);
                    }
                }
                if (c == 1 && obs) {
                    nu = wave_sum32(nu);
                    na = wave_sum32(na);
                    // (one wave: its sums stay in registers for the tail below)
                    if (lane == 0 && nwaves > 1) {
                        part[(s & 1) * 32 + wave * 2] = nu;
                        part[(s & 1) * 32 + wave * 2 + 1] = na;
                    }
                }
                __syncthreads();
            }
            if (obs && tid == 0) {
                uint32_t tu = nu, ta = na;
                if (nwaves > 1) {
                    tu = 0;
                    ta = 0;
                    for (int k = 0; k < nwaves; ++k) {
                        tu += part[(s & 1) * 32 + k * 2];
                        ta += part[(s & 1) * 32 + k * 2 + 1];
                    }
                }
                const int32_t E = (int32_t)(2 * (int64_t)ta - (int64_t)DIM * N);
                const int32_t M = (int32_t)(2 * (int64_t)tu - N);
                ring[s & 63] = E;
                ring[64 + (s & 63)] = M;
                if (a.record) {
                    const double e = (double)E, m = (double)M;
                    mE += e;
                    mE2 += e * e;
                    mAM += fabs(m);
                    mM2 += m * m;
                    mM4 += (m * m) * (m * m);
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
                }
            }
        }
        if (a.last_E && nsw > 0 && tid == 0) {
            a.last_E[ch] = ring[(nsw - 1) & 63];
            a.last_M[ch] = ring[64 + ((nsw - 1) & 63)];
            a.last_Q[ch] = (int32_t)N;
        }

        // ---- bit planes -> int8 lattice (or left as planes for the next launch) ----
        if (a.planes_out) {
            uint32_t* dst = a.planes + a.plane_off[ch];
            for (int w = tid; w < 2 * words; w += B)
                dst[w] = P[w];
        } else if (flat && BASE != 3) {
            for (int w = tid; w < 2 * words; w += B)
                ising_unpack_flat<DIM>(sp, P, w, L, half, words);
        } else {
            for (int w = tid; w < words; w += B)
                ising_unpack_word<DIM>(sp, P, w, L, nw, half, words);
        }
        acc = wave_sum(acc);
        if (lane == 0 && acc)
            atomicAdd(&a.accepted[ch], acc);
        if (tid == 0) {
            if (a.record) {
                double* mo = a.moments + oslot * 6;
                mo[0] += (double)nsw;
                mo[1] = mE;
                mo[2] = mE2;
                mo[3] = mAM;
                mo[4] = mM2;
                mo[5] = mM4;
            }
            a.chains[ch].sweep = cr.sweep + (uint64_t)nsw;
        }
        __syncthreads();
        if (a.item_ticks && tid == 0)
            a.item_ticks[item] = wall_clock64() - t_item;
    }
}

// ---------------------------------------------------------------------------
// Plain path: one thread per site of one colour of one chain.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t bitserial_less_site(uint32_t thr, uint32_t group, int t,
                                                        uint32_t sw_lo, uint32_t w2, uint32_t c3,
                                                        uint32_t k0, uint32_t k1)
{
    // the 32 bits of chunks c, c + 1 against thr (MSB first); 2 = still tied
    auto two = [&](const uint4& A, const uint4& Bk, uint32_t chunk) -> uint32_t {
        const uint32_t Rb[8] = {A.x, A.y, A.z, A.w, Bk.x, Bk.y, Bk.z, Bk.w};
        for (int b = 0; b < 8; ++b) {
            const uint32_t bit = (Rb[b] >> t) & 1u;
            const uint32_t tbit = (thr >> (31u - (4u * chunk + (uint32_t)b))) & 1u;
            if (bit != tbit)
                return tbit; // w < thr iff the first differing bit is 0 in w
        }
        return 2u;
    };
    // two chunks at a time: independent Philox chains the scheduler
    // interleaves (a wavefront of 64 sites nearly always needs two)
    for (uint32_t chunk = 0; chunk < 8; chunk += 2) {
        const uint4 R0 = philox4x32_10(group, sw_lo, w2 | (chunk << 16), c3, k0, k1);
        const uint4 R1 = philox4x32_10(group, sw_lo, w2 | ((chunk + 1) << 16), c3, k0, k1);
        const uint32_t r = two(R0, R1, chunk);
        if (r != 2u)
            return r;
    }
    return 0u; // w == thr
}

template <int DIM, int ALG>
__global__ void __launch_bounds__(1024) k_ising_plain(PlainArgs a, int n_col)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const ChainRec cr = a.chains[a.chain];
    const int L = a.L;
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int c = a.colour;
    int e_part = 0, m_part = 0;
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
        const int8_t si = sp[rowb + x];
        const IsingTable& tb = a.itab[cr.table];
        // the sweep stream's (group, bit) of the site: flat or row words
        const int f = r * half + j;
        const uint32_t group = ising_flat(L) ? (uint32_t)(2 * (f >> 5) + c)
                                             : (uint32_t)((r * 2 + c) * nw + (j >> 5));
        const int t = ising_flat(L) ? (f & 31) : (j & 31);
        const uint32_t sw_lo = (uint32_t)a.sweep;
        const uint32_t k0 = (uint32_t)cr.seed, c3 = (uint32_t)(cr.seed >> 32);
        int8_t snew = si;
        if (ALG == 0) {
            const int n_anti = (2 * DIM - si * nb) >> 1;
            bool ok;
            if (n_anti >= DIM)
                ok = true;
            else if ((tb.metro_always >> n_anti) & 1u)
                ok = true;
            else if ((tb.metro_never >> n_anti) & 1u)
                ok = false;
            else
                ok = bitserial_less_site(tb.metro_thr[n_anti], group, t, sw_lo,
                                         ctr_word2(a.sweep, 0, P_METRO), c3, k0, cr.key) != 0u;
            if (ok)
                snew = (int8_t)-si;
        } else if (ALG == 2) {
            const int n_anti = (2 * DIM - si * nb) >> 1;
            bool ok;
            if ((tb.gl_always >> n_anti) & 1u)
                ok = true;
            else if ((tb.gl_never >> n_anti) & 1u)
                ok = false;
            else
                ok = bitserial_less_site(tb.gl_thr[n_anti], group, t, sw_lo,
                                         ctr_word2(a.sweep, 0, P_METRO), c3, k0, cr.key) != 0u;
            if (ok)
                snew = (int8_t)-si;
        } else {
            const int n_up = (nb + 2 * DIM) >> 1;
            bool up;
            if ((tb.hb_always >> n_up) & 1u)
                up = true;
            else if ((tb.hb_never >> n_up) & 1u)
                up = false;
            else
                up = bitserial_less_site(tb.hb_thr[n_up], group, t, sw_lo,
                                         ctr_word2(a.sweep, 0, P_HB), c3, k0, cr.key) != 0u;
            snew = up ? 1 : -1;
        }
        if (snew != si) {
            sp[rowb + x] = snew;
            flips = 1;
        }
        if (c == 1 && a.rec_index >= 0) {
            e_part = -(int)snew * nb;
            m_part = snew + sp[rowb + (p ? xl : xr)];
        }
    }
    flips = wave_sum32(flips);
    if ((threadIdx.x & 63) == 0 && flips)
        atomicAdd(&a.accepted[a.chain], (unsigned long long)flips);
    if (c == 1 && a.rec_index >= 0) {
        e_part = wave_sum32(e_part);
        m_part = wave_sum32(m_part);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.ser_E[cr.series_off + a.rec_index], e_part);
            atomicAdd(&a.ser_M[cr.series_off + a.rec_index], m_part);
        }
    }
}

// ---------------------------------------------------------------------------
// Tile kernel: one colour of one sweep of every work item, one thread per plane
// word, the planes in device memory.  Grid (word blocks, items).  The word
// update and the observables are those of k_ising_sweeps, so a lattice gives
// the same chain on either path.
// ---------------------------------------------------------------------------

template <int DIM, int ALG>
__global__ void __launch_bounds__(256) k_ising_tiles(TileArgs a)
{
    const int item = blockIdx.y;
    if (a.s >= a.item_nsw[item])
        return;
    constexpr uint32_t purpose = ALG == 1 ? P_HB : P_METRO;
    const int ch = a.items[item];
    const ChainRec cr = a.chains[ch];
    const int L = cr.L;
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int rows = DIM == 3 ? L * L : L;
    const int words = rows * nw;
    const int c = a.colour;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t flips = 0, nu = 0, na = 0;
    if (w < words) {
        uint32_t* P = reinterpret_cast<uint32_t*>(a.base + a.item_plane[item]);
        const IsingTable tb = a.itab[cr.table];
        const uint64_t sweep = cr.sweep + (uint64_t)a.s;
        const int r = nw == 1 ? w : w / nw;
        const int q = w - r * nw;
        const int y = DIM == 3 ? r % L : r;
        const int z = DIM == 3 ? r / L : 0;
        const int p = (c + y + z) & 1;
        const int W = word_width(q, half);
        const uint32_t valid = width_mask(W);
        const uint32_t s0 = P[c * words + w];
        NbrWords n;
        load_nbr_words<DIM>(P + (c ^ 1) * words, L, nw, half, r, q, y, z, p, W, valid, n);
        const uint32_t group = (uint32_t)((r * 2 + c) * nw + q);
        const uint32_t snew =
            ising_word<DIM, ALG>(s0, n, valid, tb, group, (uint32_t)sweep,
                                 ctr_word2(sweep, 0, purpose), (uint32_t)(cr.seed >> 32),
                                 (uint32_t)cr.seed, cr.key);
        P[c * words + w] = snew;
        flips = (uint32_t)__popc(snew ^ s0);
        if (c == 1) {
            na = __popc(snew ^ n.U) + __popc(snew ^ n.Dn) + __popc(snew ^ n.A)
               + __popc(snew ^ n.Bw);
            if (DIM == 3)
                na += __popc(snew ^ n.Zu) + __popc(snew ^ n.Zd);
            nu = __popc(snew) + __popc(n.A);
        }
    }
    tile_block_add(a, item, (int32_t)nu, (int32_t)na, 0, (int32_t)flips);
}

template <int DIM>
__global__ void __launch_bounds__(256) k_ising_pack(TileArgs a, int unpack)
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
        ising_unpack_word<DIM>(sp, P, w, L, nw, half, words);
    else
        ising_pack_word<DIM>(sp, P, w, L, nw, half, words);
}

// Multi-workgroup path with temporal blocking.  A work item is one band of
// a.band slabs (rows in 2D, z-planes in 3D) of one lattice; the workgroup
// loads the band and H = 2T slabs on either side (periodic) into LDS and runs
// T sweeps there.  Half-sweep k = 2t + colour updates the local slabs
// [k + 1, R - 2 - k]: a slab outside that range has a neighbour slab that is
// stale or was never loaded, so after the T sweeps exactly the band
// [H, H + B) is right for both colours.  Halo sites are updated by several
// workgroups with the same counter-based random words, so they agree with
// their owners; only the band is written back.  Per-sweep observables and
// flips are summed over the band into the item's per-sweep replicas, as the
// per-colour kernel does.  LDS: [2 colours][R slabs][rows per slab][nw].
template <int DIM, int ALG>
__global__ void __launch_bounds__(kBandThreads) k_ising_band(TileArgs a)
{
    extern __shared__ uint32_t bq[];
    const int item = blockIdx.y;
    const int64_t nsw = a.item_nsw[item];
    if (a.s >= nsw)
        return;
    constexpr uint32_t purpose = ALG == 1 ? P_HB : P_METRO;
    const int tid = threadIdx.x;
    const int nthr = (int)blockDim.x;
    const int ch = a.items[item];
    const ChainRec cr = a.chains[ch];
    const int L = cr.L;
    const int B = a.band;
    const int s0 = blockIdx.x * B;
    if (s0 >= L)
        return;
    const int half = L >> 1;
    const int nw = (half + 31) >> 5;
    const int spr = DIM == 3 ? L : 1; // plane rows per slab
    const int H = 2 * a.T;
    const int Bb = B < L - s0 ? B : L - s0;
    const int R = Bb + 2 * H;
    const int rw = spr * nw; // words per slab and colour
    const int nloc = R * rw;
    const int words = (DIM == 3 ? L * L : L) * nw;
    uint32_t* P = reinterpret_cast<uint32_t*>(a.base + a.item_plane[item]);
    // global slab of local slab 0, in [0, L): the halo may be longer than the lattice
    const int g0 = ((s0 - H) % L + L) % L;
    for (int i = tid; i < 2 * nloc; i += nthr) {
        const int c = i >= nloc ? 1 : 0;
        const int li = i - c * nloc;
        const int ls = li / rw;
        bq[i] = P[c * words + ((g0 + ls) % L) * rw + (li - ls * rw)];
    }
    __syncthreads();
    const IsingTable tb = a.itab[cr.table];
    const uint32_t k0 = (uint32_t)cr.seed, c3 = (uint32_t)(cr.seed >> 32);
    const int64_t left = nsw - a.s;
    const int Tn = left < (int64_t)a.T ? (int)left : a.T;
    for (int t = 0; t < Tn; ++t) {
        const uint64_t sweep = cr.sweep + (uint64_t)(a.s + t);
        const uint32_t w2 = ctr_word2(sweep, 0, purpose);
        uint32_t flips = 0, nu = 0, na = 0;
        for (int c = 0; c < 2; ++c) {
            const int k = 2 * t + c;
            const int lo = k + 1, n_upd = (R - 2 - k - lo + 1) * rw;
            uint32_t* Qc = bq + c * nloc;
            const uint32_t* Qo = bq + (c ^ 1) * nloc;
            for (int i = tid; i < n_upd; i += nthr) {
                const int ls = lo + i / rw;
                const int rem = i - (ls - lo) * rw;
                const int yl = DIM == 3 ? rem / nw : 0;
                const int q = rem - yl * nw;
                const int gs = (g0 + ls) % L;
                const int y = DIM == 3 ? yl : gs;
                const int z = DIM == 3 ? gs : 0;
                const int r = DIM == 3 ? z * L + y : y; // global plane row
                const int lr = ls * spr + yl;           // local plane row
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
                const uint32_t sw0 = Qc[lr * nw + q];
                NbrWords n;
                nbr_words_rows<DIM>(Qo, nw, half, lr, iU, iD, iZu, iZd, q, p, W, valid, n);
                const uint32_t group = (uint32_t)((r * 2 + c) * nw + q);
                const uint32_t snew = ising_word<DIM, ALG>(sw0, n, valid, tb, group, (uint32_t)sweep,
                                                           w2, c3, k0, cr.key);
                Qc[lr * nw + q] = snew;
                if (ls >= H && ls < H + Bb) {
                    flips += (uint32_t)__popc(snew ^ sw0);
                    if (c == 1) {
                        na += __popc(snew ^ n.U) + __popc(snew ^ n.Dn) + __popc(snew ^ n.A) +
                              __popc(snew ^ n.Bw);
                        if (DIM == 3)
                            na += __popc(snew ^ n.Zu) + __popc(snew ^ n.Zd);
                        nu += __popc(snew) + __popc(n.A);
                    }
                }
            }
            __syncthreads();
        }
        tile_block_add_s(a, item, a.s + t, (int32_t)nu, (int32_t)na, 0, (int32_t)flips);
    }
    for (int i = tid; i < 2 * Bb * rw; i += nthr) {
        const int c = i >= Bb * rw ? 1 : 0;
        const int li = i - c * Bb * rw;
        const int b = li / rw;
        P[c * words + ((s0 + b) % L) * rw + (li - b * rw)] = bq[c * nloc + (H + b) * rw + (li - b * rw)];
    }
}

// algorithm code -> kernel template: 0 Metropolis, 1 heat bath, 2 Glauber
static int ising_alg(int alg)
{
    return alg == MCX_HEATBATH ? 1 : alg == MCX_GLAUBER ? 2 : 0;
}

hipError_t launch_ising_tiles(int dim, int alg, const TileArgs& a, int max_words, hipStream_t st)
{
    const dim3 grid((max_words + 255) / 256, a.n_items);
    const int k = ising_alg(alg) + (dim == 3 ? 3 : 0);
    switch (k) {
    case 0: hipLaunchKernelGGL((k_ising_tiles<2, 0>), grid, dim3(256), 0, st, a); break;
    case 1: hipLaunchKernelGGL((k_ising_tiles<2, 1>), grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_ising_tiles<2, 2>), grid, dim3(256), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_ising_tiles<3, 0>), grid, dim3(256), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_ising_tiles<3, 1>), grid, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((k_ising_tiles<3, 2>), grid, dim3(256), 0, st, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_ising_band(int dim, int alg, const TileArgs& a, int max_bands, int threads,
                             size_t lds, hipStream_t st)
{
    const dim3 grid(max_bands, a.n_items), bl(threads);
    const int k = ising_alg(alg) + (dim == 3 ? 3 : 0);
    switch (k) {
    case 0: hipLaunchKernelGGL((k_ising_band<2, 0>), grid, bl, lds, st, a); break;
    case 1: hipLaunchKernelGGL((k_ising_band<2, 1>), grid, bl, lds, st, a); break;
    case 2: hipLaunchKernelGGL((k_ising_band<2, 2>), grid, bl, lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_ising_band<3, 0>), grid, bl, lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_ising_band<3, 1>), grid, bl, lds, st, a); break;
    default: hipLaunchKernelGGL((k_ising_band<3, 2>), grid, bl, lds, st, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_ising_pack(int dim, int unpack, const TileArgs& a, int max_words, hipStream_t st)
{
    const dim3 grid((max_words + 255) / 256, a.n_items);
    if (dim == 2)
        hipLaunchKernelGGL((k_ising_pack<2>), grid, dim3(256), 0, st, a, unpack);
    else
        hipLaunchKernelGGL((k_ising_pack<3>), grid, dim3(256), 0, st, a, unpack);
    return hipGetLastError();
}

hipError_t launch_ising_sweeps(int dim, int alg, int layout, const SweepArgs& a, int block,
                               int grid, size_t lds_bytes, hipStream_t st)
{
    const dim3 g(grid), bl(block);
    const int k = ising_alg(alg) + (dim == 3 ? 3 : 0) + 6 * layout;
#define MCX_IS(K, D, A, Y)                                                                  \
    case K: hipLaunchKernelGGL((k_ising_sweeps<D, A, Y>), g, bl, lds_bytes, st, a); break;
    switch (k) {
    MCX_IS(0, 2, 0, 0) MCX_IS(1, 2, 1, 0) MCX_IS(2, 2, 2, 0)
    MCX_IS(3, 3, 0, 0) MCX_IS(4, 3, 1, 0) MCX_IS(5, 3, 2, 0)
    MCX_IS(6, 2, 0, 1) MCX_IS(7, 2, 1, 1) MCX_IS(8, 2, 2, 1)
    MCX_IS(9, 3, 0, 1) MCX_IS(10, 3, 1, 1) MCX_IS(11, 3, 2, 1)
    MCX_IS(12, 2, 0, 2) MCX_IS(13, 2, 1, 2) MCX_IS(14, 2, 2, 2)
    MCX_IS(15, 3, 0, 2) MCX_IS(16, 3, 1, 2) MCX_IS(17, 3, 2, 2)
    MCX_IS(18, 2, 0, 3) MCX_IS(19, 2, 1, 3) MCX_IS(20, 2, 2, 3)
    MCX_IS(21, 3, 0, 3) MCX_IS(22, 3, 1, 3) MCX_IS(23, 3, 2, 3)
    MCX_IS(24, 2, 0, 4) MCX_IS(25, 2, 1, 4) MCX_IS(26, 2, 2, 4)
    MCX_IS(27, 3, 0, 4) MCX_IS(28, 3, 1, 4) MCX_IS(29, 3, 2, 4)
    MCX_IS(30, 2, 0, 5) MCX_IS(31, 2, 1, 5) MCX_IS(32, 2, 2, 5)
    MCX_IS(33, 3, 0, 5) MCX_IS(34, 3, 1, 5) MCX_IS(35, 3, 2, 5)
    MCX_IS(36, 2, 0, 6) MCX_IS(37, 2, 1, 6) MCX_IS(38, 2, 2, 6)
    MCX_IS(39, 3, 0, 6) MCX_IS(40, 3, 1, 6) MCX_IS(41, 3, 2, 6)
    MCX_IS(42, 2, 0, 7) MCX_IS(43, 2, 1, 7) MCX_IS(44, 2, 2, 7)
    MCX_IS(45, 3, 0, 7) MCX_IS(46, 3, 1, 7)
    default: hipLaunchKernelGGL((k_ising_sweeps<3, 2, 7>), g, bl, lds_bytes, st, a); break;
    }
#undef MCX_IS
    return hipGetLastError();
}

hipError_t launch_ising_plain(int dim, int alg, const PlainArgs& a, int n_col, hipStream_t st)
{
    // one workgroup for a colour of up to 1024 sites (a single dispatch), else 256-thread ones
    const int blk = n_col <= 1024 ? (n_col + 63) / 64 * 64 : 256;
    const dim3 g((n_col + blk - 1) / blk), bl(blk);
    const int k = ising_alg(alg) + (dim == 3 ? 3 : 0);
    switch (k) {
    case 0: hipLaunchKernelGGL((k_ising_plain<2, 0>), g, bl, 0, st, a, n_col); break;
    case 1: hipLaunchKernelGGL((k_ising_plain<2, 1>), g, bl, 0, st, a, n_col); break;
    case 2: hipLaunchKernelGGL((k_ising_plain<2, 2>), g, bl, 0, st, a, n_col); break;
    case 3: hipLaunchKernelGGL((k_ising_plain<3, 0>), g, bl, 0, st, a, n_col); break;
    case 4: hipLaunchKernelGGL((k_ising_plain<3, 1>), g, bl, 0, st, a, n_col); break;
    default: hipLaunchKernelGGL((k_ising_plain<3, 2>), g, bl, 0, st, a, n_col); break;
    }
    return hipGetLastError();
}

} // namespace mcx
