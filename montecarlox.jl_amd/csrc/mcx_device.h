// mcx_device.h -- device helpers shared by the sweep kernels: Philox4x32-10,
// the checkerboard word geometry and wavefront reductions.
#pragma once

#include "mcx_internal.h"

namespace mcx {

// Three-input XOR in one VALU instruction (gfx950 v_bitop3_b32, truth table
// 0x96 = a ^ b ^ c).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  The 32x32->64 products map to v_mad_u64_u32 and each
// round's two "hi ^ c ^ key" to one v_bitop3_b32 each, so a round is four VALU
// instructions; round keys depend only on (k0, k1) and fold into scalar
// registers when the key is uniform across the wavefront.  In the first two
// rounds the counter words of the sweep, the seed and the chunk are still
// wave-uniform in every stream here, so those XORs are written plainly: the
// compiler keeps their uniform parts on the scalar unit (v_bitop3 has no
// scalar form).
__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t h1 = (uint32_t)(p1 >> 32), h0 = (uint32_t)(p0 >> 32);
        const uint32_t n0 = r < 2 ? (h1 ^ (c1 ^ k0)) : xor3(h1, c1, k0);
        const uint32_t n2 = r < 2 ? (h0 ^ (c3 ^ k1)) : xor3(h0, c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// Counter word 2 of a site stream: sweep bits 32..47, chunk, purpose.
__device__ __forceinline__ uint32_t ctr_word2(uint64_t sweep, uint32_t chunk, uint32_t purpose)
{
    return (uint32_t)((sweep >> 32) & 0xffffu) | (chunk << 16) | (purpose << 24);
}

// Site word of site (group, bit): word (bit & 3) of the block at counter
// (8*group + bit/4, sweep, purpose).
__device__ __forceinline__ uint32_t site_word(uint64_t seed, uint32_t key, uint32_t group, int bit,
                                              uint64_t sweep, uint32_t purpose)
{
    const uint4 r = philox4x32_10(8u * group + (uint32_t)(bit >> 2), (uint32_t)sweep,
                                  ctr_word2(sweep, 0, purpose), (uint32_t)(seed >> 32),
                                  (uint32_t)seed, key);
    const int k = bit & 3;
    return k == 0 ? r.x : (k == 1 ? r.y : (k == 2 ? r.z : r.w));
}

__device__ __forceinline__ int word_width(int q, int half)
{
    const int w = half - 32 * q;
    return w > 32 ? 32 : w;
}

__device__ __forceinline__ uint32_t width_mask(int W)
{
    return W >= 32 ? 0xffffffffu : ((1u << W) - 1u);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of a 32-bit integer over the 64 lanes of a wavefront, through DPP
// lane moves instead of LDS permutes: pairs and quads (quad_perm), halves of
// a row and rows of 16 (half-row and row mirrors: every lane then holds its
// row's sum), the last lane of rows 0 and 2 added into rows 1 and 3
// (row_bcast:15, rows 1 and 3 only), and lanes 31 and 63 read out.  Every
// lane of the wavefront must be active; the result is wave-uniform.
__device__ __forceinline__ int32_t wave_sum32(int32_t v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false); // row_half_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false); // row_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); // row_bcast:15
    return __builtin_amdgcn_readlane(v, 31) + __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint32_t wave_sum32(uint32_t v)
{
    return (uint32_t)wave_sum32((int32_t)v);
}

// Bit-sliced population count of 4 or 6 one-bit-per-site words:
// count = b0 + 2*b1 + 4*b2 per bit position.
__device__ __forceinline__ void count4(uint32_t x1, uint32_t x2, uint32_t x3, uint32_t x4,
                                       uint32_t& b0, uint32_t& b1, uint32_t& b2)
{
    const uint32_t s = x1 ^ x2 ^ x3;
    const uint32_t c = (x1 & x2) | (x3 & (x1 | x2));
    b0 = s ^ x4;
    const uint32_t c2 = s & x4;
    b1 = c ^ c2;
    b2 = c & c2;
}

__device__ __forceinline__ void count6(uint32_t x1, uint32_t x2, uint32_t x3, uint32_t x4,
                                       uint32_t x5, uint32_t x6, uint32_t& b0, uint32_t& b1,
                                       uint32_t& b2)
{
    const uint32_t s1 = x1 ^ x2 ^ x3;
    const uint32_t c1 = (x1 & x2) | (x3 & (x1 | x2));
    const uint32_t s2 = x4 ^ x5 ^ x6;
    const uint32_t c2 = (x4 & x5) | (x6 & (x4 | x5));
    b0 = s1 ^ s2;
    const uint32_t k = s1 & s2;
    b1 = c1 ^ c2 ^ k;
    b2 = (c1 & c2) | (k & (c1 | c2));
}

// Mask of positions whose 3-bit count equals K.
template <int K>
__device__ __forceinline__ uint32_t count_eq(uint32_t b0, uint32_t b1, uint32_t b2)
{
    const uint32_t m0 = (K & 1) ? b0 : ~b0;
    const uint32_t m1 = (K & 2) ? b1 : ~b1;
    const uint32_t m2 = (K & 4) ? b2 : ~b2;
    return m0 & m1 & m2;
}

// Block-level sum of four per-lane values, then one atomic per field into this
// block's replica of the item's per-sweep sums (same-address atomics from every
// wave of a big lattice would serialise).  blockDim.x = 256.
// tile_block_add for sweep s of the quantum, any block of up to 16 waves
// (the band kernels run several sweeps per launch)
__device__ __forceinline__ void tile_block_add_s(const TileArgs& a, int item, int64_t s, int32_t f0,
                                                 int32_t f1, int32_t f2, int32_t f3)
{
    __shared__ int32_t red_s[16][4];
    const int wave = threadIdx.x >> 6;
    const int nwaves = (int)(blockDim.x >> 6);
    f0 = wave_sum32(f0);
    f1 = wave_sum32(f1);
    f2 = wave_sum32(f2);
    f3 = wave_sum32(f3);
    if ((threadIdx.x & 63) == 0) {
        red_s[wave][0] = f0;
        red_s[wave][1] = f1;
        red_s[wave][2] = f2;
        red_s[wave][3] = f3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        int32_t v = 0;
        for (int w = 0; w < nwaves; ++w)
            v += red_s[w][k];
        if (v != 0) {
            int32_t* o = reinterpret_cast<int32_t*>(a.base + a.item_obs[item]);
            atomicAdd(&o[(s * 4 + k) * kTileObsReplicas + (int)(blockIdx.x % kTileObsReplicas)], v);
        }
    }
    // no barrier here: the callers' next writes of red_s come after the two
    // half-sweep barriers of the next sweep (or never), so the atomics' latency
    // overlaps the next half-sweep instead of stalling the workgroup
}

__device__ __forceinline__ void tile_block_add(const TileArgs& a, int item, int32_t f0, int32_t f1,
                                               int32_t f2, int32_t f3)
{
    __shared__ int32_t red[4][4];
    const int wave = threadIdx.x >> 6;
    f0 = wave_sum32(f0);
    f1 = wave_sum32(f1);
    f2 = wave_sum32(f2);
    f3 = wave_sum32(f3);
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = f0;
        red[wave][1] = f1;
        red[wave][2] = f2;
        red[wave][3] = f3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        const int32_t v = red[0][k] + red[1][k] + red[2][k] + red[3][k];
        if (v != 0) {
            int32_t* o = reinterpret_cast<int32_t*>(a.base + a.item_obs[item]);
            atomicAdd(&o[(a.s * 4 + k) * kTileObsReplicas + (int)(blockIdx.x % kTileObsReplicas)], v);
        }
    }
}

} // namespace mcx
