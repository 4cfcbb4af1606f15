// Wavefront (64-lane) helpers shared by the walk kernels.
#pragma once
#include "common.hpp"

namespace lhip {

__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Inclusive prefix sum over the wave: Hillis-Steele within each row of 16 lanes by DPP row shifts
// (a lane whose source lies outside its row adds 0), then the totals of the rows below.
__device__ __forceinline__ int prefix_sum_incl(int v, int lane) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31);
    const int r2 = __builtin_amdgcn_readlane(v, 47);
    const int row = lane >> 4;
    return v + (row == 0 ? 0 : row == 1 ? r0 : row == 2 ? r0 + r1 : r0 + r1 + r2);
}

// Exclusive scan of one value per thread over a block of W waves (wave scans, then the wave totals
// through wsum[W], one barrier); *total gets the block sum. wsum must not be rewritten before every
// thread has passed the next barrier.
template <int W>
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* wsum, int lane, int wave, uint32_t* total) {
    const uint32_t incl = uint32_t(prefix_sum_incl(int(v), lane));
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t off = 0, all = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint32_t t = wsum[w];
        off += w < wave ? t : 0u;
        all += t;
    }
    if (total) *total = all;
    return off + incl - v;
}

// Sum of a float over all 64 lanes (every lane active), returned wave-uniform: butterflies within
// each row of 16 (quad swaps, half-row and row mirrors), then the four row sums.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<0xB1>(v);   // quad_perm(1, 0, 3, 2)
    v += dpp_f<0x4E>(v);   // quad_perm(2, 3, 0, 1)
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    const float a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (a + b) + (c + d);
}

// Sums over the wave of four values at once: lanes 0-31 of v0 + v1 after a 32-lane swap hold v0's
// half sums and lanes 32-63 v1's (likewise v2, v3), a 16-lane swap of the two leaves one target per
// row, and butterflies within the row finish it. Row r (lanes 16r..16r+15) of the result holds
// the sum of v_k with k = sum4_target(r).
__device__ __forceinline__ float wave_sum4(float v0, float v1, float v2, float v3) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v0), __float_as_uint(v1), false, false);
    const float a0 = __uint_as_float(p[0]) + __uint_as_float(p[1]);
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(v2), __float_as_uint(v3), false, false);
    const float a1 = __uint_as_float(q[0]) + __uint_as_float(q[1]);
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a0), __float_as_uint(a1), false, false);
    float b = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    b += dpp_f<0xB1>(b);   // quad_perm(1, 0, 3, 2)
    b += dpp_f<0x4E>(b);   // quad_perm(2, 3, 0, 1)
    b += dpp_f<0x141>(b);  // row_half_mirror
    b += dpp_f<0x140>(b);  // row_mirror
    return b;
}
// The value index whose sum wave_sum4 leaves in row r: {0, 2, 1, 3}[r].
__device__ __forceinline__ int sum4_target(int row) { return ((row & 1) << 1) | (row >> 1); }

__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace lhip
