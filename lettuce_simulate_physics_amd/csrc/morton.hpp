// Morton key recipe shared by the first sort and the key restarts.
#pragma once
#include "common.hpp"

namespace lhip {

// Bit dilation for 21-bit coordinates: the standard mask ladder that moves bit b to bit 3b.
__device__ __forceinline__ uint64_t dilate21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Quantisation in double with a clamp, as src/cpu_backend.cpp:419-425: (v - o) * 2^21/side, clamped
// to [0, 2^21 - 1], truncated. Sub and mul stay separate IEEE operations.
template <class R>
__device__ __forceinline__ uint64_t quantize(R v, double o, double s) {
    double t = __dmul_rn(__dsub_rn(double(v), o), s);
    t = t < 0.0 ? 0.0 : t;
    t = t > 2097151.0 ? 2097151.0 : t;
    return uint64_t(t);
}

// The 63-bit key x << 2 | y << 1 | z of a position in the cube with origin (ox, oy, oz) and scale s.
template <class R>
__device__ __forceinline__ uint64_t morton_key(R x, R y, R z, double ox, double oy, double oz, double s) {
    return dilate21(quantize(x, ox, s)) << 2 | dilate21(quantize(y, oy, s)) << 1 | dilate21(quantize(z, oz, s));
}

}  // namespace lhip
