// Elementwise state kernels and FP64 reductions: kick, drift, packing, bounding box, conservation
// diagnostics and the exact O(N^2) potential energy. All are HBM-bound except exact_pe (FP64 VALU).
#include <cfloat>

#include "kernels.hpp"

namespace lhip {
namespace {

// ---- elementwise: one body per lane, fully coalesced SoA streams ----

__global__ void __launch_bounds__(kBS) kick_kernel(float* __restrict__ vx, float* __restrict__ vy,
                                                   float* __restrict__ vz, const float* __restrict__ ax,
                                                   const float* __restrict__ ay, const float* __restrict__ az,
                                                   float dt, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        // v += a dt as one fused multiply-add: the rounding the reference CPU backend's loop
        // (src/cpu_backend.cpp:284-287) gets from its x86-64-v3 build.
        vx[i] = __fmaf_rn(ax[i], dt, vx[i]);
        vy[i] = __fmaf_rn(ay[i], dt, vy[i]);
        vz[i] = __fmaf_rn(az[i], dt, vz[i]);
    }
}

__global__ void __launch_bounds__(kBS) drift_kernel(float* __restrict__ x, float* __restrict__ y,
                                                    float* __restrict__ z, const float* __restrict__ vx,
                                                    const float* __restrict__ vy, const float* __restrict__ vz,
                                                    float dt, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        x[i] = __fmaf_rn(vx[i], dt, x[i]);  // as the kick (src/cpu_backend.cpp:297-300)
        y[i] = __fmaf_rn(vy[i], dt, y[i]);
        z[i] = __fmaf_rn(vz[i], dt, z[i]);
    }
}

__global__ void __launch_bounds__(kBS) pack_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                   const float* __restrict__ z, const float* __restrict__ m,
                                                   float4* __restrict__ posm, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) posm[i] = make_float4(x[i], y[i], z[i], m[i]);
}

__global__ void __launch_bounds__(kBS) fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, long long n) {
    const long long i = (long long)blockIdx.x * kBS + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kBS) zero_f32_kernel(float* __restrict__ p, long long n) {
    const long long i = (long long)blockIdx.x * kBS + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ---- wave64 + block reductions ----

template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    return v;
}

struct MinOp { __device__ float operator()(float a, float b) const { return fminf(a, b); } };
struct MaxOp { __device__ float operator()(float a, float b) const { return fmaxf(a, b); } };
struct AddD { __device__ double operator()(double a, double b) const { return a + b; } };

// Block-wide reduction of K values; the result is valid in thread 0. kBS threads = 4 waves.
template <int K, class T, class Op>
__device__ void block_reduce(T (&v)[K], Op op, T (*sh)[K]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_reduce(v[k], op);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            T a = sh[0][k];
            for (int w = 1; w < kBS / 64; ++w) a = op(a, sh[w][k]);
            v[k] = a;
        }
}

// ---- bounding box -> Morton cube ----

__global__ void __launch_bounds__(kBS) bbox_partial(const float* __restrict__ x, const float* __restrict__ y,
                                                    const float* __restrict__ z, int n, float* __restrict__ part) {
    __shared__ float shlo[kBS / 64][3], shhi[kBS / 64][3];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const float p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) { lo[c] = fminf(lo[c], p[c]); hi[c] = fmaxf(hi[c], p[c]); }
    }
    block_reduce<3>(lo, MinOp{}, shlo);
    __syncthreads();
    block_reduce<3>(hi, MaxOp{}, shhi);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) { part[blockIdx.x * 6 + c] = lo[c]; part[blockIdx.x * 6 + 3 + c] = hi[c]; }
}

// One block. The cube follows the reference CPU (src/cpu_backend.cpp:398-406): origin at the float
// minimum, side = largest extent (in double) times (1 + 1e-6), or 1 if every body coincides;
// scale = 2^21 / side (src/cpu_backend.cpp:419). IEEE double arithmetic, no contraction.
__global__ void __launch_bounds__(kBS) bbox_final(const float* __restrict__ part, int parts, Cube* __restrict__ cube) {
    __shared__ float shlo[kBS / 64][3], shhi[kBS / 64][3];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int b = threadIdx.x; b < parts; b += kBS)
        for (int c = 0; c < 3; ++c) {
            lo[c] = fminf(lo[c], part[b * 6 + c]);
            hi[c] = fmaxf(hi[c], part[b * 6 + 3 + c]);
        }
    block_reduce<3>(lo, MinOp{}, shlo);
    __syncthreads();
    block_reduce<3>(hi, MaxOp{}, shhi);
    if (threadIdx.x == 0) {
        double side = double(hi[0]) - double(lo[0]);
        side = fmax(side, double(hi[1]) - double(lo[1]));
        side = fmax(side, double(hi[2]) - double(lo[2]));
        side = side > 0 ? __dmul_rn(side, 1.0 + 1e-6) : 1.0;  // the constant as the reference spells it
        Cube c;
        c.ox = lo[0]; c.oy = lo[1]; c.oz = lo[2];
        c.side = side;
        c.scale = __ddiv_rn(2097152.0, side);
        for (int k = 0; k < 3; ++k) { c.lo[k] = lo[k]; c.hi[k] = hi[k]; }
        *cube = c;
    }
}

// ---- conservation diagnostics in FP64 (reference src/cpu_backend.cpp:307-329) ----

constexpr int kDiag = 12;

__global__ void __launch_bounds__(kBS) diag_partial(const float* __restrict__ x, const float* __restrict__ y,
                                                    const float* __restrict__ z, const float* __restrict__ vx,
                                                    const float* __restrict__ vy, const float* __restrict__ vz,
                                                    const float* __restrict__ m, const float* __restrict__ pot,
                                                    int use_pot, int n, double* __restrict__ part) {
    __shared__ double sh[kBS / 64][kDiag];
    double v[kDiag] = {};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const double mi = m[i], px = x[i], py = y[i], pz = z[i], ux = vx[i], uy = vy[i], uz = vz[i];
        v[0] += mi;
        v[1] += 0.5 * mi * (ux * ux + uy * uy + uz * uz);
        if (use_pot) v[2] += mi * double(pot[i]);
        v[3] += mi * ux; v[4] += mi * uy; v[5] += mi * uz;
        v[6] += mi * (py * uz - pz * uy);
        v[7] += mi * (pz * ux - px * uz);
        v[8] += mi * (px * uy - py * ux);
        v[9] += mi * px; v[10] += mi * py; v[11] += mi * pz;
    }
    block_reduce<kDiag>(v, AddD{}, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < kDiag; ++k) part[blockIdx.x * kDiag + k] = v[k];
}

template <int K>
__global__ void __launch_bounds__(kBS) sum_partials(const double* __restrict__ part, int parts, double* __restrict__ out) {
    __shared__ double sh[kBS / 64][K];
    double v[K] = {};
    for (int b = threadIdx.x; b < parts; b += kBS)
        for (int k = 0; k < K; ++k) v[k] += part[b * K + k];
    block_reduce<K>(v, AddD{}, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) out[k] = v[k];
}

// ---- exact potential energy (reference src/cpu_backend.cpp:331-347) ----
// Thread per target i; sources j are wave-uniform (scalar loads). Pair terms in double.
__global__ void __launch_bounds__(kBS) exact_pe_kernel(const float4* __restrict__ posm, int n, double eps2,
                                                       double* __restrict__ part) {
    __shared__ double sh[kBS / 64][1];
    const int i = blockIdx.x * kBS + threadIdx.x;
    const bool valid = i < n;
    const float4 p = posm[valid ? i : 0];
    double phi = 0;
    for (int j = 0; j < n; ++j) {
        const float4 q = posm[j];
        const double dx = double(q.x) - p.x, dy = double(q.y) - p.y, dz = double(q.z) - p.z;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        phi += (j != i && r2 > 0) ? double(q.w) / sqrt(r2) : 0.0;
    }
    double v[1] = {valid ? double(p.w) * phi : 0.0};
    block_reduce<1>(v, AddD{}, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

}  // namespace

void launch_kick(hipStream_t s, float* vx, float* vy, float* vz, const float* ax, const float* ay, const float* az,
                 float dt, int n) {
    if (n <= 0) return;
    kick_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(vx, vy, vz, ax, ay, az, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_drift(hipStream_t s, float* x, float* y, float* z, const float* vx, const float* vy, const float* vz,
                  float dt, int n) {
    if (n <= 0) return;
    drift_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, vx, vy, vz, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_pack_posm(hipStream_t s, const float* x, const float* y, const float* z, const float* m, float4* posm,
                      int n) {
    if (n <= 0) return;
    pack_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, m, posm, n);
    LHIP_LAUNCH_CHECK();
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, long long n) {
    if (n <= 0) return;
    fill_u32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(p, v, n);
    LHIP_LAUNCH_CHECK();
}

void launch_zero_f32(hipStream_t s, float* p, long long n) {
    if (n <= 0) return;
    zero_f32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(p, n);
    LHIP_LAUNCH_CHECK();
}

void launch_bbox(hipStream_t s, const float* x, const float* y, const float* z, int n, float* part, Cube* cube) {
    const int pb = n > 0 ? int(std::min<long long>(grid_for(n, kBS), kReduceBlocks)) : 1;
    bbox_partial<<<pb, kBS, 0, s>>>(x, y, z, n, part);
    LHIP_LAUNCH_CHECK();
    bbox_final<<<1, kBS, 0, s>>>(part, pb, cube);
    LHIP_LAUNCH_CHECK();
}

void launch_diag(hipStream_t s, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                 const float* vz, const float* m, const float* pot, int use_pot, int n, double* part, double* out) {
    const int pb = n > 0 ? int(std::min<long long>(grid_for(n, kBS), kReduceBlocks)) : 1;
    diag_partial<<<pb, kBS, 0, s>>>(x, y, z, vx, vy, vz, m, pot, use_pot, n, part);
    LHIP_LAUNCH_CHECK();
    sum_partials<kDiag><<<1, kBS, 0, s>>>(part, pb, out);
    LHIP_LAUNCH_CHECK();
}

void launch_exact_pe(hipStream_t s, const float4* posm, int n, double eps2, double* part, double* out) {
    const int g = int(grid_for(n, kBS));
    exact_pe_kernel<<<g, kBS, 0, s>>>(posm, n, eps2, part);
    LHIP_LAUNCH_CHECK();
    sum_partials<1><<<1, kBS, 0, s>>>(part, g, out);
    LHIP_LAUNCH_CHECK();
}

}  // namespace lhip
