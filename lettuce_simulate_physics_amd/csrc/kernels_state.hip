// Elementwise state kernels and FP64 reductions: kick, drift, packing, bounding box, conservation
// diagnostics and the exact O(N^2) potential energy. All are HBM-bound except exact_pe (FP64 VALU).
#include <cfloat>

#include "kernels.hpp"

namespace lhip {
namespace {

// ---- elementwise: one body per lane, fully coalesced SoA streams ----

__device__ __forceinline__ float fmadd(float a, float b, float c) { return __fmaf_rn(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return fma(a, b, c); }

// R = float: the plain state; R = double: the mixed-precision form's FP64 state.
template <class R>
__global__ void __launch_bounds__(kBS) kick_kernel(R* __restrict__ vx, R* __restrict__ vy, R* __restrict__ vz,
                                                   const R* __restrict__ ax, const R* __restrict__ ay,
                                                   const R* __restrict__ az, R dt, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        // v += a dt as one fused multiply-add: the rounding the reference CPU backend's loop
        // (src/cpu_backend.cpp:284-287) gets from its x86-64-v3 build.
        vx[i] = fmadd(ax[i], dt, vx[i]);
        vy[i] = fmadd(ay[i], dt, vy[i]);
        vz[i] = fmadd(az[i], dt, vz[i]);
    }
}

template <class R>
__global__ void __launch_bounds__(kBS) drift_kernel(R* __restrict__ x, R* __restrict__ y, R* __restrict__ z,
                                                    const R* __restrict__ vx, const R* __restrict__ vy,
                                                    const R* __restrict__ vz, R dt, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        x[i] = fmadd(vx[i], dt, x[i]);  // as the kick (src/cpu_backend.cpp:297-300)
        y[i] = fmadd(vy[i], dt, y[i]);
        z[i] = fmadd(vz[i], dt, z[i]);
    }
}

// FP64 <-> FP32 state conversion (switching a context between the plain and mixed forms).
__global__ void __launch_bounds__(kBS) f32_to_f64_kernel(const float* __restrict__ a, double* __restrict__ b, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) b[i] = double(a[i]);
}
__global__ void __launch_bounds__(kBS) f64_to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}

__global__ void __launch_bounds__(kBS) pack_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                   const float* __restrict__ z, const float* __restrict__ m,
                                                   float4* __restrict__ posm, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) posm[i] = make_float4(x[i], y[i], z[i], m[i]);
}

__global__ void __launch_bounds__(kBS) pack_posd_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const float* __restrict__ m,
                                                        double4* __restrict__ posd, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) posd[i] = make_double4(x[i], y[i], z[i], double(m[i]));
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

template <class R>
struct MinR { __device__ R operator()(R a, R b) const { return a < b ? a : b; } };
template <class R>
struct MaxR { __device__ R operator()(R a, R b) const { return a > b ? a : b; } };
template <class R>
constexpr R kBig = R(1e300);
template <>
constexpr float kBig<float> = FLT_MAX;

template <class R>
__global__ void __launch_bounds__(kBS) bbox_partial(const R* __restrict__ x, const R* __restrict__ y,
                                                    const R* __restrict__ z, int n, R* __restrict__ part) {
    __shared__ R shlo[kBS / 64][3], shhi[kBS / 64][3];
    R lo[3] = {kBig<R>, kBig<R>, kBig<R>}, hi[3] = {-kBig<R>, -kBig<R>, -kBig<R>};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const R p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) { lo[c] = fmin(lo[c], p[c]); hi[c] = fmax(hi[c], p[c]); }
    }
    block_reduce<3>(lo, MinR<R>{}, shlo);
    __syncthreads();
    block_reduce<3>(hi, MaxR<R>{}, shhi);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) { part[blockIdx.x * 6 + c] = lo[c]; part[blockIdx.x * 6 + 3 + c] = hi[c]; }
}

// One block. The cube follows the reference CPU (src/cpu_backend.cpp:398-406): origin at the
// minimum, side = largest extent (in double) times (1 + 1e-6), or 1 if every body coincides;
// scale = 2^21 / side (src/cpu_backend.cpp:419). IEEE double arithmetic, no contraction. For FP64
// positions the same recipe runs on the FP64 minimum and maximum.
template <class R>
__global__ void __launch_bounds__(kBS) bbox_final(const R* __restrict__ part, int parts, Cube* __restrict__ cube) {
    __shared__ R shlo[kBS / 64][3], shhi[kBS / 64][3];
    R lo[3] = {kBig<R>, kBig<R>, kBig<R>}, hi[3] = {-kBig<R>, -kBig<R>, -kBig<R>};
    for (int b = threadIdx.x; b < parts; b += kBS)
        for (int c = 0; c < 3; ++c) {
            lo[c] = fmin(lo[c], part[b * 6 + c]);
            hi[c] = fmax(hi[c], part[b * 6 + 3 + c]);
        }
    block_reduce<3>(lo, MinR<R>{}, shlo);
    __syncthreads();
    block_reduce<3>(hi, MaxR<R>{}, shhi);
    if (threadIdx.x == 0) {
        double side = __dsub_rn(double(hi[0]), double(lo[0]));
        side = fmax(side, __dsub_rn(double(hi[1]), double(lo[1])));
        side = fmax(side, __dsub_rn(double(hi[2]), double(lo[2])));
        side = side > 0 ? __dmul_rn(side, 1.0 + 1e-6) : 1.0;  // the constant as the reference spells it
        Cube c;
        c.ox = lo[0]; c.oy = lo[1]; c.oz = lo[2];
        c.side = side;
        c.scale = __ddiv_rn(2097152.0, side);
        for (int k = 0; k < 3; ++k) { c.lo[k] = float(lo[k]); c.hi[k] = float(hi[k]); }
        *cube = c;
    }
}

// ---- conservation diagnostics in FP64 (reference src/cpu_backend.cpp:307-329) ----

constexpr int kDiag = 12;

template <class R>
__global__ void __launch_bounds__(kBS) diag_partial(const R* __restrict__ x, const R* __restrict__ y,
                                                    const R* __restrict__ z, const R* __restrict__ vx,
                                                    const R* __restrict__ vy, const R* __restrict__ vz,
                                                    const float* __restrict__ m, const R* __restrict__ pot,
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

// The same on FP64 positions (the mixed form's state).
__global__ void __launch_bounds__(kBS) exact_pe64_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z, const float* __restrict__ m,
                                                         int n, double eps2, double* __restrict__ part) {
    __shared__ double sh[kBS / 64][1];
    const int i = blockIdx.x * kBS + threadIdx.x;
    const bool valid = i < n;
    const int c = valid ? i : 0;
    const double px = x[c], py = y[c], pz = z[c];
    double phi = 0;
    for (int j = 0; j < n; ++j) {
        const double dx = x[j] - px, dy = y[j] - py, dz = z[j] - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        phi += (j != i && r2 > 0) ? double(m[j]) / sqrt(r2) : 0.0;
    }
    double v[1] = {valid ? double(m[c]) * phi : 0.0};
    block_reduce<1>(v, AddD{}, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

}  // namespace

void launch_kick(hipStream_t s, float* vx, float* vy, float* vz, const float* ax, const float* ay, const float* az,
                 float dt, int n) {
    if (n <= 0) return;
    kick_kernel<float><<<grid_for(n, kBS), kBS, 0, s>>>(vx, vy, vz, ax, ay, az, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_drift(hipStream_t s, float* x, float* y, float* z, const float* vx, const float* vy, const float* vz,
                  float dt, int n) {
    if (n <= 0) return;
    drift_kernel<float><<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, vx, vy, vz, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_kick(hipStream_t s, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                 const double* az, double dt, int n) {
    if (n <= 0) return;
    kick_kernel<double><<<grid_for(n, kBS), kBS, 0, s>>>(vx, vy, vz, ax, ay, az, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_drift(hipStream_t s, double* x, double* y, double* z, const double* vx, const double* vy,
                  const double* vz, double dt, int n) {
    if (n <= 0) return;
    drift_kernel<double><<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, vx, vy, vz, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_f32_to_f64(hipStream_t s, const float* a, double* b, int n) {
    if (n <= 0) return;
    f32_to_f64_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(a, b, n);
    LHIP_LAUNCH_CHECK();
}

void launch_f64_to_f32(hipStream_t s, const double* a, float* b, int n) {
    if (n <= 0) return;
    f64_to_f32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(a, b, n);
    LHIP_LAUNCH_CHECK();
}

void launch_pack_posm(hipStream_t s, const float* x, const float* y, const float* z, const float* m, float4* posm,
                      int n) {
    if (n <= 0) return;
    pack_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, m, posm, n);
    LHIP_LAUNCH_CHECK();
}

void launch_pack_posd(hipStream_t s, const double* x, const double* y, const double* z, const float* m, double4* posd,
                      int n) {
    if (n <= 0) return;
    pack_posd_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, m, posd, n);
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

template <class R>
void bbox_impl(hipStream_t s, const R* x, const R* y, const R* z, int n, void* part, Cube* cube) {
    const int pb = n > 0 ? int(std::min<long long>(grid_for(n, kBS), kReduceBlocks)) : 1;
    bbox_partial<R><<<pb, kBS, 0, s>>>(x, y, z, n, static_cast<R*>(part));
    LHIP_LAUNCH_CHECK();
    bbox_final<R><<<1, kBS, 0, s>>>(static_cast<const R*>(part), pb, cube);
    LHIP_LAUNCH_CHECK();
}

void launch_bbox(hipStream_t s, const float* x, const float* y, const float* z, int n, void* part, Cube* cube) {
    bbox_impl(s, x, y, z, n, part, cube);
}

void launch_bbox(hipStream_t s, const double* x, const double* y, const double* z, int n, void* part, Cube* cube) {
    bbox_impl(s, x, y, z, n, part, cube);
}

template <class R>
void diag_impl(hipStream_t s, const R* x, const R* y, const R* z, const R* vx, const R* vy, const R* vz,
               const float* m, const R* pot, int use_pot, int n, double* part, double* out) {
    const int pb = n > 0 ? int(std::min<long long>(grid_for(n, kBS), kReduceBlocks)) : 1;
    diag_partial<R><<<pb, kBS, 0, s>>>(x, y, z, vx, vy, vz, m, pot, use_pot, n, part);
    LHIP_LAUNCH_CHECK();
    sum_partials<kDiag><<<1, kBS, 0, s>>>(part, pb, out);
    LHIP_LAUNCH_CHECK();
}

void launch_diag(hipStream_t s, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                 const float* vz, const float* m, const float* pot, int use_pot, int n, double* part, double* out) {
    diag_impl(s, x, y, z, vx, vy, vz, m, pot, use_pot, n, part, out);
}

void launch_diag(hipStream_t s, const double* x, const double* y, const double* z, const double* vx,
                 const double* vy, const double* vz, const float* m, const double* pot, int use_pot, int n,
                 double* part, double* out) {
    diag_impl(s, x, y, z, vx, vy, vz, m, pot, use_pot, n, part, out);
}

void launch_exact_pe(hipStream_t s, const double* x, const double* y, const double* z, const float* m, int n,
                     double eps2, double* part, double* out) {
    const int g = int(grid_for(n, kBS));
    exact_pe64_kernel<<<g, kBS, 0, s>>>(x, y, z, m, n, eps2, part);
    LHIP_LAUNCH_CHECK();
    sum_partials<1><<<1, kBS, 0, s>>>(part, g, out);
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
