// Direct summation, FP32 VALU-bound (reference: CpuBackend::direct, src/cpu_backend.cpp:353-392).
//
// Layout: sources are the packed posm array (float4 x, y, z, m). Every lane of a wave walks the same
// source j at the same time, so j is wave-uniform and its float4 arrives through the scalar cache
// into SGPRs: the pair arithmetic reads it as a scalar operand, with no LDS staging and no vector
// memory traffic in the inner loop. Each lane keeps T targets in registers (register blocking), so
// one scalar load feeds 64*T pair evaluations.
//
// Parallelism: a block of 256 lanes holds 256*T targets; the source range is split S ways (grid.y)
// so that small N still launches thousands of waves. With S > 1 every split writes raw partial sums
// and a second pass adds them in split order (deterministic) and applies G.
//
// Self pair: skipped by index (reference src/cpu_backend.cpp:375-377). Only sources inside the
// block's own target range can be the target itself, so only that stretch pays for the compare.
// With eps = 0 the r^2 > 0 guard of the reference applies everywhere (GUARD).
#include "kernels.hpp"

namespace lhip {
namespace {

constexpr int kDirectBS = 256;

template <bool POT>
__device__ __forceinline__ void pair_plain(float px, float py, float pz, float4 q, float e2, float& sx, float& sy,
                                           float& sz, float& sp) {
    const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float mi = q.w * inv;
    const float f = mi * (inv * inv);
    sx = __fmaf_rn(f, dx, sx);
    sy = __fmaf_rn(f, dy, sy);
    sz = __fmaf_rn(f, dz, sz);
    if (POT) sp += mi;
}

// Masked variant: contributes nothing when j == i, or (GUARD) when r^2 is not positive.
template <bool POT, bool GUARD>
__device__ __forceinline__ void pair_masked(float px, float py, float pz, int i, int j, float4 q, float e2, float& sx,
                                            float& sy, float& sz, float& sp) {
    const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    const bool ok = (j != i) && (!GUARD || r2 > 0.f);
    const float inv = ok ? __builtin_amdgcn_rsqf(r2) : 0.f;
    const float mi = q.w * inv;
    const float f = mi * (inv * inv);
    sx = __fmaf_rn(f, dx, sx);
    sy = __fmaf_rn(f, dy, sy);
    sz = __fmaf_rn(f, dz, sz);
    if (POT) sp += mi;
}

template <int T, bool POT, bool GUARD, bool SPLIT>
__global__ void __launch_bounds__(kDirectBS) direct_kernel(const float4* __restrict__ posm, int n, float e2, float G,
                                                           int jchunk, float* __restrict__ ax, float* __restrict__ ay,
                                                           float* __restrict__ az, float* __restrict__ pot,
                                                           float4* __restrict__ part) {
    const int base = blockIdx.x * (kDirectBS * T);
    const int split = blockIdx.y;
    float px[T], py[T], pz[T], sx[T], sy[T], sz[T], sp[T];
    int ii[T];
#pragma unroll
    for (int k = 0; k < T; ++k) {
        ii[k] = base + k * kDirectBS + int(threadIdx.x);
        const float4 p = posm[ii[k] < n ? ii[k] : n - 1];
        px[k] = p.x; py[k] = p.y; pz[k] = p.z;
        sx[k] = sy[k] = sz[k] = sp[k] = 0.f;
    }
    const int j0 = split * jchunk;
    const int j1 = min(n, j0 + jchunk);
    // Sources that can equal one of this block's targets.
    const int sa = max(j0, base), sb = min(j1, base + kDirectBS * T);

    auto plain = [&](int a, int b) {
        int j = a;
        auto four = [&](float4 q0, float4 q1, float4 q2, float4 q3) {
#pragma unroll
            for (int k = 0; k < T; ++k) {
                pair_plain<POT>(px[k], py[k], pz[k], q0, e2, sx[k], sy[k], sz[k], sp[k]);
                pair_plain<POT>(px[k], py[k], pz[k], q1, e2, sx[k], sy[k], sz[k], sp[k]);
                pair_plain<POT>(px[k], py[k], pz[k], q2, e2, sx[k], sy[k], sz[k], sp[k]);
                pair_plain<POT>(px[k], py[k], pz[k], q3, e2, sx[k], sy[k], sz[k], sp[k]);
            }
        };
        if (j + 4 <= b) {
            // The next four sources load (scalar loads) while the current four are evaluated, so
            // the scalar cache's latency hides behind the pair arithmetic.
            float4 q0 = posm[j], q1 = posm[j + 1], q2 = posm[j + 2], q3 = posm[j + 3];
            for (; j + 8 <= b; j += 4) {
                const float4 n0 = posm[j + 4], n1 = posm[j + 5], n2 = posm[j + 6], n3 = posm[j + 7];
                four(q0, q1, q2, q3);
                q0 = n0; q1 = n1; q2 = n2; q3 = n3;
            }
            four(q0, q1, q2, q3);
            j += 4;
        }
        for (; j < b; ++j) {
            const float4 q = posm[j];
#pragma unroll
            for (int k = 0; k < T; ++k) pair_plain<POT>(px[k], py[k], pz[k], q, e2, sx[k], sy[k], sz[k], sp[k]);
        }
    };
    auto masked = [&](int a, int b) {
        for (int j = a; j < b; ++j) {
            const float4 q = posm[j];
#pragma unroll
            for (int k = 0; k < T; ++k)
                pair_masked<POT, GUARD>(px[k], py[k], pz[k], ii[k], j, q, e2, sx[k], sy[k], sz[k], sp[k]);
        }
    };
    if (GUARD) {
        masked(j0, j1);
    } else if (sa >= sb) {
        plain(j0, j1);
    } else {
        plain(j0, sa);
        masked(sa, sb);
        plain(sb, j1);
    }
#pragma unroll
    for (int k = 0; k < T; ++k) {
        const int i = ii[k];
        if (i >= n) continue;
        if (SPLIT) {
            part[(long long)split * n + i] = make_float4(sx[k], sy[k], sz[k], sp[k]);
        } else {
            ax[i] = G * sx[k];
            ay[i] = G * sy[k];
            az[i] = G * sz[k];
            if (POT) pot[i] = -G * sp[k];
        }
    }
}

template <bool POT>
__global__ void __launch_bounds__(kBS) direct_reduce(const float4* __restrict__ part, int n, int S, float G,
                                                     float* __restrict__ ax, float* __restrict__ ay,
                                                     float* __restrict__ az, float* __restrict__ pot) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    float sx = 0, sy = 0, sz = 0, sp = 0;
    for (int s = 0; s < S; ++s) {
        const float4 p = part[(long long)s * n + i];
        sx += p.x; sy += p.y; sz += p.z; sp += p.w;
    }
    ax[i] = G * sx;
    ay[i] = G * sy;
    az[i] = G * sz;
    if (POT) pot[i] = -G * sp;
}

int targets_per_lane(int n) { return n >= 65536 ? 4 : 2; }

template <int T, bool POT, bool GUARD>
void launch_t(hipStream_t s, const float4* posm, int n, float e2, float G, int S, float* ax, float* ay, float* az,
              float* pot, float4* part) {
    const int jchunk = (n + S - 1) / S;
    dim3 grid(grid_for(n, kDirectBS * T), S);
    if (S > 1)
        direct_kernel<T, POT, GUARD, true><<<grid, kDirectBS, 0, s>>>(posm, n, e2, G, jchunk, ax, ay, az, pot, part);
    else
        direct_kernel<T, POT, GUARD, false><<<grid, kDirectBS, 0, s>>>(posm, n, e2, G, jchunk, ax, ay, az, pot, part);
    LHIP_LAUNCH_CHECK();
    if (S > 1) {
        direct_reduce<POT><<<grid_for(n, kBS), kBS, 0, s>>>(part, n, S, G, ax, ay, az, pot);
        LHIP_LAUNCH_CHECK();
    }
}

// FP64 direct sum for a list of targets (validation of the FP32 paths): every body is a source,
// the target itself skipped by index, pair terms in double with a correctly rounded square root.
// Thread per target, sources split over blockIdx.y; partials (a_x, a_y, a_z, sum m/s) per split.
template <class R>
__global__ void __launch_bounds__(kBS) direct_f64_targets_kernel(const R* __restrict__ x, const R* __restrict__ y,
                                                                 const R* __restrict__ z, const float* __restrict__ m,
                                                                 int n, const uint32_t* __restrict__ tgt, int nt,
                                                                 double e2, int jchunk, double* __restrict__ part) {
    const int t = blockIdx.x * kBS + threadIdx.x;
    const int split = blockIdx.y;
    const int i = t < nt ? int(tgt[t]) : 0;
    const double px = x[i], py = y[i], pz = z[i];
    double sx = 0, sy = 0, sz = 0, sp = 0;
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    for (int j = j0; j < j1; ++j) {
        const double dx = double(x[j]) - px, dy = double(y[j]) - py, dz = double(z[j]) - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        if (j == i || !(r2 > 0)) continue;
        const double inv = 1.0 / sqrt(r2), mj = m[j];
        const double f = mj * inv * inv * inv;
        sx += f * dx; sy += f * dy; sz += f * dz; sp += mj * inv;
    }
    if (t >= nt) return;
    part[std::size_t(4 * split + 0) * nt + t] = sx;
    part[std::size_t(4 * split + 1) * nt + t] = sy;
    part[std::size_t(4 * split + 2) * nt + t] = sz;
    part[std::size_t(4 * split + 3) * nt + t] = sp;
}

__global__ void __launch_bounds__(kBS) f64_targets_finish(const double* __restrict__ part, int nt, int S, double G,
                                                          double* __restrict__ ax, double* __restrict__ ay,
                                                          double* __restrict__ az, double* __restrict__ pot) {
    const int t = blockIdx.x * kBS + threadIdx.x;
    if (t >= nt) return;
    double s[4] = {0, 0, 0, 0};
    for (int sp = 0; sp < S; ++sp)
        for (int c = 0; c < 4; ++c) s[c] += part[std::size_t(4 * sp + c) * nt + t];
    ax[t] = G * s[0];
    ay[t] = G * s[1];
    az[t] = G * s[2];
    pot[t] = -G * s[3];
}

}  // namespace

int direct_split(int n) {
    if (n <= 0) return 1;
    const int T = targets_per_lane(n);
    const long long waves = (static_cast<long long>(n) + 64 * T - 1) / (64 * T);
    // About eight waves per SIMD across the chip: the pair loop is VALU-bound, and a SIMD issues FP32
    // faster with more waves to interleave (16,384 bodies 0.131 -> 0.124 ms, 65,536 bodies 1.67 ->
    // 1.57 ms against four waves per SIMD).
    long long S = (8192 + waves - 1) / waves;
    S = std::min<long long>(S, std::max(1, n / 256));  // at least 256 sources per split
    S = std::max<long long>(1, std::min<long long>(S, 64));
    return static_cast<int>(S);
}

void launch_direct(hipStream_t s, const float4* posm, int n, float eps2, float G, bool potential, float* ax, float* ay,
                   float* az, float* pot, float4* part) {
    if (n <= 0) return;
    const int S = direct_split(n);
    const bool guard = !(eps2 > 0.f);
    const int T = targets_per_lane(n);
#define LHIP_DIRECT(TT, P, GD) launch_t<TT, P, GD>(s, posm, n, eps2, G, S, ax, ay, az, pot, part)
    if (T == 4) {
        if (potential) { if (guard) LHIP_DIRECT(4, true, true); else LHIP_DIRECT(4, true, false); }
        else { if (guard) LHIP_DIRECT(4, false, true); else LHIP_DIRECT(4, false, false); }
    } else {
        if (potential) { if (guard) LHIP_DIRECT(2, true, true); else LHIP_DIRECT(2, true, false); }
        else { if (guard) LHIP_DIRECT(2, false, true); else LHIP_DIRECT(2, false, false); }
    }
#undef LHIP_DIRECT
}

}  // namespace lhip

namespace lhip {

int direct_f64_split(int n, int nt) {
    const long long blocks = (static_cast<long long>(nt) + kBS - 1) / kBS;
    long long S = (2048 + blocks - 1) / blocks;
    S = std::min<long long>(S, std::max(1, n / 256));
    return int(std::max<long long>(1, std::min<long long>(S, 64)));
}

void launch_direct_f64_targets(hipStream_t s, const void* x, const void* y, const void* z, bool f64,
                               const float* m, int n, const uint32_t* tgt, int nt, double eps, double G, double* ax,
                               double* ay, double* az, double* pot, double* part) {
    if (n <= 0 || nt <= 0) return;
    const int S = direct_f64_split(n, nt);
    const int jchunk = (n + S - 1) / S;
    const dim3 grid(grid_for(nt, kBS), S);
    if (f64)
        direct_f64_targets_kernel<double><<<grid, kBS, 0, s>>>(static_cast<const double*>(x),
                                                               static_cast<const double*>(y),
                                                               static_cast<const double*>(z), m, n, tgt, nt,
                                                               eps * eps, jchunk, part);
    else
        direct_f64_targets_kernel<float><<<grid, kBS, 0, s>>>(static_cast<const float*>(x),
                                                              static_cast<const float*>(y),
                                                              static_cast<const float*>(z), m, n, tgt, nt, eps * eps,
                                                              jchunk, part);
    LHIP_LAUNCH_CHECK();
    f64_targets_finish<<<grid_for(nt, kBS), kBS, 0, s>>>(part, nt, S, G, ax, ay, az, pot);
    LHIP_LAUNCH_CHECK();
}

}  // namespace lhip
