// Kernels of the differentiable FP64-state simulation: FP64 kick/drift, the mixed-precision direct
// force, an all-FP64 direct force (the finite-difference reference), and the direct-sum
// vector-Jacobian product.
//
// Mixed precision. Positions are FP64. A pair's separation d = x_j - x_i is formed in FP64 and
// rounded to FP32 (relative coordinates: exact to FP32 precision however far the pair is from the
// origin); the pair arithmetic is FP32; each lane's FP32 partial sums are folded into FP64 totals
// every 64 sources, which bounds the FP32 accumulation error by 64 terms.
//
// VJP. For a_i = G sum_{j != i} m_j d_ij / s_ij^3 (d_ij = x_j - x_i, s^2 = |d|^2 + eps^2) and a
// cotangent u_i on each a_i, the cotangent on positions is
//   w_k = G sum_{j != k} T_kj (m_k u_j - m_j u_k),   T = s^-3 I - 3 s^-5 d d^T  (tidal tensor),
// and on masses g_k = -G sum_{j != k} s_kj^-3 (u_j . d_kj). One pass over pairs, no transposes.
#include "adjoint.hpp"

namespace lhip {
namespace {

constexpr int kAdjBS = 256;
constexpr int kFold = 64;  // FP32 partial sums are folded into FP64 every kFold sources

__global__ void __launch_bounds__(kBS) kick64_kernel(double* __restrict__ vx, double* __restrict__ vy,
                                                     double* __restrict__ vz, const double* __restrict__ ax,
                                                     const double* __restrict__ ay, const double* __restrict__ az,
                                                     double dt, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        vx[i] = fma(ax[i], dt, vx[i]);
        vy[i] = fma(ay[i], dt, vy[i]);
        vz[i] = fma(az[i], dt, vz[i]);
    }
}

// y += c * x for three arrays (drift: x += dt v; cotangent updates).
__global__ void __launch_bounds__(kBS) axpy3_kernel(double* __restrict__ y0, double* __restrict__ y1,
                                                    double* __restrict__ y2, const double* __restrict__ x0,
                                                    const double* __restrict__ x1, const double* __restrict__ x2,
                                                    double c, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        y0[i] = fma(x0[i], c, y0[i]);
        y1[i] = fma(x1[i], c, y1[i]);
        y2[i] = fma(x2[i], c, y2[i]);
    }
}

__global__ void __launch_bounds__(kBS) axpy1_kernel(double* __restrict__ y, const double* __restrict__ x, double c,
                                                    int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) y[i] = fma(x[i], c, y[i]);
}

__global__ void __launch_bounds__(kBS) fill64_kernel(double* __restrict__ y, double v, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) y[i] = v;
}

// y = c * x for three arrays.
__global__ void __launch_bounds__(kBS) scale3_kernel(double* __restrict__ y0, double* __restrict__ y1,
                                                     double* __restrict__ y2, const double* __restrict__ x0,
                                                     const double* __restrict__ x1, const double* __restrict__ x2,
                                                     double c, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        y0[i] = c * x0[i];
        y1[i] = c * x1[i];
        y2[i] = c * x2[i];
    }
}

// Mixed-precision direct force for targets [blockIdx.x * 256 * T ...), sources split S ways.
template <int T, bool SPLIT>
__global__ void __launch_bounds__(kAdjBS) force_mixed_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                             const double* __restrict__ z, const double* __restrict__ m,
                                                             int n, float e2, double G, int jchunk,
                                                             double* __restrict__ ax, double* __restrict__ ay,
                                                             double* __restrict__ az, double* __restrict__ part) {
    const int base = blockIdx.x * (kAdjBS * T);
    const int split = blockIdx.y;
    double px[T], py[T], pz[T], sx[T], sy[T], sz[T];
    float fx[T], fy[T], fz[T];
    int ii[T];
#pragma unroll
    for (int k = 0; k < T; ++k) {
        ii[k] = base + k * kAdjBS + int(threadIdx.x);
        const int c = ii[k] < n ? ii[k] : n - 1;
        px[k] = x[c]; py[k] = y[c]; pz[k] = z[c];
        sx[k] = sy[k] = sz[k] = 0.0;
        fx[k] = fy[k] = fz[k] = 0.f;
    }
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    for (int jb = j0; jb < j1; jb += kFold) {
        const int je = min(j1, jb + kFold);
        for (int j = jb; j < je; ++j) {
            const double qx = x[j], qy = y[j], qz = z[j];
            const float qm = float(m[j]);
#pragma unroll
            for (int k = 0; k < T; ++k) {
                const float dx = float(qx - px[k]), dy = float(qy - py[k]), dz = float(qz - pz[k]);
                const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
                const float inv = (j != ii[k] && r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;
                const float f = qm * inv * (inv * inv);
                fx[k] = __fmaf_rn(f, dx, fx[k]);
                fy[k] = __fmaf_rn(f, dy, fy[k]);
                fz[k] = __fmaf_rn(f, dz, fz[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < T; ++k) {
            sx[k] += fx[k]; sy[k] += fy[k]; sz[k] += fz[k];
            fx[k] = fy[k] = fz[k] = 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < T; ++k) {
        const int i = ii[k];
        if (i >= n) continue;
        if (SPLIT) {
            part[(long long)(3 * split + 0) * n + i] = sx[k];
            part[(long long)(3 * split + 1) * n + i] = sy[k];
            part[(long long)(3 * split + 2) * n + i] = sz[k];
        } else {
            ax[i] = G * sx[k];
            ay[i] = G * sy[k];
            az[i] = G * sz[k];
        }
    }
}

// All-FP64 direct force (pair arithmetic in double, correctly rounded sqrt and division).
template <bool SPLIT>
__global__ void __launch_bounds__(kAdjBS) force_f64_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                           const double* __restrict__ z, const double* __restrict__ m,
                                                           int n, double e2, double G, int jchunk,
                                                           double* __restrict__ ax, double* __restrict__ ay,
                                                           double* __restrict__ az, double* __restrict__ part) {
    const int i = blockIdx.x * kAdjBS + threadIdx.x;
    const int split = blockIdx.y;
    const int c = i < n ? i : n - 1;
    const double px = x[c], py = y[c], pz = z[c];
    double sx = 0, sy = 0, sz = 0;
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    for (int j = j0; j < j1; ++j) {
        const double dx = x[j] - px, dy = y[j] - py, dz = z[j] - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        if (j == i || !(r2 > 0)) continue;
        const double inv = 1.0 / sqrt(r2);
        const double f = m[j] * inv * inv * inv;
        sx += f * dx; sy += f * dy; sz += f * dz;
    }
    if (i >= n) return;
    if (SPLIT) {
        part[(long long)(3 * split + 0) * n + i] = sx;
        part[(long long)(3 * split + 1) * n + i] = sy;
        part[(long long)(3 * split + 2) * n + i] = sz;
    } else {
        ax[i] = G * sx; ay[i] = G * sy; az[i] = G * sz;
    }
}

// VJP of the direct force (see the file comment). u = cotangent on accelerations.
template <int T, bool SPLIT>
__global__ void __launch_bounds__(kAdjBS) vjp_direct_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                            const double* __restrict__ z, const double* __restrict__ m,
                                                            const double* __restrict__ ux, const double* __restrict__ uy,
                                                            const double* __restrict__ uz, int n, float e2, double G,
                                                            int jchunk, double* __restrict__ wx, double* __restrict__ wy,
                                                            double* __restrict__ wz, double* __restrict__ gm,
                                                            double* __restrict__ part) {
    const int base = blockIdx.x * (kAdjBS * T);
    const int split = blockIdx.y;
    double px[T], py[T], pz[T], sx[T], sy[T], sz[T], sm[T];
    float mk[T], ukx[T], uky[T], ukz[T], fx[T], fy[T], fz[T], fm[T];
    int ii[T];
#pragma unroll
    for (int k = 0; k < T; ++k) {
        ii[k] = base + k * kAdjBS + int(threadIdx.x);
        const int c = ii[k] < n ? ii[k] : n - 1;
        px[k] = x[c]; py[k] = y[c]; pz[k] = z[c];
        mk[k] = float(m[c]);
        ukx[k] = float(ux[c]); uky[k] = float(uy[c]); ukz[k] = float(uz[c]);
        sx[k] = sy[k] = sz[k] = sm[k] = 0.0;
        fx[k] = fy[k] = fz[k] = fm[k] = 0.f;
    }
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    for (int jb = j0; jb < j1; jb += kFold) {
        const int je = min(j1, jb + kFold);
        for (int j = jb; j < je; ++j) {
            const double qx = x[j], qy = y[j], qz = z[j];
            const float qm = float(m[j]);
            const float ujx = float(ux[j]), ujy = float(uy[j]), ujz = float(uz[j]);
#pragma unroll
            for (int k = 0; k < T; ++k) {
                const float dx = float(qx - px[k]), dy = float(qy - py[k]), dz = float(qz - pz[k]);
                const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
                const float inv = (j != ii[k] && r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;
                const float is2 = inv * inv, is3 = inv * is2, is5 = is3 * is2;
                const float vx = mk[k] * ujx - qm * ukx[k];
                const float vy = mk[k] * ujy - qm * uky[k];
                const float vz = mk[k] * ujz - qm * ukz[k];
                const float dv = dx * vx + dy * vy + dz * vz;
                const float c3 = 3.f * is5 * dv;
                fx[k] += is3 * vx - c3 * dx;
                fy[k] += is3 * vy - c3 * dy;
                fz[k] += is3 * vz - c3 * dz;
                fm[k] -= is3 * (ujx * dx + ujy * dy + ujz * dz);
            }
        }
#pragma unroll
        for (int k = 0; k < T; ++k) {
            sx[k] += fx[k]; sy[k] += fy[k]; sz[k] += fz[k]; sm[k] += fm[k];
            fx[k] = fy[k] = fz[k] = fm[k] = 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < T; ++k) {
        const int i = ii[k];
        if (i >= n) continue;
        if (SPLIT) {
            part[(long long)(4 * split + 0) * n + i] = sx[k];
            part[(long long)(4 * split + 1) * n + i] = sy[k];
            part[(long long)(4 * split + 2) * n + i] = sz[k];
            part[(long long)(4 * split + 3) * n + i] = sm[k];
        } else {
            wx[i] = G * sx[k]; wy[i] = G * sy[k]; wz[i] = G * sz[k]; gm[i] = G * sm[k];
        }
    }
}

// Adds the S split partials of `comps` components in split order and applies G.
__global__ void __launch_bounds__(kBS) reduce_split_kernel(const double* __restrict__ part, int n, int S, int comps,
                                                           double G, double* __restrict__ o0, double* __restrict__ o1,
                                                           double* __restrict__ o2, double* __restrict__ o3) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    double acc[4] = {0, 0, 0, 0};
    for (int s = 0; s < S; ++s)
        for (int c = 0; c < comps; ++c) acc[c] += part[(long long)(comps * s + c) * n + i];
    o0[i] = G * acc[0];
    o1[i] = G * acc[1];
    o2[i] = G * acc[2];
    if (comps == 4) o3[i] = G * acc[3];
}

constexpr int kT = 2;

int split_for(int n, int per_lane) {
    const long long waves = (static_cast<long long>(n) + 64 * per_lane - 1) / (64 * per_lane);
    long long S = (4096 + waves - 1) / waves;
    S = std::min<long long>(S, std::max(1, n / 256));
    return int(std::max<long long>(1, std::min<long long>(S, 64)));
}

}  // namespace

int adj_split(int n) { return std::max(split_for(n, kT), split_for(n, 1)); }

void launch_kick64(hipStream_t s, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                   const double* az, double dt, int n) {
    if (n <= 0) return;
    kick64_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(vx, vy, vz, ax, ay, az, dt, n);
    LHIP_LAUNCH_CHECK();
}

void launch_axpy3(hipStream_t s, double* y0, double* y1, double* y2, const double* x0, const double* x1,
                  const double* x2, double c, int n) {
    if (n <= 0) return;
    axpy3_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(y0, y1, y2, x0, x1, x2, c, n);
    LHIP_LAUNCH_CHECK();
}

void launch_axpy1(hipStream_t s, double* y, const double* x, double c, int n) {
    if (n <= 0) return;
    axpy1_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(y, x, c, n);
    LHIP_LAUNCH_CHECK();
}

void launch_fill64(hipStream_t s, double* y, double v, int n) {
    if (n <= 0) return;
    fill64_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(y, v, n);
    LHIP_LAUNCH_CHECK();
}

void launch_scale3(hipStream_t s, double* y0, double* y1, double* y2, const double* x0, const double* x1,
                   const double* x2, double c, int n) {
    if (n <= 0) return;
    scale3_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(y0, y1, y2, x0, x1, x2, c, n);
    LHIP_LAUNCH_CHECK();
}

void launch_force64(hipStream_t s, const double* x, const double* y, const double* z, const double* m, int n,
                    double eps, double G, bool pure_f64, double* ax, double* ay, double* az, double* part) {
    if (n <= 0) return;
    if (pure_f64) {
        const int S = split_for(n, 1);
        const int jchunk = (n + S - 1) / S;
        dim3 grid(grid_for(n, kAdjBS), S);
        if (S > 1) force_f64_kernel<true><<<grid, kAdjBS, 0, s>>>(x, y, z, m, n, eps * eps, G, jchunk, ax, ay, az, part);
        else force_f64_kernel<false><<<grid, kAdjBS, 0, s>>>(x, y, z, m, n, eps * eps, G, jchunk, ax, ay, az, part);
        LHIP_LAUNCH_CHECK();
        if (S > 1) {
            reduce_split_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(part, n, S, 3, G, ax, ay, az, nullptr);
            LHIP_LAUNCH_CHECK();
        }
        return;
    }
    const int S = split_for(n, kT);
    const int jchunk = (n + S - 1) / S;
    dim3 grid(grid_for(n, kAdjBS * kT), S);
    const float e2 = float(eps * eps);
    if (S > 1) force_mixed_kernel<kT, true><<<grid, kAdjBS, 0, s>>>(x, y, z, m, n, e2, G, jchunk, ax, ay, az, part);
    else force_mixed_kernel<kT, false><<<grid, kAdjBS, 0, s>>>(x, y, z, m, n, e2, G, jchunk, ax, ay, az, part);
    LHIP_LAUNCH_CHECK();
    if (S > 1) {
        reduce_split_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(part, n, S, 3, G, ax, ay, az, nullptr);
        LHIP_LAUNCH_CHECK();
    }
}

void launch_vjp_direct(hipStream_t s, const double* x, const double* y, const double* z, const double* m,
                       const double* ux, const double* uy, const double* uz, int n, double eps, double G, double* wx,
                       double* wy, double* wz, double* gm, double* part) {
    if (n <= 0) return;
    const int S = split_for(n, kT);
    const int jchunk = (n + S - 1) / S;
    dim3 grid(grid_for(n, kAdjBS * kT), S);
    const float e2 = float(eps * eps);
    if (S > 1)
        vjp_direct_kernel<kT, true><<<grid, kAdjBS, 0, s>>>(x, y, z, m, ux, uy, uz, n, e2, G, jchunk, wx, wy, wz, gm, part);
    else
        vjp_direct_kernel<kT, false><<<grid, kAdjBS, 0, s>>>(x, y, z, m, ux, uy, uz, n, e2, G, jchunk, wx, wy, wz, gm,
                                                             part);
    LHIP_LAUNCH_CHECK();
    if (S > 1) {
        reduce_split_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(part, n, S, 4, G, wx, wy, wz, gm);
        LHIP_LAUNCH_CHECK();
    }
}

}  // namespace lhip
