// Kernels of the differentiable FP64-state simulation: fused KDK pieces, the mixed-precision direct
// force, an all-FP64 direct force (the finite-difference reference), the direct-sum vector-Jacobian
// product and the fused adjoint update.
//
// Sources are read from SimRec records (FP64 position, FP32 mass, FP32 cotangent), one 48-byte
// record per body, rewritten by the element-wise kernels each step. A wave loads the record of
// source j with scalar loads (j is wave-uniform), so the pair loops spend no vector instructions
// on source data.
//
// Mixed precision. A pair's separation d = x_j - x_i is formed in FP64 and rounded to FP32 (relative
// coordinates: exact to FP32 precision however far the pair is from the origin); the pair
// arithmetic is FP32; each lane's FP32 partial sums are folded into FP64 totals every 64 sources,
// which bounds the FP32 accumulation error by 64 terms.
//
// Self pairs. With eps > 0 the self pair has d = 0 exactly and adds exactly zero to the force, to
// Q and to the mass cotangent below, so the force loop has no mask at all; only the VJP's P term
// (weight s^-3 on m_k u_j - m_j u_k, a rounding residue when j = k) is masked, in the one block of
// 64 sources that holds the wave's own bodies. eps = 0 takes a guarded variant that skips j = i and
// coincident pairs everywhere.
//
// VJP. For a_i = G sum_{j != i} m_j d_ij / s_ij^3 (d_ij = x_j - x_i, s^2 = |d|^2 + eps^2) and a
// cotangent u_i on each a_i, the cotangent on positions is
//   w_k = G sum_{j != k} T_kj (m_k u_j - m_j u_k),   T = s^-3 I - 3 s^-5 d d^T  (tidal tensor),
// accumulated as P = sum s^-3 v and Q = sum s^-5 (d.v) d with v = m_k u_j - m_j u_k, w = G (P - 3Q);
// on masses g_k = -G sum_{j != k} s_kj^-3 (u_j . d_kj). One pass over pairs, no transposes.
#include "adjoint.hpp"

namespace lhip {
namespace {

constexpr int kAdjBS = 256;
constexpr int kWaves = kAdjBS / 64;
constexpr int kFold = 64;  // FP32 partial sums are folded into FP64 every kFold sources
constexpr int kT = 2;      // targets per lane

__device__ __forceinline__ int lane_id() { return int(threadIdx.x) & 63; }

// ------------------------------------------------------------------ element-wise kernels

// rec <- (x, m) for the force at the current positions.
__global__ void __launch_bounds__(kBS) pack_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, const double* __restrict__ m,
                                                   SimRec* __restrict__ rec, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        SimRec r;
        r.x = x[i]; r.y = y[i]; r.z = z[i];
        r.m = float(m[i]); r.ux = r.uy = r.uz = 0.f; r.pad[0] = r.pad[1] = 0.f;
        rec[i] = r;
    }
}

// rec <- (x, m) with float masses (the mixed-precision backend state).
__global__ void __launch_bounds__(kBS) pack32_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const float* __restrict__ m,
                                                     SimRec* __restrict__ rec, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) {
        SimRec r;
        r.x = x[i]; r.y = y[i]; r.z = z[i];
        r.m = m[i]; r.ux = r.uy = r.uz = 0.f; r.pad[0] = r.pad[1] = 0.f;
        rec[i] = r;
    }
}

// Opening half of a KDK step: v += h a; x += dt v; rec <- (x, m); optionally a copy of x.
__global__ void __launch_bounds__(kBS) kick_drift_kernel(double* __restrict__ x, double* __restrict__ y,
                                                         double* __restrict__ z, double* __restrict__ vx,
                                                         double* __restrict__ vy, double* __restrict__ vz,
                                                         const double* __restrict__ ax, const double* __restrict__ ay,
                                                         const double* __restrict__ az, const double* __restrict__ m,
                                                         double h, double dt, SimRec* __restrict__ rec,
                                                         double* __restrict__ copy, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const double wx = fma(ax[i], h, vx[i]), wy = fma(ay[i], h, vy[i]), wz = fma(az[i], h, vz[i]);
    vx[i] = wx; vy[i] = wy; vz[i] = wz;
    const double px = fma(wx, dt, x[i]), py = fma(wy, dt, y[i]), pz = fma(wz, dt, z[i]);
    x[i] = px; y[i] = py; z[i] = pz;
    SimRec r;
    r.x = px; r.y = py; r.z = pz;
    r.m = float(m[i]); r.ux = r.uy = r.uz = 0.f; r.pad[0] = r.pad[1] = 0.f;
    rec[i] = r;
    if (copy) {
        copy[i] = px;
        copy[std::size_t(n) + i] = py;
        copy[2 * std::size_t(n) + i] = pz;
    }
}

// Closing half of a KDK step: a = G * (sum of the S split partials, in split order); v += h a.
__global__ void __launch_bounds__(kBS) finish_kernel(const double* __restrict__ part, int n, int S, double G, double h,
                                                     double* __restrict__ ax, double* __restrict__ ay,
                                                     double* __restrict__ az, double* __restrict__ vx,
                                                     double* __restrict__ vy, double* __restrict__ vz) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    double sx = 0, sy = 0, sz = 0;
    for (int s = 0; s < S; ++s) {
        sx += part[std::size_t(3 * s + 0) * n + i];
        sy += part[std::size_t(3 * s + 1) * n + i];
        sz += part[std::size_t(3 * s + 2) * n + i];
    }
    const double a0 = G * sx, a1 = G * sy, a2 = G * sz;
    ax[i] = a0; ay[i] = a1; az[i] = a2;
    if (h != 0.0) {
        vx[i] = fma(a0, h, vx[i]);
        vy[i] = fma(a1, h, vy[i]);
        vz[i] = fma(a2, h, vz[i]);
    }
}

// rec <- (pos, m, c * lambda_v): the sources of a VJP.
__global__ void __launch_bounds__(kBS) pack_vjp_kernel(const double* __restrict__ pos, const double* __restrict__ m,
                                                       const double* __restrict__ lvx, const double* __restrict__ lvy,
                                                       const double* __restrict__ lvz, double c,
                                                       SimRec* __restrict__ rec, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    SimRec r;
    r.x = pos[i]; r.y = pos[std::size_t(n) + i]; r.z = pos[2 * std::size_t(n) + i];
    r.m = float(m[i]);
    r.ux = float(c * lvx[i]); r.uy = float(c * lvy[i]); r.uz = float(c * lvz[i]);
    r.pad[0] = r.pad[1] = 0.f;
    rec[i] = r;
}

// After a VJP: lambda_x += G w, lambda_m += G g (sums of the S split partials, in split order); then,
// if drift != 0, the drift's adjoint lambda_v += drift * lambda_x; then, if npos, the sources of the
// next VJP: rec <- (npos, m, c_next * lambda_v).
__global__ void __launch_bounds__(kBS) adj_update_kernel(const double* __restrict__ part, int n, int S, double G,
                                                         double* __restrict__ lx, double* __restrict__ ly,
                                                         double* __restrict__ lz, double* __restrict__ lm,
                                                         double* __restrict__ lvx, double* __restrict__ lvy,
                                                         double* __restrict__ lvz, double drift,
                                                         const double* __restrict__ npos,
                                                         const double* __restrict__ m, double c_next,
                                                         SimRec* __restrict__ rec) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    double s[4] = {0, 0, 0, 0};
    for (int sp = 0; sp < S; ++sp)
        for (int c = 0; c < 4; ++c) s[c] += part[std::size_t(4 * sp + c) * n + i];
    const double gx = lx[i] + G * s[0], gy = ly[i] + G * s[1], gz = lz[i] + G * s[2];
    lx[i] = gx; ly[i] = gy; lz[i] = gz;
    lm[i] += G * s[3];
    double vx = lvx[i], vy = lvy[i], vz = lvz[i];
    if (drift != 0.0) {
        vx = fma(gx, drift, vx);
        vy = fma(gy, drift, vy);
        vz = fma(gz, drift, vz);
        lvx[i] = vx; lvy[i] = vy; lvz[i] = vz;
    }
    if (npos) {
        SimRec r;
        r.x = npos[i]; r.y = npos[std::size_t(n) + i]; r.z = npos[2 * std::size_t(n) + i];
        r.m = float(m[i]);
        r.ux = float(c_next * vx); r.uy = float(c_next * vy); r.uz = float(c_next * vz);
        r.pad[0] = r.pad[1] = 0.f;
        rec[i] = r;
    }
}

__global__ void __launch_bounds__(kBS) axpy_f64_kernel(double* __restrict__ y, const double* __restrict__ x, double c,
                                                       int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) y[i] = fma(x[i], c, y[i]);
}

__global__ void __launch_bounds__(kBS) fill64_kernel(double* __restrict__ y, double v, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) y[i] = v;
}

// ------------------------------------------------------------------ pair kernels
//
// Wave w of the grid owns targets [w * 64 T, (w + 1) * 64 T): target k of lane l is w * 64 T + 64 k + l.
// blockIdx.y selects the source range [split * jchunk, ...), jchunk a multiple of 64, so every block
// of 64 sources either holds all of the wave's own bodies for some k or none of them.

template <int T, bool GUARD, bool POT, bool MASK>
__device__ __forceinline__ void force_block(const SimRec* __restrict__ rec, int jb, int je, float e2,
                                            const double (&px)[T], const double (&py)[T], const double (&pz)[T],
                                            const int (&ii)[T], float (&fx)[T], float (&fy)[T], float (&fz)[T],
                                            float (&fp)[T]) {
#pragma unroll 4
    for (int j = jb; j < je; ++j) {
        const double qx = rec[j].x, qy = rec[j].y, qz = rec[j].z;
        const float qm = rec[j].m;
#pragma unroll
        for (int k = 0; k < T; ++k) {
            const float dx = float(qx - px[k]), dy = float(qy - py[k]), dz = float(qz - pz[k]);
            const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
            float inv = __builtin_amdgcn_rsqf(r2);
            if (MASK) inv = (j != ii[k] && (!GUARD || r2 > 0.f)) ? inv : 0.f;
            const float mi = qm * inv;
            const float f = mi * (inv * inv);
            fx[k] = __fmaf_rn(f, dx, fx[k]);
            fy[k] = __fmaf_rn(f, dy, fy[k]);
            fz[k] = __fmaf_rn(f, dz, fz[k]);
            if (POT) fp[k] += mi;
        }
    }
}

// Mixed-precision direct force (and with POT the potential sum); writes the split's partial sums
// (without G): 3 components, or 4 with POT.
template <int T, bool GUARD, bool POT>
__global__ void __launch_bounds__(kAdjBS) force_mixed_kernel(const SimRec* __restrict__ rec, int n, float e2,
                                                             int jchunk, double* __restrict__ part) {
    constexpr int C = POT ? 4 : 3;
    const int lane = lane_id();
    const int t0 = (blockIdx.x * kWaves + int(threadIdx.x >> 6)) * 64 * T;
    if (t0 >= n) return;  // wave-uniform
    const int split = blockIdx.y;
    double px[T], py[T], pz[T], sx[T], sy[T], sz[T], sp[T];
    float fx[T], fy[T], fz[T], fp[T];
    int ii[T];
#pragma unroll
    for (int k = 0; k < T; ++k) {
        ii[k] = t0 + 64 * k + lane;
        const int c = ii[k] < n ? ii[k] : n - 1;
        px[k] = rec[c].x; py[k] = rec[c].y; pz[k] = rec[c].z;
        sx[k] = sy[k] = sz[k] = sp[k] = 0.0;
        fx[k] = fy[k] = fz[k] = fp[k] = 0.f;
    }
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    const int t1 = t0 + 64 * T;
    for (int jb = j0; jb < j1; jb += kFold) {
        const int je = min(j1, jb + kFold);
        // The potential's self term is not zero: mask the block holding the wave's own bodies.
        if (GUARD || (POT && jb < t1 && je > t0))
            force_block<T, GUARD, POT, true>(rec, jb, je, e2, px, py, pz, ii, fx, fy, fz, fp);
        else
            force_block<T, false, POT, false>(rec, jb, je, e2, px, py, pz, ii, fx, fy, fz, fp);
#pragma unroll
        for (int k = 0; k < T; ++k) {
            sx[k] += fx[k]; sy[k] += fy[k]; sz[k] += fz[k]; sp[k] += fp[k];
            fx[k] = fy[k] = fz[k] = fp[k] = 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < T; ++k) {
        const int i = ii[k];
        if (i >= n) continue;
        part[std::size_t(C * split + 0) * n + i] = sx[k];
        part[std::size_t(C * split + 1) * n + i] = sy[k];
        part[std::size_t(C * split + 2) * n + i] = sz[k];
        if (POT) part[std::size_t(C * split + 3) * n + i] = sp[k];
    }
}

// a = G * sum of the split partials, phi = -G * sum (with the potential); in split order.
__global__ void __launch_bounds__(kBS) direct64_finish_kernel(const double* __restrict__ part, int n, int S, int C,
                                                              double G, double* __restrict__ ax,
                                                              double* __restrict__ ay, double* __restrict__ az,
                                                              double* __restrict__ pot) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    double s[4] = {0, 0, 0, 0};
    for (int sp = 0; sp < S; ++sp)
        for (int c = 0; c < C; ++c) s[c] += part[std::size_t(C * sp + c) * n + i];
    ax[i] = G * s[0];
    ay[i] = G * s[1];
    az[i] = G * s[2];
    if (C == 4) pot[i] = -G * s[3];
}

// All-FP64 direct force (pair arithmetic in double, correctly rounded sqrt and division; FP64 masses).
__global__ void __launch_bounds__(kAdjBS) force_f64_kernel(const SimRec* __restrict__ rec,
                                                           const double* __restrict__ m, int n, double e2, int jchunk,
                                                           double* __restrict__ part) {
    const int i = blockIdx.x * kAdjBS + threadIdx.x;
    const int split = blockIdx.y;
    const int c = i < n ? i : n - 1;
    const double px = rec[c].x, py = rec[c].y, pz = rec[c].z;
    double sx = 0, sy = 0, sz = 0;
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    for (int j = j0; j < j1; ++j) {
        const double dx = rec[j].x - px, dy = rec[j].y - py, dz = rec[j].z - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        if (j == i || !(r2 > 0)) continue;
        const double inv = 1.0 / sqrt(r2);
        const double f = m[j] * inv * inv * inv;
        sx += f * dx; sy += f * dy; sz += f * dz;
    }
    if (i >= n) return;
    part[std::size_t(3 * split + 0) * n + i] = sx;
    part[std::size_t(3 * split + 1) * n + i] = sy;
    part[std::size_t(3 * split + 2) * n + i] = sz;
}

template <int T, bool MASK, bool GUARD>
__device__ __forceinline__ void vjp_block(const SimRec* __restrict__ rec, int jb, int je, float e2,
                                          const double (&px)[T], const double (&py)[T], const double (&pz)[T],
                                          const float (&mk)[T], const float (&ukx)[T], const float (&uky)[T],
                                          const float (&ukz)[T], const int (&ii)[T], float (&Px)[T], float (&Py)[T],
                                          float (&Pz)[T], float (&Qx)[T], float (&Qy)[T], float (&Qz)[T],
                                          float (&Gm)[T]) {
#pragma unroll 2
    for (int j = jb; j < je; ++j) {
        const double qx = rec[j].x, qy = rec[j].y, qz = rec[j].z;
        const float qm = rec[j].m, ujx = rec[j].ux, ujy = rec[j].uy, ujz = rec[j].uz;
#pragma unroll
        for (int k = 0; k < T; ++k) {
            const float dx = float(qx - px[k]), dy = float(qy - py[k]), dz = float(qz - pz[k]);
            const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
            float inv = __builtin_amdgcn_rsqf(r2);
            if (MASK) inv = (j != ii[k] && (!GUARD || r2 > 0.f)) ? inv : 0.f;
            const float is2 = inv * inv, is3 = inv * is2, is5 = is3 * is2;
            const float vx = __fmaf_rn(mk[k], ujx, -(qm * ukx[k]));
            const float vy = __fmaf_rn(mk[k], ujy, -(qm * uky[k]));
            const float vz = __fmaf_rn(mk[k], ujz, -(qm * ukz[k]));
            const float c = is5 * __fmaf_rn(dz, vz, __fmaf_rn(dy, vy, dx * vx));
            Px[k] = __fmaf_rn(is3, vx, Px[k]);
            Py[k] = __fmaf_rn(is3, vy, Py[k]);
            Pz[k] = __fmaf_rn(is3, vz, Pz[k]);
            Qx[k] = __fmaf_rn(c, dx, Qx[k]);
            Qy[k] = __fmaf_rn(c, dy, Qy[k]);
            Qz[k] = __fmaf_rn(c, dz, Qz[k]);
            Gm[k] = __fmaf_rn(is3, __fmaf_rn(dz, ujz, __fmaf_rn(dy, ujy, dx * ujx)), Gm[k]);
        }
    }
}

// VJP of the direct force (see the file comment); writes the split's partial sums (without G):
// components w_x, w_y, w_z, g.
template <int T, bool GUARD>
__global__ void __launch_bounds__(kAdjBS) vjp_direct_kernel(const SimRec* __restrict__ rec, int n, float e2,
                                                            int jchunk, double* __restrict__ part) {
    const int lane = lane_id();
    const int t0 = (blockIdx.x * kWaves + int(threadIdx.x >> 6)) * 64 * T;
    if (t0 >= n) return;  // wave-uniform
    const int split = blockIdx.y;
    double px[T], py[T], pz[T], sx[T], sy[T], sz[T], sm[T];
    float mk[T], ukx[T], uky[T], ukz[T], Px[T], Py[T], Pz[T], Qx[T], Qy[T], Qz[T], Gm[T];
    int ii[T];
#pragma unroll
    for (int k = 0; k < T; ++k) {
        ii[k] = t0 + 64 * k + lane;
        const int c = ii[k] < n ? ii[k] : n - 1;
        px[k] = rec[c].x; py[k] = rec[c].y; pz[k] = rec[c].z;
        mk[k] = rec[c].m; ukx[k] = rec[c].ux; uky[k] = rec[c].uy; ukz[k] = rec[c].uz;
        sx[k] = sy[k] = sz[k] = sm[k] = 0.0;
        Px[k] = Py[k] = Pz[k] = Qx[k] = Qy[k] = Qz[k] = Gm[k] = 0.f;
    }
    const int j0 = split * jchunk, j1 = min(n, j0 + jchunk);
    const int t1 = t0 + 64 * T;
    for (int jb = j0; jb < j1; jb += kFold) {
        const int je = min(j1, jb + kFold);
        if (GUARD || (jb < t1 && je > t0))
            vjp_block<T, true, GUARD>(rec, jb, je, e2, px, py, pz, mk, ukx, uky, ukz, ii, Px, Py, Pz, Qx, Qy, Qz, Gm);
        else
            vjp_block<T, false, false>(rec, jb, je, e2, px, py, pz, mk, ukx, uky, ukz, ii, Px, Py, Pz, Qx, Qy, Qz,
                                       Gm);
#pragma unroll
        for (int k = 0; k < T; ++k) {
            sx[k] = fma(-3.0, double(Qx[k]), sx[k] + double(Px[k]));
            sy[k] = fma(-3.0, double(Qy[k]), sy[k] + double(Py[k]));
            sz[k] = fma(-3.0, double(Qz[k]), sz[k] + double(Pz[k]));
            sm[k] -= double(Gm[k]);
            Px[k] = Py[k] = Pz[k] = Qx[k] = Qy[k] = Qz[k] = Gm[k] = 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < T; ++k) {
        const int i = ii[k];
        if (i >= n) continue;
        part[std::size_t(4 * split + 0) * n + i] = sx[k];
        part[std::size_t(4 * split + 1) * n + i] = sy[k];
        part[std::size_t(4 * split + 2) * n + i] = sz[k];
        part[std::size_t(4 * split + 3) * n + i] = sm[k];
    }
}

// Source split: enough waves to give every SIMD of the chip about eight, at least 256 sources per
// split, at most 64 splits; the chunk is a multiple of 64 sources.
struct Split {
    int S, jchunk;
};

Split split_for(int n, int per_lane) {
    const long long waves = (static_cast<long long>(n) + 64 * per_lane - 1) / (64 * per_lane);
    long long S = (8192 + waves - 1) / waves;
    S = std::min<long long>(S, std::max(1, n / 256));
    S = std::max<long long>(1, std::min<long long>(S, 64));
    long long chunk = (n + S - 1) / S;
    chunk = (chunk + 63) / 64 * 64;
    S = (n + chunk - 1) / chunk;
    return {int(S), int(chunk)};
}

int wave_blocks(int n, int per_lane) {
    const long long waves = (static_cast<long long>(n) + 64 * per_lane - 1) / (64 * per_lane);
    return int((waves + kWaves - 1) / kWaves);
}

}  // namespace

int adj_split(int n) { return std::max(split_for(n, kT).S, split_for(n, 1).S); }

void launch_pack(hipStream_t s, const double* x, const double* y, const double* z, const double* m, SimRec* rec,
                 int n) {
    if (n <= 0) return;
    pack_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, m, rec, n);
    LHIP_LAUNCH_CHECK();
}

void launch_kick_drift(hipStream_t s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                       const double* ax, const double* ay, const double* az, const double* m, double h, double dt,
                       SimRec* rec, double* copy, int n) {
    if (n <= 0) return;
    kick_drift_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, vx, vy, vz, ax, ay, az, m, h, dt, rec, copy, n);
    LHIP_LAUNCH_CHECK();
}

void launch_force_finish(hipStream_t s, const SimRec* rec, const double* m, int n, double eps, double G, bool pure_f64,
                         double h, double* ax, double* ay, double* az, double* vx, double* vy, double* vz,
                         double* part) {
    if (n <= 0) return;
    int S;
    if (pure_f64) {
        const Split sp = split_for(n, 1);
        S = sp.S;
        force_f64_kernel<<<dim3(grid_for(n, kAdjBS), S), kAdjBS, 0, s>>>(rec, m, n, eps * eps, sp.jchunk, part);
    } else {
        const Split sp = split_for(n, kT);
        S = sp.S;
        const float e2 = float(eps * eps);
        const dim3 grid(wave_blocks(n, kT), S);
        if (e2 > 0.f) force_mixed_kernel<kT, false, false><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
        else force_mixed_kernel<kT, true, false><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
    }
    LHIP_LAUNCH_CHECK();
    finish_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(part, n, S, G, h, ax, ay, az, vx, vy, vz);
    LHIP_LAUNCH_CHECK();
}

int direct64_split(int n) { return split_for(n, kT).S; }

void launch_direct64(hipStream_t s, const double* x, const double* y, const double* z, const float* m, int n,
                     double eps, double G, bool potential, double* ax, double* ay, double* az, double* pot,
                     SimRec* rec, double* part) {
    if (n <= 0) return;
    pack32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, m, rec, n);
    LHIP_LAUNCH_CHECK();
    const Split sp = split_for(n, kT);
    const float e2 = float(eps * eps);
    const dim3 grid(wave_blocks(n, kT), sp.S);
    const bool guard = !(e2 > 0.f);
    if (potential) {
        if (guard) force_mixed_kernel<kT, true, true><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
        else force_mixed_kernel<kT, false, true><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
    } else {
        if (guard) force_mixed_kernel<kT, true, false><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
        else force_mixed_kernel<kT, false, false><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
    }
    LHIP_LAUNCH_CHECK();
    direct64_finish_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(part, n, sp.S, potential ? 4 : 3, G, ax, ay, az, pot);
    LHIP_LAUNCH_CHECK();
}

void launch_pack_vjp(hipStream_t s, const double* pos, const double* m, const double* lvx, const double* lvy,
                     const double* lvz, double c, SimRec* rec, int n) {
    if (n <= 0) return;
    pack_vjp_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(pos, m, lvx, lvy, lvz, c, rec, n);
    LHIP_LAUNCH_CHECK();
}

void launch_vjp_update(hipStream_t s, const SimRec* rec, int n, double eps, double G, double* part, double* lx,
                       double* ly, double* lz, double* lm, double* lvx, double* lvy, double* lvz, double drift,
                       const double* npos, const double* m, double c_next, SimRec* next_rec) {
    if (n <= 0) return;
    const Split sp = split_for(n, kT);
    const float e2 = float(eps * eps);
    const dim3 grid(wave_blocks(n, kT), sp.S);
    if (e2 > 0.f) vjp_direct_kernel<kT, false><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
    else vjp_direct_kernel<kT, true><<<grid, kAdjBS, 0, s>>>(rec, n, e2, sp.jchunk, part);
    LHIP_LAUNCH_CHECK();
    adj_update_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(part, n, sp.S, G, lx, ly, lz, lm, lvx, lvy, lvz, drift, npos, m,
                                                       c_next, next_rec);
    LHIP_LAUNCH_CHECK();
}

void launch_axpy_f64(hipStream_t s, double* y, const double* x, double c, int n) {
    if (n <= 0) return;
    axpy_f64_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(y, x, c, n);
    LHIP_LAUNCH_CHECK();
}

void launch_fill64(hipStream_t s, double* y, double v, int n) {
    if (n <= 0) return;
    fill64_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(y, v, n);
    LHIP_LAUNCH_CHECK();
}

}  // namespace lhip
