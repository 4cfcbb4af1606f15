// Accumulation probe for the mixed-precision form (run by bench.py beside the force path, not part of
// it): the same FP32 pair arithmetic as the mixed direct sum (separation formed in FP64 and rounded
// to FP32, then r^2, v_rsq_f32, m r^-3 d in FP32), summed five ways over every source j for each
// target i, so that the accumulation's share of the force error and of the time can be told apart:
//
//   mode 0  the FP64 reference: FP64 separation, FP64 pair arithmetic with correctly rounded sqrt and
//           division, FP64 sums;
//   mode 1  FP32 terms summed in FP32 (one fma per component and term);
//   mode 2  FP32 terms summed with Kahan compensation in FP32 (sum and compensation per component);
//   mode 3  FP32 terms summed in FP32 over runs of 64 sources, each run's sum carried into FP64 (the
//           library's mixed direct sum and walk accumulate this way);
//   mode 4  every FP32 term converted and summed in FP64 (the accumulation error left: none beyond
//           FP64 rounding, so mode 4's error is the FP32 pair arithmetic's alone).
//
// One thread per target; the source index is wave-uniform, so sources arrive by scalar loads.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdlib>

namespace {

constexpr int kThreads = 256;
constexpr int kRun = 64;  // sources per FP32 run of mode 3

template <int MODE>
__global__ void __launch_bounds__(kThreads) accum_kernel(const double4* __restrict__ src, int n, double e2d,
                                                         double* __restrict__ out) {
    const int i = blockIdx.x * kThreads + int(threadIdx.x);
    const int c = i < n ? i : n - 1;
    const double4 p = src[c];
    const float e2 = float(e2d);
    double dx64 = 0, dy64 = 0, dz64 = 0;        // modes 0, 3, 4
    float sx = 0.f, sy = 0.f, sz = 0.f;         // modes 1, 2, 3
    float cx = 0.f, cy = 0.f, cz = 0.f;         // mode 2: compensations
    auto kahan = [](float& s, float& comp, float term) {
        const float y = term - comp;
        const float t = s + y;
        comp = (t - s) - y;
        s = t;
    };
    for (int jb = 0; jb < n; jb += kRun) {
        const int je = jb + kRun < n ? jb + kRun : n;
        for (int j = jb; j < je; ++j) {
            const double4 q = src[j];
            if constexpr (MODE == 0) {
                const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
                const double r2 = dx * dx + dy * dy + dz * dz + e2d;
                const double inv = 1.0 / sqrt(r2);
                const double f = q.w * inv * inv * inv;
                dx64 += f * dx;
                dy64 += f * dy;
                dz64 += f * dz;
            } else {
                const float dx = float(q.x - p.x), dy = float(q.y - p.y), dz = float(q.z - p.z);
                const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float mi = float(q.w) * inv;
                const float f = mi * (inv * inv);
                if constexpr (MODE == 1 || MODE == 3) {
                    sx = __fmaf_rn(f, dx, sx);
                    sy = __fmaf_rn(f, dy, sy);
                    sz = __fmaf_rn(f, dz, sz);
                } else if constexpr (MODE == 2) {
                    kahan(sx, cx, f * dx);
                    kahan(sy, cy, f * dy);
                    kahan(sz, cz, f * dz);
                } else {
                    dx64 += double(f * dx);
                    dy64 += double(f * dy);
                    dz64 += double(f * dz);
                }
            }
        }
        if constexpr (MODE == 3) {
            dx64 += double(sx);
            dy64 += double(sy);
            dz64 += double(sz);
            sx = sy = sz = 0.f;
        }
    }
    if (i >= n) return;
    if constexpr (MODE == 1 || MODE == 2) {
        out[3 * std::size_t(i) + 0] = sx;
        out[3 * std::size_t(i) + 1] = sy;
        out[3 * std::size_t(i) + 2] = sz;
    } else {
        out[3 * std::size_t(i) + 0] = dx64;
        out[3 * std::size_t(i) + 1] = dy64;
        out[3 * std::size_t(i) + 2] = dz64;
    }
}

template <int MODE>
void launch(const double4* src, int n, double e2, double* out) {
    accum_kernel<MODE><<<(n + kThreads - 1) / kThreads, kThreads>>>(src, n, e2, out);
}

}  // namespace

// Sums a(i) = sum_j m_j d_ij / (|d_ij|^2 + eps^2)^1.5 (without G) for every body i of the n bodies
// (x, y, z, m) in the given mode (eps > 0: a body's own term is then exactly zero); out receives
// 3 n doubles (x, y, z per body) and ms the kernel time of the second of two launches. Returns 0 on success, -1 on a bad argument, else the HIP error.
extern "C" int accum_probe_run(const double* x, const double* y, const double* z, const double* m, int n, double eps,
                               int mode, double* out, double* ms) {
    if (n < 1 || !(eps > 0) || mode < 0 || mode > 4 || !x || !y || !z || !m || !out || !ms) return -1;
    double4* host = static_cast<double4*>(std::malloc(std::size_t(n) * sizeof(double4)));
    if (!host) return -1;
    for (int i = 0; i < n; ++i) host[i] = make_double4(x[i], y[i], z[i], m[i]);
    double4* src = nullptr;
    double* dout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(reinterpret_cast<void**>(&src), std::size_t(n) * sizeof(double4));
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&dout), std::size_t(n) * 3 * sizeof(double));
    if (err == hipSuccess) err = hipMemcpy(src, host, std::size_t(n) * sizeof(double4), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    const double e2 = eps * eps;
    for (int rep = 0; rep < 2 && err == hipSuccess; ++rep) {
        err = hipEventRecord(e0, nullptr);
        switch (mode) {
            case 0: launch<0>(src, n, e2, dout); break;
            case 1: launch<1>(src, n, e2, dout); break;
            case 2: launch<2>(src, n, e2, dout); break;
            case 3: launch<3>(src, n, e2, dout); break;
            default: launch<4>(src, n, e2, dout); break;
        }
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
    }
    float t = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    if (err == hipSuccess) err = hipMemcpy(out, dout, std::size_t(n) * 3 * sizeof(double), hipMemcpyDeviceToHost);
    *ms = t;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dout) (void)hipFree(dout);
    std::free(host);
    return err == hipSuccess ? 0 : int(err);
}
