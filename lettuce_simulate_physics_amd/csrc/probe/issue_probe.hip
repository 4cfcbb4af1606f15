// Issue-rate probe for the roofline model of the pair loops (run by bench.py beside the force
// path, not part of it): times, over the whole device, wave64 streams of independent FP32 FMAs
// (12 chains per lane), of independent v_rsq_f32 (12 chains), and of 12 FMAs plus one v_rsq_f32
// per step (the pair loops' mix: about 12 VALU instructions per reciprocal square root). The FMA
// stream gives the device's VALU lane-instruction rate; the mixed stream's extra time over the FMA
// stream, per reciprocal square root, is what an rsq costs the pair loops in FMA slots.
#include <hip/hip_runtime.h>

namespace {

constexpr int kThreads = 256;
constexpr int kChains = 12;

template <int KIND>  // 0: FMAs; 1: rsq; 2: 12 FMAs + 1 rsq per step
__global__ void __launch_bounds__(kThreads) issue_probe_kernel(float* __restrict__ out, int steps) {
    const float t = float(threadIdx.x) * 1e-7f;
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = 1.0f + t + float(k) * 1e-3f;
    float r = 1.5f + t;
    for (int i = 0; i < steps; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k)
            a[k] = KIND == 1 ? __builtin_amdgcn_rsqf(a[k]) : __fmaf_rn(a[k], 0.9999999f, 1e-7f);
        if (KIND == 2) r = __builtin_amdgcn_rsqf(r);
    }
    float s = r;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[std::size_t(blockIdx.x) * kThreads + threadIdx.x] = s;
}

template <int KIND>
void launch(float* out, int blocks, int steps, hipStream_t s) {
    issue_probe_kernel<KIND><<<blocks, kThreads, 0, s>>>(out, steps);
}

}  // namespace

// ms[0..3) = milliseconds of the FMA, rsq and mixed streams (each launch timed after a warm-up),
// over `blocks` blocks of 256 threads and `steps` steps. Returns 0, or -1 on a runtime error.
extern "C" int issue_probe_run(int blocks, int steps, double* ms) {
    if (blocks <= 0 || steps <= 0 || !ms) return -1;
    float* out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto ok = [&rc](hipError_t e) {
        if (e != hipSuccess && rc == 0) rc = -1;
        return rc == 0;
    };
    if (ok(hipMalloc(reinterpret_cast<void**>(&out), std::size_t(blocks) * kThreads * sizeof(float))) &&
        ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1))) {
        void (*fn[3])(float*, int, int, hipStream_t) = {launch<0>, launch<1>, launch<2>};
        for (int k = 0; k < 3 && rc == 0; ++k) {
            fn[k](out, blocks, steps, s);  // warm-up
            ok(hipEventRecord(e0, s));
            fn[k](out, blocks, steps, s);
            ok(hipEventRecord(e1, s));
            ok(hipGetLastError());
            float t = 0.f;
            if (ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&t, e0, e1))) ms[k] = double(t);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    return rc;
}
