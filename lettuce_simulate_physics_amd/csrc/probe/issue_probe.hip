// Issue-rate probe for the roofline model of the pair loops (run by bench.py beside the force
// path, not part of it): times, over the whole device, wave64 streams of independent FP32 FMAs
// (12 chains per lane), of independent v_rsq_f32 (12 chains), and of 12 FMAs plus one v_rsq_f32
// per step (the pair loops' mix: about 12 VALU instructions per reciprocal square root). The FMA
// stream gives the device's VALU lane-instruction rate; the mixed stream's extra time over the FMA
// stream, per reciprocal square root, is what an rsq costs the pair loops in FMA slots.
//
// A second entry point sweeps the issue rate of single-form streams (FMA with a literal, FMA
// accumulating two VGPR operands, three-VGPR FMA, multiply, add) over the chain count per lane and
// the waves per SIMD (the grid size), which is what bounds the pair loops at a given occupancy.
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

// Forms: 0 v_fmamk_f32 (a = a * k + k'), 1 v_fmac_f32 (a += b * c, b and c in VGPRs), 2 v_fma_f32
// (a = a * b + c, three VGPR operands), 3 v_mul_f32 (a *= b), 4 v_add_f32 (a += b); C independent
// chains per lane.
template <int FORM, int C>
__global__ void __launch_bounds__(kThreads) form_kernel(float* __restrict__ out, int steps, float b0, float c0) {
    const float t = float(threadIdx.x) * 1e-7f;
    float a[C];
#pragma unroll
    for (int k = 0; k < C; ++k) a[k] = 1.0f + t + float(k) * 1e-3f;
    const float b = b0 + t, c = c0 + 0.5f * t;
    for (int i = 0; i < steps; ++i) {
#pragma unroll
        for (int k = 0; k < C; ++k) {
            if constexpr (FORM == 0) a[k] = __fmaf_rn(a[k], 0.9999999f, 1e-7f);
            if constexpr (FORM == 1) a[k] = __fmaf_rn(b, c, a[k]);
            if constexpr (FORM == 2) a[k] = __fmaf_rn(a[k], b, c);
            if constexpr (FORM == 3) a[k] = a[k] * b;
            if constexpr (FORM == 4) a[k] = a[k] + b;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < C; ++k) s += a[k];
    out[std::size_t(blockIdx.x) * kThreads + threadIdx.x] = s;
}

template <int FORM>
bool launch_form(int chains, float* out, int blocks, int steps, hipStream_t s) {
    switch (chains) {
        case 8: form_kernel<FORM, 8><<<blocks, kThreads, 0, s>>>(out, steps, 0.9999f, 1e-7f); return true;
        case 12: form_kernel<FORM, 12><<<blocks, kThreads, 0, s>>>(out, steps, 0.9999f, 1e-7f); return true;
        case 24: form_kernel<FORM, 24><<<blocks, kThreads, 0, s>>>(out, steps, 0.9999f, 1e-7f); return true;
        default: return false;
    }
}

bool launch_any_form(int form, int chains, float* out, int blocks, int steps, hipStream_t s) {
    switch (form) {
        case 0: return launch_form<0>(chains, out, blocks, steps, s);
        case 1: return launch_form<1>(chains, out, blocks, steps, s);
        case 2: return launch_form<2>(chains, out, blocks, steps, s);
        case 3: return launch_form<3>(chains, out, blocks, steps, s);
        case 4: return launch_form<4>(chains, out, blocks, steps, s);
        default: return false;
    }
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

// *ms = milliseconds of one launch (after a warm-up) of form `form` (0-4) with `chains` (8, 12 or
// 24) chains per lane over `blocks` blocks of 256 threads and `steps` steps. Returns 0, or -1 on an
// unsupported form or chain count or a runtime error.
extern "C" int issue_probe_sweep(int form, int chains, int blocks, int steps, double* ms) {
    if (form < 0 || form > 4 || (chains != 8 && chains != 12 && chains != 24) || blocks <= 0 || steps <= 0 || !ms)
        return -1;
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
        launch_any_form(form, chains, out, blocks, steps, s);  // warm-up
        ok(hipEventRecord(e0, s));
        launch_any_form(form, chains, out, blocks, steps, s);
        ok(hipEventRecord(e1, s));
        ok(hipGetLastError());
        float t = 0.f;
        if (ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&t, e0, e1))) *ms = double(t);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    return rc;
}
