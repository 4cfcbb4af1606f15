// Direct summation that allocates nothing: lhip_direct_forces_noalloc, the direct sum of
// lhip_compute_forces (method direct) on caller-owned device arrays, for callers that cannot call
// hipMalloc (code under stream capture, tight serving loops, harnesses that forbid allocation).
//
// The kernels are the context's (launch_direct, kernels_direct.hip), with the same split count S
// for every n, so the results are the context's bit for bit. Their one scratch buffer, the S raw
// partial sums per target when the source range is split (S > 1), lives in `scratch`, a file-scope
// device array of 4,194,304 float4 (64 MiB) that every device this library runs on holds until the
// process ends. The split count direct_split(n) keeps S * n at or below 4,193,792 for every n below
// 2^31 - 1 (largest at n = 2,096,896; S = 1, no scratch, from n = 2^21 on), so every size fits.
//
// The array's name is a host-side stand-in in host code: its device address is taken once per
// device with hipGetSymbolAddress (first call on that device) and only that address reaches the
// kernels. The partial sums are written in full before they are read, so nothing needs clearing.
// Calls that overlap on one device share the array and race.
#include <mutex>

#include "kernels.hpp"
#include "../../include/lettuce_hip.h"

namespace lhip {
namespace {

constexpr std::size_t kScratchFloat4 = std::size_t(1) << 22;  // 4,194,304 partial sums, 64 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float4 scratch[kScratchFloat4];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_on(int dev, float4** out) {
    if (dev < 0 || dev >= kMaxDevices) return LHIP_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return LHIP_RUNTIME;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<float4*>(g_scratch_ptr[dev]);
    return LHIP_OK;
}

}  // namespace
}  // namespace lhip

extern "C" int lhip_direct_forces_noalloc(const float* posm, uint64_t n, double G, double eps, int potential, float* ax,
                                          float* ay, float* az, float* pot, void* stream) {
    using namespace lhip;
    if (n >= uint64_t(INT32_MAX)) return LHIP_INVALID;
    if (n == 0) return LHIP_OK;
    if (!posm || !ax || !ay || !az || (potential && !pot)) return LHIP_INVALID;
    if (reinterpret_cast<uintptr_t>(posm) % alignof(float4) != 0) return LHIP_INVALID;
    const int N = int(n);
    const int S = direct_split(N);
    if (S > 1 && std::size_t(S) * std::size_t(N) > kScratchFloat4) return LHIP_INVALID;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return LHIP_UNAVAILABLE;
    }
    float4* part = nullptr;
    const int rc = scratch_on(dev, &part);
    if (rc != LHIP_OK) return rc;
    const float Gf = float(G), eps2 = float(eps * eps);
    try {
        launch_direct(static_cast<hipStream_t>(stream), reinterpret_cast<const float4*>(posm), N, eps2, Gf,
                      potential != 0, ax, ay, az, potential ? pot : nullptr, part);
    } catch (const Error&) {
        return LHIP_RUNTIME;
    }
    return LHIP_OK;
}
