// Shared definitions for the lettuce HIP backend library (gfx950 / MI355X).
//
// Device memory is owned by a Context (one per backend instance, one HIP stream each). Kernels live
// in per-phase translation units and are launched through the host functions declared here, so the
// host code never needs a kernel's definition.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace lhip {

// Status codes of the C ABI (include/lettuce_hip.h).
enum Status : int {
    kOk = 0,
    kRuntime = 1,      // device or runtime failure (reference: std::runtime_error)
    kLogic = 2,        // protocol misuse (reference: std::logic_error)
    kUnavailable = 3,  // no usable device
    kOutOfMemory = 4,
    kInvalid = 5,      // bad argument
    kNotProvided = 6,  // optional query without data (e.g. body costs before a force call)
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        const int code = e == hipErrorOutOfMemory ? kOutOfMemory : kRuntime;
        throw Error(code, std::string(file) + ":" + std::to_string(line) + ": " + what + ": " + hipGetErrorString(e));
    }
}
#define LHIP_CHECK(call) ::lhip::check((call), #call, __FILE__, __LINE__)
#define LHIP_LAUNCH_CHECK() ::lhip::check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Grow-only device buffer; contents are not preserved across growth.
class DevBuf {
public:
    DevBuf() = default;
    ~DevBuf() {
        if (p_) (void)hipFree(p_);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void reserve(std::size_t bytes) {
        if (bytes <= cap_) return;
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        cap_ = 0;
        const std::size_t want = bytes < 256 ? 256 : bytes;
        hipError_t e = hipMalloc(&p_, want);
        if (e != hipSuccess) {
            p_ = nullptr;
            (void)hipGetLastError();
            throw Error(kOutOfMemory, "hipMalloc(" + std::to_string(want) + " bytes) failed: " + hipGetErrorString(e));
        }
        cap_ = want;
    }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    std::size_t capacity() const { return cap_; }
    friend void swap(DevBuf& a, DevBuf& b) noexcept {
        void* p = a.p_; a.p_ = b.p_; b.p_ = p;
        std::size_t c = a.cap_; a.cap_ = b.cap_; b.cap_ = c;
    }

private:
    void* p_ = nullptr;
    std::size_t cap_ = 0;
};

constexpr int kBS = 256;            // elementwise / build block size (4 waves)
constexpr int kReduceBlocks = 1024; // partial-reduction grid cap
constexpr int kLeafMax = 16;        // bodies per leaf (reference src/cpu_backend.cpp:24)
constexpr int kGroupMax = 128;      // walk group size (reference src/cpu_backend.cpp:27)
constexpr int kMaxLevel = 21;       // Morton bits per axis (reference src/cpu_backend.cpp:28)
constexpr int kSmallLeaf = 4;       // accepted leaves this small are summed directly (src/cpu_backend.cpp:29)
constexpr int kLevels = kMaxLevel + 1;
// Key restarts (reference src/cpu_backend.cpp:501-544): an over-full level-21 cell whose bodies do
// not coincide restarts the keys in its own bounding cube, one more 21-level key frame. At most
// kMaxFrames frames (tree depth kMaxDepth); deeper over-full cells stay leaves and are counted.
constexpr int kMaxFrames = 4;
constexpr int kMaxDepth = kMaxLevel * kMaxFrames;
constexpr int kDepthLevels = kMaxDepth + 1;

inline unsigned grid_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

// Morton cube of the current build, written on the device by the bounding-box reduction:
// origin = float minimum (as double), side = max extent * (1 + 1e-6) (1 if zero), scale = 2^21 / side.
struct Cube {
    double ox, oy, oz, side, scale;
    float lo[3], hi[3];
};

// Node info word: level (depth) in bits 0-6, leaf flag in bit 7, child count in bits 8-11.
constexpr int32_t kInfoLeaf = 1 << 7;
__host__ __device__ inline int info_level(int32_t info) { return info & 127; }
__host__ __device__ inline bool info_leaf(int32_t info) { return (info & kInfoLeaf) != 0; }
__host__ __device__ inline int info_nchild(int32_t info) { return (info >> 8) & 15; }

// 64-byte multipole record, one scalar-cache load in the walk: com, mass, traceless quadrupole.
struct alignas(64) NodeMom {
    float cx, cy, cz, m;
    float qxx, qxy, qxz, qyy;
    float qyz, qzz, rcrit2, qrel;  // qrel = 1.5 |Q|_F / M: quadrupole/monopole term ratio times dist^2
    float gx, gy, gz, half;  // geometric cell centre and half side
};
static_assert(sizeof(NodeMom) == 64, "NodeMom is one 64-byte record");

}  // namespace lhip
