// Host-side launchers for the device kernels. Each launcher enqueues on the given stream and
// checks the launch; none synchronises.
#pragma once
#include "common.hpp"

namespace lhip {

// ---- state (kernels_state.hip) ----
// float: the plain state; double: the mixed-precision form's FP64 state.
void launch_kick(hipStream_t s, float* vx, float* vy, float* vz, const float* ax, const float* ay, const float* az,
                 float dt, int n);
void launch_drift(hipStream_t s, float* x, float* y, float* z, const float* vx, const float* vy, const float* vz,
                  float dt, int n);
void launch_kick(hipStream_t s, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                 const double* az, double dt, int n);
void launch_drift(hipStream_t s, double* x, double* y, double* z, const double* vx, const double* vy,
                  const double* vz, double dt, int n);
void launch_f32_to_f64(hipStream_t s, const float* a, double* b, int n);
void launch_f64_to_f32(hipStream_t s, const double* a, float* b, int n);
void launch_pack_posm(hipStream_t s, const float* x, const float* y, const float* z, const float* m, float4* posm,
                      int n);
// FP64 positions and float masses into (x, y, z, m) records.
void launch_pack_posd(hipStream_t s, const double* x, const double* y, const double* z, const float* m, double4* posd,
                      int n);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, long long n);
void launch_zero_f32(hipStream_t s, float* p, long long n);
// Bounding box of (x, y, z) into cube (float lo/hi, and the Morton cube in double).
// part: kReduceBlocks * 6 values of scratch (float or double as the positions).
void launch_bbox(hipStream_t s, const float* x, const float* y, const float* z, int n, void* part, Cube* cube);
void launch_bbox(hipStream_t s, const double* x, const double* y, const double* z, int n, void* part, Cube* cube);
// FP64 reductions: mass, ke, sum m*phi, p, L, sum m*x. part: kReduceBlocks * 12 doubles, out: 12 doubles.
void launch_diag(hipStream_t s, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                 const float* vz, const float* m, const float* pot, int use_pot, int n, double* part, double* out);
void launch_diag(hipStream_t s, const double* x, const double* y, const double* z, const double* vx,
                 const double* vy, const double* vz, const float* m, const double* pot, int use_pot, int n,
                 double* part, double* out);
// Exact potential energy sum_i m_i sum_{j != i} m_j / s_ij in double. part: grid doubles; out: 1 double.
void launch_exact_pe(hipStream_t s, const float4* posm, int n, double eps2, double* part, double* out);
void launch_exact_pe(hipStream_t s, const double* x, const double* y, const double* z, const float* m, int n,
                     double eps2, double* part, double* out);

// ---- direct summation (kernels_direct.hip) ----
// Accelerations G sum_j m_j d / (d^2 + eps^2)^1.5 (and phi = -G sum m_j / s), self skipped by index.
// part: scratch of direct_split(n) * n float4 when the split is > 1.
int direct_split(int n);
void launch_direct(hipStream_t s, const float4* posm, int n, float eps2, float G, bool potential, float* ax, float* ay,
                   float* az, float* pot, float4* part);

// FP64 direct sum for nt target indices (validation): a = G sum m_j d / s^3, phi = -G sum m_j / s,
// double pair arithmetic on the current positions (float, or double when f64). part: scratch of
// direct_f64_split(n, nt) * 4 * nt doubles.
int direct_f64_split(int n, int nt);
void launch_direct_f64_targets(hipStream_t s, const void* x, const void* y, const void* z, bool f64,
                               const float* m, int n, const uint32_t* tgt, int nt, double eps, double G, double* ax,
                               double* ay, double* az, double* pot, double* part);

// ---- tree (kernels_tree.hip) ----
constexpr int kRadix = 256;      // 8-bit digits
constexpr int kRadixDigits = 8;  // digits of a 64-bit key
// Work memory of radix_sort, in device words: the tile counts of one digit, [256][tiles], then the
// digit totals [256].
struct RadixWork {
    uint32_t* words = nullptr;
};
// Device words of radix_sort's work memory for up to n elements.
std::size_t radix_work_words(int n);
// Stable LSD radix sort of n (key, value) pairs by the 8-bit digits d0 <= d < d1 of the keys (bits
// [8 d, 8 d + 8)): per digit a count, a scan and a scatter launch. The input is (k[0], v[0]), or
// k[0] with the values 0, 1, ..., n - 1 when iota_values (v[0] is then not read, but later passes
// still write it). Pass j reads buffer j & 1 and writes the other, so the result is in buffer
// (d1 - d0) & 1. n_dev (optional) holds the element count on the device, n is then the capacity.
// first_counted: the kernel that made the keys already wrote digit d0's tile counts (the body
// sort's keys kernel). The last pass can also write inv[value] = sorted position for every element
// and its digit starts (exclusive sums of the digit totals) to starts[0 .. nstarts), nstarts <= 256;
// with starts_stride s > 1, only the starts of digits 0, s, 2s, ... < nstarts, to starts[d / s].
// Returns the digit totals of the last pass, valid until the next sort on the same work memory.
struct RadixBufs {
    uint64_t* k[2];
    uint32_t* v[2];
};
const uint32_t* radix_sort(hipStream_t s, RadixWork& w, RadixBufs b, bool iota_values, int n,
                           const uint32_t* n_dev, int d0, int d1, bool first_counted = false,
                           uint32_t* inv = nullptr, uint32_t* starts = nullptr, int nstarts = 0,
                           int starts_stride = 1);

struct TreeBufs {
    bool f64;  // positions and velocities are double (the mixed form), else float
    // bodies (sorted order after the permute)
    void *x, *y, *z, *vx, *vy, *vz;
    float* m;
    uint32_t* id;
    float* cost;  // per-body cost of the last walk (permuted with the bodies), may be null
    // alternates for the permute (same sizes)
    void *x2, *y2, *z2, *vx2, *vy2, *vz2;
    float* m2;
    uint32_t* id2;
    float* cost2;
    float4* posm;  // float (position, mass) in sorted order, may be null
    double4* posd; // FP64 (position, mass) in sorted order (mixed form), may be null
    uint64_t *key, *key2;
    uint32_t *idx, *idx2;
    RadixWork* radix;
    Cube* cube;
    void* bbox_part;      // kReduceBlocks * 6 values of the position type
};

// Morton keys, stable radix sort, permute of every body array into key order (swaps the caller's
// pointers: after the call b.x etc. hold the sorted arrays). Keys stay in b.key. top_digits < 8
// (odd): radix passes over the top digits only, then each run of equal top digits ordered by its low
// bits (the same stable order; fast when the runs are short).
void tree_sort(hipStream_t s, TreeBufs& b, int n, bool permute_cost, int top_digits = kRadixDigits);
// Every body array into the order perm gives (new i <- old perm[i]); swaps like tree_sort.
void tree_permute(hipStream_t s, TreeBufs& b, const uint32_t* perm, int n, bool permute_cost);
// Keys of FP64 positions already in key order (a recorded sorted state): bounding cube and keys,
// no sort (idx = identity).
void tree_keys_sorted(hipStream_t s, const double* x, const double* y, const double* z, int n, void* bbox_part,
                      Cube* cube, uint64_t* key, uint32_t* idx);

}  // namespace lhip
