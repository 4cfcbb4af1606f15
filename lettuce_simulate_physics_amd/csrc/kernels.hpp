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
    uint32_t* tile_hist;  // radix tile histograms
    Cube* cube;
    void* bbox_part;      // kReduceBlocks * 6 values of the position type
};

// Morton keys, stable radix sort, permute of every body array into key order (swaps the caller's
// pointers: after the call b.x etc. hold the sorted arrays). Keys stay in b.key.
void tree_sort(hipStream_t s, TreeBufs& b, int n, bool permute_cost);
// Every body array into the order perm gives (new i <- old perm[i]); swaps like tree_sort.
void tree_permute(hipStream_t s, TreeBufs& b, const uint32_t* perm, int n, bool permute_cost);
// Keys of FP64 positions already in key order (a recorded sorted state): bounding cube and keys,
// no sort (idx = identity).
void tree_keys_sorted(hipStream_t s, const double* x, const double* y, const double* z, int n, void* bbox_part,
                      Cube* cube, uint64_t* key, uint32_t* idx);
std::size_t radix_tile_hist_entries(int n);
// One stable 8-bit LSD pass over bits [shift, shift + 8) of n (key, value) pairs; with n_dev the
// count is read on the device and n is the capacity. The digit totals follow the tile histograms:
// tile_hist[ntiles * 256 + d], ntiles = ceil(n / 2048).
void radix_pass(hipStream_t s, const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, int n,
                const uint32_t* n_dev, int shift, uint32_t* tile_hist);
constexpr int kRadixTile = 2048;

}  // namespace lhip
