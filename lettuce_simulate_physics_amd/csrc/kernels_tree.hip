// Barnes-Hut tree on the device, first stage: Morton keys, the stable LSD radix sort and the permute.
// The octree itself (pre-order enumeration, level order, multipoles, walk groups) is built in
// kernels_build.hip and walked in kernels_walk.hip; the overview below covers all three.
//
// Reference algorithm: CpuBackend::sort_bodies / make_keys / radix_sort / permute / build_node /
// build_octree / build_list / eval_group (src/cpu_backend.cpp:395-882). The device form differs in
// how the same octree is found and stored:
//
//  * Leaf rule by windows. The cell at level L containing sorted body i holds more than W bodies iff
//    some window [j, j+W] containing i has keys sharing their top 3L bits. So the leaf level of i is
//    1 + the largest common-octal-prefix length over those windows (W = 16), and the walk-group level
//    is the same with W = 128 (a group is the topmost cell with <= 128 bodies, or a leaf).
//  * Pre-order numbering. Every octree cell is identified by (first body, level). Ordering cells by
//    (first body, level) is a depth-first pre-order, so body i starts the cells of levels
//    (c_i, L_i], where c_i is the common prefix length of keys i-1 and i. A scan of those counts
//    numbers every node; a node's subtree is the index range [k, skip) with skip = the first node of
//    the body just past the cell, and its first child (if any) is k + 1.
//  * Multipoles level by level, deepest first, each level one launch (kernel boundaries order the
//    dependent levels without cross-XCD cache maintenance inside a kernel).
//  * Walk: one wavefront per group, a parallel depth-first traversal with an LDS stack
//    (kernels_walk.hip).
//
// Positions are float (the plain state, bit-exact keys with the reference) or double (the
// mixed-precision form); the key recipe is the same.
#include <cfloat>

#include "kernels.hpp"
#include "morton.hpp"
#include "tree.hpp"

namespace lhip {
namespace {

// ---------------------------------------------------------------- keys

template <class R>
__global__ void __launch_bounds__(kBS) keys_kernel(const R* __restrict__ x, const R* __restrict__ y,
                                                   const R* __restrict__ z, const Cube* __restrict__ cube,
                                                   uint64_t* __restrict__ key, uint32_t* __restrict__ idx, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    key[i] = morton_key(x[i], y[i], z[i], cube->ox, cube->oy, cube->oz, cube->scale);
    idx[i] = uint32_t(i);
}

// ---------------------------------------------------------------- radix sort (stable LSD, 8-bit digits)
//
// Per pass: (1) tile histograms, (2) one block per digit scans that digit's row over the tiles,
// (3) scatter: each tile ranks its keys stably in LDS (ballot matching per wave, running per-digit
// counts per round), stores them tile-sorted in LDS, then writes runs of equal digits to
// consecutive global addresses.

constexpr int kSortThreads = 256;
constexpr int kSortKPT = 8;                           // keys per thread per tile
constexpr int kSortTile = kSortThreads * kSortKPT;    // 2048 keys per tile
constexpr int kRadix = 256;

// n_dev (optional) holds the element count on the device; n is then only the capacity the grid
// was sized for.
__global__ void __launch_bounds__(kSortThreads) radix_count(const uint64_t* __restrict__ key, int n,
                                                            const uint32_t* __restrict__ n_dev, int shift, int ntiles,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRadix];
    if (n_dev) n = min(n, int(*n_dev));
    const int base = blockIdx.x * kSortTile;
    if (base >= n) return;  // a tile past the elements (n_dev below the capacity): never scanned
    // All of the tile's loads are issued before the first use, so their latencies overlap.
    uint64_t kk[kSortKPT];
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int i = base + r * kSortThreads + threadIdx.x;
        kk[r] = i < n ? key[i] : 0ull;
    }
    h[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(kk[r] >> shift) & 0xff], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan of hist[d][0..tiles) in place, over the tiles that hold
// elements (all ntiles, or those below n_dev's count); total into tot[d].
__global__ void __launch_bounds__(kSortThreads) radix_row_scan(uint32_t* __restrict__ hist, int ntiles,
                                                               uint32_t* __restrict__ tot, int n,
                                                               const uint32_t* __restrict__ n_dev) {
    __shared__ uint32_t sh[kSortThreads];
    uint32_t* row = hist + blockIdx.x * ntiles;
    if (n_dev) ntiles = min(ntiles, int((min(n, int(*n_dev)) + kSortTile - 1) / kSortTile));
    uint32_t carry = 0;
    for (int b = 0; b < ntiles; b += kSortThreads) {
        const int t = b + threadIdx.x;
        const uint32_t v = t < ntiles ? row[t] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        // Hillis-Steele inclusive scan over 256 entries.
        for (int o = 1; o < kSortThreads; o <<= 1) {
            const uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        if (t < ntiles) row[t] = carry + sh[threadIdx.x] - v;
        carry += sh[kSortThreads - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kSortThreads) radix_scatter(const uint64_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              int n, const uint32_t* __restrict__ n_dev, int shift,
                                                              int ntiles, const uint32_t* __restrict__ hist,
                                                              const uint32_t* __restrict__ tot) {
    __shared__ uint64_t sk[kSortTile];
    __shared__ uint32_t sv[kSortTile];
    __shared__ uint32_t cnt[kSortThreads / 64][kRadix];
    __shared__ uint32_t pre[kSortThreads / 64][kRadix];
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t dstart[kRadix];   // tile-local start of each digit
    __shared__ uint32_t gbase[kRadix];    // global start of each digit for this tile
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int base = blockIdx.x * kSortTile;
    if (n_dev) n = min(n, int(*n_dev));
    if (base >= n) return;  // nothing to place (the row scans skipped this tile)
    // All global loads of the block are issued here, before the first barrier, so that their
    // latencies overlap each other and the digit-base scan.
    uint64_t k[kSortKPT];
    uint32_t v[kSortKPT], pos[kSortKPT];
    int dg[kSortKPT];
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int i = base + r * kSortThreads + tid;
        const bool ok = i < n;
        k[r] = ok ? kin[i] : ~0ull;
        v[r] = ok ? vin[i] : 0u;
    }
    const uint32_t my_tot = tot[tid];
    const uint32_t my_hist = hist[tid * ntiles + blockIdx.x];
    run[tid] = 0;
    for (int w = 0; w < kSortThreads / 64; ++w) cnt[w][tid] = 0;
    // Digit base = exclusive scan of the digit totals (block scan of 256 values).
    {
        dstart[tid] = my_tot;
        __syncthreads();
        for (int o = 1; o < kRadix; o <<= 1) {
            const uint32_t a = tid >= o ? dstart[tid - o] : 0u;
            __syncthreads();
            dstart[tid] += a;
            __syncthreads();
        }
        gbase[tid] = dstart[tid] - my_tot + my_hist;
        __syncthreads();
    }
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const bool ok = base + r * kSortThreads + tid < n;
        // Out-of-range slots take digit 255 and sort after every real key of the tile (they are last
        // in tile order), so real keys keep their ranks.
        dg[r] = ok ? int((k[r] >> shift) & 0xff) : 0xff;
        // Lanes of this wave with the same digit.
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((dg[r] >> b) & 1);
            peers &= ((dg[r] >> b) & 1) ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt);
        if (rank == 0) cnt[wave][dg[r]] = __popcll(peers);
        __syncthreads();
        {
            uint32_t o = run[tid];
#pragma unroll
            for (int w = 0; w < kSortThreads / 64; ++w) {
                pre[w][tid] = o;
                o += cnt[w][tid];
                cnt[w][tid] = 0;
            }
            run[tid] = o;
        }
        __syncthreads();
        pos[r] = pre[wave][dg[r]] + rank;
        __syncthreads();
    }
    // Tile-local digit starts: exclusive scan of run[].
    {
        const uint32_t mine = run[tid];
        __syncthreads();
        dstart[tid] = mine;
        __syncthreads();
        for (int o = 1; o < kRadix; o <<= 1) {
            const uint32_t a = tid >= o ? dstart[tid - o] : 0u;
            __syncthreads();
            dstart[tid] += a;
            __syncthreads();
        }
        dstart[tid] -= mine;
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const uint32_t p = dstart[dg[r]] + pos[r];
        sk[p] = k[r];
        sv[p] = v[r];
    }
    __syncthreads();
    const int valid = max(0, min(kSortTile, n - base));
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int p = r * kSortThreads + tid;
        if (p < valid) {
            const uint64_t kk = sk[p];
            const int d = int((kk >> shift) & 0xff);
            const uint32_t dst = gbase[d] + (uint32_t(p) - dstart[d]);
            kout[dst] = kk;
            vout[dst] = sv[p];
        }
    }
}

// ---------------------------------------------------------------- permute

template <class R>
__global__ void __launch_bounds__(kBS) permute_kernel(const uint32_t* __restrict__ perm, int n,
                                                      const R* __restrict__ x, const R* __restrict__ y,
                                                      const R* __restrict__ z, const R* __restrict__ vx,
                                                      const R* __restrict__ vy, const R* __restrict__ vz,
                                                      const float* __restrict__ m, const uint32_t* __restrict__ id,
                                                      const float* __restrict__ cost, R* __restrict__ x2,
                                                      R* __restrict__ y2, R* __restrict__ z2, R* __restrict__ vx2,
                                                      R* __restrict__ vy2, R* __restrict__ vz2,
                                                      float* __restrict__ m2, uint32_t* __restrict__ id2,
                                                      float* __restrict__ cost2, float4* __restrict__ posm,
                                                      double4* __restrict__ posd) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    const R px = x[j], py = y[j], pz = z[j];
    const float pm = m[j];
    x2[i] = px; y2[i] = py; z2[i] = pz; m2[i] = pm;
    vx2[i] = vx[j]; vy2[i] = vy[j]; vz2[i] = vz[j];
    id2[i] = id[j];
    if (cost) cost2[i] = cost[j];
    if (posm) posm[i] = make_float4(float(px), float(py), float(pz), pm);
    if (posd) posd[i] = make_double4(double(px), double(py), double(pz), double(pm));
}

template <class R>
void sort_impl(hipStream_t s, TreeBufs& b, int n, bool permute_cost) {
    const R *x = static_cast<const R*>(b.x), *y = static_cast<const R*>(b.y), *z = static_cast<const R*>(b.z);
    launch_bbox(s, x, y, z, n, b.bbox_part, b.cube);
    keys_kernel<R><<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, b.cube, b.key, b.idx, n);
    LHIP_LAUNCH_CHECK();
    uint64_t *ka = b.key, *kb = b.key2;
    uint32_t *va = b.idx, *vb = b.idx2;
    for (int shift = 0; shift < 63; shift += 8) {
        radix_pass(s, ka, va, kb, vb, n, nullptr, shift, b.tile_hist);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // Eight passes: the sorted keys and permutation are back in key/idx.
    auto P = [](void* p) { return static_cast<R*>(p); };
    permute_kernel<R><<<grid_for(n, kBS), kBS, 0, s>>>(va, n, x, y, z, P(b.vx), P(b.vy), P(b.vz), b.m, b.id,
                                                      permute_cost ? b.cost : nullptr, P(b.x2), P(b.y2), P(b.z2),
                                                      P(b.vx2), P(b.vy2), P(b.vz2), b.m2, b.id2,
                                                      permute_cost ? b.cost2 : nullptr, b.posm, b.posd);
    LHIP_LAUNCH_CHECK();
}

}  // namespace

void radix_pass(hipStream_t s, const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, int n,
                const uint32_t* n_dev, int shift, uint32_t* tile_hist) {
    const int ntiles = int((static_cast<long long>(n) + kSortTile - 1) / kSortTile);
    uint32_t* tot = tile_hist + static_cast<std::size_t>(ntiles) * kRadix;
    radix_count<<<ntiles, kSortThreads, 0, s>>>(kin, n, n_dev, shift, ntiles, tile_hist);
    LHIP_LAUNCH_CHECK();
    radix_row_scan<<<kRadix, kSortThreads, 0, s>>>(tile_hist, ntiles, tot, n, n_dev);
    LHIP_LAUNCH_CHECK();
    radix_scatter<<<ntiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, n_dev, shift, ntiles, tile_hist, tot);
    LHIP_LAUNCH_CHECK();
}

std::size_t radix_tile_hist_entries(int n) {
    const std::size_t ntiles = (static_cast<std::size_t>(n) + kSortTile - 1) / kSortTile;
    return ntiles * kRadix + kRadix;  // + digit totals
}

void tree_keys_sorted(hipStream_t s, const double* x, const double* y, const double* z, int n, void* bbox_part,
                      Cube* cube, uint64_t* key, uint32_t* idx) {
    if (n <= 0) return;
    launch_bbox(s, x, y, z, n, bbox_part, cube);
    keys_kernel<double><<<grid_for(n, kBS), kBS, 0, s>>>(x, y, z, cube, key, idx, n);
    LHIP_LAUNCH_CHECK();
}

void tree_permute(hipStream_t s, TreeBufs& b, const uint32_t* perm, int n, bool permute_cost) {
    if (n <= 0) return;
    auto run = [&](auto tag) {
        using R = decltype(tag);
        auto P = [](void* p) { return static_cast<R*>(p); };
        permute_kernel<R><<<grid_for(n, kBS), kBS, 0, s>>>(perm, n, P(b.x), P(b.y), P(b.z), P(b.vx), P(b.vy), P(b.vz),
                                                          b.m, b.id, permute_cost ? b.cost : nullptr, P(b.x2), P(b.y2),
                                                          P(b.z2), P(b.vx2), P(b.vy2), P(b.vz2), b.m2, b.id2,
                                                          permute_cost ? b.cost2 : nullptr, b.posm, b.posd);
        LHIP_LAUNCH_CHECK();
    };
    if (b.f64) run(double{});
    else run(float{});
    std::swap(b.x, b.x2); std::swap(b.y, b.y2); std::swap(b.z, b.z2);
    std::swap(b.vx, b.vx2); std::swap(b.vy, b.vy2); std::swap(b.vz, b.vz2);
    std::swap(b.m, b.m2); std::swap(b.id, b.id2);
    if (permute_cost) std::swap(b.cost, b.cost2);
}

void tree_sort(hipStream_t s, TreeBufs& b, int n, bool permute_cost) {
    if (n <= 0) return;
    if (b.f64) sort_impl<double>(s, b, n, permute_cost);
    else sort_impl<float>(s, b, n, permute_cost);
    std::swap(b.x, b.x2); std::swap(b.y, b.y2); std::swap(b.z, b.z2);
    std::swap(b.vx, b.vx2); std::swap(b.vy, b.vy2); std::swap(b.vz, b.vz2);
    std::swap(b.m, b.m2); std::swap(b.id, b.id2);
    if (permute_cost) std::swap(b.cost, b.cost2);
}

}  // namespace lhip
