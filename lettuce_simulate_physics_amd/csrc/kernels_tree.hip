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
#include "wave.hpp"

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
// Per digit: (1) count: each tile of 2048 keys counts its digits; (2) one block per digit scans
// that digit's counts over the tiles (exclusive, in place) and records the digit's total; (3)
// scatter: each tile ranks its keys stably in LDS and writes runs of equal digits to consecutive
// global addresses. The body sort's keys kernel counts digit 0 of its tiles as it makes the keys,
// so the first pass has no count launch.
//
// Tile layout: wave w of a block holds keys [512 w, 512 w + 512) of the tile, lane l of round r
// the key 512 w + 64 r + l, so each load is 64 consecutive keys and the tile order is (wave,
// round, lane). A wave ranks its own keys round by round with a running count per digit in its own
// LDS row (lanes of a round with the same digit are found by matching the digit's 8 bits with
// ballots), so the rounds need no block barrier; two block scans over the 256 digits (wave scans,
// then the wave totals) place each wave's run of every digit.

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortKPT = 8;                           // keys per thread per tile
constexpr int kSortTile = kSortThreads * kSortKPT;    // 2048 keys per tile
__host__ __device__ inline int sort_tiles(int n) { return int((static_cast<long long>(n) + kSortTile - 1) / kSortTile); }

__device__ __forceinline__ int tile_slot(int wave, int r, int lane) { return wave * (kSortKPT * 64) + r * 64 + lane; }

// Lanes of the wave whose digit equals this lane's (8 ballots on its bits).
__device__ __forceinline__ uint64_t digit_peers(int d) {
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint64_t m = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? m : ~m;
    }
    return peers;
}

// Stable ranks of a wave's keys by digit: pos[r] is the rank of the lane's round-r key among the
// wave's keys of the same digit in wave order, and wc[d] ends as the wave's count of digit d. wc
// is the wave's own LDS row, zeroed before; it is read and written by this wave only, in order.
__device__ __forceinline__ void wave_rank(const int (&dg)[kSortKPT], uint32_t (&pos)[kSortKPT], uint32_t* wc,
                                          int lane) {
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const uint64_t peers = digit_peers(dg[r]);
        const uint32_t rank = __popcll(peers & lt);
        const uint32_t before = wc[dg[r]];
        pos[r] = before + rank;
        if (rank == 0) wc[dg[r]] = before + uint32_t(__popcll(peers));
    }
}

// n_dev (optional) holds the element count on the device; n is then only the capacity the grid
// was sized for.
// The digit being counted is effectively random in the order a pass reads (the keys are sorted by
// the lower digits by then, or near their previous order for the first), so LDS atomics on one
// histogram rarely collide within a wave.
__global__ void __launch_bounds__(kSortThreads) radix_count(const uint64_t* __restrict__ key, int n,
                                                            const uint32_t* __restrict__ n_dev, int shift, int ntiles,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRadix];
    const int tid = int(threadIdx.x);
    if (n_dev) n = min(n, int(*n_dev));
    const int base = blockIdx.x * kSortTile;
    if (base >= n) return;  // a tile past the elements (n_dev below the capacity): never scanned
    uint64_t k[kSortKPT];
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int i = base + r * kSortThreads + tid;
        k[r] = i < n ? key[i] : 0ull;
    }
    h[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r)
        if (base + r * kSortThreads + tid < n) atomicAdd(&h[(k[r] >> shift) & 0xff], 1u);
    __syncthreads();
    hist[std::size_t(tid) * ntiles + blockIdx.x] = h[tid];
}

// One block per digit: exclusive scan of hist[d][0..tiles) in place, over the tiles that hold
// elements (all ntiles, or those below n_dev's count); total into tot[d].
__global__ void __launch_bounds__(kSortThreads) radix_row_scan(uint32_t* __restrict__ hist, int ntiles,
                                                               uint32_t* __restrict__ tot, int n,
                                                               const uint32_t* __restrict__ n_dev) {
    __shared__ uint32_t wsum[2][kSortWaves];
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    uint32_t* row = hist + std::size_t(blockIdx.x) * ntiles;
    if (n_dev) ntiles = min(ntiles, sort_tiles(min(n, int(*n_dev))));
    constexpr int K = 4;  // consecutive entries per thread per chunk
    uint32_t carry = 0;
    int parity = 0;
    for (int b = 0; b < ntiles; b += kSortThreads * K) {
        uint32_t v[K], sum = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int t = b + tid * K + j;
            v[j] = t < ntiles ? row[t] : 0u;
            sum += v[j];
        }
        uint32_t chunk = 0;
        uint32_t run = carry + block_scan_excl<kSortWaves>(sum, wsum[parity], lane, wave, &chunk);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int t = b + tid * K + j;
            if (t < ntiles) row[t] = run;
            run += v[j];
        }
        carry += chunk;
        parity ^= 1;
    }
    if (tid == 0) tot[blockIdx.x] = carry;
}

// vin null: the values are the input positions. inv (optional): inv[value] = its sorted position;
// starts (optional): block 0 writes the starts of digits 0, ss, 2 ss, ... < nstarts to starts[d / ss].
__global__ void __launch_bounds__(kSortThreads) radix_scatter(const uint64_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              int n, const uint32_t* __restrict__ n_dev, int shift,
                                                              int ntiles, const uint32_t* __restrict__ hist,
                                                              const uint32_t* __restrict__ tot,
                                                              uint32_t* __restrict__ inv, uint32_t* __restrict__ starts,
                                                              int nstarts, int ss) {
    __shared__ uint64_t sk[kSortTile];
    __shared__ uint32_t sv[kSortTile];
    __shared__ uint32_t wc[kSortWaves][kRadix];   // the waves' digit counts, then their tile offsets
    __shared__ uint32_t dstart[kRadix];           // tile-local start of each digit
    __shared__ uint32_t gbase[kRadix];            // global start of each digit for this tile
    __shared__ uint32_t wsum[2][kSortWaves];
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    if (n_dev) n = min(n, int(*n_dev));
    const int base = blockIdx.x * kSortTile;
    if (base >= n) return;  // nothing to place (the row scans skipped this tile)
    // All global loads of the block are issued before their first use.
    uint64_t k[kSortKPT];
    uint32_t v[kSortKPT], pos[kSortKPT];
    int dg[kSortKPT];
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int i = base + tile_slot(wave, r, lane);
        const bool ok = i < n;
        k[r] = ok ? kin[i] : ~0ull;
        v[r] = ok ? (vin ? vin[i] : uint32_t(i)) : 0u;
    }
    const uint32_t my_tot = tot[tid];
    const uint32_t my_row = hist[std::size_t(tid) * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) wc[w][tid] = 0;
    // Out-of-range slots take digit 255; they are last in tile order, so they rank after every real
    // key of the tile and land past the valid part of the tile-sorted LDS copy.
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r)
        dg[r] = base + tile_slot(wave, r, lane) < n ? int((k[r] >> shift) & 0xff) : 0xff;
    __syncthreads();
    wave_rank(dg, pos, wc[wave], lane);
    // Global digit starts: exclusive scan of the digit totals (its barrier also completes the ranks).
    const uint32_t dbase = block_scan_excl<kSortWaves>(my_tot, wsum[0], lane, wave, nullptr);
    if (starts && blockIdx.x == 0 && tid < nstarts && tid % ss == 0) starts[tid / ss] = dbase;
    uint32_t c[kSortWaves], all = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        c[w] = wc[w][tid];
        all += c[w];
    }
    const uint32_t ds = block_scan_excl<kSortWaves>(all, wsum[1], lane, wave, nullptr);
    {
        uint32_t o = ds;  // this thread's column of wc only: no other thread reads it before the barrier
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            wc[w][tid] = o;
            o += c[w];
        }
    }
    dstart[tid] = ds;
    gbase[tid] = dbase + my_row;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const uint32_t p = wc[wave][dg[r]] + pos[r];
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
            if (inv) inv[sv[p]] = dst;
        }
    }
}

// Morton keys of the body sort, tile by tile, with the digit at `shift` of each tile counted for its
// first pass (LDS atomics, as radix_count).
template <class R>
__global__ void __launch_bounds__(kSortThreads) keys_count_kernel(const R* __restrict__ x, const R* __restrict__ y,
                                                                  const R* __restrict__ z,
                                                                  const Cube* __restrict__ cube,
                                                                  uint64_t* __restrict__ key, int n, int ntiles,
                                                                  uint32_t* __restrict__ hist, int shift) {
    __shared__ uint32_t h[kRadix];
    const int tid = int(threadIdx.x);
    const int base = blockIdx.x * kSortTile;
    const double ox = cube->ox, oy = cube->oy, oz = cube->oz, scale = cube->scale;
    R px[kSortKPT], py[kSortKPT], pz[kSortKPT];
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int j = min(base + r * kSortThreads + tid, n - 1);
        px[r] = x[j]; py[r] = y[j]; pz[r] = z[j];
    }
    h[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortKPT; ++r) {
        const int i = base + r * kSortThreads + tid;
        if (i < n) {
            const uint64_t kk = morton_key(px[r], py[r], pz[r], ox, oy, oz, scale);
            key[i] = kk;
            atomicAdd(&h[(kk >> shift) & 0xff], 1u);
        }
    }
    __syncthreads();
    hist[std::size_t(tid) * ntiles + blockIdx.x] = h[tid];
}

// ---------------------------------------------------------------- runs of equal top digits
//
// A sort by the top digits only (the passes d0 .. 7) leaves the keys ordered by key >> 8 d0, each
// run of equal top digits in the input order (the passes are stable). Ordering every run by the low
// bits, ties by position, then gives the full stable sort. run_fixup does it out of place, one
// element per thread: a singleton is copied; an element of a run of at most kShortRun finds the
// run's bounds and its rank in it by scanning it; a longer run's first element records the run's
// start for run_sort_long and the run's elements are left to it.

constexpr int kShortRun = 64;
constexpr int kLongRunBlocks = 64;  // blocks that take the long runs in turn

__global__ void __launch_bounds__(kBS) run_fixup(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                 uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                                                 int shift, uint32_t* __restrict__ long_runs, int long_cap) {
    const int e = blockIdx.x * kBS + threadIdx.x;
    if (e >= n) return;
    const uint64_t k = kin[e];
    const uint64_t hi = k >> shift;
    const bool prev_same = e > 0 && (kin[e - 1] >> shift) == hi;
    const bool next_same = e + 1 < n && (kin[e + 1] >> shift) == hi;
    if (!prev_same && !next_same) {
        kout[e] = k;
        vout[e] = vin[e];
        return;
    }
    // The run's start, at most kShortRun back; further back it is a long run.
    int s = e;
    while (s > 0 && e - s < kShortRun && (kin[s - 1] >> shift) == hi) --s;
    if (s > 0 && (kin[s - 1] >> shift) == hi) return;  // a long run (its start is further back)
    int t = e + 1;
    while (t < n && t - s <= kShortRun && (kin[t] >> shift) == hi) ++t;
    if (t - s > kShortRun) {
        // A long run: its first element records it (long_runs[0] counts, starts from [1]).
        if (e == s) {
            const uint32_t slot = atomicAdd(&long_runs[0], 1u);
            if (int(slot) < long_cap) long_runs[1 + slot] = uint32_t(s);
        }
        return;
    }
    const uint64_t lmask = (1ull << shift) - 1ull;
    const uint64_t lo = k & lmask;
    int rank = 0;
    for (int f = s; f < t; ++f) {
        const uint64_t lf = kin[f] & lmask;
        rank += (lf < lo || (lf == lo && f < e)) ? 1 : 0;
    }
    kout[s + rank] = k;
    vout[s + rank] = vin[e];
}

// Long runs (more than kShortRun equal top digits; none unless bodies crowd a cell far beyond the
// previous tree's leaves): each block takes recorded runs in turn and sorts its run stably by the
// low digits 0 .. d0 - 1 with a block-wide LSD counting sort, passes alternating between the
// input and the output arrays over the run's positions (d0 odd: the last pass writes the output).
__global__ void __launch_bounds__(kSortThreads) run_sort_long(uint64_t* __restrict__ kin, uint32_t* __restrict__ vin,
                                                              uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              int n, int d0, const uint32_t* __restrict__ long_runs,
                                                              int long_cap) {
    __shared__ uint32_t hist[kRadix];
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wc[kSortWaves][kRadix];
    __shared__ uint32_t wsum[2][kSortWaves];
    __shared__ int run_end;
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int runs = min(int(long_runs[0]), long_cap);
    const int shift = 8 * d0;
    for (int r = blockIdx.x; r < runs; r += gridDim.x) {
        const int s = int(long_runs[1 + r]);
        const uint64_t hi = kin[s] >> shift;
        // The run's end: the first position past s with other top digits.
        if (tid == 0) run_end = n;
        __syncthreads();
        for (int c = s; c < n; c += kSortThreads) {
            const int i = c + tid;
            const bool other = i < n && (kin[i] >> shift) != hi;
            const unsigned long long m = __ballot(other);
            if (m && lane == int(__ffsll((long long)m) - 1)) atomicMin(&run_end, i);
            __syncthreads();
            const bool done = run_end < n;
            __syncthreads();
            if (done) break;
        }
        const int t = run_end;
        for (int p = 0; p < d0; ++p) {
            uint64_t* ks = (p & 1) ? kout : kin;
            uint32_t* vs = (p & 1) ? vout : vin;
            uint64_t* kd = (p & 1) ? kin : kout;
            uint32_t* vd = (p & 1) ? vin : vout;
            const int sh = 8 * p;
            hist[tid] = 0;
            __syncthreads();
            for (int i = s + tid; i < t; i += kSortThreads) atomicAdd(&hist[(ks[i] >> sh) & 0xff], 1u);
            __syncthreads();
            const uint32_t b = block_scan_excl<kSortWaves>(hist[tid], wsum[0], lane, wave, nullptr);
            base[tid] = b;  // running start of each digit (advanced chunk by chunk)
            __syncthreads();
            for (int c = s; c < t; c += kSortThreads) {
                const int i = c + tid;
                const bool ok = i < t;
                const uint64_t kk = ok ? ks[i] : 0ull;
                const uint32_t vv = ok ? vs[i] : 0u;
                const int d = ok ? int((kk >> sh) & 0xff) : -1;
#pragma unroll
                for (int w = 0; w < kSortWaves; ++w) wc[w][tid] = 0;
                __syncthreads();
                // Rank among the wave's lanes of the same digit; the wave's count of each digit.
                uint64_t peers = ok ? ~0ull : 0ull;
#pragma unroll
                for (int bit = 0; bit < 8; ++bit) {
                    const uint64_t m = __ballot(ok && ((d >> bit) & 1));
                    peers &= ((d >> bit) & 1) ? m : ~m;
                }
                peers &= __ballot(ok);
                const uint32_t rank = ok ? uint32_t(__popcll(peers & ((1ull << lane) - 1ull))) : 0u;
                if (ok && rank == 0) wc[wave][d] = uint32_t(__popcll(peers));
                __syncthreads();
                if (ok) {
                    uint32_t off = base[d] + rank;
                    for (int w = 0; w < wave; ++w) off += wc[w][d];
                    kd[s + int(off)] = kk;
                    vd[s + int(off)] = vv;
                }
                __syncthreads();
                uint32_t add = 0;
#pragma unroll
                for (int w = 0; w < kSortWaves; ++w) add += wc[w][tid];
                base[tid] += add;
                __syncthreads();
            }
        }
        __syncthreads();
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
void sort_impl(hipStream_t s, TreeBufs& b, int n, bool permute_cost, int top) {
    const R *x = static_cast<const R*>(b.x), *y = static_cast<const R*>(b.y), *z = static_cast<const R*>(b.z);
    launch_bbox(s, x, y, z, n, b.bbox_part, b.cube);
    const int ntiles = sort_tiles(n);
    const int d0 = kRadixDigits - top;
    keys_count_kernel<R><<<ntiles, kSortThreads, 0, s>>>(x, y, z, b.cube, b.key, n, ntiles, b.radix->words, 8 * d0);
    LHIP_LAUNCH_CHECK();
    // Passes d0 .. 7. All eight: the sorted keys and permutation end in key/idx. The top (odd) ones:
    // they end in key2/idx2, and the run fix-up writes the full order back to key/idx.
    radix_sort(s, *b.radix, RadixBufs{{b.key, b.key2}, {b.idx, b.idx2}}, true, n, nullptr, d0, kRadixDigits, true);
    if (d0 > 0) {
        // The radix work memory is free until the next sort: it holds the long runs' starts.
        uint32_t* long_runs = b.radix->words;
        const int long_cap = int(std::min<std::size_t>(radix_work_words(n) - 1, std::size_t(n / kShortRun + 1)));
        launch_fill_u32(s, long_runs, 0u, 1);
        run_fixup<<<grid_for(n, kBS), kBS, 0, s>>>(b.key2, b.idx2, b.key, b.idx, n, 8 * d0, long_runs, long_cap);
        LHIP_LAUNCH_CHECK();
        run_sort_long<<<kLongRunBlocks, kSortThreads, 0, s>>>(b.key2, b.idx2, b.key, b.idx, n, d0, long_runs,
                                                             long_cap);
        LHIP_LAUNCH_CHECK();
    }
    auto P = [](void* p) { return static_cast<R*>(p); };
    permute_kernel<R><<<grid_for(n, kBS), kBS, 0, s>>>(b.idx, n, x, y, z, P(b.vx), P(b.vy), P(b.vz), b.m, b.id,
                                                      permute_cost ? b.cost : nullptr, P(b.x2), P(b.y2), P(b.z2),
                                                      P(b.vx2), P(b.vy2), P(b.vz2), b.m2, b.id2,
                                                      permute_cost ? b.cost2 : nullptr, b.posm, b.posd);
    LHIP_LAUNCH_CHECK();
}

}  // namespace

std::size_t radix_work_words(int n) { return std::size_t(std::max(sort_tiles(n), 1)) * kRadix + kRadix; }

const uint32_t* radix_sort(hipStream_t s, RadixWork& w, RadixBufs b, bool iota_values, int n, const uint32_t* n_dev,
                           int d0, int d1, bool first_counted, uint32_t* inv, uint32_t* starts, int nstarts,
                           int starts_stride) {
    if (d0 < 0 || d1 > kRadixDigits || d0 > d1 || nstarts < 0 || nstarts > kRadix || starts_stride < 1)
        throw Error(kInvalid, "radix_sort: digits or starts out of range");
    const int ntiles = sort_tiles(n);
    uint32_t* hist = w.words;
    uint32_t* tot = w.words + std::size_t(ntiles) * kRadix;
    if (n <= 0) return tot;
    for (int d = d0; d < d1; ++d) {
        const int j = d - d0, src = j & 1, dst = src ^ 1;
        if (j > 0 || !first_counted) {
            radix_count<<<ntiles, kSortThreads, 0, s>>>(b.k[src], n, n_dev, 8 * d, ntiles, hist);
            LHIP_LAUNCH_CHECK();
        }
        radix_row_scan<<<kRadix, kSortThreads, 0, s>>>(hist, ntiles, tot, n, n_dev);
        LHIP_LAUNCH_CHECK();
        const bool last = d + 1 == d1;
        radix_scatter<<<ntiles, kSortThreads, 0, s>>>(b.k[src], j == 0 && iota_values ? nullptr : b.v[src], b.k[dst],
                                                      b.v[dst], n, n_dev, 8 * d, ntiles, hist, tot,
                                                      last ? inv : nullptr, last ? starts : nullptr, nstarts,
                                                      starts_stride);
        LHIP_LAUNCH_CHECK();
    }
    return tot;
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

void tree_sort(hipStream_t s, TreeBufs& b, int n, bool permute_cost, int top_digits) {
    if (n <= 0) return;
    if (top_digits != kRadixDigits && (top_digits < 1 || top_digits > kRadixDigits || top_digits % 2 == 0))
        throw Error(kInvalid, "tree_sort: top digits must be 8 or odd");
    if (b.f64) sort_impl<double>(s, b, n, permute_cost, top_digits);
    else sort_impl<float>(s, b, n, permute_cost, top_digits);
    std::swap(b.x, b.x2); std::swap(b.y, b.y2); std::swap(b.z, b.z2);
    std::swap(b.vx, b.vx2); std::swap(b.vy, b.vy2); std::swap(b.vz, b.vz2);
    std::swap(b.m, b.m2); std::swap(b.id, b.id2);
    if (permute_cost) std::swap(b.cost, b.cost2);
}

}  // namespace lhip
