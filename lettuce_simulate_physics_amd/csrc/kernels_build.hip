// Octree build on sorted Morton keys: leaf and group levels by windows, pre-order node numbering by
// a scan, node records (range, skip pointer, geometric cell), and bottom-up multipoles by level.
//
// Reference: build_node / build_octree (src/cpu_backend.cpp:534-664). The octree is the same one
// (same leaf rule, same cells, single-child cells kept, same walk groups); only its numbering
// (depth-first pre-order instead of per-parent child blocks) and construction order differ.
#include <cfloat>

#include "kernels.hpp"
#include "morton.hpp"
#include "tree.hpp"
#include "wave.hpp"

namespace lhip {
namespace {

// Number of leading octal digits two 63-bit keys share (21 when equal).
__device__ __forceinline__ int common_octs(uint64_t a, uint64_t b) {
    return a == b ? 21 : (__clzll(static_cast<long long>(a ^ b)) - 1) / 3;
}

// Key frames of a build: frame 0 keys, frame f >= 1 keys at fkey[(f - 1) n + i].
struct Frames {
    const uint64_t* key;
    const uint64_t* fkey;
    int frames;
    int n;
};

// Leading octal digits sorted bodies a and b share over all key frames: a later frame counts only
// where the earlier ones agree entirely (both bodies then lie in the same restarted cell).
__device__ __forceinline__ int ext_common(const Frames& F, int a, int b) {
    int c = common_octs(F.key[a], F.key[b]);
    for (int f = 1; f < F.frames && c == kMaxLevel * f; ++f)
        c = kMaxLevel * f + common_octs(F.fkey[std::size_t(f - 1) * F.n + a], F.fkey[std::size_t(f - 1) * F.n + b]);
    return c;
}

// ---------------------------------------------------------------- state reset

__global__ void reset_build(WalkCounters* __restrict__ c) {
    if (threadIdx.x == 0) *c = WalkCounters{};
}

// ---------------------------------------------------------------- leaf and group levels

constexpr int kLvTile = 256;
constexpr int kHalo = kGroupMax;  // 128

// Per sorted body i: L = leaf level (1 + max common prefix over 17-body windows holding i, capped at
// 22 = "more than 16 identical keys"), G = group level (the same with (gmax + 1)-body windows, never
// deeper than the leaf; gmax = 128 is the reference's group), c = common prefix with body i - 1
// (-1 for i = 0).
// cnt[i] = number of octree cells that start at body i = levels (c, min(L, 21)].
// gflag[i] = 1 iff the group cell of body i starts at i.
// Per tile of kLvTile bodies it also writes what the scan's top step reduces: the sums of cnt and
// gflag over the tile (tsum[0][tile], tsum[1][tile]), its deepest leaf level (bdeep[tile]) and its
// count of bodies in over-full finest cells (bdeep[tiles + tile]).
__global__ void __launch_bounds__(kLvTile) levels_kernel(const uint64_t* __restrict__ key, int n, int gmax,
                                                         uint32_t* __restrict__ bl, uint32_t* __restrict__ cnt,
                                                         uint32_t* __restrict__ gflag, uint32_t* __restrict__ tsum0,
                                                         uint32_t* __restrict__ tsum1, int* __restrict__ host_over,
                                                         uint32_t* __restrict__ bdeep) {
    constexpr int kA = kHalo + kLvTile;  // window starts held per tile
    __shared__ uint64_t sk[kHalo + kLvTile + kHalo];
    // Window values as 1 + common octal prefix (0 where the window leaves the bodies), so that a
    // window maximum is directly 1 + the largest prefix, and then range maxima over them.
    __shared__ uint8_t d16[kA];
    __shared__ uint8_t d128[kA];
    const int tid = threadIdx.x;
    const int base = blockIdx.x * kLvTile;
    {
        // Both loads of the tile are issued before either is stored, so their latencies overlap.
        static_assert(kHalo + kLvTile + kHalo == 2 * kLvTile, "the key tile is two loads per thread");
        uint64_t kv[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const long long j = (long long)base - kHalo + tid + r * kLvTile;
            kv[r] = (j >= 0 && j < n) ? key[j] : 0ull;
        }
        sk[tid] = kv[0];
        sk[tid + kLvTile] = kv[1];
    }
    __syncthreads();
    for (int t = tid; t < kA; t += kLvTile) {
        const long long j = (long long)base - kHalo + t;
        d16[t] = (j >= 0 && j + kLeafMax < n) ? uint8_t(1 + common_octs(sk[t], sk[t + kLeafMax])) : uint8_t(0);
        d128[t] = (j >= 0 && j + gmax < n) ? uint8_t(1 + common_octs(sk[t], sk[t + gmax])) : uint8_t(0);
    }
    // Range maxima by doubling, in place: after step s, entry t holds the maximum over
    // [t, t + 2^s). The leaf windows (17 starts) need 16-wide ranges, the group windows (gmax + 1
    // starts) the widest power of two not above gmax + 1. A query only reads ranges inside the
    // array, so the entries whose range runs past its end (read clamped) are never used.
    const int k128 = 31 - __builtin_clz(unsigned(gmax + 1));
    for (int step = 0; step < max(4, k128); ++step) {
        const int h = 1 << step;
        uint8_t a16[2], b16[2], a128[2], b128[2];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int t = tid + r * kLvTile;
            if (t < kA) {
                const int u = min(t + h, kA - 1);
                a16[r] = d16[t]; b16[r] = d16[u];
                a128[r] = d128[t]; b128[r] = d128[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int t = tid + r * kLvTile;
            if (t < kA) {
                if (step < 4) d16[t] = max(a16[r], b16[r]);
                if (step < k128) d128[t] = max(a128[r], b128[r]);
            }
        }
    }
    __syncthreads();
    const int i = base + tid;
    bool over = false;
    int deepest = 0;
    uint32_t ci = 0, gi = 0;
    if (i < n) {
        // Leaf window starts [i - 16, i]: two overlapping 16-wide ranges.
        const int a = tid + kHalo - kLeafMax;
        const int L = max(int(d16[a]), int(d16[a + 1]));
        // Group window starts [i - gmax, i]: two overlapping 2^k128-wide ranges.
        const int ag = tid + kHalo - gmax;
        const int Gw = max(int(d128[ag]), int(d128[ag + gmax + 1 - (1 << k128)]));
        const int Lc = min(L, kMaxLevel);
        const int Gc = min(Gw, Lc);
        const int c = i == 0 ? -1 : common_octs(sk[tid + kHalo - 1], sk[tid + kHalo]);
        bl[i] = uint32_t(L) | uint32_t(Gc) << 8 | uint32_t(c + 1) << 16;
        ci = uint32_t(max(0, Lc - c));
        gi = c < Gc ? 1u : 0u;
        cnt[i] = ci;
        gflag[i] = gi;
        over = L > kMaxLevel;
        deepest = Lc;
    }
    const unsigned long long ov = __ballot(over);
    if ((tid & 63) == 0 && ov && host_over) __hip_atomic_store(host_over, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    // Tile reductions: wave sums and maxima, then one plain store each per tile (no shared counter
    // word that every wave would wait on); the scan's top step reduces the tiles.
    __shared__ uint32_t wred[4][kLvTile / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        deepest = max(deepest, __shfl_xor(deepest, o, 64));
        ci += __shfl_xor(ci, o, 64);
        gi += __shfl_xor(gi, o, 64);
    }
    if ((tid & 63) == 0) {
        const int w = tid >> 6;
        wred[0][w] = uint32_t(deepest);
        wred[1][w] = ci;
        wred[2][w] = gi;
        wred[3][w] = uint32_t(__popcll(ov));
    }
    __syncthreads();
    if (tid < 4) {
        uint32_t r = wred[tid][0];
#pragma unroll
        for (int w = 1; w < kLvTile / 64; ++w) r = tid == 0 ? max(r, wred[tid][w]) : r + wred[tid][w];
        if (tid == 0) bdeep[blockIdx.x] = r;
        if (tid == 1) tsum0[blockIdx.x] = r;
        if (tid == 2) tsum1[blockIdx.x] = r;
        if (tid == 3) bdeep[gridDim.x + blockIdx.x] = r;
    }
}

// The level pass over key frames (the restart path): the same windows on the extended keys, each
// body's levels capped at its deepest frame (cap[i]). Bodies whose deepest cell still holds more
// than 16 bodies are counted (coincident bodies, or restarts beyond the last frame).
__global__ void __launch_bounds__(kBS) levels_ext_kernel(Frames F, const uint32_t* __restrict__ cap, int gmax,
                                                         uint32_t* __restrict__ bl, uint32_t* __restrict__ cnt,
                                                         uint32_t* __restrict__ gflag, WalkCounters* __restrict__ ctr,
                                                         int* __restrict__ host_over) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    bool over = false;
    if (i < F.n) {
        const int n = F.n;
        int best = -1;
        for (int j = max(0, i - kLeafMax); j <= i && j + kLeafMax < n; ++j) best = max(best, ext_common(F, j, j + kLeafMax));
        const int L = best + 1;
        best = -1;
        for (int j = max(0, i - gmax); j <= i && j + gmax < n; ++j) best = max(best, ext_common(F, j, j + gmax));
        const int Lc = min(L, int(cap[i]));
        const int Gc = min(best + 1, Lc);
        const int c = i == 0 ? -1 : ext_common(F, i - 1, i);
        bl[i] = uint32_t(L) | uint32_t(Gc) << 8 | uint32_t(c + 1) << 16;
        cnt[i] = uint32_t(max(0, Lc - c));
        gflag[i] = c < Gc ? 1u : 0u;
        over = L > int(cap[i]);
    }
    const unsigned long long ov = __ballot(over);
    if ((threadIdx.x & 63) == 0 && ov) {
        atomicAdd(&ctr->overfull, __popcll(ov));
        if (host_over) __hip_atomic_store(host_over, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------- exclusive scan (u32, in place)
// out[0..n) = exclusive prefix sums, out[n] = total. Reduce, scan the block sums, downsweep.

constexpr int kScanThreads = 256;
constexpr int kScanPT = 8;
constexpr int kScanTile = kScanThreads * kScanPT;

constexpr int kScanWaves = kScanThreads / 64;

// Block-wide exclusive scan (256 threads): wave scans, then the wave totals (wsum[2][4], used in
// turn so that one barrier per scan suffices: a call's slot is rewritten two calls later, after
// every thread has passed the barrier of the call in between).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t (*wsum)[kScanWaves], int& turn,
                                                         uint32_t& total) {
    const int lane = int(threadIdx.x & 63), wave = int(threadIdx.x >> 6);
    const uint32_t r = block_scan_excl<kScanWaves>(v, wsum[turn], lane, wave, &total);
    turn ^= 1;
    return r;
}

// One or two arrays scanned at once: blockIdx.y picks the array and its block sums.
struct ScanArrays {
    uint32_t* a[2];
    uint32_t* bsum[2];
};

__global__ void __launch_bounds__(kScanThreads) scan_reduce(ScanArrays A, int n) {
    __shared__ uint32_t sh[2][kScanWaves];
    int turn = 0;
    const uint32_t* __restrict__ a = A.a[blockIdx.y];
    const long long b0 = (long long)blockIdx.x * kScanTile + threadIdx.x * kScanPT;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPT; ++k)
        if (b0 + k < n) s += a[b0 + k];
    uint32_t total;
    block_exclusive_scan(s, sh, turn, total);
    if (threadIdx.x == 0) A.bsum[blockIdx.y][blockIdx.x] = total;
}

// What the level pass's scan also finishes (null: a plain scan): the tiles' deepest levels and
// over-full counts (bdeep[0 .. tiles), bdeep[tiles .. 2 tiles)) to the counters and the pinned host
// words (host[1] deepest level; host[2] the first array's total, the node count).
struct ScanReport {
    const uint32_t* bdeep;
    int tiles;
    WalkCounters* ctr;
    int* host;
};

// One block per array: exclusive scan of the nb block sums in place (kTopPT consecutive sums per
// thread per round), the array's total at a[n].
constexpr int kTopPT = 16;
__global__ void __launch_bounds__(kScanThreads) scan_top(ScanArrays A, int nb, int n, ScanReport R) {
    __shared__ uint32_t sh[2][kScanWaves];
    __shared__ uint32_t red[kScanWaves];
    int turn = 0;
    uint32_t* __restrict__ bsum = A.bsum[blockIdx.x];
    uint32_t carry = 0;
    for (int b = 0; b < nb; b += kScanThreads * kTopPT) {
        const int t0 = b + int(threadIdx.x) * kTopPT;
        uint32_t v[kTopPT], sum = 0;
#pragma unroll
        for (int k = 0; k < kTopPT; ++k) {
            v[k] = t0 + k < nb ? bsum[t0 + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan(sum, sh, turn, total);
#pragma unroll
        for (int k = 0; k < kTopPT; ++k) {
            if (t0 + k < nb) bsum[t0 + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) A.a[blockIdx.x][n] = carry;
    if (!R.bdeep) return;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0 && R.host) __hip_atomic_store(R.host + 2, int(carry), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    uint32_t d = 0, over = 0;
    for (int b = 0; b < R.tiles; b += kScanThreads * kTopPT) {  // kTopPT independent loads per thread
        const int t0 = b + int(threadIdx.x) * kTopPT;
#pragma unroll
        for (int k = 0; k < kTopPT; ++k) {
            if (t0 + k < R.tiles) {
                d = max(d, R.bdeep[t0 + k]);
                over += R.bdeep[R.tiles + t0 + k];
            }
        }
    }
    uint32_t total;
    block_exclusive_scan(over, sh, turn, total);
    over = total;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d = max(d, __shfl_xor(d, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kScanWaves; ++w) d = max(d, red[w]);
        R.ctr->max_level = int(d);
        R.ctr->overfull = int(over);
        if (R.host) __hip_atomic_store(R.host + 1, int(d), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// stride: block sums per scan tile (1 after scan_reduce; kScanTile / kLvTile after the level pass).
__global__ void __launch_bounds__(kScanThreads) scan_down(ScanArrays A, int n, int stride) {
    __shared__ uint32_t sh[2][kScanWaves];
    int turn = 0;
    uint32_t* __restrict__ a = A.a[blockIdx.y];
    const long long b0 = (long long)blockIdx.x * kScanTile + threadIdx.x * kScanPT;
    uint32_t v[kScanPT];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPT; ++k) {
        v[k] = b0 + k < n ? a[b0 + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, sh, turn, total) + A.bsum[blockIdx.y][std::size_t(blockIdx.x) * stride];
#pragma unroll
    for (int k = 0; k < kScanPT; ++k) {
        if (b0 + k < n) a[b0 + k] = run;
        run += v[k];
    }
}

// In-place exclusive scans of `count` (1 or 2) arrays of n entries, a[n] = total, in one set of
// three launches; tmp holds count * scan_tmp_entries(n) entries.
void scan_exclusive_n(hipStream_t s, uint32_t* a0, uint32_t* a1, int count, int n, uint32_t* tmp) {
    const int nb = int((static_cast<long long>(n) + kScanTile - 1) / kScanTile);
    ScanArrays A{{a0, a1}, {tmp, tmp + scan_tmp_entries(n)}};
    scan_reduce<<<dim3(nb, count), kScanThreads, 0, s>>>(A, n);
    LHIP_LAUNCH_CHECK();
    scan_top<<<count, kScanThreads, 0, s>>>(A, nb, n, ScanReport{});
    LHIP_LAUNCH_CHECK();
    scan_down<<<dim3(nb, count), kScanThreads, 0, s>>>(A, n, 1);
    LHIP_LAUNCH_CHECK();
}

void scan_exclusive(hipStream_t s, uint32_t* a, int n, uint32_t* tmp) { scan_exclusive_n(s, a, a, 1, n, tmp); }

// ---------------------------------------------------------------- nodes

// Pre-order enumeration: every cell that starts at body i (levels c+1 .. min(L, 21)) is node
// k = scan[i] + (level - c - 1). Writes the sort key (its level) and the value k, so a stable sort by
// level yields the level order (within a level, pre-order = first-body order).
// Nodes at or past ncap (the node arrays' capacity) are not written: the host sees the count and
// builds again with larger arrays.
__global__ void __launch_bounds__(kBS) enumerate_kernel(int n, const uint32_t* __restrict__ bl,
                                                        const uint32_t* __restrict__ scan, uint64_t* __restrict__ pkey,
                                                        uint32_t* __restrict__ pval, const uint32_t* __restrict__ cap,
                                                        uint32_t ncap) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = bl[i];
    const int L = int(b & 0xff), c = int((b >> 16) & 0xff) - 1;
    const int Lc = min(L, cap ? int(cap[i]) : kMaxLevel);
    const uint32_t first = scan[i];
    for (int l = c + 1; l <= Lc; ++l) {
        const uint32_t k = first + uint32_t(l - c - 1);
        if (k >= ncap) break;
        pkey[k] = uint64_t(l);
        pval[k] = k;
    }
}

// Level-order node records: body range start, level, leaf flag, first child, key prefix and the
// geometric cell. The centre follows the reference's halving in double (src/cpu_backend.cpp:576-581):
// child centre = parent centre +- half/2 per axis bit, from the root centre at origin + side/2.
// EXT (key frames): a restarted level-21 cell takes its own cube's centre and half side, and its
// descendants halve from there by the next frame's key digits (src/cpu_backend.cpp:541-545).
template <bool EXT>
__global__ void __launch_bounds__(kBS) nodes_kernel(const uint64_t* __restrict__ key, int n,
                                                    const uint32_t* __restrict__ bl, const uint32_t* __restrict__ scan,
                                                    const uint32_t* __restrict__ inv, const Cube* __restrict__ cube,
                                                    WalkNode* __restrict__ node, NodeMom* __restrict__ mom,
                                                    uint64_t* __restrict__ prefix, Frames F,
                                                    const uint32_t* __restrict__ cap, const uint32_t* __restrict__ fseg,
                                                    const double4* __restrict__ fcube, uint32_t ncap) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = bl[i];
    const int L = int(b & 0xff), c = int((b >> 16) & 0xff) - 1;
    const int top = EXT ? int(cap[i]) : kMaxLevel;
    const int Lc = min(L, top);
    if (Lc <= c) return;
    uint64_t kw = key[i];  // key word of the current frame
    const uint32_t first = scan[i], M = scan[n];
    const Cube cb = *cube;
    double half = 0.5 * cb.side;
    double gx = cb.ox + half, gy = cb.oy + half, gz = cb.oz + half;
    for (int l = 0; l <= Lc; ++l) {
        // the prefix within the frame (a frame's level-21 cell carries that frame's whole key)
        const int jl = l - kMaxLevel * ((l - 1) / kMaxLevel);
        const uint64_t pre = l == 0 ? 0ull : (kw >> (63 - 3 * jl));
        if (EXT && l > 0 && l % kMaxLevel == 0 && top > l) {
            const int f = l / kMaxLevel;
            const double4 q = fcube[fseg[std::size_t(f - 1) * n + i]];
            half = 0.5 * q.w;
            gx = q.x + half; gy = q.y + half; gz = q.z + half;
            kw = F.fkey[std::size_t(f - 1) * n + i];
        }
        if (l > c) {
            const uint32_t k = first + uint32_t(l - c - 1);
            if (k >= ncap) return;  // a tree larger than the arrays: the host builds it again
            const uint32_t p = inv[k];
            if (p >= ncap) return;
            // Every field but hi: a leaf's hi is written by the body that starts the next pre-order
            // node (below), an internal node's by the multipole pass.
            WalkNode* const np = node + p;
            *reinterpret_cast<float4*>(np) = make_float4(0.f, 0.f, 0.f, 0.f);  // com, rcrit2
            np->lo = i;
            // The first child is the next pre-order node (same first body, one level deeper).
            *reinterpret_cast<int2*>(&np->info) =
                make_int2(l | (l == Lc ? kInfoLeaf : 0), l == Lc || k + 1 >= ncap ? -1 : int(inv[k + 1]));
            // The deepest cell body i starts is a leaf, and the pre-order node after a leaf starts at
            // the body just past it: this body ends the leaf numbered first - 1, and the last leaf
            // ends at n.
            if (l == c + 1 && first > 0 && first - 1 < ncap) {
                const uint32_t q = inv[first - 1];
                if (q < ncap) node[q].hi = i;
            }
            if (l == Lc && k + 1 == M) node[p].hi = n;
            prefix[p] = pre;
            NodeMom& md = mom[p];
            md.gx = float(gx);
            md.gy = float(gy);
            md.gz = float(gz);
            md.half = float(half);
        }
        if (l < Lc) {
            const int j = l % kMaxLevel;
            const int oct = int((kw >> (3 * (kMaxLevel - 1 - j))) & 7);
            const double q = 0.5 * half;
            gx = gx + ((oct & 4) ? q : -q);
            gy = gy + ((oct & 2) ? q : -q);
            gz = gz + ((oct & 1) ? q : -q);
            half = q;
        }
    }
}

// Group starts in body order; gstart[count] = n.
__global__ void __launch_bounds__(kBS) groups_kernel(const uint32_t* __restrict__ gflag_scan, int n,
                                                     const uint32_t* __restrict__ bl, uint32_t* __restrict__ gstart) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = bl[i];
    const int c = int((b >> 16) & 0xff) - 1, G = int((b >> 8) & 0xff);
    if (c < G) gstart[gflag_scan[i]] = uint32_t(i);
    if (i == n - 1) gstart[gflag_scan[n]] = uint32_t(n);
}

// ---------------------------------------------------------------- multipoles, one level per launch
// Reference: leaf moments and the internal reverse sweep with the parallel-axis theorem for the
// traceless quadrupole, accumulated in double and stored as float (src/cpu_backend.cpp:615-663);
// rcrit = 2 half / theta + |com - cell centre| (src/cpu_backend.cpp:596-600).

__device__ __forceinline__ float rcrit2_of(float cx, float cy, float cz, float gx, float gy, float gz, float half,
                                          float inv_theta) {
    const float dx = cx - gx, dy = cy - gy, dz = cz - gz;
    // Rounded as the reference build evaluates it: fma(dz, dz, fma(dx, dx, dy * dy)), then the
    // product and the square root added in double.
    const float t = __fmaf_rn(dz, dz, __fmaf_rn(dx, dx, __fmul_rn(dy, dy)));
    const double r = __dadd_rn(__dmul_rn(double(2.f * half), double(inv_theta)), __dsqrt_rn(double(t)));
    return float(fmin(__dmul_rn(r, r), double(FLT_MAX)));
}

// Nodes of one level are contiguous in level order: [levstart[level], levstart[level + 1]). A leaf
// ends at the first body whose boundary level c is below the leaf's level (at most 16 steps except in
// an over-full finest cell); an internal node's children are the contiguous run of next-level nodes
// from `first` whose key prefix extends the node's, and the node ends where its last child ends.
// MIXED: bodies are FP64 positions (xd, yd, zd) with float masses, and each node's FP64 centre of
// mass, mass and quadrupole go to momd as well; internal nodes combine their children's FP64
// moments, so the quadrupoles of the root levels are accumulated in FP64 throughout.
// KEEP: a second pass on other positions over a built tree (frozen-topology replay): node records,
// rcrit2 and qrel (the walk's decisions) stay as the first pass wrote them; com, mass and
// quadrupole are recomputed.
// part: kMomLeaves takes the leaves of levels 0..level in one launch (a leaf needs only its bodies);
// kMomInternal takes the internal nodes of `level` (their children are done by then). Each node's
// sums are the same either way.
constexpr int kMomLeaves = 1, kMomInternal = 2;
struct MomArgs {
    int n;
    const float4* __restrict__ posm;
    const double *__restrict__ xd, *__restrict__ yd, *__restrict__ zd;
    const float* __restrict__ bm;
    const uint32_t* __restrict__ levstart;
    const uint64_t* __restrict__ prefix;
    WalkNode* __restrict__ node;
    NodeMom* __restrict__ mom;
    NodeMomD* __restrict__ momd;
    float inv_theta;
    Frames F;
};

// The multipoles of node p (a leaf from its bodies, an internal node of `level` from its children).
template <bool MIXED, bool KEEP, bool EXT>
__device__ __forceinline__ void moment_node(const MomArgs& A, int p, bool leaves, int level, int cend) {
    const float4* __restrict__ posm = A.posm;
    const double *__restrict__ xd = A.xd, *__restrict__ yd = A.yd, *__restrict__ zd = A.zd;
    const float* __restrict__ bm = A.bm;
    const uint64_t* __restrict__ prefix = A.prefix;
    WalkNode* __restrict__ node = A.node;
    NodeMom* __restrict__ mom = A.mom;
    NodeMomD* __restrict__ momd = A.momd;
    const float inv_theta = A.inv_theta;
    const Frames& F = A.F;
    {
        WalkNode nd = node[p];
        if (info_leaf(nd.info) != leaves) return;
        NodeMom md = mom[p];
        double M = 0, sx = 0, sy = 0, sz = 0;
        double q[6] = {0, 0, 0, 0, 0, 0};
        double cx, cy, cz;
        int nchild = 0;
        auto body = [&](int j, double& x, double& y, double& z, double& m) {
            if (MIXED) {
                x = xd[j]; y = yd[j]; z = zd[j]; m = bm[j];
            } else {
                const float4 b = posm[j];
                x = b.x; y = b.y; z = b.z; m = b.w;
            }
        };
        // The loads that need only the node record are issued together (a leaf's bodies; a node's
        // eight child candidates: prefixes, body ends and records), so a level costs two dependent
        // round trips; the sums keep the one-at-a-time order.
        if (leaves) {
            // The leaf's bodies [lo, hi) (hi from the node pass), the first KB of them loaded at once.
            const int hi = nd.hi;
            constexpr int KB = MIXED ? 8 : kLeafMax;  // bodies held in registers
            float4 pb[MIXED ? 1 : KB];
            double rx[MIXED ? KB : 1], ry[MIXED ? KB : 1], rz[MIXED ? KB : 1];
            float rm[MIXED ? KB : 1];
#pragma unroll
            for (int t = 0; t < KB; ++t) {
                const int j = min(nd.lo + t, hi - 1);
                if constexpr (MIXED) {
                    rx[t] = xd[j]; ry[t] = yd[j]; rz[t] = zd[j]; rm[t] = bm[j];
                } else {
                    pb[t] = posm[j];
                }
            }
            auto held = [&](int t, double& x, double& y, double& z, double& m) {
                if constexpr (MIXED) {
                    x = rx[t]; y = ry[t]; z = rz[t]; m = rm[t];
                } else {
                    x = pb[t].x; y = pb[t].y; z = pb[t].z; m = pb[t].w;
                }
            };
#pragma unroll
            for (int t = 0; t < KB; ++t)
                if (nd.lo + t < hi) {
                    double x, y, z, m;
                    held(t, x, y, z, m);
                    M += m;
                    sx = fma(m, x, sx);
                    sy = fma(m, y, sy);
                    sz = fma(m, z, sz);
                }
            for (int j0 = nd.lo + KB; j0 < hi; j0 += 4) {
                double bx[4], by[4], bz[4], m[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) body(min(j0 + t, hi - 1), bx[t], by[t], bz[t], m[t]);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (j0 + t < hi) {
                        M += m[t];
                        sx = fma(m[t], bx[t], sx);
                        sy = fma(m[t], by[t], sy);
                        sz = fma(m[t], bz[t], sz);
                    }
            }
            cx = M > 0 ? sx / M : double(md.gx);
            cy = M > 0 ? sy / M : double(md.gy);
            cz = M > 0 ? sz / M : double(md.gz);
            auto quad = [&](double x, double y, double z, double m) {
                const double dx = x - cx, dy = y - cy, dz = z - cz;
                const double r2 = fma(dz, dz, fma(dx, dx, dy * dy));
                q[0] = fma(m, fma(3 * dx, dx, -r2), q[0]);
                q[3] = fma(m, fma(3 * dy, dy, -r2), q[3]);
                q[5] += m * fma(3 * dz, dz, -r2);
                const double m3 = m * 3;
                q[1] = fma(m3 * dx, dy, q[1]);
                q[2] = fma(m3 * dx, dz, q[2]);
                q[4] += m3 * dy * dz;
            };
#pragma unroll
            for (int t = 0; t < KB; ++t)
                if (nd.lo + t < hi) {
                    double x, y, z, m;
                    held(t, x, y, z, m);
                    quad(x, y, z, m);
                }
            for (int j0 = nd.lo + KB; j0 < hi; j0 += 4) {
                double bx[4], by[4], bz[4], mm[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) body(min(j0 + t, hi - 1), bx[t], by[t], bz[t], mm[t]);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (j0 + t < hi) quad(bx[t], by[t], bz[t], mm[t]);
            }
        } else {
            const uint64_t pre = prefix[p];
            int ch = nd.first;
            // Children: the run of next-level nodes from `first` inside this cell (key prefix; over
            // key frames, the first bodies' common extended prefix). At most 8.
            auto shifted = [&](double m, double ccx, double ccy, double ccz, const double (&cq)[6]) {
                const double dx = ccx - cx, dy = ccy - cy, dz = ccz - cz;
                const double r2 = fma(dz, dz, fma(dx, dx, dy * dy));
                q[0] += cq[0] + m * (3 * dx * dx - r2);
                q[3] += cq[3] + m * (3 * dy * dy - r2);
                q[5] += cq[5] + m * (3 * dz * dz - r2);
                q[1] += cq[1] + m * 3 * dx * dy;
                q[2] += cq[2] + m * 3 * dx * dz;
                q[4] += cq[4] + m * 3 * dy * dz;
            };
            auto child_at = [&](int k, double& m, double& ccx, double& ccy, double& ccz, double (&cq)[6]) {
                if (MIXED) {
                    const NodeMomD& c = momd[k];
                    m = c.m; ccx = c.cx; ccy = c.cy; ccz = c.cz;
                    for (int u = 0; u < 6; ++u) cq[u] = c.q[u];
                } else {
                    const NodeMom& c = mom[k];
                    m = c.m; ccx = c.cx; ccy = c.cy; ccz = c.cz;
                    cq[0] = c.qxx; cq[1] = c.qxy; cq[2] = c.qxz; cq[3] = c.qyy; cq[4] = c.qyz; cq[5] = c.qzz;
                }
            };
            if (EXT) {
                while (ch < cend && ext_common(F, nd.lo, node[ch].lo) >= level) ++ch;
                for (int k = nd.first; k < ch; ++k) {
                    double m, x, y, z, cq[6];
                    child_at(k, m, x, y, z, cq);
                    M += m;
                    sx = fma(m, x, sx);
                    sy = fma(m, y, sy);
                    sz = fma(m, z, sz);
                }
                nd.hi = node[ch - 1].hi;
                cx = M > 0 ? sx / M : double(md.gx);
                cy = M > 0 ? sy / M : double(md.gy);
                cz = M > 0 ? sz / M : double(md.gz);
                for (int k = nd.first; k < ch; ++k) {
                    double m, x, y, z, cq[6];
                    child_at(k, m, x, y, z, cq);
                    shifted(m, x, y, z, cq);
                }
            } else {
                // The eight candidates' prefixes, body ends and whole records, loaded at once.
                uint64_t pr[8];
                int hi8[8];
                double cm[8], ccx[8], ccy[8], ccz[8], cq[8][6];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int k = min(nd.first + t, cend - 1);
                    pr[t] = nd.first + t < cend ? prefix[k] : 0ull;
                    hi8[t] = node[k].hi;
                    child_at(k, cm[t], ccx[t], ccy[t], ccz[t], cq[t]);
                }
                bool run = true;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    run = run && nd.first + t < cend && (pr[t] >> 3) == pre;
                    ch += run ? 1 : 0;
                }
                const int nc = ch - nd.first;
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (t < nc) {
                        M += cm[t];
                        sx = fma(cm[t], ccx[t], sx);
                        sy = fma(cm[t], ccy[t], sy);
                        sz = fma(cm[t], ccz[t], sz);
                    }
                int h = hi8[0];
#pragma unroll
                for (int t = 1; t < 8; ++t) h = t < nc ? hi8[t] : h;
                nd.hi = h;
                cx = M > 0 ? sx / M : double(md.gx);
                cy = M > 0 ? sy / M : double(md.gy);
                cz = M > 0 ? sz / M : double(md.gz);
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (t < nc) shifted(cm[t], ccx[t], ccy[t], ccz[t], cq[t]);
            }
            nchild = ch - nd.first;
        }
        if (MIXED) {
            NodeMomD d;
            d.cx = cx; d.cy = cy; d.cz = cz; d.m = M;
            for (int u = 0; u < 6; ++u) d.q[u] = q[u];
            momd[p] = d;
        }
        md.cx = float(cx); md.cy = float(cy); md.cz = float(cz); md.m = float(M);
        md.qxx = float(q[0]); md.qxy = float(q[1]); md.qxz = float(q[2]);
        md.qyy = float(q[3]); md.qyz = float(q[4]); md.qzz = float(q[5]);
        if (KEEP) {
            mom[p] = md;
            return;
        }
        md.rcrit2 = rcrit2_of(md.cx, md.cy, md.cz, md.gx, md.gy, md.gz, md.half, inv_theta);
        {
            const double qf = sqrt(q[0] * q[0] + q[3] * q[3] + q[5] * q[5] +
                                   2.0 * (q[1] * q[1] + q[2] * q[2] + q[4] * q[4]));
            md.qrel = M > 0 ? float(1.5 * qf / M) : 0.f;
        }
        mom[p] = md;
        nd.cx = md.cx; nd.cy = md.cy; nd.cz = md.cz; nd.rcrit2 = md.rcrit2;
        nd.info |= nchild << 8;
        node[p] = nd;
    }
}

template <bool MIXED, bool KEEP = false, bool EXT = false>
__global__ void __launch_bounds__(kBS) moments_level(int part, int level, MomArgs A) {
    const bool leaves = part == kMomLeaves;
    const int p0 = leaves ? 0 : int(A.levstart[level]), p1 = int(A.levstart[level + 1]);
    const int cend = int(A.levstart[level + 2]);  // end of the next level
    for (int p = p0 + blockIdx.x * kBS + threadIdx.x; p < p1; p += gridDim.x * kBS)
        moment_node<MIXED, KEEP, EXT>(A, p, leaves, level, cend);
}


// Keys that take the costly walk units first. A unit's bodies share one interaction list, so its
// work is estimated as (list length + 1) x (bodies + 8), the list length taken from the previous
// walk's cost of its first body (costs travel with the bodies through the sort; before any walk
// they may hold anything, which only changes the order). Units estimated at thr or more get
// 255 - 8 log2(estimate) (8 buckets per octave), the others 255, so one stable 8-bit radix pass
// puts the costly units first, longest first, and leaves the rest in their natural (Morton) order,
// which keeps neighbouring groups, and the tree nodes they read, together in time.
// ranges: the key is (range << 5) | (31 - bucket) with range = the unit's eighth of the Morton order
// and a coarser bucket (units at or above thr: 1 + 2.5 log2(est / thr), at most 31), so that each
// range holds its costly units first, longest first, then the others in Morton order.
__global__ void __launch_bounds__(kBS) unit_order_keys(const uint32_t* __restrict__ ustart,
                                                       const uint32_t* __restrict__ ucount, int n,
                                                       const float* __restrict__ cost, float thr, int ranges,
                                                       uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int g = blockIdx.x * kBS + threadIdx.x;
    const int units = int(ucount[n]);
    if (g >= units) return;
    const int lo = int(ustart[g]), hi = int(ustart[g + 1]);
    float est = (fmaxf(cost[lo], 0.f) + 1.f) * float(hi - lo + 8);
    est = est >= 1.f ? fminf(est, 1e30f) : 1.f;  // NaN and garbage included
    if (ranges) {
        const int range = int((long long)g * kWalkRanges / units);
        const int bucket = est >= thr ? min(31, 1 + int(2.5f * log2f(est / thr))) : 0;
        key[g] = uint64_t((range << 5) | (31 - bucket));
    } else {
        const int bucket = est >= thr ? min(255, int(8.f * log2f(est))) : 0;
        key[g] = uint64_t(255 - bucket);
    }
    val[g] = uint32_t(g);
}

// ---------------------------------------------------------------- key restarts (the rare path)
// src/cpu_backend.cpp:501-544: an over-full level-21 cell (of the deepest key frame so far) whose
// bodies do not all coincide gets keys in its own bounding cube (float minimum, side = largest
// extent x (1 + 1e-6)); its bodies are re-sorted stably by them and the octree continues below it
// in that frame.

__device__ __forceinline__ bool over_full(const uint32_t* bl, const uint32_t* cap, int i) {
    return int(bl[i] & 0xff) > (cap ? int(cap[i]) : kMaxLevel);
}

// seg[i] = 1 at the first body of each over-full deepest cell.
__global__ void __launch_bounds__(kBS) restart_cells_kernel(Frames F, const uint32_t* __restrict__ cap,
                                                            const uint32_t* __restrict__ bl, uint32_t* __restrict__ seg) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= F.n) return;
    bool start = over_full(bl, cap, i);
    if (start && i > 0 && over_full(bl, cap, i - 1))
        start = ext_common(F, i - 1, i) < (cap ? int(cap[i]) : kMaxLevel);
    seg[i] = start ? 1u : 0u;
}

// Order-preserving 64-bit image of a double, for atomic min and max.
__device__ __forceinline__ unsigned long long ord64(double v) {
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double unord64(unsigned long long o) {
    const unsigned long long u = (o >> 63) ? (o & ~(1ull << 63)) : ~o;
    return __longlong_as_double(static_cast<long long>(u));
}

__global__ void __launch_bounds__(kBS) restart_bbox_init(unsigned long long* __restrict__ bbox, int ncell) {
    const int t = blockIdx.x * kBS + threadIdx.x;
    if (t < 6 * ncell) bbox[t] = (t % 6) < 3 ? ~0ull : 0ull;
}

// bbox[6 c .. 6 c + 6) = min x, y, z, max x, y, z of cell c (seg is the scanned start flags).
template <class R>
__global__ void __launch_bounds__(kBS) restart_bbox_kernel(int n, const uint32_t* __restrict__ cap,
                                                           const uint32_t* __restrict__ bl,
                                                           const uint32_t* __restrict__ seg, const R* __restrict__ x,
                                                           const R* __restrict__ y, const R* __restrict__ z,
                                                           unsigned long long* __restrict__ bbox) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n || !over_full(bl, cap, i)) return;
    unsigned long long* q = bbox + 6 * std::size_t(seg[i + 1] - 1);
    const unsigned long long ox = ord64(double(x[i])), oy = ord64(double(y[i])), oz = ord64(double(z[i]));
    atomicMin(q + 0, ox); atomicMin(q + 1, oy); atomicMin(q + 2, oz);
    atomicMax(q + 3, ox); atomicMax(q + 4, oy); atomicMax(q + 5, oz);
}

// Cube of each cell: origin = the minimum, side = largest extent x (1 + 1e-6) (0: all coincide).
__global__ void __launch_bounds__(kBS) restart_cube_kernel(const unsigned long long* __restrict__ bbox, int ncell,
                                                           double4* __restrict__ cubes) {
    const int c = blockIdx.x * kBS + threadIdx.x;
    if (c >= ncell) return;
    const unsigned long long* q = bbox + 6 * std::size_t(c);
    const double mnx = unord64(q[0]), mny = unord64(q[1]), mnz = unord64(q[2]);
    const double sx = __dsub_rn(unord64(q[3]), mnx), sy = __dsub_rn(unord64(q[4]), mny), sz = __dsub_rn(unord64(q[5]), mnz);
    double side = fmax(sx, fmax(sy, sz));
    side = side == 0.0 ? 0.0 : __dmul_rn(side, 1.0 + 1e-6);
    cubes[c] = make_double4(mnx, mny, mnz, side);
}

// Frame keys of the bodies of restarted cells (side > 0), their cube index and cap; gkey = the
// first body of the body's restarted cell, else the body itself.
template <class R>
__global__ void __launch_bounds__(kBS) restart_keys_kernel(int n, const uint32_t* __restrict__ bl,
                                                           const uint32_t* __restrict__ seg,
                                                           const double4* __restrict__ cubes, int cube0,
                                                           const R* __restrict__ x, const R* __restrict__ y,
                                                           const R* __restrict__ z, uint64_t* __restrict__ fkey,
                                                           uint32_t* __restrict__ fseg, uint32_t* __restrict__ cap,
                                                           uint64_t* __restrict__ gkey,
                                                           const uint32_t* __restrict__ cstart) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    uint64_t k = 0, g = uint64_t(i);
    uint32_t fs = ~0u;
    if (over_full(bl, cap, i)) {
        const int c = int(seg[i + 1]) - 1;
        const double4 q = cubes[c];
        if (q.w > 0.0) {
            k = morton_key(x[i], y[i], z[i], q.x, q.y, q.z, __ddiv_rn(2097152.0, q.w));
            fs = uint32_t(cube0 + c);
            g = uint64_t(cstart[c]);
            cap[i] += kMaxLevel;
        }
    }
    fkey[i] = k;
    fseg[i] = fs;
    gkey[i] = g;
}

// cstart[c] = the first body of cell c.
__global__ void __launch_bounds__(kBS) restart_starts_kernel(int n, const uint32_t* __restrict__ seg,
                                                             uint32_t* __restrict__ cstart) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n && seg[i + 1] != seg[i]) cstart[seg[i]] = uint32_t(i);
}

__global__ void __launch_bounds__(kBS) gather_u64_kernel(const uint32_t* __restrict__ perm,
                                                         const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                         int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}
__global__ void __launch_bounds__(kBS) gather_u32_kernel(const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                         int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}
__global__ void __launch_bounds__(kBS) iota_u32_kernel(uint32_t* __restrict__ a, int n) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) a[i] = uint32_t(i);
}
__global__ void __launch_bounds__(kBS) fill_u32_kernel(uint32_t* __restrict__ a, uint32_t v, long long n) {
    const long long i = (long long)blockIdx.x * kBS + threadIdx.x;
    if (i < n) a[i] = v;
}

Frames frames_of(const BuildBufs& b) { return Frames{b.key, b.fkey, b.frames > 1 ? b.frames : 1, b.n}; }

}  // namespace

void tree_levels(hipStream_t s, const BuildBufs& b) {
    levels_ext_kernel<<<grid_for(b.n, kBS), kBS, 0, s>>>(frames_of(b), b.cap, b.gmax, b.bl, b.cnt, b.gflag, b.counters,
                                                        b.host_over);
    LHIP_LAUNCH_CHECK();
}

int tree_restart_cells(hipStream_t s, const BuildBufs& b, uint32_t* seg) {
    const int n = b.n;
    restart_cells_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(frames_of(b), b.cap, b.bl, seg);
    LHIP_LAUNCH_CHECK();
    scan_exclusive(s, seg, n, b.scan_tmp);
    uint32_t count = 0;
    LHIP_CHECK(hipMemcpyAsync(&count, seg + n, 4, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipStreamSynchronize(s));
    return int(count);
}

int tree_node_count(hipStream_t s, const BuildBufs& b) {
    scan_exclusive(s, b.cnt, b.n, b.scan_tmp);
    uint32_t M = 0;
    LHIP_CHECK(hipMemcpyAsync(&M, b.cnt + b.n, 4, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipStreamSynchronize(s));
    return int(M);
}

void tree_restart_cubes(hipStream_t s, const BuildBufs& b, const void* x, const void* y, const void* z, bool f64,
                        const uint32_t* seg, int ncell, unsigned long long* bbox, double4* cubes) {
    if (ncell <= 0) return;
    const int n = b.n;
    restart_bbox_init<<<grid_for(6ll * ncell, kBS), kBS, 0, s>>>(bbox, ncell);
    LHIP_LAUNCH_CHECK();
    if (f64)
        restart_bbox_kernel<double><<<grid_for(n, kBS), kBS, 0, s>>>(
            n, b.cap, b.bl, seg, static_cast<const double*>(x), static_cast<const double*>(y),
            static_cast<const double*>(z), bbox);
    else
        restart_bbox_kernel<float><<<grid_for(n, kBS), kBS, 0, s>>>(n, b.cap, b.bl, seg, static_cast<const float*>(x),
                                                                    static_cast<const float*>(y),
                                                                    static_cast<const float*>(z), bbox);
    LHIP_LAUNCH_CHECK();
    restart_cube_kernel<<<grid_for(ncell, kBS), kBS, 0, s>>>(bbox, ncell, cubes);
    LHIP_LAUNCH_CHECK();
}

void tree_restart_keys(hipStream_t s, const BuildBufs& b, const void* x, const void* y, const void* z, bool f64,
                       const uint32_t* seg, const double4* cubes, int cube0, uint64_t* fkey_frame,
                       uint32_t* fseg_frame, uint32_t* cap, uint64_t* gkey, uint32_t* cstart) {
    const int n = b.n;
    restart_starts_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(n, seg, cstart);
    LHIP_LAUNCH_CHECK();
    if (f64)
        restart_keys_kernel<double><<<grid_for(n, kBS), kBS, 0, s>>>(
            n, b.bl, seg, cubes, cube0, static_cast<const double*>(x), static_cast<const double*>(y),
            static_cast<const double*>(z), fkey_frame, fseg_frame, cap, gkey, cstart);
    else
        restart_keys_kernel<float><<<grid_for(n, kBS), kBS, 0, s>>>(
            n, b.bl, seg, cubes, cube0, static_cast<const float*>(x), static_cast<const float*>(y),
            static_cast<const float*>(z), fkey_frame, fseg_frame, cap, gkey, cstart);
    LHIP_LAUNCH_CHECK();
}

void gather_u64(hipStream_t s, const uint32_t* perm, const uint64_t* in, uint64_t* out, int n) {
    gather_u64_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(perm, in, out, n);
    LHIP_LAUNCH_CHECK();
}
void gather_u32(hipStream_t s, const uint32_t* perm, const uint32_t* in, uint32_t* out, int n) {
    gather_u32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(perm, in, out, n);
    LHIP_LAUNCH_CHECK();
}
void iota_u32(hipStream_t s, uint32_t* a, int n) {
    if (n <= 0) return;
    iota_u32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(a, n);
    LHIP_LAUNCH_CHECK();
}
void fill_u32(hipStream_t s, uint32_t* a, uint32_t v, long long n) {
    if (n <= 0) return;
    fill_u32_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(a, v, n);
    LHIP_LAUNCH_CHECK();
}

// Launch grids of the multipole pass: the leaves of every level at once, then one level of internal
// nodes per launch (grid-stride loops, so any grid covers the nodes). Level l holds at most 8^l
// nodes, so the top levels take a block or a few.
inline unsigned leaf_grid(int n) { return std::min<unsigned>(grid_for(n, kBS), 2048u); }
inline unsigned level_grid(int n, int level) {
    const long long most = level < 10 ? 1ll << (3 * level) : (long long)n;
    return std::min<unsigned>(grid_for(std::min<long long>(n, most), kBS), 512u);
}

// ranges (optional): the order is kWalkRanges Morton ranges (costly units first in each), their
// starts written to ranges[0 .. kWalkRanges) (the digit starts of digits 0, 32, ..., 224).
void unit_order(hipStream_t s, const BuildBufs& b, const uint32_t* ustart, const uint32_t* ucount, float thr,
                uint32_t* ranges, uint32_t* order) {
    unit_order_keys<<<grid_for(b.n, kBS), kBS, 0, s>>>(ustart, ucount, b.n, b.wcost, thr, ranges ? 1 : 0, b.pkey,
                                                       b.pval);
    LHIP_LAUNCH_CHECK();
    radix_sort(s, *b.radix, RadixBufs{{b.pkey, b.pkey2}, {b.pval, order}}, false, b.n, ucount + b.n, 0, 1, false,
               nullptr, ranges, ranges ? kRadix : 0, kRadix / kWalkRanges);
}

// The multipole launches over a built level order: every leaf in one launch, then the internal
// levels deepest first, one launch each (at 2^20 bodies the ten internal levels hold 30,653 nodes,
// at most 12,105 in one level, so each launch costs its two dependent round trips and its start).
// Measured against the alternatives at 2^20 bodies: levels 4..0 finished by one workgroup with
// barriers between them took 51.6 us against about 35 us as five launches; and a single climb
// kernel (every leaf thread arrives at its parent through an acquire-release atomic at agent
// scope, the last child computes the parent and climbs on) took 356 us against about 95 us for
// these eleven launches: every arrival writes back and invalidates its XCD's L2 (buffer_wbl2 sc1,
// buffer_inv sc1), and thousands of those serialise.
template <bool MIXED, bool KEEP, bool EXT>
void launch_moments(hipStream_t s, const MomArgs& A, int depth) {
    moments_level<MIXED, KEEP, EXT><<<leaf_grid(A.n), kBS, 0, s>>>(kMomLeaves, depth, A);
    LHIP_LAUNCH_CHECK();
    for (int l = depth - 1; l >= 0; --l) {
        moments_level<MIXED, KEEP, EXT><<<level_grid(A.n, l), kBS, 0, s>>>(kMomInternal, l, A);
        LHIP_LAUNCH_CHECK();
    }
}

MomArgs mom_args(const BuildBufs& b, bool mixed, float inv_theta) {
    MomArgs A{};
    A.n = b.n;
    A.posm = mixed ? nullptr : b.posm;
    if (mixed) {
        A.xd = b.xd; A.yd = b.yd; A.zd = b.zd; A.bm = b.m;
    }
    A.levstart = b.levstart;
    A.prefix = b.prefix;
    A.node = b.node;
    A.mom = b.mom;
    A.momd = mixed ? b.momd : nullptr;
    A.inv_theta = inv_theta;
    A.F = frames_of(b);
    return A;
}

void tree_moments_keep(hipStream_t s, const BuildBufs& b) {
    const bool ext = b.frames > 1;
    const int full = tree_depth(ext ? b.frames : 1);
    const int depth = b.depth_hint >= 0 ? std::min(b.depth_hint, full) : full;
    const MomArgs A = mom_args(b, true, 1.f);
    if (ext) launch_moments<true, true, true>(s, A, depth);
    else launch_moments<true, true, false>(s, A, depth);
}


// Block sums of a scan: one per kLvTile entries (the level pass's tiles; a plain scan uses one per
// kScanTile), plus the total.
std::size_t scan_tmp_entries(int n) {
    return std::size_t((static_cast<long long>(n) + kLvTile) / kLvTile) + 1;
}

void scan_u32(hipStream_t s, uint32_t* a, int n, uint32_t* tmp) { scan_exclusive(s, a, n, tmp); }

int tree_build(hipStream_t s, const BuildBufs& b, float inv_theta, hipEvent_t levels_done) {
    const int n = b.n;
    const bool ext = b.frames > 1;
    if (b.gmax < kLeafMax || b.gmax > kGroupMax) throw Error(kInvalid, "walk group size must be in [16, 128]");
    const bool wait_levels = levels_done && b.host_over && !ext;
    if (ext) {
        reset_build<<<1, 64, 0, s>>>(b.counters);
        LHIP_LAUNCH_CHECK();
        tree_levels(s, b);
        scan_exclusive_n(s, b.cnt, b.gflag, 2, n, b.scan_tmp);  // node counts and group starts
    } else {
        // The level pass with the first step of the scans of node counts and group starts (its tile
        // sums), then the scans' top step, which also reports the depth and the node count.
        const int tiles = int(grid_for(n, kLvTile));
        const ScanArrays A{{b.cnt, b.gflag}, {b.scan_tmp, b.scan_tmp + scan_tmp_entries(n)}};
        levels_kernel<<<tiles, kLvTile, 0, s>>>(b.key, n, b.gmax, b.bl, b.cnt, b.gflag, A.bsum[0], A.bsum[1],
                                                b.host_over, b.gstart);
        LHIP_LAUNCH_CHECK();
        scan_top<<<2, kScanThreads, 0, s>>>(A, tiles, n,
                                            ScanReport{b.gstart, tiles, b.counters, wait_levels ? b.host_over : nullptr});
        LHIP_LAUNCH_CHECK();
        scan_down<<<dim3(grid_for(n, kScanTile), 2), kScanThreads, 0, s>>>(A, n, kScanTile / kLvTile);
        LHIP_LAUNCH_CHECK();
    }
    const uint32_t* M_dev = b.cnt + n;  // node count
    if (levels_done) LHIP_CHECK(hipEventRecord(levels_done, s));
    const uint32_t ncap = uint32_t(b.cap_nodes);
    enumerate_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(n, b.bl, b.cnt, b.pkey, b.pval, ext ? b.cap : nullptr, ncap);
    LHIP_LAUNCH_CHECK();
    // Stable sort of the pre-order nodes by level (one 8-bit pass: levels <= kMaxDepth < 256).
    // Its scatter also writes the inverse (inv[pre-order index] = level-order index) and the level
    // starts (the digit starts: levels run 0 .. kMaxDepth, so starts kMaxDepth + 1 on are the count).
    radix_sort(s, *b.radix, RadixBufs{{b.pkey, b.pkey2}, {b.pval, b.pval2}}, false, b.cap_nodes, M_dev, 0, 1, false,
               b.inv, b.levstart, kDepthLevels + 2);
    if (ext)
        nodes_kernel<true><<<grid_for(n, kBS), kBS, 0, s>>>(b.key, n, b.bl, b.cnt, b.inv, b.cube, b.node, b.mom,
                                                            b.prefix, frames_of(b), b.cap, b.fseg, b.fcube, ncap);
    else
        nodes_kernel<false><<<grid_for(n, kBS), kBS, 0, s>>>(b.key, n, b.bl, b.cnt, b.inv, b.cube, b.node, b.mom,
                                                             b.prefix, frames_of(b), nullptr, nullptr, nullptr, ncap);
    LHIP_LAUNCH_CHECK();
    groups_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(b.gflag, n, b.bl, b.gstart);
    LHIP_LAUNCH_CHECK();
    if (b.worder && b.wcost) {
        // The walk order of this tree's units (queued before the host waits on the level pass).
        unit_order(s, b, b.gstart, b.gflag, b.order_thr,
                   range_starts_of(b.counters), b.worder);
    }
    int depth = tree_depth(ext ? b.frames : 1);
    if (wait_levels) {
        // The rest of the build above is queued; the multipoles wait for the level pass's report.
        LHIP_CHECK(hipEventSynchronize(levels_done));
        if (*reinterpret_cast<volatile int*>(b.host_over)) return -1;  // key restarts: the caller rebuilds
        // More nodes than the arrays hold: nothing past them was written; the caller grows them and
        // builds again.
        if (*reinterpret_cast<volatile int*>(b.host_over + 2) > b.cap_nodes) return kBuildGrow;
        const int deepest = *reinterpret_cast<volatile int*>(b.host_over + 1);
        depth = std::min(depth, std::max(0, deepest));
    }
    // The deepest level present holds only leaves, so the internal levels start one above it.
    const MomArgs A = mom_args(b, b.momd != nullptr, inv_theta);
    if (b.momd) ext ? launch_moments<true, false, true>(s, A, depth) : launch_moments<true, false, false>(s, A, depth);
    else ext ? launch_moments<false, false, true>(s, A, depth) : launch_moments<false, false, false>(s, A, depth);
    return depth;
}

}  // namespace lhip
