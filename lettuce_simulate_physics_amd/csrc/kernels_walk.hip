// Tree walk: one wavefront per walk group, stackless traversal of the pre-order octree, interaction
// lists buffered in LDS and evaluated by every lane for its one or two bodies.
//
// Reference: walk / build_list / eval_group (src/cpu_backend.cpp:667-689, 816-882). Same
// acceptance test against the group's bounding box (a node is accepted when it does not contain
// the group and dist^2(box, com) >= rcrit2), same treatment of leaves (opened leaves and accepted
// leaves of <= 4 bodies contribute bodies, other accepted nodes their monopole + quadrupole), same
// self exclusion by index. The list a group gets is therefore the reference's list.
//
// Traversal. Nodes are in depth-first pre-order: the subtree of node k is [k, skip_k). A window of
// 64 consecutive nodes is tested at once, one per lane (coalesced loads). A node is visited iff no
// earlier node of the window that ends the descent (accepted, or a leaf) skips past it: with
// s_j = skip_j for such nodes and 0 otherwise, node k is visited iff max_{j<k} s_j <= k, an
// exclusive prefix-max across the wave. The next window starts at max(window end, max_j s_j).
//
// Evaluation. Accepted cells are staged in LDS with their moments (loaded by the accepting lane);
// opened leaves as body ranges, whose bodies are then read as wave-uniform scalar loads. Lists are
// evaluated whenever the next append would overflow them and once at the end.
//
// Scheduling. A persistent grid; each wave takes groups from an atomic counter until none are left.
#include <cfloat>

#include "kernels.hpp"
#include "tree.hpp"

namespace lhip {
namespace {

constexpr int kWalkWaves = 4;
constexpr int kWalkBS = 64 * kWalkWaves;
constexpr int kCellCap = 128;
constexpr int kRangeCap = 64;

struct Body {
    float x, y, z;
    int i;
    bool on;
    float sx, sy, sz, sp;
    uint32_t self;
};

template <bool POT, bool GUARD>
__device__ __forceinline__ void p2p(Body& b, float4 q, float e2) {
    const float dx = q.x - b.x, dy = q.y - b.y, dz = q.z - b.z;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    const float inv = (!GUARD || r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;
    const float mi = q.w * inv;
    const float f = mi * (inv * inv);
    b.sx = __fmaf_rn(f, dx, b.sx);
    b.sy = __fmaf_rn(f, dy, b.sy);
    b.sz = __fmaf_rn(f, dz, b.sz);
    if (POT) b.sp += mi;
}

template <bool POT, bool GUARD>
__device__ __forceinline__ void p2p_self(Body& b, float4 q, int j, float e2) {
    const float dx = q.x - b.x, dy = q.y - b.y, dz = q.z - b.z;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    const bool ok = j != b.i && (!GUARD || r2 > 0.f);
    const float inv = ok ? __builtin_amdgcn_rsqf(r2) : 0.f;
    const float mi = q.w * inv;
    const float f = mi * (inv * inv);
    b.sx = __fmaf_rn(f, dx, b.sx);
    b.sy = __fmaf_rn(f, dy, b.sy);
    b.sz = __fmaf_rn(f, dz, b.sz);
    if (POT) b.sp += mi;
    b.self += j == b.i ? 1u : 0u;
}

// Monopole + quadrupole about the com, d = com - x (reference src/cpu_backend.cpp:855-871):
//   a += M d/r^3 - Q d/r^5 + 5/2 (d.Q.d) d/r^7,  phi_sum += M/r + 1/2 d.Q.d/r^5.
template <bool POT, bool GUARD>
__device__ __forceinline__ void p2c(Body& b, float4 A, float4 B, float4 C, float e2) {
    const float dx = A.x - b.x, dy = A.y - b.y, dz = A.z - b.z;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    const float inv = (!GUARD || r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;
    const float inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
    // B = (qxx, qxy, qxz, qyy), C = (qyz, qzz, -, -)
    const float qx = B.x * dx + B.y * dy + B.z * dz;
    const float qy = B.y * dx + B.w * dy + C.x * dz;
    const float qz = B.z * dx + C.x * dy + C.y * dz;
    const float dqd = dx * qx + dy * qy + dz * qz;
    const float f = A.w * inv3 + 2.5f * dqd * inv5 * inv2;
    b.sx += f * dx - qx * inv5;
    b.sy += f * dy - qy * inv5;
    b.sz += f * dz - qz * inv5;
    if (POT) b.sp += A.w * inv + 0.5f * dqd * inv5;
}

__device__ __forceinline__ int prefix_max_incl(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v = max(v, u);
    }
    return v;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

template <bool POT, bool GUARD>
__global__ void __launch_bounds__(kWalkBS) walk_kernel(WalkArgs a) {
    __shared__ float4 sA[kWalkWaves][kCellCap];
    __shared__ float4 sB[kWalkWaves][kCellCap];
    __shared__ float4 sC[kWalkWaves][kCellCap];
    __shared__ int2 sR[kWalkWaves][kRangeCap];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int M = int(a.cnt_scan[a.n]);
    const int ngroups = int(a.gcount[a.n]);
    const float e2 = a.eps2;
    unsigned long long t_inter = 0, t_p2p = 0, t_p2c = 0, t_listed = 0;

    while (true) {
        int g = 0;
        if (lane == 0) g = int(atomicAdd(&a.counters->next_group, 1u));
        g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
        if (g >= ngroups) break;
        const int glo = int(a.gstart[g]), ghi = int(a.gstart[g + 1]);
        const bool chunked = ghi - glo > kGroupMax;
        for (int clo = glo; clo < ghi; clo += kGroupMax) {
            const int chi = min(ghi, clo + kGroupMax);
            const bool two = chi - clo > 64;
            Body b0, b1;
            b0.i = clo + lane;
            b0.on = b0.i < chi;
            b1.i = clo + 64 + lane;
            b1.on = two && b1.i < chi;
            {
                const float4 p = a.posm[b0.on ? b0.i : clo];
                b0.x = p.x; b0.y = p.y; b0.z = p.z;
                const float4 q = a.posm[b1.on ? b1.i : clo];
                b1.x = q.x; b1.y = q.y; b1.z = q.z;
            }
            b0.sx = b0.sy = b0.sz = b0.sp = 0.f;
            b1.sx = b1.sy = b1.sz = b1.sp = 0.f;
            b0.self = b1.self = 0;
            // Group box as the reference forms it (src/cpu_backend.cpp:606-612): centre and half
            // extent from the float min/max.
            const float mnx = wave_min(fminf(b0.x, b1.x)), mxx = wave_max(fmaxf(b0.x, b1.x));
            const float mny = wave_min(fminf(b0.y, b1.y)), mxy = wave_max(fmaxf(b0.y, b1.y));
            const float mnz = wave_min(fminf(b0.z, b1.z)), mxz = wave_max(fmaxf(b0.z, b1.z));
            const float bx = 0.5f * (mnx + mxx), by = 0.5f * (mny + mxy), bz = 0.5f * (mnz + mxz);
            const float hx = 0.5f * (mxx - mnx), hy = 0.5f * (mxy - mny), hz = 0.5f * (mxz - mnz);

            int nc = 0, nr = 0;
            uint32_t ncell_total = 0, nbody_total = 0;

            auto flush_cells = [&]() {
                for (int c = 0; c < nc; ++c) {
                    const float4 A = sA[w][c], B = sB[w][c], C = sC[w][c];
                    p2c<POT, GUARD>(b0, A, B, C, e2);
                    if (two) p2c<POT, GUARD>(b1, A, B, C, e2);
                }
                ncell_total += uint32_t(nc);
                nc = 0;
            };
            auto flush_ranges = [&]() {
                for (int r = 0; r < nr; ++r) {
                    const int2 R = sR[w][r];
                    const int lo = __builtin_amdgcn_readfirstlane(R.x), hi = __builtin_amdgcn_readfirstlane(R.y);
                    nbody_total += uint32_t(hi - lo);
                    if (lo < chi && clo < hi) {
                        for (int j = lo; j < hi; ++j) {
                            const float4 q = a.posm[j];
                            p2p_self<POT, GUARD>(b0, q, j, e2);
                            if (two) p2p_self<POT, GUARD>(b1, q, j, e2);
                        }
                    } else {
                        int j = lo;
                        for (; j + 4 <= hi; j += 4) {
                            const float4 q0 = a.posm[j], q1 = a.posm[j + 1], q2 = a.posm[j + 2], q3 = a.posm[j + 3];
                            p2p<POT, GUARD>(b0, q0, e2);
                            p2p<POT, GUARD>(b0, q1, e2);
                            p2p<POT, GUARD>(b0, q2, e2);
                            p2p<POT, GUARD>(b0, q3, e2);
                            if (two) {
                                p2p<POT, GUARD>(b1, q0, e2);
                                p2p<POT, GUARD>(b1, q1, e2);
                                p2p<POT, GUARD>(b1, q2, e2);
                                p2p<POT, GUARD>(b1, q3, e2);
                            }
                        }
                        for (; j < hi; ++j) {
                            const float4 q = a.posm[j];
                            p2p<POT, GUARD>(b0, q, e2);
                            if (two) p2p<POT, GUARD>(b1, q, e2);
                        }
                    }
                }
                nr = 0;
            };

            int cur = 0;
            while (cur < M) {
                const int k = cur + lane;
                const bool act = k < M;
                float4 T = make_float4(0.f, 0.f, 0.f, 0.f);
                NodeLink L{0, 0, 0, 0};
                if (act) {
                    T = a.trav[k];
                    L = a.link[k];
                }
                // A node holding bodies of this (sub)group is always opened: the reference's
                // "contains" test for whole groups; any overlap for chunks of an over-full group.
                const bool keep_open = chunked ? (L.lo < chi && clo < L.hi) : (L.lo <= clo && chi <= L.hi);
                const float dx = fmaxf(fabsf(T.x - bx) - hx, 0.f);
                const float dy = fmaxf(fabsf(T.y - by) - hy, 0.f);
                const float dz = fmaxf(fabsf(T.z - bz) - hz, 0.f);
                // Rounded as the reference build evaluates it: three products, then two sums.
                const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
                const bool acc = act && !keep_open && d2 >= T.w;
                const bool leaf = act && info_leaf(L.info);
                const bool term = acc || leaf;
                const int incl = prefix_max_incl(term ? L.skip : 0, lane);
                int excl = __shfl_up(incl, 1, 64);
                if (lane == 0) excl = 0;
                const bool visited = act && excl <= k;
                const bool small = leaf && (L.hi - L.lo) <= kSmallLeaf;
                const bool is_cell = visited && acc && !small;
                const bool is_range = visited && leaf && (!acc || small);

                const unsigned long long mc = __ballot(is_cell);
                const int ncnew = __popcll(mc);
                if (ncnew) {
                    if (nc + ncnew > kCellCap) flush_cells();
                    if (is_cell) {
                        const int slot = nc + int(lanes_below(mc));
                        const NodeMom& md = a.mom[k];
                        sA[w][slot] = make_float4(md.cx, md.cy, md.cz, md.m);
                        sB[w][slot] = make_float4(md.qxx, md.qxy, md.qxz, md.qyy);
                        sC[w][slot] = make_float4(md.qyz, md.qzz, 0.f, 0.f);
                    }
                    nc += ncnew;
                }
                const unsigned long long mr = __ballot(is_range);
                const int nrnew = __popcll(mr);
                if (nrnew) {
                    if (nr + nrnew > kRangeCap) flush_ranges();
                    if (is_range) sR[w][nr + int(lanes_below(mr))] = make_int2(L.lo, L.hi);
                    nr += nrnew;
                }
                const int wmax = __builtin_amdgcn_readlane(incl, 63);
                cur = max(cur + 64, wmax);
            }
            flush_cells();
            flush_ranges();

            const float G = a.G;
            const uint32_t listed = ncell_total + nbody_total;
            if (b0.on) {
                a.ax[b0.i] = G * b0.sx;
                a.ay[b0.i] = G * b0.sy;
                a.az[b0.i] = G * b0.sz;
                if (POT) a.pot[b0.i] = -G * b0.sp;
                if (a.cost) a.cost[b0.i] = float(listed - b0.self);
            }
            if (b1.on) {
                a.ax[b1.i] = G * b1.sx;
                a.ay[b1.i] = G * b1.sy;
                a.az[b1.i] = G * b1.sz;
                if (POT) a.pot[b1.i] = -G * b1.sp;
                if (a.cost) a.cost[b1.i] = float(listed - b1.self);
            }
            const uint32_t nb = uint32_t(b0.on) + uint32_t(b1.on);
            const uint32_t self = (b0.on ? b0.self : 0u) + (b1.on ? b1.self : 0u);
            t_listed += (unsigned long long)listed * nb;
            t_p2c += (unsigned long long)ncell_total * nb;
            t_p2p += (unsigned long long)nbody_total * nb - self;
            t_inter += (unsigned long long)listed * nb - self;
        }
    }
    t_inter = wave_sum_u64(t_inter);
    t_p2p = wave_sum_u64(t_p2p);
    t_p2c = wave_sum_u64(t_p2c);
    t_listed = wave_sum_u64(t_listed);
    if (lane == 0) {
        atomicAdd(&a.counters->interactions, t_inter);
        atomicAdd(&a.counters->p2p, t_p2p);
        atomicAdd(&a.counters->p2c, t_p2c);
        atomicAdd(&a.counters->listed, t_listed);
    }
}

__global__ void reset_walk(WalkCounters* __restrict__ c) {
    c->interactions = c->p2p = c->p2c = c->listed = 0;
    c->error = 0;
    c->next_group = 0;
}

}  // namespace

int walk_blocks_per_cu() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(walk_kernel<false, false>),
                                                     kWalkBS, 0) != hipSuccess || blocks < 1) {
        (void)hipGetLastError();
        blocks = 2;
    }
    return blocks;
}

void tree_walk(hipStream_t s, const WalkArgs& a, bool potential) {
    if (a.n <= 0) return;
    const bool guard = !(a.eps2 > 0.f);
    const int blocks = std::max(1, a.blocks);
    reset_walk<<<1, 1, 0, s>>>(a.counters);
    LHIP_LAUNCH_CHECK();
    if (potential) {
        if (guard) walk_kernel<true, true><<<blocks, kWalkBS, 0, s>>>(a);
        else walk_kernel<true, false><<<blocks, kWalkBS, 0, s>>>(a);
    } else {
        if (guard) walk_kernel<false, true><<<blocks, kWalkBS, 0, s>>>(a);
        else walk_kernel<false, false><<<blocks, kWalkBS, 0, s>>>(a);
    }
    LHIP_LAUNCH_CHECK();
}

}  // namespace lhip
