// Vector-Jacobian product of the tree force with frozen topology (no reference counterpart: the
// reference has no differentiation).
//
// The force of the mixed-precision walk on body i is a sum over its group's list:
//   body or monopole entries y (mass m):  a_i += m K(y - x_i),     K(d) = d / r^3, r^2 = |d|^2 + eps^2
//   multipole cells (com C, mass M, Q):   a_i += -grad Psi(d),     d = C - x_i,
//                                         Psi(d) = M / r + (1/2) d.Q.d / r^5 (the walk's p2c).
// With the lists, the tree and the decisions frozen, and a cotangent u_i on each a_i, the VJP is
//   point entries:  x_i gets -m T u_i,  y gets +m T u_i,   T = I / r^3 - 3 d d^T / r^5
//   cells:          x_i gets  H u_i,    C gets -H u_i,     H = Hessian of Psi,
//                   Q gets G = -(u d^T + d u^T) / (2 r^5) + (5/2) (d.u) d d^T / r^7
// with H u = M (3 d (d.u) / r^5 - u / r^3) + Q u / r^5 - 5 Qd (d.u) / r^7 - 5 d (Qd.u) / r^7
//            - (5/2) (d.Q.d) u / r^7 + (35/2) (d.Q.d) (d.u) d / r^9.
// A cell's com and quadrupole depend on its bodies: com = sum m_j x_j / M, and
// Q = sum m_j (3 s s^T - |s|^2 I), s = x_j - com (whose com derivative vanishes), so body j of cell c
// gets m_j [gC / M + (6 G - 2 tr(G) I) (x_j - com)]. Summed over the cells containing j, that is
// m_j (A + B x_j) with A, B accumulated from the root down (down_level) and applied at the leaves.
//
// Contention. The top of the tree (the first kHot nodes in level order) sits in nearly every
// group's list, so its cotangents go to a private copy per block (gpriv, blocks x kHot x 12) that
// hot_reduce adds up afterwards; deeper cells and bodies take global atomics directly.
//
// Evaluation is entry-major: each lane owns a list entry (one cell, or three bodies) and loops over
// the group's targets four at a time, whose positions and cotangents come from the slot registers
// by readlane. The lane keeps its entry's source-side sums; the target sides of four targets are
// reduced over the wave together (wave_sum4: 32- and 16-lane swaps, then DPP butterflies) and
// carried into FP64 once per list batch. The source sums go out as FP64 atomic adds per
// (group, entry), component-major through LDS so that neighbouring lanes hit neighbouring
// doubles (hardware float atomics on this library's coarse-grained device buffers).
#include <cfloat>
#include <type_traits>

#include "kernels.hpp"
#include "tree.hpp"
#include "wave.hpp"

namespace lhip {
namespace {

constexpr int kWaves = 4;
constexpr int kBS4 = 64 * kWaves;
constexpr int kCells = 64;     // cells per batch (one per lane)
constexpr int kPer = 3;        // body entries per lane in a batch
constexpr int kBodies = 64 * kPer;  // bodies (and monopole cells) per batch
constexpr int kStack = 896;

struct VjpLds {
    float4 cA[kCells];  // com relative to the group centre, mass
    float4 cB[kCells];  // qxx, qxy, qxz, qyy
    float2 cC[kCells];  // qyz, qzz
    int cI[kCells];     // node index
    float bP[3 * kBodies];  // scratch of a batch: the target sums (3 x 128), the monopole cotangents
    int bI[kBodies];        // body index; -(k + 2) for a monopole cell k
};

struct Target {
    float x, y, z, ux, uy, uz;  // position relative to the group centre, cotangent
    int i;
    double gx, gy, gz;          // FP64 totals of the target-side terms
};

__global__ void __launch_bounds__(kBS4, 4) tree_vjp_kernel(TreeVjpArgs a) {
    __shared__ VjpLds lds[kWaves];
    __shared__ int stk[kWaves][kStack];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    VjpLds& L = lds[w];
    int* st = stk[w];
    const int ngroups = int(a.gcount[a.n]);
    const float e2 = a.eps2;
    bool overflow = false;
    // the walk's stack budget for this tree depth (kernels_walk.hip)
    const int thr = kStack - 7 * (a.depth + 1) - 8;

    double* const priv = a.gpriv + std::size_t(blockIdx.x) * kHot * 12;
    auto cell_out = [&](int k) -> double* {  // quadrupole cell k: gC (3), G (6)
        return k < kHot ? priv + std::size_t(k) * 9 : a.gcell + std::size_t(k) * 9;
    };
    auto mono_out = [&](int k) -> double* {  // monopole cell k: gC (3)
        return k < kHot ? priv + 9 * kHot + std::size_t(k) * 3 : a.gmono + std::size_t(k) * 3;
    };

    // The walk's XCD-local scheduling over the order's ranges (kernels_walk.hip), or one counter.
    const bool ranged = a.ranges && a.order;
    const int rstart = ranged && lane < kWalkRanges ? min(ngroups, int(a.ranges[lane])) : ngroups;
    const int xcd = int(blockIdx.x) & (kWalkRanges - 1);
    int tried = 0;
    while (true) {
        int g = 0;
        if (ranged) {
            bool got = false;
            while (!got && tried < kWalkRanges) {
                const int r = (xcd + tried) & (kWalkRanges - 1);
                const int lo = __builtin_amdgcn_readlane(rstart, r), hi = __builtin_amdgcn_readlane(rstart, r + 1);
                int t = 0;
                if (lane == 0) t = int(atomicAdd(&a.counters->next_range[r][0], 1u));
                t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
                got = lo + t < hi;
                if (got) g = lo + t;
                else ++tried;
            }
            if (!got) break;
        } else {
            if (lane == 0) g = int(atomicAdd(&a.counters->next_group, 1u));
            g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
            if (g >= ngroups) break;
        }
        const int gu = a.order ? int(a.order[g]) : g;  // the g-th group to take
        const int glo = int(a.gstart[gu]), ghi = int(a.gstart[gu + 1]);
        const bool chunked = ghi - glo > kGroupMax;
        for (int clo = glo; clo < ghi; clo += kGroupMax) {
            const int chi = min(ghi, clo + kGroupMax);
            const int nbody = chi - clo;  // 1..128, wave-uniform
            Target tg[2];
            float mnx = FLT_MAX, mny = FLT_MAX, mnz = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX, mxz = -FLT_MAX;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int t = s * 64 + lane;
                tg[s].i = clo + min(t, nbody - 1);  // lanes past the chunk repeat its last body
                const double4 d = a.posd[tg[s].i];
                mnx = fminf(mnx, float(d.x)); mny = fminf(mny, float(d.y)); mnz = fminf(mnz, float(d.z));
                mxx = fmaxf(mxx, float(d.x)); mxy = fmaxf(mxy, float(d.y)); mxz = fmaxf(mxz, float(d.z));
                tg[s].gx = tg[s].gy = tg[s].gz = 0.0;
            }
            // The forward walk's group box (same float rounding), and its centre for relative
            // coordinates.
            mnx = wave_min(mnx); mny = wave_min(mny); mnz = wave_min(mnz);
            mxx = wave_max(mxx); mxy = wave_max(mxy); mxz = wave_max(mxz);
            const float bx = 0.5f * (mnx + mxx), by = 0.5f * (mny + mxy), bz = 0.5f * (mnz + mxz);
            const float hx = 0.5f * (mxx - mnx), hy = 0.5f * (mxy - mny), hz = 0.5f * (mxz - mnz);
            const double ccx = bx, ccy = by, ccz = bz;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const double4 d = a.posd[tg[s].i];
                tg[s].x = float(d.x - ccx); tg[s].y = float(d.y - ccy); tg[s].z = float(d.z - ccz);
                tg[s].ux = float(a.ux[tg[s].i]); tg[s].uy = float(a.uy[tg[s].i]); tg[s].uz = float(a.uz[tg[s].i]);
            }
            // Targets go four at a time (never across the two 64-target slots); wave_sum4 leaves
            // their sums in one row each, and lanes 0-2 of the row write them to out[3 t + c]. out
            // is the flush's own LDS, already unpacked into registers. Each target's owner lane
            // then adds its sums to its FP64 totals.
            auto put_sums = [&](float* out, int t, float Sx, float Sy, float Sz) {
                const int c = lane & 15;
                if (c < 3) out[3 * (t + sum4_target(lane >> 4)) + c] = c == 0 ? Sx : (c == 1 ? Sy : Sz);
            };
            auto take_sums = [&](const float* out) {
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) {
                    const int t = sl * 64 + lane;
                    if (t < nbody) {
                        tg[sl].gx += double(out[3 * t]);
                        tg[sl].gy += double(out[3 * t + 1]);
                        tg[sl].gz += double(out[3 * t + 2]);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            };

            int nc = 0, nb = 0;
            auto flush_cells = [&]() {
                __builtin_amdgcn_wave_barrier();
                // Cells entered the batch as node indices: their FP64 com and mass (relative to the
                // group centre) and quadrupole load here, one cell per lane.
                const bool valid = lane < nc;
                float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = A;
                float2 C = make_float2(0.f, 0.f);
                if (valid) {
                    const int k = L.cI[lane];
                    const NodeMomD& dm = a.momd[k];
                    const NodeMom& md = a.mom[k];
                    A = make_float4(float(dm.cx - ccx), float(dm.cy - ccy), float(dm.cz - ccz), float(dm.m));
                    B = make_float4(md.qxx, md.qxy, md.qxz, md.qyy);
                    C = make_float2(md.qyz, md.qzz);
                }
                const float qxx = B.x, qxy = B.y, qxz = B.z, qyy = B.w, qyz = C.x, qzz = C.y;
                float sCx = 0.f, sCy = 0.f, sCz = 0.f;
                float Gxx = 0.f, Gxy = 0.f, Gxz = 0.f, Gyy = 0.f, Gyz = 0.f, Gzz = 0.f;
                float* out = reinterpret_cast<float*>(L.cA);  // 640 floats (cA..cC)
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) {
                    const int tend = min(nbody - sl * 64, 64);
                    for (int t0 = 0; t0 < tend; t0 += 4) {
                        float Hx[4], Hy[4], Hz[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            Hx[k] = Hy[k] = Hz[k] = 0.f;
                            if (t0 + k >= tend) continue;
                            const Target& T = tg[sl];
                            const float x = readlane_f(T.x, t0 + k), y = readlane_f(T.y, t0 + k);
                            const float z = readlane_f(T.z, t0 + k), ux = readlane_f(T.ux, t0 + k);
                            const float uy = readlane_f(T.uy, t0 + k), uz = readlane_f(T.uz, t0 + k);
                            const float dx = A.x - x, dy = A.y - y, dz = A.z - z;
                            const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
                            const float inv = __builtin_amdgcn_rsqf(r2);
                            const float i2 = inv * inv, i3 = inv * i2, i5 = i3 * i2, i7 = i5 * i2, i9 = i7 * i2;
                            const float Qdx = qxx * dx + qxy * dy + qxz * dz;
                            const float Qdy = qxy * dx + qyy * dy + qyz * dz;
                            const float Qdz = qxz * dx + qyz * dy + qzz * dz;
                            const float Qux = qxx * ux + qxy * uy + qxz * uz;
                            const float Quy = qxy * ux + qyy * uy + qyz * uz;
                            const float Quz = qxz * ux + qyz * uy + qzz * uz;
                            const float dQd = dx * Qdx + dy * Qdy + dz * Qdz;
                            const float du = dx * ux + dy * uy + dz * uz;
                            const float Qdu = Qdx * ux + Qdy * uy + Qdz * uz;
                            // H u = c_d d + c_u u + i5 Q u - 5 i7 (d.u) Q d
                            const float c_d = 3.f * A.w * du * i5 - 5.f * Qdu * i7 + 17.5f * dQd * du * i9;
                            const float c_u = -A.w * i3 - 2.5f * dQd * i7;
                            const float c_q = -5.f * du * i7;
                            Hx[k] = c_d * dx + c_u * ux + i5 * Qux + c_q * Qdx;
                            Hy[k] = c_d * dy + c_u * uy + i5 * Quy + c_q * Qdy;
                            Hz[k] = c_d * dz + c_u * uz + i5 * Quz + c_q * Qdz;
                            sCx -= Hx[k]; sCy -= Hy[k]; sCz -= Hz[k];
                            const float h5 = -0.5f * i5, k7 = 2.5f * du * i7;
                            Gxx += 2.f * h5 * ux * dx + k7 * dx * dx;
                            Gyy += 2.f * h5 * uy * dy + k7 * dy * dy;
                            Gzz += 2.f * h5 * uz * dz + k7 * dz * dz;
                            Gxy += h5 * (ux * dy + dx * uy) + k7 * dx * dy;
                            Gxz += h5 * (ux * dz + dx * uz) + k7 * dx * dz;
                            Gyz += h5 * (uy * dz + dy * uz) + k7 * dy * dz;
                        }
                        put_sums(out, sl * 64 + t0, wave_sum4(Hx[0], Hx[1], Hx[2], Hx[3]),
                                 wave_sum4(Hy[0], Hy[1], Hy[2], Hy[3]), wave_sum4(Hz[0], Hz[1], Hz[2], Hz[3]));
                    }
                }
                take_sums(out);
                // Component-major atomics: the 9 values of each cell go through LDS (cA..cC) so that
                // neighbouring lanes add neighbouring doubles of one cell.
                float* tb = out;
                if (valid) {
                    float* q = tb + lane * 9;
                    q[0] = sCx; q[1] = sCy; q[2] = sCz; q[3] = Gxx; q[4] = Gxy;
                    q[5] = Gxz; q[6] = Gyy; q[7] = Gyz; q[8] = Gzz;
                }
                __builtin_amdgcn_wave_barrier();
                for (int item = lane; item < 9 * nc; item += 64) {
                    const int c = item / 9;
                    unsafeAtomicAdd(cell_out(L.cI[c]) + (item - 9 * c), double(tb[item]));
                }
                __builtin_amdgcn_wave_barrier();
                nc = 0;
            };
            auto flush_bodies = [&]() {
                __builtin_amdgcn_wave_barrier();
                float4 P[kPer];
                int J[kPer];
#pragma unroll
                for (int h = 0; h < kPer; ++h) {
                    const int e = h * 64 + lane;
                    const bool valid = e < nb;
                    J[h] = valid ? L.bI[e] : -1;
                    P[h] = make_float4(0.f, 0.f, 0.f, 0.f);
                }
                // Bodies and monopole cells entered the batch as indices (as in the walk): positions,
                // FP64 coms and masses load here, the batch's in one round trip.
#pragma unroll
                for (int h = 0; h < kPer; ++h) {
                    if (J[h] >= 0) {
                        const double4 d = a.posd[J[h]];
                        P[h] = make_float4(float(d.x - ccx), float(d.y - ccy), float(d.z - ccz), float(d.w));
                    } else if (J[h] < -1) {
                        const NodeMomD& dm = a.momd[-J[h] - 2];
                        P[h] = make_float4(float(dm.cx - ccx), float(dm.cy - ccy), float(dm.cz - ccz), float(dm.m));
                    }
                }
                float s[kPer][3];
#pragma unroll
                for (int h = 0; h < kPer; ++h) s[h][0] = s[h][1] = s[h][2] = 0.f;
                float* out = L.bP;
                __builtin_amdgcn_wave_barrier();
                // MASK: the batch holds bodies of the group itself, whose own entry adds nothing.
                auto run = [&](auto mask) {
                    constexpr bool MASK = decltype(mask)::value;
#pragma unroll
                    for (int sl = 0; sl < 2; ++sl) {
                        const int tend = min(nbody - sl * 64, 64);
                        for (int t0 = 0; t0 < tend; t0 += 4) {
                            float tx[4], ty[4], tz[4];
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                tx[k] = ty[k] = tz[k] = 0.f;
                                if (t0 + k >= tend) continue;
                                const Target& T = tg[sl];
                                const float x = readlane_f(T.x, t0 + k), y = readlane_f(T.y, t0 + k);
                                const float z = readlane_f(T.z, t0 + k), ux = readlane_f(T.ux, t0 + k);
                                const float uy = readlane_f(T.uy, t0 + k), uz = readlane_f(T.uz, t0 + k);
                                const int it = clo + sl * 64 + t0 + k;
#pragma unroll
                                for (int h = 0; h < kPer; ++h) {
                                    const float dx = P[h].x - x, dy = P[h].y - y, dz = P[h].z - z;
                                    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
                                    const float inv = __builtin_amdgcn_rsqf(r2);
                                    const float i2 = inv * inv, i3 = i2 * inv;
                                    const float m = (MASK && J[h] == it) ? 0.f : P[h].w;
                                    const float mi3 = m * i3;
                                    // w = m r^-3 (u - 3 r^-2 (d.u) d)
                                    const float kd = (-3.f * i2) * __fmaf_rn(dz, uz, __fmaf_rn(dy, uy, dx * ux));
                                    const float wx = mi3 * __fmaf_rn(kd, dx, ux), wy = mi3 * __fmaf_rn(kd, dy, uy),
                                                wz = mi3 * __fmaf_rn(kd, dz, uz);
                                    s[h][0] += wx; s[h][1] += wy; s[h][2] += wz;
                                    tx[k] -= wx; ty[k] -= wy; tz[k] -= wz;
                                }
                            }
                            put_sums(out, sl * 64 + t0, wave_sum4(tx[0], tx[1], tx[2], tx[3]),
                                     wave_sum4(ty[0], ty[1], ty[2], ty[3]), wave_sum4(tz[0], tz[1], tz[2], tz[3]));
                        }
                    }
                };
                bool own = false;
#pragma unroll
                for (int h = 0; h < kPer; ++h) own = own || (J[h] >= clo && J[h] < chi);
                if (__ballot(own)) run(std::true_type{});
                else run(std::false_type{});
                take_sums(out);
                // Bodies: consecutive entries are consecutive bodies of a leaf, so the adds coalesce.
#pragma unroll
                for (int h = 0; h < kPer; ++h)
                    if (J[h] >= 0) {
                        unsafeAtomicAdd(a.gsx + J[h], double(s[h][0]));
                        unsafeAtomicAdd(a.gsy + J[h], double(s[h][1]));
                        unsafeAtomicAdd(a.gsz + J[h], double(s[h][2]));
                    }
                // Monopole cells: compacted through LDS (bP, bI) and added component-major.
                unsigned long long mm[kPer];
                int mbase[kPer], nm = 0;
#pragma unroll
                for (int h = 0; h < kPer; ++h) {
                    mm[h] = __ballot(J[h] <= -2);
                    mbase[h] = nm;
                    nm += __popcll(mm[h]);
                }
                if (nm) {
                    float* mb = out;
#pragma unroll
                    for (int h = 0; h < kPer; ++h)
                        if (J[h] <= -2) {
                            const int slot = mbase[h] + int(lanes_below(mm[h]));
                            L.bI[slot] = -J[h] - 2;
                            mb[3 * slot] = s[h][0]; mb[3 * slot + 1] = s[h][1]; mb[3 * slot + 2] = s[h][2];
                        }
                    __builtin_amdgcn_wave_barrier();
                    for (int item = lane; item < 3 * nm; item += 64) {
                        const int c = item / 3;
                        unsafeAtomicAdd(mono_out(L.bI[c]) + (item - 3 * c), double(mb[item]));
                    }
                }
                __builtin_amdgcn_wave_barrier();
                nb = 0;
            };

            // Traversal: the forward walk's (kernels_walk.hip), same decisions.
            if (lane == 0) st[0] = 0;
            int top = 1;
            while (top > 0) {
                const int P = top <= thr - 7 ? min(64, top) : 1;
                __builtin_amdgcn_wave_barrier();
                const bool act = lane < P;
                const int k = act ? st[top - P + lane] : 0;
                __builtin_amdgcn_wave_barrier();
                top -= P;
                WalkNode nd{0.f, 0.f, 0.f, 0.f, 0, 0, 0, -1};
                // qrel loads with the node (one round trip), as in the walk; cells and monopoles enter
                // the batches as node indices.
                float qr = 0.f;
                if (act) {
                    nd = a.node[k];
                    if (a.quad_tau > 0.f) qr = a.mom[k].qrel;
                }
                const bool keep_open = chunked ? (nd.lo < chi && clo < nd.hi) : (nd.lo <= clo && chi <= nd.hi);
                const float dx = fmaxf(fabsf(nd.cx - bx) - hx, 0.f);
                const float dy = fmaxf(fabsf(nd.cy - by) - hy, 0.f);
                const float dz = fmaxf(fabsf(nd.cz - bz) - hz, 0.f);
                const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
                const bool acc = act && !keep_open && d2 >= nd.rcrit2;
                const bool leaf = act && info_leaf(nd.info);
                const bool small = leaf && (nd.hi - nd.lo) <= a.small_leaf;
                bool is_cell = acc && !small;
                const bool is_mono = is_cell && a.quad_tau > 0.f && qr < a.quad_tau * d2;
                is_cell = is_cell && !is_mono;
                const bool is_range = leaf && (!acc || small);
                const bool open_int = act && !acc && !leaf;

                // Pushes as the walk's: a wide step stays at or below thr, pushing back the opened
                // nodes whose children would pass it.
                const int nch = open_int ? info_nchild(nd.info) : 0;
                const int cend = prefix_sum_incl(nch, lane);
                const int pushed = __builtin_amdgcn_readlane(cend, 63);
                if (P > 1 && top + pushed > thr) {
                    const unsigned long long om = __ballot(open_int);
                    const int nopen = __popcll(om);
                    const int after = nopen - int(lanes_below(om)) - 1;
                    const bool fit = open_int && top + cend + after <= thr;
                    const unsigned long long fm = __ballot(fit);
                    const int nfit = __popcll(fm);
                    const int cfit = nfit ? __builtin_amdgcn_readlane(cend, 63 - __clzll(fm)) : 0;
                    if (fit)
                        for (int c = 0; c < nch; ++c) st[top + cend - nch + c] = nd.first + c;
                    else if (open_int)
                        st[top + cfit + int(lanes_below(om)) - nfit] = k;
                    top += cfit + nopen - nfit;
                } else {
                    if (top + pushed > kStack) {
                        overflow = true;
                        break;
                    }
                    for (int c = 0; c < nch; ++c) st[top + cend - nch + c] = nd.first + c;
                    top += pushed;
                }

                const unsigned long long mc = __ballot(is_cell);
                const int ncnew = __popcll(mc);
                if (ncnew) {
                    if (nc + ncnew > kCells) flush_cells();
                    if (is_cell) L.cI[nc + int(lanes_below(mc))] = k;
                    nc += ncnew;
                }
                const unsigned long long mm = __ballot(is_mono);
                const int nmnew = __popcll(mm);
                if (nmnew) {
                    if (nb + nmnew > kBodies) flush_bodies();
                    if (is_mono) L.bI[nb + int(lanes_below(mm))] = -(k + 2);
                    nb += nmnew;
                }
                const unsigned long long mr = __ballot(is_range);
                if (mr) {
                    const int len = is_range ? nd.hi - nd.lo : 0;
                    const int end = prefix_sum_incl(len, lane);
                    const int off = end - len;
                    const int total = __builtin_amdgcn_readlane(end, 63);
                    int done = 0;
                    while (done < total) {
                        if (nb == kBodies) flush_bodies();
                        const int take = min(kBodies - nb, total - done);
                        const int u0 = max(off, done), u1 = min(off + len, done + take);
                        for (int u = u0; u < u1; ++u) L.bI[nb + (u - done)] = nd.lo + (u - off);
                        nb += take;
                        done += take;
                    }
                }
            }
            if (nc) flush_cells();
            if (nb) flush_bodies();
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int t = s * 64 + lane;
                if (t < nbody) {
                    a.gtx[tg[s].i] = tg[s].gx;
                    a.gty[tg[s].i] = tg[s].gy;
                    a.gtz[tg[s].i] = tg[s].gz;
                }
            }
        }
    }
    if (overflow && lane == 0) atomicOr(&a.counters->error, 1);
}

// The per-block copies of the first kHot nodes (9 quadrupole-cell then 3 monopole-cell values
// per node) added into gcell and gmono: grid.y splits the copies into runs of kHotRun, each summed
// by one thread and added atomically; node_count bounds the nodes.
constexpr int kHotRun = 32;
__global__ void __launch_bounds__(kBS) hot_reduce(const double* __restrict__ gpriv, int blocks,
                                                  const uint32_t* __restrict__ node_count,
                                                  double* __restrict__ gcell, double* __restrict__ gmono) {
    const int t = blockIdx.x * kBS + threadIdx.x;
    if (t >= kHot * 12) return;
    const int k = t < 9 * kHot ? t / 9 : (t - 9 * kHot) / 3;
    if (k >= int(*node_count)) return;
    const int b0 = blockIdx.y * kHotRun, b1 = min(blocks, b0 + kHotRun);
    double s = 0;
#pragma unroll 8
    for (int b = b0; b < b1; ++b) s += gpriv[std::size_t(b) * kHot * 12 + t];
    if (s == 0.0) return;
    unsafeAtomicAdd(t < 9 * kHot ? gcell + t : gmono + (t - 9 * kHot), s);
}

__global__ void reset_vjp(WalkCounters* __restrict__ c) {
    c->interactions = c->p2p = c->p2c = c->listed = c->p2m = 0;
    c->error = 0;
    c->next_group = 0;
    for (int r = 0; r < kWalkRanges; ++r) c->next_range[r][0] = 0;
}

// Per-node A (3) and B (6: xx, xy, xz, yy, yz, zz) of the downward pass: the node's own terms
// B = 6 G - 2 tr(G) I and A = gC / M - B com, gC summed over the quadrupole (gc) and monopole
// (gm) entries of the node.
__device__ __forceinline__ void own_terms(const double* __restrict__ gc, const double* __restrict__ gm,
                                          const NodeMomD& md, double o[9]) {
    if (!(md.m > 0)) {
        for (int c = 0; c < 9; ++c) o[c] = 0.0;
        return;
    }
    const double t = gc[3] + gc[6] + gc[8];
    const double Bxx = 6 * gc[3] - 2 * t, Bxy = 6 * gc[4], Bxz = 6 * gc[5];
    const double Byy = 6 * gc[6] - 2 * t, Byz = 6 * gc[7], Bzz = 6 * gc[8] - 2 * t;
    o[0] = (gc[0] + gm[0]) / md.m - (Bxx * md.cx + Bxy * md.cy + Bxz * md.cz);
    o[1] = (gc[1] + gm[1]) / md.m - (Bxy * md.cx + Byy * md.cy + Byz * md.cz);
    o[2] = (gc[2] + gm[2]) / md.m - (Bxz * md.cx + Byz * md.cy + Bzz * md.cz);
    o[3] = Bxx; o[4] = Bxy; o[5] = Bxz; o[6] = Byy; o[7] = Byz; o[8] = Bzz;
}

__global__ void __launch_bounds__(kBS) down_root(const double* __restrict__ gcell, const double* __restrict__ gmono,
                                                 const NodeMomD* __restrict__ momd, double* __restrict__ cum) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double o[9];
        own_terms(gcell, gmono, momd[0], o);
        for (int c = 0; c < 9; ++c) cum[c] = o[c];
    }
}

// Children of the nodes of one level: cum(child) = cum(parent) + own(child). Eight lanes per parent
// (thread t: parent p0 + t / 8, child slots t mod 8, t mod 8 + 8, ...), so the children of a parent
// are computed side by side and their records read and written as contiguous runs.
__global__ void __launch_bounds__(kBS) down_level(int level, const uint32_t* __restrict__ levstart,
                                                  const WalkNode* __restrict__ node,
                                                  const double* __restrict__ gcell,
                                                  const double* __restrict__ gmono,
                                                  const NodeMomD* __restrict__ momd, double* __restrict__ cum) {
    const int p0 = int(levstart[level]), p1 = int(levstart[level + 1]);
    const long long slots = 8ll * (p1 - p0);
    for (long long t = (long long)blockIdx.x * kBS + threadIdx.x; t < slots; t += (long long)gridDim.x * kBS) {
        const int p = p0 + int(t >> 3);
        const WalkNode nd = node[p];
        const int nch = info_nchild(nd.info);
        if (info_leaf(nd.info)) continue;
        for (int ch = nd.first + int(t & 7); ch < nd.first + nch; ch += 8) {
            double o[9];
            own_terms(gcell + std::size_t(ch) * 9, gmono + std::size_t(ch) * 3, momd[ch], o);
            for (int c = 0; c < 9; ++c) cum[std::size_t(ch) * 9 + c] = cum[std::size_t(p) * 9 + c] + o[c];
        }
    }
}

// Each leaf's bodies j: the cell-moment term m_j (A + B x_j) of the leaf's cum, added to the
// body's target and source sides, and the whole scattered into lambda_x by id (the last step of
// the VJP; every body is in exactly one leaf; lambda holds (x, y, z) per body). A wave takes 64 consecutive nodes and lists its
// leaves in LDS (node, first body, exclusive prefix of the body counts); its lanes then take the
// listed leaves' bodies one each, lane f finding its leaf by a binary search over the prefixes, so
// every lane holds a body whatever the leaf sizes.
__global__ void __launch_bounds__(kBS) down_leaves(const uint32_t* __restrict__ node_count,
                                                   const WalkNode* __restrict__ node,
                                                   const double* __restrict__ cum, const double4* __restrict__ posd,
                                                   TreeVjpArgs a) {
    __shared__ int lq[kBS], llo[kBS], lex[kBS];
    const int lane = int(threadIdx.x & 63), base = int(threadIdx.x >> 6) * 64;
    const int p = int(blockIdx.x) * kBS + int(threadIdx.x);
    WalkNode nd{};
    if (p < int(*node_count)) nd = node[p];
    const bool leaf = p < int(*node_count) && info_leaf(nd.info);
    const unsigned long long m = __ballot(leaf);
    const int cnt = leaf ? nd.hi - nd.lo : 0;
    const int incl = prefix_sum_incl(cnt, lane);  // over lanes, so in leaf rank order
    if (leaf) {
        const int k = base + int(lanes_below(m));
        lq[k] = p;
        llo[k] = nd.lo;
        lex[k] = incl - cnt;
    }
    const int nleaf = __popcll(m), total = __builtin_amdgcn_readlane(incl, 63);
    __syncthreads();
    const uint32_t* __restrict__ id = a.id;
    const double *__restrict__ gtx = a.gtx, *__restrict__ gty = a.gty, *__restrict__ gtz = a.gtz;
    const double *__restrict__ gsx = a.gsx, *__restrict__ gsy = a.gsy, *__restrict__ gsz = a.gsz;
    double* __restrict__ lam = a.lam;
    for (int f = lane; f < total; f += 64) {
        int lo = 0, hi = nleaf;  // the last listed leaf whose prefix is <= f (it holds body f)
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (lex[base + mid] <= f) lo = mid;
            else hi = mid;
        }
        const int j = llo[base + lo] + (f - lex[base + lo]);
        const double* c = cum + std::size_t(lq[base + lo]) * 9;
        const double4 b = posd[j];
        double* l = lam + 3 * std::size_t(id[j]);
        l[0] += gtx[j] + (gsx[j] + b.w * (c[0] + c[3] * b.x + c[4] * b.y + c[5] * b.z));
        l[1] += gty[j] + (gsy[j] + b.w * (c[1] + c[4] * b.x + c[6] * b.y + c[7] * b.z));
        l[2] += gtz[j] + (gsz[j] + b.w * (c[2] + c[5] * b.x + c[7] * b.y + c[8] * b.z));
    }
}

}  // namespace

void tree_vjp(hipStream_t s, const TreeVjpArgs& a, const uint32_t* levstart, const uint32_t* node_count, int nodes,
              double* cum) {
    if (a.n <= 0) return;
    reset_vjp<<<1, 1, 0, s>>>(a.counters);
    LHIP_LAUNCH_CHECK();
    tree_vjp_kernel<<<std::max(1, a.blocks), kBS4, 0, s>>>(a);
    LHIP_LAUNCH_CHECK();
    const int nb = std::max(1, a.blocks);
    hot_reduce<<<dim3(grid_for(kHot * 12, kBS), (nb + kHotRun - 1) / kHotRun), kBS, 0, s>>>(a.gpriv, nb, node_count,
                                                                                         a.gcell, a.gmono);
    LHIP_LAUNCH_CHECK();
    down_root<<<1, 64, 0, s>>>(a.gcell, a.gmono, a.momd, cum);
    LHIP_LAUNCH_CHECK();
    const unsigned g = std::min<unsigned>(grid_for(8ll * a.n, kBS), 2048u);
    for (int l = 0; l < a.depth; ++l) {  // the deepest level holds leaves only
        down_level<<<g, kBS, 0, s>>>(l, levstart, a.node, a.gcell, a.gmono, a.momd, cum);
        LHIP_LAUNCH_CHECK();
    }
    down_leaves<<<grid_for(nodes, kBS), kBS, 0, s>>>(node_count, a.node, cum, a.posd, a);
    LHIP_LAUNCH_CHECK();
}

int tree_vjp_blocks_per_cu() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(tree_vjp_kernel), kBS4,
                                                     0) != hipSuccess ||
        blocks < 1) {
        (void)hipGetLastError();
        blocks = 2;
    }
    return blocks;
}

}  // namespace lhip
