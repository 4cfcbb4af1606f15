// Tree walk: one wavefront per walk group, a parallel depth-first traversal of the level-order
// octree, interaction lists buffered in LDS and evaluated with every lane busy.
//
// Reference: walk / build_list / eval_group (src/cpu_backend.cpp:667-689, 816-882). Same
// acceptance test against the group's bounding box (a node is accepted when it does not contain
// the group and dist^2(box, com) >= rcrit2), same treatment of leaves (opened leaves and accepted
// leaves of <= 4 bodies contribute bodies, other accepted nodes their monopole + quadrupole), same
// self exclusion by index. The list a group gets is therefore the reference's list.
//
// Traversal. Nodes are in level order, so the children of a node are one contiguous block. The wave
// keeps a stack of pending nodes in LDS; each step pops up to 64 of them (one per lane), tests them
// in parallel, compacts accepted cells and opened leaves into the lists (ballot + lane prefix), and
// pushes the child blocks of opened internal nodes (lane prefix sum of child counts). The number
// popped is limited so that the stack cannot outgrow its LDS array: a step popping P nodes grows it
// by at most 7P, so P is chosen to keep it below a threshold; above it the wave pops one node at a
// time (a plain depth-first walk), which adds at most 7 per level of the tree. The stack holds the
// same nodes the reference's depth-first build_list visits, so the lists are the same sets.
//
// Evaluation. Accepted cells are staged in LDS with their moments, bodies of opened leaves with
// their positions. Walk groups hold 1-128 bodies (39 on average for a Plummer sphere), so one body
// per lane would idle most lanes. Instead each lane holds up to 3 target slots (2 with the potential
// or the self test), and the chunk's bodies are split into at most two parts, each on its own slots
// of every lane: a part of n bodies on ns slots is taken by lps = ceil(n / ns) lanes per list share,
// lane l of a share holding the part's bodies l, l + lps, ... (one LDS read of an entry feeds all of
// them), and S = min(16, 64 / lps) shares divide the list entries between them (share s takes
// entries s, s + S, ...). The split and the slots per part come from a table by chunk size
// (walk_eval_plan: the fewest slot-steps per list entry, each loop step weighted a quarter slot-step
// more); two parts fill the lanes where one would leave many slots idle. Partial sums of the shares
// are added in LDS at the end. Lists are evaluated whenever the next append would overflow them.
// Before an evaluation 15 zero-mass entries are appended to the list (S <= 16), so the loops carry
// no bounds test; a zero-mass entry adds exactly zero. With eps > 0 a body's own entry adds exactly
// zero too (d = 0), so only the potential and eps = 0 need the per-pair self test; every body of a
// group meets itself exactly once (the leaves holding the group are always opened), which is how
// the self pairs are taken out of the interaction counts.
//
// Walk options. small_leaf (the reference's 4) is the largest accepted leaf summed body by body;
// quad_tau > 0 turns accepted cells whose quadrupole term is small, qrel = 1.5 |Q|_F / M <
// quad_tau * dist^2(box, com), into point masses in the body list (an error-budgeted far field; 0,
// the reference's choice, keeps every quadrupole).
//
// Mixed-precision form (REL). Positions and node centres of mass are FP64. The traversal is the
// same (float group box and float node records); the evaluation runs on FP32 coordinates relative
// to the group's box centre c: a body enters as float(x - c), a cell as float(com - c), both
// differences formed in FP64, so a pair's separation keeps FP32 relative precision wherever the
// group sits. Each slot's FP32 partial sums are folded into FP64 after every list evaluation (at
// most 64 cells or 128 bodies per share), and the shares are added in FP64.
//
// Remote trees (LETs from other ranks, src/backend.hpp:78-83; src/cpu_backend.cpp:667-725, 788-814).
// Each group continues its lists with every remote tree, walked with the same stack and staging: a
// terminal node (no children, no bodies) or one whose com is at rcrit2 or beyond is accepted (its
// bodies if it has 1 to small_leaf of them, else its multipole); an opened node contributes its
// bodies if it has any, else its children. Remote bodies carry index -1, so the self test never
// drops one. The direct method's remote pass runs the same kernel over blocks of kGroupMax
// consecutive bodies with the remote trees only and adds to the direct sum.
//
// Scheduling. A persistent grid; each wave takes groups from an atomic counter until none are left.
#include <array>
#include <cfloat>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"
#include "tree.hpp"
#include "wave.hpp"

namespace lhip {
namespace {

constexpr int kWalkWaves = kWalkBlockWaves;
constexpr int kWalkBS = 64 * kWalkWaves;
constexpr int kCellCap = 64;    // cells staged per wave
constexpr int kBodyCap = 128;   // bodies staged per wave (>= 128: also the share-reduction buffer)
constexpr int kMaxShare = 16;

struct Slot {
    float x, y, z;
    int i;       // body index
    int bi;      // body within the chunk
    bool on;
    float sx, sy, sz, sp_pot;
    double dx, dy, dz, dp;  // FP64 carry of the mixed form
};

struct D4 {
    double x, y, z, w;
};

constexpr int kPad = kMaxShare - 1;  // zero entries appended to a list before it is evaluated

struct CellL {
    float4 A;  // com, mass
    float4 B;  // qxx, qxy, qxz, qyy
    float2 C;  // qyz, qzz
    float2 pad;
};

struct WaveLds {
    CellL c[kCellCap + kPad];
    float4 bP[kBodyCap + kPad];  // position, mass
    int bI[kBodyCap + kPad];     // body index (-1 for padding)
};

// MASK: exclude j == i (and, with GUARD, coincident pairs); needed for the potential and eps = 0.
template <bool POT, bool MASK, bool GUARD>
__device__ __forceinline__ void p2p(Slot& b, float4 q, int j, float e2) {
    const float dx = q.x - b.x, dy = q.y - b.y, dz = q.z - b.z;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    float inv = __builtin_amdgcn_rsqf(r2);
    if (MASK) inv = (j != b.i && (!GUARD || r2 > 0.f)) ? inv : 0.f;
    const float mi = q.w * inv;
    const float f = mi * (inv * inv);
    b.sx = __fmaf_rn(f, dx, b.sx);
    b.sy = __fmaf_rn(f, dy, b.sy);
    b.sz = __fmaf_rn(f, dz, b.sz);
    if (POT) b.sp_pot += mi;
}

// Monopole + quadrupole about the com, d = com - x (reference src/cpu_backend.cpp:855-871):
//   a += M d/r^3 - Q d/r^5 + 5/2 (d.Q.d) d/r^7,  phi_sum += M/r + 1/2 d.Q.d/r^5.
template <bool POT, bool GUARD>
__device__ __forceinline__ void p2c(Slot& b, float4 A, float4 B, float2 C, float e2) {
    const float dx = A.x - b.x, dy = A.y - b.y, dz = A.z - b.z;
    const float r2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, __fmaf_rn(dx, dx, e2)));
    float inv = __builtin_amdgcn_rsqf(r2);
    if (GUARD) inv = r2 > 0.f ? inv : 0.f;
    const float inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
    const float qx = B.x * dx + B.y * dy + B.z * dz;
    const float qy = B.y * dx + B.w * dy + C.x * dz;
    const float qz = B.z * dx + C.x * dy + C.y * dz;
    const float dqd = dx * qx + dy * qy + dz * qz;
    const float t = dqd * inv5;
    const float f = __fmaf_rn(2.5f * t, inv2, A.w * inv3);
    b.sx = __fmaf_rn(-inv5, qx, __fmaf_rn(f, dx, b.sx));
    b.sy = __fmaf_rn(-inv5, qy, __fmaf_rn(f, dy, b.sy));
    b.sz = __fmaf_rn(-inv5, qz, __fmaf_rn(f, dz, b.sz));
    if (POT) b.sp_pot += __fmaf_rn(0.5f, t, A.w * inv);
}

// The lane's slots OFF .. OFF + NS - 1 (one part of the chunk's plan) evaluate entries sp, sp + S,
// ... < steps * S of their list share sp (the list is padded to that length): one LDS read of an
// entry feeds every slot of the part.
template <int NS, int OFF, bool POT, bool GUARD, int NX>
__device__ __forceinline__ void eval_cells(Slot (&sl)[NX], int sp, int S, int nc, const WaveLds& L, float e2) {
    static_assert(OFF + NS <= NX, "slots of the part");
    const int steps = (nc + S - 1) / S;
    for (int t = 0; t < steps; ++t) {
        const CellL& c = L.c[t * S + sp];
        const float4 A = c.A, B = c.B;
        const float2 C = c.C;
#pragma unroll
        for (int s = 0; s < NS; ++s) p2c<POT, GUARD>(sl[OFF + s], A, B, C, e2);
    }
}

template <int NS, int OFF, bool POT, bool MASK, bool GUARD, int NX>
__device__ __forceinline__ void eval_bodies(Slot (&sl)[NX], int sp, int S, int nb, const WaveLds& L, float e2) {
    static_assert(OFF + NS <= NX, "slots of the part");
    const int steps = (nb + S - 1) / S;
    for (int t = 0; t < steps; ++t) {
        const int e = t * S + sp;
        const float4 q = L.bP[e];
        const int j = MASK ? L.bI[e] : 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) p2p<POT, MASK, GUARD>(sl[OFF + s], q, j, e2);
    }
}

constexpr int kStackCap = kWalkStack;
// Stack budget: a step popping P nodes pushes at most 8P, so P is held to (thr - top) / 7 with thr
// = capacity - 7 (depth + 1) - 8; at or above thr one node per step adds at most 7 per tree level.
// depth is the deepest level of the tree walked (the build reports it; 21 x key frames bounds it).
__device__ __forceinline__ int stack_threshold(int depth, int cap = kStackCap) {
    return cap - 7 * (depth + 1) - 8;
}

// What a local node contributes to the list of the bodies [clo, chi) with box centre b and half
// extent h (reference build_list, src/cpu_backend.cpp:816-849): a quadrupole cell, a monopole (the
// error-budgeted far field), the bodies of a leaf, or its children (opened internal node). A node
// holding bodies of this (sub)group is always opened: the reference's "contains" test for whole
// groups; any overlap for chunks of an over-full group. The distance is rounded as the reference
// build evaluates it: three products, then two sums. Both the list traversal and the fused walk
// decide through this one function, so their lists are the same.
struct Pick {
    bool cell, mono, range, open;
};
__device__ __forceinline__ Pick decide(bool act, const WalkNode& nd, float qr, float bx, float by, float bz, float hx,
                                       float hy, float hz, int clo, int chi, bool chunked, int small_leaf,
                                       float quad_tau) {
    const bool keep_open = chunked ? (nd.lo < chi && clo < nd.hi) : (nd.lo <= clo && chi <= nd.hi);
    const float dx = fmaxf(fabsf(nd.cx - bx) - hx, 0.f);
    const float dy = fmaxf(fabsf(nd.cy - by) - hy, 0.f);
    const float dz = fmaxf(fabsf(nd.cz - bz) - hz, 0.f);
    const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    const bool acc = act && !keep_open && d2 >= nd.rcrit2;
    const bool leaf = act && info_leaf(nd.info);
    const bool small = leaf && (nd.hi - nd.lo) <= small_leaf;
    const bool cell = acc && !small;
    Pick r;
    r.mono = cell && quad_tau > 0.f && qr < quad_tau * d2;
    r.cell = cell && !r.mono;
    r.range = leaf && (!acc || small);
    r.open = act && !acc && !leaf;
    return r;
}

// ---------------------------------------------------------------- interaction lists
//
// walk_lists_kernel traverses like walk_kernel (same stack discipline, same decisions) but only
// writes each group's list: accepted cells, monopoles and opened leaves as tagged node indices. The
// walk then evaluates the lists. Split this way, the traversal runs at twice the waves per SIMD of
// the evaluation (it holds no list buffers and few registers), so its dependent node loads
// overlap each other, and the evaluation streams its lists with no traversal latency between the
// pair loops.

constexpr int kListWaves = 4;
constexpr int kListBS = 64 * kListWaves;
constexpr int kListStack = kStackCap;  // traversal stack entries per wave

// Pool chunks a wave takes from the shared counter at a time (one atomic per slab, not per chunk:
// every wave of the chip adds to that one counter).
constexpr uint32_t kListSlab = 8;

// Appends the entries of the lanes with `has` to the unit list being written (chunk, fill), taking
// a new pool chunk whenever the current one fills up, from the wave's slab [slab, slab_end);
// writes beyond the pool are dropped (the chunk counter then exceeds the pool's and the host runs
// the walk again with a larger pool).
struct ListWriter {
    uint32_t chunk, fill, slab, slab_end;
    __device__ __forceinline__ uint32_t take(const WalkArgs& a, int lane) {
        if (slab == slab_end) {
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(&a.counters->list_chunks, kListSlab);
            slab = uint32_t(__builtin_amdgcn_readfirstlane(__shfl(int(c), 0, 64)));
            slab_end = slab + kListSlab;
        }
        return slab++;
    }
    __device__ __forceinline__ void append(const WalkArgs& a, int lane, bool has, uint32_t ent) {
        const unsigned long long m = __ballot(has);
        const uint32_t cnt = uint32_t(__popcll(m));
        if (!cnt) return;
        uint32_t next = chunk;
        if (fill + cnt >= uint32_t(kListChunk)) {
            next = take(a, lane);
            if (lane == 0 && chunk < a.list_cap && next < a.list_cap) a.lnext[chunk] = next;
        }
        if (has) {
            const uint32_t slot = fill + lanes_below(m);
            const uint32_t c = slot < uint32_t(kListChunk) ? chunk : next;
            if (c < a.list_cap) a.lpool[std::size_t(c) * kListChunk + (slot & (kListChunk - 1))] = ent;
        }
        fill += cnt;
        if (fill >= uint32_t(kListChunk)) {
            fill -= kListChunk;
            chunk = next;
        }
    }
};

template <bool REL>
__global__ void __launch_bounds__(kListBS, 8) walk_lists_kernel(WalkArgs a) {
    __shared__ int stk[kListWaves][kListStack];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int* st = stk[w];
    const int ngroups = int(a.gcount[a.n]);
    const int stack_thr = stack_threshold(a.depth, kListStack);
    const bool need_qrel = a.quad_tau > 0.f;
    bool overflow = false;
    ListWriter wr;
    wr.slab = wr.slab_end = 0;
    while (!overflow) {
        int g = 0;
        if (lane == 0) g = int(atomicAdd(&a.counters->next_list, 1u));
        g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
        if (g >= ngroups) break;
        const int gu = a.order ? int(a.order[g]) : g;
        const int glo = int(a.gstart[gu]), ghi = int(a.gstart[gu + 1]);
        const bool chunked = ghi - glo > kGroupMax;
        {
            wr.chunk = wr.take(a, lane);
            wr.fill = 0;
            if (lane == 0 && wr.chunk < a.list_cap) a.lhead[gu] = wr.chunk;
        }
        for (int clo = glo; clo < ghi && !overflow; clo += kGroupMax) {
            const int chi = min(ghi, clo + kGroupMax);
            const int nbody = chi - clo;
            // The group box of walk_kernel: float min/max over the chunk's bodies.
            float mnx = FLT_MAX, mny = FLT_MAX, mnz = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX, mxz = -FLT_MAX;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int v = t * 64 + lane;
                if (v < nbody) {
                    float px, py, pz;
                    if (REL) {
                        const double4 d = a.posd_box[clo + v];
                        px = float(d.x); py = float(d.y); pz = float(d.z);
                    } else {
                        const float4 q = a.posm[clo + v];
                        px = q.x; py = q.y; pz = q.z;
                    }
                    mnx = fminf(mnx, px); mny = fminf(mny, py); mnz = fminf(mnz, pz);
                    mxx = fmaxf(mxx, px); mxy = fmaxf(mxy, py); mxz = fmaxf(mxz, pz);
                }
            }
            mnx = wave_min(mnx); mny = wave_min(mny); mnz = wave_min(mnz);
            mxx = wave_max(mxx); mxy = wave_max(mxy); mxz = wave_max(mxz);
            const float bx = 0.5f * (mnx + mxx), by = 0.5f * (mny + mxy), bz = 0.5f * (mnz + mxz);
            const float hx = 0.5f * (mxx - mnx), hy = 0.5f * (mxy - mny), hz = 0.5f * (mxz - mnz);
            // The unit's count entry, filled in at its end.
            const uint32_t hchunk = wr.chunk, hslot = wr.fill;
            wr.append(a, lane, lane == 0, 0u);
            uint32_t count = 0;
            if (lane == 0) st[0] = 0;  // the root
            int top = 1;
            while (top > 0) {
                const int room = (stack_thr - top) / 7;
                const int P = min(64, min(top, max(1, room)));
                __builtin_amdgcn_wave_barrier();
                const bool act = lane < P;
                const int k = act ? st[top - P + lane] : 0;
                __builtin_amdgcn_wave_barrier();
                top -= P;
                WalkNode nd{0.f, 0.f, 0.f, 0.f, 0, 0, 0, -1};
                float qr = 0.f;
                if (act) {
                    nd = a.node[k];
                    if (need_qrel) qr = a.mom[k].qrel;
                }
                const Pick pk = decide(act, nd, qr, bx, by, bz, hx, hy, hz, clo, chi, chunked, a.small_leaf,
                                       a.quad_tau);
                const int nch = pk.open ? info_nchild(nd.info) : 0;
                const int cend = prefix_sum_incl(nch, lane);
                const int pushed = __builtin_amdgcn_readlane(cend, 63);
                if (top + pushed > kListStack) {
                    overflow = true;
                    break;
                }
                for (int c = 0; c < nch; ++c) st[top + cend - nch + c] = nd.first + c;
                top += pushed;
                const bool has = pk.cell || pk.mono || pk.range;
                const uint32_t tag = pk.cell ? kEntCell : (pk.mono ? kEntMono : kEntRange);
                count += uint32_t(__popcll(__ballot(has)));
                wr.append(a, lane, has, tag | uint32_t(k));
            }
            if (lane == 0 && hchunk < a.list_cap) a.lpool[std::size_t(hchunk) * kListChunk + hslot] = count;
        }
    }
    if (overflow && lane == 0) atomicOr(&a.counters->error, 1);
}

// LET: every group also walks every remote tree (a.let) after the local one, into the same lists
// (the reference's order: src/cpu_backend.cpp:667-689). With direct_blocks the groups are blocks of
// kGroupMax consecutive bodies that walk the remote trees only, and the forces, costs and counts are
// added to the direct sum's.
template <bool POT, bool GUARD, bool REL, bool LET, bool LISTS>
__global__ void __launch_bounds__(kWalkBS, 4) walk_kernel(WalkArgs a) {
    static_assert(!(LISTS && LET), "remote trees are walked, not listed");
    constexpr bool MASK = POT || GUARD;
    // Slots per lane at most: 3 (4 would spill at 4 waves per SIMD, the occupancy the bounds ask
    // for); 2 with the potential or the self test, whose extra registers would spill a third.
    constexpr int NX = MASK ? 2 : 3;
    // Plain form: a wave traverses at issue priority 1 and evaluates its lists at 0, so that the
    // traversal's short dependent steps issue as soon as their loads return while the other waves
    // of the SIMD evaluate (1M Plummer walk 1-1.5% faster; the mixed form measured 0.5-1% slower).
    constexpr bool kPrio = !REL && !LISTS;
    static_assert(sizeof(WaveLds) >= 64 * NX * (REL ? sizeof(D4) : sizeof(float4)), "share reduction buffer");
    __shared__ WaveLds lds[kWalkWaves];
    __shared__ int stk[kWalkWaves][LISTS ? 1 : kStackCap];  // the lists need no traversal stack
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    WaveLds& L = lds[w];
    int* st = stk[w];
    // A list traversal that overflowed its stack or the pool left no usable lists: the host raises,
    // or runs the walk again with a larger pool.
    if (LISTS && (a.counters->error != 0 || a.counters->list_chunks > a.list_cap)) return;
    const bool blocks = LET && a.direct_blocks;
    const int ngroups = blocks ? (a.n + kGroupMax - 1) / kGroupMax : int(a.gcount[a.n]);
    const float e2 = a.eps2;
    const int stack_thr = stack_threshold(a.depth);
    const int let_thr = LET ? a.let_thr : 0;
    unsigned long long t_inter = 0, t_p2p = 0, t_p2c = 0, t_listed = 0, t_p2m = 0;
    bool overflow = false;

    // With ranges, the waves of XCD x (blocks are placed on the XCDs round robin) take the positions
    // of range x first (its costly units, then the others in Morton order), so that the groups one
    // XCD walks at a time are neighbours and share its L2; then the other ranges. Lane r holds the
    // start of range r (loaded once: the counters' line is busy with the atomics).
    const bool ranged = a.ranges && !blocks && a.order;
    const int rstart = ranged && lane < kWalkRanges ? min(ngroups, int(a.counters->range_start[lane])) : ngroups;
    const int xcd = int(blockIdx.x) & (kWalkRanges - 1);
    int tried = 0;  // ranges found empty (wave-uniform)
    while (true) {
        if constexpr (kPrio) __builtin_amdgcn_s_setprio(1);
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
        const int gu = !blocks && a.order ? int(a.order[g]) : g;  // the g-th group to take
        const int glo = blocks ? g * kGroupMax : int(a.gstart[gu]);
        const int ghi = blocks ? min(a.n, glo + kGroupMax) : int(a.gstart[gu + 1]);
        const bool chunked = ghi - glo > kGroupMax;
        // List reader: chunk and offset in the pool (indices clamped into the pool).
        uint32_t lc = 0, loff = 0;
        if (LISTS) lc = min(a.lhead[gu], a.list_cap - 1u);
        for (int clo = glo; clo < ghi; clo += kGroupMax) {
            const int chi = min(ghi, clo + kGroupMax);
            const int nbody = chi - clo;  // 1..128, wave-uniform
            // The chunk's evaluation plan (walk_eval_plan): part A, bodies [0, na), on slots
            // 0 .. nsA - 1 with SA list shares; part B, bodies [na, nbody), on slots nsA .. nsA + nsB - 1
            // with SB shares (nsB = 0: no part B). A share of part P is taken by lps_P =
            // ceil(n_P / ns_P) lanes, lane l of it holding the part's bodies l, l + lps_P, ... in its
            // slots of the part, so those slots read the same entries.
            const uint32_t pw = a.plan[nbody - 1];
            const int na = int(pw & 0xffu), nsA = int((pw >> 8) & 0xfu), SA = int((pw >> 12) & 0x3fu);
            const int nsB = int((pw >> 18) & 0xfu), SB = int((pw >> 22) & 0x3fu);
            const int nbB = nbody - na;
            const int lpsA = (na + nsA - 1) / nsA, lpsB = nsB ? (nbB + nsB - 1) / nsB : 64;
            // The lane's share in each part; lanes past a part's shares read share 0 and their
            // slots of that part are discarded.
            int spA = lane / lpsA, spB = lane / lpsB;
            const bool onA = spA < SA, onB = nsB > 0 && spB < SB;
            if (!onA) spA = 0;
            if (!onB) spB = 0;
            Slot sl[NX];
            float mnx = FLT_MAX, mny = FLT_MAX, mnz = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX, mxz = -FLT_MAX;
#pragma unroll
            for (int s = 0; s < NX; ++s) {
                // Slot s < nsA of lane l: body (l mod lpsA) + s lpsA; slot nsA + t: body na +
                // (l mod lpsB) + t lpsB. Off slots take body (s 64 + l) mod nbody (a valid load whose
                // sums are dropped). Share 0 of each part holds every body of the part, so the box
                // below takes every body of the chunk.
                const bool inA = s < nsA;
                const int b = inA ? lane % lpsA + s * lpsA : na + lane % lpsB + (s - nsA) * lpsB;
                sl[s].on = inA ? onA && b < na : onB && s < nsA + nsB && b < nbody;
                sl[s].bi = sl[s].on ? b : (s * 64 + lane) % nbody;
                sl[s].i = clo + sl[s].bi;
                float4 p;
                if (REL) {
                    const double4 d = a.posd_box[sl[s].i];
                    p = make_float4(float(d.x), float(d.y), float(d.z), 0.f);
                } else {
                    p = a.posm[sl[s].i];
                }
                sl[s].x = p.x; sl[s].y = p.y; sl[s].z = p.z;
                sl[s].sx = sl[s].sy = sl[s].sz = sl[s].sp_pot = 0.f;
                sl[s].dx = sl[s].dy = sl[s].dz = sl[s].dp = 0.0;
                mnx = fminf(mnx, p.x); mny = fminf(mny, p.y); mnz = fminf(mnz, p.z);
                mxx = fmaxf(mxx, p.x); mxy = fmaxf(mxy, p.y); mxz = fmaxf(mxz, p.z);
            }
            // Group box as the reference forms it (src/cpu_backend.cpp:606-612): centre and half
            // extent from the float min/max (every slot holds a body of the chunk).
            mnx = wave_min(mnx); mny = wave_min(mny); mnz = wave_min(mnz);
            mxx = wave_max(mxx); mxy = wave_max(mxy); mxz = wave_max(mxz);
            const float bx = 0.5f * (mnx + mxx), by = 0.5f * (mny + mxy), bz = 0.5f * (mnz + mxz);
            const float hx = 0.5f * (mxx - mnx), hy = 0.5f * (mxy - mny), hz = 0.5f * (mxz - mnz);
            // Mixed form: coordinates relative to c = the box centre.
            const double ccx = bx, ccy = by, ccz = bz;
            if (REL) {
#pragma unroll
                for (int s = 0; s < NX; ++s) {
                    const double4 d = a.posd[sl[s].i];
                    sl[s].x = float(d.x - ccx);
                    sl[s].y = float(d.y - ccy);
                    sl[s].z = float(d.z - ccz);
                }
            }
            auto fold = [&]() {
                if (REL) {
#pragma unroll
                    for (int s = 0; s < NX; ++s) {
                        sl[s].dx += double(sl[s].sx); sl[s].dy += double(sl[s].sy);
                        sl[s].dz += double(sl[s].sz); sl[s].dp += double(sl[s].sp_pot);
                        sl[s].sx = sl[s].sy = sl[s].sz = sl[s].sp_pot = 0.f;
                    }
                }
            };

            // A local body's list entry: its position (relative to c in the mixed form) and mass.
            auto body_at = [&](int j) {
                if (REL) {
                    const double4 d = a.posd[j];
                    return make_float4(float(d.x - ccx), float(d.y - ccy), float(d.z - ccz), float(d.w));
                }
                return a.posm[j];
            };
            // Mixed form: a node's FP64 com and mass relative to c, as a monopole or a cell enters.
            auto com_at = [&](int k) {
                const NodeMomD& dm = a.momd[k];
                return make_float4(float(dm.cx - ccx), float(dm.cy - ccy), float(dm.cz - ccz), float(dm.m));
            };

            int nc = 0, nb = 0;
            uint32_t ncell_total = 0, nbody_total = 0, nmono_total = 0;  // body list = bodies + monopoles
            // Mixed form: a local cell enters the cell list as its node index (pad.x); its FP64 com,
            // mass and quadrupole are loaded here, one cell per lane. Remote cells (pad.x = -1) were
            // written whole.
            auto flush_cells = [&]() {
                if constexpr (REL) {
                    const int k = lane < nc ? __float_as_int(L.c[lane].pad.x) : -1;
                    if (k >= 0) {
                        const float4 A = com_at(k);
                        const NodeMom& md = a.mom[k];
                        const float4 B = make_float4(md.qxx, md.qxy, md.qxz, md.qyy);
                        const float2 C = make_float2(md.qyz, md.qzz);
                        CellL& d = L.c[lane];
                        d.A = A;
                        d.B = B;
                        d.C = C;
                    }
                }
                if (lane < kPad) {
                    L.c[nc + lane].A = make_float4(0.f, 0.f, 0.f, 0.f);
                    L.c[nc + lane].B = make_float4(0.f, 0.f, 0.f, 0.f);
                    L.c[nc + lane].C = make_float2(0.f, 0.f);
                }
                __builtin_amdgcn_wave_barrier();
                if constexpr (kPrio) __builtin_amdgcn_s_setprio(0);
                if constexpr (NX >= 3) {
                    if (nsA == 3) eval_cells<3, 0, POT, GUARD>(sl, spA, SA, nc, L, e2);
                    else if (nsA == 2) eval_cells<2, 0, POT, GUARD>(sl, spA, SA, nc, L, e2);
                    else eval_cells<1, 0, POT, GUARD>(sl, spA, SA, nc, L, e2);
                    if (nsB == 2) eval_cells<2, 1, POT, GUARD>(sl, spB, SB, nc, L, e2);
                    else if (nsB == 1 && nsA == 2) eval_cells<1, 2, POT, GUARD>(sl, spB, SB, nc, L, e2);
                    else if (nsB == 1) eval_cells<1, 1, POT, GUARD>(sl, spB, SB, nc, L, e2);
                } else {
                    if (nsA == 2) eval_cells<2, 0, POT, GUARD>(sl, spA, SA, nc, L, e2);
                    else eval_cells<1, 0, POT, GUARD>(sl, spA, SA, nc, L, e2);
                    if (nsB == 1) eval_cells<1, 1, POT, GUARD>(sl, spB, SB, nc, L, e2);
                }
                fold();
                if constexpr (kPrio) __builtin_amdgcn_s_setprio(1);
                __builtin_amdgcn_wave_barrier();
                ncell_total += uint32_t(nc);
                nc = 0;
            };
            // A local body enters the body list as its index only (bI); its position is loaded here,
            // for the whole list in one round trip, just before the list is evaluated. So does a local
            // monopole of the mixed form (bI = -(node + 2)). Other monopoles and remote bodies
            // (bI = -1) were written with their positions.
            auto flush_bodies = [&]() {
                {
                    const int j0 = lane < nb ? L.bI[lane] : -1;
                    const int j1 = lane + 64 < nb ? L.bI[lane + 64] : -1;
                    float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
                    if (j0 >= 0) q0 = body_at(j0);
                    else if (REL && j0 < -1) q0 = com_at(-j0 - 2);
                    if (j1 >= 0) q1 = body_at(j1);
                    else if (REL && j1 < -1) q1 = com_at(-j1 - 2);
                    if (j0 >= 0 || (REL && j0 < -1)) L.bP[lane] = q0;
                    if (j1 >= 0 || (REL && j1 < -1)) L.bP[lane + 64] = q1;
                }
                if (lane < kPad) {
                    L.bP[nb + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
                    L.bI[nb + lane] = -1;
                }
                __builtin_amdgcn_wave_barrier();
                if constexpr (kPrio) __builtin_amdgcn_s_setprio(0);
                if constexpr (NX >= 3) {
                    if (nsA == 3) eval_bodies<3, 0, POT, MASK, GUARD>(sl, spA, SA, nb, L, e2);
                    else if (nsA == 2) eval_bodies<2, 0, POT, MASK, GUARD>(sl, spA, SA, nb, L, e2);
                    else eval_bodies<1, 0, POT, MASK, GUARD>(sl, spA, SA, nb, L, e2);
                    if (nsB == 2) eval_bodies<2, 1, POT, MASK, GUARD>(sl, spB, SB, nb, L, e2);
                    else if (nsB == 1 && nsA == 2) eval_bodies<1, 2, POT, MASK, GUARD>(sl, spB, SB, nb, L, e2);
                    else if (nsB == 1) eval_bodies<1, 1, POT, MASK, GUARD>(sl, spB, SB, nb, L, e2);
                } else {
                    if (nsA == 2) eval_bodies<2, 0, POT, MASK, GUARD>(sl, spA, SA, nb, L, e2);
                    else eval_bodies<1, 0, POT, MASK, GUARD>(sl, spA, SA, nb, L, e2);
                    if (nsB == 1) eval_bodies<1, 1, POT, MASK, GUARD>(sl, spB, SB, nb, L, e2);
                }
                fold();
                if constexpr (kPrio) __builtin_amdgcn_s_setprio(1);
                __builtin_amdgcn_wave_barrier();
                nbody_total += uint32_t(nb);
                nb = 0;
            };

            // Staging into the LDS lists; get() and load() run on the staging lanes only.
            auto stage_cells = [&](bool is_cell, auto&& get) {
                const unsigned long long mc = __ballot(is_cell);
                const int ncnew = __popcll(mc);
                if (ncnew) {
                    if (nc + ncnew > kCellCap) flush_cells();
                    if (is_cell) {
                        const CellL c = get();
                        CellL& d = L.c[nc + int(lanes_below(mc))];
                        d.A = c.A;
                        d.B = c.B;
                        d.C = c.C;
                        if (REL) d.pad.x = __int_as_float(-1);
                    }
                    nc += ncnew;
                }
            };
            // Mixed form: local cells and monopoles enter as node indices (loaded by the flushes).
            auto stage_cells_at = [&](bool is_cell, int k) {
                const unsigned long long mc = __ballot(is_cell);
                const int ncnew = __popcll(mc);
                if (ncnew) {
                    if (nc + ncnew > kCellCap) flush_cells();
                    if (is_cell) L.c[nc + int(lanes_below(mc))].pad.x = __int_as_float(k);
                    nc += ncnew;
                }
            };
            auto stage_monos_at = [&](bool is_mono, int k) {
                const unsigned long long mm = __ballot(is_mono);
                const int nmnew = __popcll(mm);
                if (nmnew) {
                    if (nb + nmnew > kBodyCap) flush_bodies();
                    if (is_mono) L.bI[nb + int(lanes_below(mm))] = -(k + 2);
                    nb += nmnew;
                    nmono_total += uint32_t(nmnew);
                }
            };
            auto stage_monos = [&](bool is_mono, auto&& get) {
                const unsigned long long mm = __ballot(is_mono);
                const int nmnew = __popcll(mm);
                if (nmnew) {
                    if (nb + nmnew > kBodyCap) flush_bodies();
                    if (is_mono) {
                        const int slot = nb + int(lanes_below(mm));
                        L.bP[slot] = get();
                        L.bI[slot] = -1;
                    }
                    nb += nmnew;
                    nmono_total += uint32_t(nmnew);
                }
            };
            // Body ranges [lo, lo + len) of the staging lanes, in chunks of at most the free buffer
            // space. Local bodies (load null) enter as their indices, loaded when the list is
            // evaluated; remote ones (index -1: never the target) are loaded here.
            auto stage_ranges = [&](bool is_range, int lo, int len_in, auto&& load) {
                constexpr bool local = std::is_same_v<std::decay_t<decltype(load)>, std::nullptr_t>;
                const unsigned long long mr = __ballot(is_range);
                if (mr) {
                    const int len = is_range ? len_in : 0;
                    const int end = prefix_sum_incl(len, lane);
                    const int off = end - len;
                    const int total = __builtin_amdgcn_readlane(end, 63);
                    int done = 0;
                    while (done < total) {
                        if (nb == kBodyCap) flush_bodies();
                        const int take = min(kBodyCap - nb, total - done);
                        const int u0 = max(off, done), u1 = min(off + len, done + take);
                        if constexpr (local) {
                            for (int u = u0; u < u1; ++u) L.bI[nb + (u - done)] = lo + (u - off);
                        } else {
                            // Four loads in flight per lane before their LDS stores (a load past u1
                            // reads the range's last body again and is not stored).
                            for (int u = u0; u < u1; u += 4) {
                                float4 q[4];
#pragma unroll
                                for (int t = 0; t < 4; ++t) q[t] = load(lo + (min(u + t, u1 - 1) - off));
#pragma unroll
                                for (int t = 0; t < 4; ++t)
                                    if (u + t < u1) {
                                        L.bP[nb + (u + t - done)] = q[t];
                                        L.bI[nb + (u + t - done)] = -1;
                                    }
                            }
                        }
                        nb += take;
                        done += take;
                    }
                }
            };

            if constexpr (LISTS) {
                // This unit's list, 64 entries at a time, staged as the traversal would stage them.
                const uint32_t last = a.list_cap - 1u;
                const uint32_t units = uint32_t(__builtin_amdgcn_readfirstlane(
                    int(a.lpool[std::size_t(lc) * kListChunk + loff])));
                if (++loff == uint32_t(kListChunk)) {
                    lc = min(a.lnext[lc], last);
                    loff = 0;
                }
                for (uint32_t done = 0; done < units; done += 64) {
                    const uint32_t take = min(64u, units - done);
                    const bool act = uint32_t(lane) < take;
                    uint32_t nc = lc;
                    if (loff + take >= uint32_t(kListChunk)) nc = min(a.lnext[lc], last);
                    uint32_t e = 0;
                    if (act) {
                        const uint32_t o = loff + uint32_t(lane);
                        e = o < uint32_t(kListChunk) ? a.lpool[std::size_t(lc) * kListChunk + o]
                                                     : a.lpool[std::size_t(nc) * kListChunk + (o - kListChunk)];
                    }
                    loff += take;
                    if (loff >= uint32_t(kListChunk)) {
                        loff -= kListChunk;
                        lc = nc;
                    }
                    const uint32_t tag = e & ~kEntIndex;
                    const int k = int(e & kEntIndex);
                    const bool is_cell = act && tag == kEntCell, is_mono = act && tag == kEntMono;
                    const bool is_range = act && tag == kEntRange;
                    float4 cA = make_float4(0.f, 0.f, 0.f, 0.f), cB = cA;
                    float2 cC = make_float2(0.f, 0.f);
                    int rlo = 0, rhi = 0;
                    if (!REL && (is_cell || is_mono)) {
                        const float4* r = reinterpret_cast<const float4*>(a.mom + k);
                        cA = r[0];
                        if (is_cell) {
                            cB = r[1];
                            const float4 r2 = r[2];
                            cC = make_float2(r2.x, r2.y);
                        }
                    }
                    if (is_range) {
                        const int2 lh = *reinterpret_cast<const int2*>(&a.node[k].lo);
                        rlo = lh.x;
                        rhi = lh.y;
                    }
                    if constexpr (REL) {
                        stage_cells_at(is_cell, k);
                        stage_monos_at(is_mono, k);
                    } else {
                        stage_cells(is_cell, [&]() {
                            CellL c;
                            c.A = cA;
                            c.B = cB;
                            c.C = cC;
                            return c;
                        });
                        stage_monos(is_mono, [&]() { return cA; });
                    }
                    stage_ranges(is_range, rlo, rhi - rlo, nullptr);
                }
            } else if (!blocks) {
                if (lane == 0) st[0] = 0;  // the root
                int top = 1;
                while (top > 0) {
                    // Wide steps pop up to 64 nodes while the stack is at least 7 below thr; their
                    // pushes are held to thr (below). Above that, one node per step.
                    const int P = top <= stack_thr - 7 ? min(64, top) : 1;
                    __builtin_amdgcn_wave_barrier();
                    const bool act = lane < P;
                    const int k = act ? st[top - P + lane] : 0;
                    __builtin_amdgcn_wave_barrier();  // the popped entries are read before pushes overwrite them
                    top -= P;
                    WalkNode nd{0.f, 0.f, 0.f, 0.f, 0, 0, 0, -1};
                    // Plain form: the mass, quadrupole and qrel (the com is the node record's) load
                    // together with the node, so that deciding and staging wait for one round trip.
                    // Mixed form: qrel (the far-field decision) only; accepted cells and monopoles
                    // enter the lists as node indices and their FP64 moments load when a list is
                    // evaluated.
                    float qr = 0.f, pm = 0.f;
                    float4 pq = make_float4(0.f, 0.f, 0.f, 0.f);
                    float2 pq2 = make_float2(0.f, 0.f);
                    if (act) {
                        nd = a.node[k];
                        if (REL) {
                            if (a.quad_tau > 0.f) qr = a.mom[k].qrel;
                        } else {
                            const float4* r = reinterpret_cast<const float4*>(a.mom + k);
                            const float4 r2 = r[2];
                            pm = r[0].w;
                            pq = r[1];
                            pq2 = make_float2(r2.x, r2.y);
                            qr = r2.w;
                        }
                    }
                    // Error-budgeted far field: an accepted cell whose quadrupole term is below
                    // quad_tau of its monopole at the group box's distance enters the body list as a
                    // point mass.
                    const Pick pk = decide(act, nd, qr, bx, by, bz, hx, hy, hz, clo, chi, chunked, a.small_leaf,
                                           a.quad_tau);
                    const bool is_cell = pk.cell, is_mono = pk.mono, is_range = pk.range, open_int = pk.open;

                    // Push the child blocks of opened internal nodes. A wide step keeps the stack at
                    // or below thr: the opened nodes (in lane order) whose children would pass it go
                    // back on the stack themselves and are popped and opened again later (the same
                    // decision: the lists stay the same sets). Popped at 7 or more below thr, the
                    // first opened node's children always fit, so every step makes progress.
                    const int nch = open_int ? info_nchild(nd.info) : 0;
                    const int cend = prefix_sum_incl(nch, lane);
                    const int pushed = __builtin_amdgcn_readlane(cend, 63);
                    if (P > 1 && top + pushed > stack_thr) {
                        const unsigned long long om = __ballot(open_int);
                        const int nopen = __popcll(om);
                        const int after = nopen - int(lanes_below(om)) - 1;  // opened lanes above this one
                        const bool fit = open_int && top + cend + after <= stack_thr;
                        const unsigned long long fm = __ballot(fit);
                        const int nfit = __popcll(fm);
                        const int cfit = nfit ? __builtin_amdgcn_readlane(cend, 63 - __clzll(fm)) : 0;
                        if (fit)
                            for (int c = 0; c < nch; ++c) st[top + cend - nch + c] = nd.first + c;
                        else if (open_int)
                            st[top + cfit + int(lanes_below(om)) - nfit] = k;
                        top += cfit + nopen - nfit;
                    } else {
                        if (top + pushed > kStackCap) {
                            overflow = true;
                            break;
                        }
                        for (int c = 0; c < nch; ++c) st[top + cend - nch + c] = nd.first + c;
                        top += pushed;
                    }

                    if constexpr (REL) {
                        stage_cells_at(is_cell, k);
                        stage_monos_at(is_mono, k);
                    } else {
                        stage_cells(is_cell, [&]() {
                            CellL c;
                            c.A = make_float4(nd.cx, nd.cy, nd.cz, pm);
                            c.B = pq;
                            c.C = pq2;
                            return c;
                        });
                        stage_monos(is_mono, [&]() { return make_float4(nd.cx, nd.cy, nd.cz, pm); });
                    }
                    stage_ranges(is_range, nd.lo, nd.hi - nd.lo, nullptr);
                }
            }
            if constexpr (LET) {
                // Remote trees (src/backend.hpp:78-83): a terminal node (no children, no bodies) or
                // one whose com is at rcrit2 or beyond is accepted: its bodies if it has 1 to
                // small_leaf of them, else its multipole; an opened node contributes its bodies if it
                // has any, else its children. Remote bodies are never the target (index -1).
                for (int t = 0; t < a.let_trees && !overflow; ++t) {
                    if (lane == 0) st[0] = a.let_root[t];
                    int top = 1;
                    while (top > 0) {
                        const int room = (let_thr - top) / a.let_fan;
                        const int P = min(64, min(top, max(1, room)));
                        __builtin_amdgcn_wave_barrier();
                        const bool act = lane < P;
                        const int k = act ? st[top - P + lane] : 0;
                        __builtin_amdgcn_wave_barrier();
                        top -= P;
                        LetRec r;
                        r.A = r.B = r.C = make_float4(0.f, 0.f, 0.f, 0.f);
                        r.D = make_int4(-1, 0, 0, 0);
                        if (act) r = a.let[k];
                        const bool terminal = r.D.x < 0 && r.D.w == 0;
                        const float dx = fmaxf(fabsf(r.A.x - bx) - hx, 0.f);
                        const float dy = fmaxf(fabsf(r.A.y - by) - hy, 0.f);
                        const float dz = fmaxf(fabsf(r.A.z - bz) - hz, 0.f);
                        const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
                        const bool acc = act && (terminal || d2 >= r.C.z);
                        const bool small = r.D.w > 0 && r.D.w <= a.small_leaf;
                        bool is_cell = acc && !small;
                        const bool is_mono = is_cell && a.quad_tau > 0.f && r.C.w < a.quad_tau * d2;
                        is_cell = is_cell && !is_mono;
                        const bool is_range = act && r.D.w > 0 && (!acc || small);
                        const int nch = (act && !acc && r.D.w <= 0) ? r.D.y : 0;
                        const int cend = prefix_sum_incl(nch, lane);
                        const int pushed = __builtin_amdgcn_readlane(cend, 63);
                        if (top + pushed > kStackCap) {
                            overflow = true;
                            break;
                        }
                        for (int c = 0; c < nch; ++c) st[top + cend - nch + c] = r.D.x + c;
                        top += pushed;

                        auto rel = [&](float4 p) {
                            return REL ? make_float4(float(double(p.x) - ccx), float(double(p.y) - ccy),
                                                     float(double(p.z) - ccz), p.w)
                                       : p;
                        };
                        stage_cells(is_cell, [&]() {
                            CellL c;
                            c.A = rel(r.A);
                            c.B = r.B;
                            c.C = make_float2(r.C.x, r.C.y);
                            return c;
                        });
                        stage_monos(is_mono, [&]() { return rel(r.A); });
                        stage_ranges(is_range, r.D.z, r.D.w, [&](int j) { return rel(a.let_body[j]); });
                    }
                }
            }
            flush_cells();
            flush_bodies();

            // Add the list shares of each body through LDS: part A's share sp of body v at
            // sp na + v, part B's at SA na + sp nbB + (v - na).
            auto red_slot = [&](int s, int v) { return s < nsA ? spA * na + v : SA * na + spB * nbB + (v - na); };
            auto red_base = [&](int v) { return v < na ? v : SA * na + (v - na); };
            const uint32_t listed = ncell_total + nbody_total;  // nbody_total includes the monopoles
            // Per-body cost, self excluded: the list less the body's own entry; remote blocks after
            // the direct sum add their list to n - 1.
            auto cost_of = [&](int, uint32_t len) {
                if (blocks) return float((unsigned long long)(a.n - 1) + len);
                return float(len - 1u);
            };
            if (REL) {
                D4* red = reinterpret_cast<D4*>(&L);  // the lists are done: their LDS holds 64 NX D4
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int s = 0; s < NX; ++s)
                    if (sl[s].on) red[red_slot(s, sl[s].bi)] = D4{sl[s].dx, sl[s].dy, sl[s].dz, sl[s].dp};
                __builtin_amdgcn_wave_barrier();
                const double G = a.Gd;
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int v = s * 64 + lane;
                    if (v >= nbody) continue;
                    const int r0 = red_base(v), rs = v < na ? na : nbB, nsh = v < na ? SA : SB;
                    D4 acc4 = red[r0];
                    for (int sp = 1; sp < nsh; ++sp) {
                        const D4 q = red[r0 + sp * rs];
                        acc4.x += q.x; acc4.y += q.y; acc4.z += q.z; acc4.w += q.w;
                    }
                    const int i = clo + v;
                    if (blocks) {
                        a.axd[i] += G * acc4.x;
                        a.ayd[i] += G * acc4.y;
                        a.azd[i] += G * acc4.z;
                        if (POT) a.potd[i] -= G * acc4.w;
                    } else {
                        a.axd[i] = G * acc4.x;
                        a.ayd[i] = G * acc4.y;
                        a.azd[i] = G * acc4.z;
                        if (POT) a.potd[i] = -G * acc4.w;
                    }
                    if (a.cost) a.cost[i] = cost_of(i, listed);
                }
            } else {
                float4* red = reinterpret_cast<float4*>(&L);  // the lists are done: 64 NX float4
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int s = 0; s < NX; ++s)
                    if (sl[s].on)
                        red[red_slot(s, sl[s].bi)] = make_float4(sl[s].sx, sl[s].sy, sl[s].sz, sl[s].sp_pot);
                __builtin_amdgcn_wave_barrier();
                const float G = a.G;
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int v = s * 64 + lane;
                    if (v >= nbody) continue;
                    const int r0 = red_base(v), rs = v < na ? na : nbB, nsh = v < na ? SA : SB;
                    float4 acc4 = red[r0];
                    for (int sp = 1; sp < nsh; ++sp) {
                        const float4 q = red[r0 + sp * rs];
                        acc4.x += q.x; acc4.y += q.y; acc4.z += q.z; acc4.w += q.w;
                    }
                    const int i = clo + v;
                    if (blocks) {
                        a.ax[i] += G * acc4.x;
                        a.ay[i] += G * acc4.y;
                        a.az[i] += G * acc4.z;
                        if (POT) a.pot[i] -= G * acc4.w;
                    } else {
                        a.ax[i] = G * acc4.x;
                        a.ay[i] = G * acc4.y;
                        a.az[i] = G * acc4.z;
                        if (POT) a.pot[i] = -G * acc4.w;
                    }
                    if (a.cost) a.cost[i] = cost_of(i, listed);
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                // Each body of the chunk meets itself once in the local body list (never in a
                // remote one).
                const unsigned long long self = blocks ? 0ull : (unsigned long long)nbody;
                t_listed += (unsigned long long)listed * nbody;
                t_p2c += (unsigned long long)(ncell_total + nmono_total) * nbody;
                t_p2m += (unsigned long long)nmono_total * nbody;
                t_p2p += (unsigned long long)(nbody_total - nmono_total) * nbody - self;
                t_inter += (unsigned long long)listed * nbody - self;
            }
        }
    }
    if (overflow && lane == 0) atomicOr(&a.counters->error, 1);
    t_inter = wave_sum_u64(t_inter);
    t_p2p = wave_sum_u64(t_p2p);
    t_p2c = wave_sum_u64(t_p2c);
    t_listed = wave_sum_u64(t_listed);
    t_p2m = wave_sum_u64(t_p2m);
    if (lane == 0) {
        atomicAdd(&a.counters->interactions, t_inter);
        atomicAdd(&a.counters->p2p, t_p2p);
        atomicAdd(&a.counters->p2c, t_p2c);
        atomicAdd(&a.counters->listed, t_listed);
        atomicAdd(&a.counters->p2m, t_p2m);
    }
}

__global__ void reset_walk(WalkCounters* __restrict__ c) {
    c->interactions = c->p2p = c->p2c = c->listed = c->p2m = 0;
    c->error = 0;
    c->next_group = 0;
    c->next_list = 0;
    c->list_chunks = 0;
    for (int r = 0; r < kWalkRanges; ++r) c->next_range[r][0] = 0;
}


}  // namespace

uint32_t walk_eval_plan(int nb, int nx) {
    if (nb < 1 || nb > kGroupMax || nx < 1 || nx > 3) throw Error(kInvalid, "walk_eval_plan: chunk or slots out of range");
    // Shares of a part of n bodies on ns slots per lane (0: the part does not fit the wave).
    auto shares = [](int n, int ns) {
        const int lps = (n + ns - 1) / ns;
        return lps > 64 ? 0 : std::min(kMaxShare, 64 / lps);
    };
    long long bnum = 0, bden = 0;  // best cost bnum / bden
    int bparts = 0, bslots = 0;
    uint32_t word = 0;
    for (int na = 1; na <= nb; ++na) {
        const int nbb = nb - na;
        for (int nsa = 1; nsa <= nx; ++nsa) {
            const int sa = shares(na, nsa);
            if (!sa) continue;
            const int lo = nbb ? 1 : 0, hi = nbb ? nx - nsa : 0;
            for (int nsb = lo; nsb <= hi; ++nsb) {
                const int sb = nsb ? shares(nbb, nsb) : 1;
                if (!sb) continue;
                // Cost per list entry, in quarter slot-steps: (4 ns_A + 1) / S_A + (4 ns_B + 1) / S_B.
                const long long ca = 4ll * nsa + 1, cb = nsb ? 4ll * nsb + 1 : 0;
                const long long num = ca * sb + cb * sa, den = (long long)sa * sb;
                const int parts = nsb ? 2 : 1, slots = nsa + nsb;
                const long long l = num * bden, r = bnum * den;
                const bool better = bden == 0 || l < r ||
                                    (l == r && (parts < bparts || (parts == bparts && slots < bslots)));
                if (better) {
                    bnum = num; bden = den; bparts = parts; bslots = slots;
                    word = uint32_t(na) | uint32_t(nsa) << 8 | uint32_t(sa) << 12 | uint32_t(nsb) << 18 |
                           uint32_t(nsb ? sb : 0) << 22;
                }
            }
        }
    }
    return word;
}

namespace {
// The plans of every chunk size for 2 and 3 slots per lane, made once.
const uint32_t* eval_plans(int nx) {
    static const auto tables = [] {
        std::array<std::array<uint32_t, kGroupMax>, 2> t{};
        for (int k = 0; k < 2; ++k)
            for (int nb = 1; nb <= kGroupMax; ++nb) t[k][nb - 1] = walk_eval_plan(nb, k + 2);
        return t;
    }();
    return tables[nx - 2].data();
}
}  // namespace

int walk_blocks_per_cu() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks,
                                                     reinterpret_cast<const void*>(walk_kernel<false, false, false, false, false>),
                                                     kWalkBS, 0) != hipSuccess ||
        blocks < 1) {
        (void)hipGetLastError();
        blocks = 2;
    }
    return blocks;
}

template <bool REL, bool LET, bool LISTS>
void walk_launch(hipStream_t s, const WalkArgs& a, bool potential, bool guard, int blocks) {
    if (potential) {
        if (guard) walk_kernel<true, true, REL, LET, LISTS><<<blocks, kWalkBS, 0, s>>>(a);
        else walk_kernel<true, false, REL, LET, LISTS><<<blocks, kWalkBS, 0, s>>>(a);
    } else {
        if (guard) walk_kernel<false, true, REL, LET, LISTS><<<blocks, kWalkBS, 0, s>>>(a);
        else walk_kernel<false, false, REL, LET, LISTS><<<blocks, kWalkBS, 0, s>>>(a);
    }
}

void tree_walk(hipStream_t s, const WalkArgs& args, bool potential) {
    if (args.n <= 0) return;
    const bool guard = !(args.eps2 > 0.f);
    // The slots per lane of the launch's kernel: 2 with the potential or the self test, else 3.
    WalkArgs a = args;
    std::memcpy(a.plan, eval_plans(potential || guard ? 2 : 3), sizeof a.plan);
    const int blocks = std::max(1, a.blocks);
    reset_walk<<<1, 1, 0, s>>>(a.counters);
    LHIP_LAUNCH_CHECK();
    const bool let = a.let_trees > 0;
    if (a.lpool && !let && !a.direct_blocks) {
        const int lblocks = std::max(1, a.list_blocks);
        if (a.posd) walk_lists_kernel<true><<<lblocks, kListBS, 0, s>>>(a);
        else walk_lists_kernel<false><<<lblocks, kListBS, 0, s>>>(a);
        LHIP_LAUNCH_CHECK();
        if (a.posd) walk_launch<true, false, true>(s, a, potential, guard, blocks);
        else walk_launch<false, false, true>(s, a, potential, guard, blocks);
        LHIP_LAUNCH_CHECK();
        return;
    }
    if (a.posd) let ? walk_launch<true, true, false>(s, a, potential, guard, blocks)
                    : walk_launch<true, false, false>(s, a, potential, guard, blocks);
    else let ? walk_launch<false, true, false>(s, a, potential, guard, blocks)
             : walk_launch<false, false, false>(s, a, potential, guard, blocks);
    LHIP_LAUNCH_CHECK();
}

int walk_lists_blocks_per_cu() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(walk_lists_kernel<true>),
                                                     kListBS, 0) != hipSuccess ||
        blocks < 1) {
        (void)hipGetLastError();
        blocks = 4;
    }
    return blocks;
}

}  // namespace lhip
