// Locally essential tree (LET) export: the part of the local octree that a target anywhere in a
// remote box needs, in the LET wire format (LetNode, src/backend.hpp:63-87).
//
// Reference: export_let (src/cpu_backend.cpp:210-258). Breadth-first from the root: a node is
// terminal when the box distance of its com reaches rcrit2 (min over the box of |x - com|^2 >=
// rcrit2, the box as centre and half extent in float); an opened leaf exports its bodies; an opened
// internal node appends its children of positive mass as one contiguous block. Nodes are emitted
// in queue order, bodies in the order their leaves are reached.
//
// Device form. The tree's nodes are in level order: by depth, and within a depth by first body. The
// breadth-first queue visits depth by depth, each depth's nodes in the order of their parents and,
// within a sibling block, in child order, which is again the order by first body. So the LET is
// exactly the subsequence of exported nodes of the level-order array, and a node's LET index is the
// number of exported nodes before it: one exclusive scan. Likewise a leaf's first body is the
// exclusive scan of the opened leaves' body counts. Whether a node is exported depends on its
// parent's decision only, so one pass per depth (top-down) marks the nodes; a scan of the marks,
// a scan of the body counts and one emit pass write the records and gather the bodies.
#include "kernels.hpp"
#include "tree.hpp"

namespace lhip {
namespace {

// Node states: not exported, exported (terminal once its depth's pass has run), opened leaf,
// opened internal node.
enum : uint32_t { kLetOut = 0, kLetTerm = 1, kLetLeaf = 2, kLetOpen = 3 };

__global__ void __launch_bounds__(kBS) let_init_kernel(const NodeMom* __restrict__ mom, int count,
                                                       uint32_t* __restrict__ state) {
    const int p = blockIdx.x * kBS + threadIdx.x;
    if (p < count) state[p] = (p == 0 && mom[0].m > 0.f) ? kLetTerm : kLetOut;
}

// One depth: decide every exported node of [levstart[level], levstart[level + 1]) and mark the
// exported children of the opened ones. Writes ex (exported flag) and bcnt (bodies exported) for
// every node of the depth.
__global__ void __launch_bounds__(kBS) let_level_kernel(int level, const uint32_t* __restrict__ levstart,
                                                        const WalkNode* __restrict__ node,
                                                        const NodeMom* __restrict__ mom, float bcx, float bcy,
                                                        float bcz, float bhx, float bhy, float bhz,
                                                        uint32_t* __restrict__ state, uint32_t* __restrict__ ex,
                                                        uint32_t* __restrict__ bcnt) {
    const int p0 = int(levstart[level]), p1 = int(levstart[level + 1]);
    for (int p = p0 + int(blockIdx.x) * kBS + int(threadIdx.x); p < p1; p += int(gridDim.x) * kBS) {
        if (state[p] == kLetOut) {
            ex[p] = 0;
            bcnt[p] = 0;
            continue;
        }
        const WalkNode nd = node[p];
        const float dx = fmaxf(fabsf(nd.cx - bcx) - bhx, 0.f);
        const float dy = fmaxf(fabsf(nd.cy - bcy) - bhy, 0.f);
        const float dz = fmaxf(fabsf(nd.cz - bcz) - bhz, 0.f);
        // Rounded as the reference build evaluates it: three products, then two sums.
        const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
        uint32_t st = kLetTerm, nb = 0;
        if (!(d2 >= nd.rcrit2)) {
            if (info_leaf(nd.info)) {
                st = kLetLeaf;
                nb = uint32_t(nd.hi - nd.lo);
            } else {
                st = kLetOpen;
                const int nc = info_nchild(nd.info);
                for (int c = nd.first; c < nd.first + nc; ++c)
                    if (mom[c].m > 0.f) state[c] = kLetTerm;
            }
        }
        state[p] = st;
        ex[p] = 1;
        bcnt[p] = nb;
    }
}

// Records and bodies: node p goes to ex[p] (scanned), its bodies to bcnt[p] (scanned).
template <typename T>
__global__ void __launch_bounds__(kBS) let_emit_kernel(int count, const WalkNode* __restrict__ node,
                                                       const NodeMom* __restrict__ mom,
                                                       const uint32_t* __restrict__ state,
                                                       const uint32_t* __restrict__ ex,
                                                       const uint32_t* __restrict__ bscan, const T* __restrict__ x,
                                                       const T* __restrict__ y, const T* __restrict__ z,
                                                       const float* __restrict__ m, LetWire* __restrict__ out,
                                                       float* __restrict__ bx, float* __restrict__ by,
                                                       float* __restrict__ bz, float* __restrict__ bm) {
    const int p = blockIdx.x * kBS + threadIdx.x;
    if (p >= count) return;
    const uint32_t st = state[p];
    if (st == kLetOut) return;
    const WalkNode nd = node[p];
    const NodeMom md = mom[p];
    LetWire w;
    w.cx = nd.cx; w.cy = nd.cy; w.cz = nd.cz; w.mass = md.m;
    w.qxx = md.qxx; w.qxy = md.qxy; w.qxz = md.qxz; w.qyy = md.qyy; w.qyz = md.qyz; w.qzz = md.qzz;
    w.rcrit2 = nd.rcrit2;
    w.first_child = -1; w.child_count = 0; w.body_first = -1; w.body_count = 0;
    if (st == kLetOpen) {
        const int nc = info_nchild(nd.info);
        w.first_child = int32_t(ex[nd.first]);
        w.child_count = int32_t(ex[nd.first + nc] - ex[nd.first]);
    } else if (st == kLetLeaf) {
        const uint32_t b0 = bscan[p];
        w.body_first = int32_t(b0);
        w.body_count = nd.hi - nd.lo;
        for (int j = nd.lo; j < nd.hi; ++j) {
            const uint32_t o = b0 + uint32_t(j - nd.lo);
            bx[o] = float(x[j]); by[o] = float(y[j]); bz[o] = float(z[j]); bm[o] = m[j];
        }
    }
    out[ex[p]] = w;
}

}  // namespace

void let_export(hipStream_t s, const LetExport& e) {
    if (e.count <= 0) return;
    let_init_kernel<<<grid_for(e.count, kBS), kBS, 0, s>>>(e.mom, e.count, e.state);
    LHIP_LAUNCH_CHECK();
    // Target box as centre and half extent, in float as the reference forms it.
    const float bcx = 0.5f * (e.lo[0] + e.hi[0]), bhx = 0.5f * (e.hi[0] - e.lo[0]);
    const float bcy = 0.5f * (e.lo[1] + e.hi[1]), bhy = 0.5f * (e.hi[1] - e.lo[1]);
    const float bcz = 0.5f * (e.lo[2] + e.hi[2]), bhz = 0.5f * (e.hi[2] - e.lo[2]);
    // A depth holds at most count nodes; a grid-stride loop covers the larger ones.
    const unsigned g = std::min<unsigned>(grid_for(e.count, kBS), 1024u);
    for (int l = 0; l <= e.depth; ++l) {
        let_level_kernel<<<g, kBS, 0, s>>>(l, e.levstart, e.node, e.mom, bcx, bcy, bcz, bhx, bhy, bhz, e.state, e.ex,
                                           e.bcnt);
        LHIP_LAUNCH_CHECK();
    }
    scan_u32(s, e.ex, e.count, e.tmp);
    scan_u32(s, e.bcnt, e.count, e.tmp);
    if (e.f64)
        let_emit_kernel<double><<<grid_for(e.count, kBS), kBS, 0, s>>>(
            e.count, e.node, e.mom, e.state, e.ex, e.bcnt, static_cast<const double*>(e.x),
            static_cast<const double*>(e.y), static_cast<const double*>(e.z), e.m, e.out, e.bx, e.by, e.bz, e.bm);
    else
        let_emit_kernel<float><<<grid_for(e.count, kBS), kBS, 0, s>>>(
            e.count, e.node, e.mom, e.state, e.ex, e.bcnt, static_cast<const float*>(e.x),
            static_cast<const float*>(e.y), static_cast<const float*>(e.z), e.m, e.out, e.bx, e.by, e.bz, e.bm);
    LHIP_LAUNCH_CHECK();
}

}  // namespace lhip
