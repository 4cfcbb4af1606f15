// Device octree build and walk: buffers and launchers (kernels_build.hip, kernels_walk.hip).
#pragma once
#include <algorithm>
#include <cstddef>

#include "common.hpp"

namespace lhip {

// Counters written by the device during one force call, read back once at its end.
struct alignas(128) WalkCounters {
    unsigned long long interactions;  // body-body + body-node, self excluded
    unsigned long long p2p;           // body-body (self excluded)
    unsigned long long p2c;           // body-node
    unsigned long long listed;        // list entries per body summed (the reference CPU count)
    unsigned long long p2m;           // body-node interactions evaluated as monopoles (in p2c)
    int error;                        // nonzero: walk failure (traversal stack overflow)
    int overfull;                     // bodies in finest-level cells with > 16 bodies
    unsigned int next_group;          // walk work counter
    int max_level;                    // build: deepest leaf level of the level pass (frames = 1)
    unsigned int next_list;           // list traversal work counter
    unsigned int list_chunks;         // list chunks the traversal asked for (> the pool's: overflow)
    // The walk order's kWalkRanges ranges (eighths of the Morton order, each with its costly units
    // first): range r holds the positions [range_start[r], range_start[r + 1]), the last ending at the
    // unit count; written by the build's order pass. next_range[r][0] counts the positions taken
    // from range r (each on its own 128-byte line: the waves of one XCD take from it), reset by
    // each walk.
    unsigned int range_start[8];
    alignas(128) unsigned int next_range[8][32];
};
constexpr int kWalkRanges = 8;
// Device address of counters->range_start (no host access to the device struct).
inline uint32_t* range_starts_of(WalkCounters* c) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c) + offsetof(WalkCounters, range_start));
}

// The walk's node record, in level order (all cells of level 0, then level 1, ...; within a level
// by first body), so that the children of a node are contiguous: [first, first + nchild).
struct alignas(32) WalkNode {
    float cx, cy, cz, rcrit2;  // centre of mass; open the node if dist^2(group box, com) < rcrit2
    int32_t lo, hi;            // body range [lo, hi) in sorted order
    int32_t info;              // level | leaf << 7 | children << 8
    int32_t first;             // first child (level order), -1 for a leaf
};
static_assert(sizeof(WalkNode) == 32, "WalkNode is one 32-byte record");

// The mixed form's FP64 node moments: centre of mass, mass, traceless quadrupole about the com
// (xx, xy, xz, yy, yz, zz).
struct NodeMomD {
    double cx, cy, cz, m;
    double q[6];
};
static_assert(sizeof(NodeMomD) == 80, "NodeMomD layout");

// A remote LET node on the device: com and mass; quadrupole; (qyz, qzz, rcrit2, qrel); first child,
// child count, first body, body count (indices into the concatenated remote arrays; first child -1
// only for a node without children, which with no bodies is terminal).
struct alignas(16) LetRec {
    float4 A;
    float4 B;
    float4 C;
    int4 D;
};
static_assert(sizeof(LetRec) == 64, "LetRec is one 64-byte record");

// The LET wire record (src/backend.hpp:63-76), as written by the export.
struct LetWire {
    float cx, cy, cz, mass;
    float qxx, qxy, qxz, qyy, qyz, qzz;
    float rcrit2;
    int32_t first_child, child_count, body_first, body_count;
};
static_assert(sizeof(LetWire) == 60, "LetWire is the 60-byte LetNode");

struct BuildBufs {
    int n;
    int cap_nodes;             // node capacity (upper bound for n bodies)
    int gmax;                  // walk group size: groups are the topmost cells with <= gmax bodies
    const uint64_t* key;       // sorted keys
    const float4* posm;        // sorted bodies (plain form)
    const double *xd, *yd, *zd;  // sorted FP64 positions (mixed form; null in the plain form)
    const float* m;            // sorted masses (mixed form)
    NodeMomD* momd;            // cap FP64 moments (mixed form; null in the plain form)
    const Cube* cube;
    uint32_t* bl;              // per body: leaf level | group level << 8 | (c + 1) << 16
    uint32_t* cnt;             // n + 1: nodes starting at each body, then their exclusive scan
    uint32_t* gflag;           // n + 1: group starts, then their exclusive scan
    uint32_t* scan_tmp;        // scan block sums
    // pre-order scratch (node k = the k-th cell by (first body, level))
    uint64_t* pkey;            // cap: level of node k (sort key)
    uint64_t* pkey2;
    uint32_t* pval;            // cap: k, then the level-order permutation
    uint32_t* pval2;
    uint32_t* inv;             // cap: pre-order -> level order
    RadixWork* radix;          // radix sort work memory for cap_nodes elements
    uint32_t* levstart;        // kLevels + 2 level starts in level order
    // level-order node data
    WalkNode* node;            // cap
    NodeMom* mom;              // cap
    uint64_t* prefix;          // cap
    uint32_t* gstart;          // n + 1 (until the group pass writes it: the level pass's per-block deepest levels)
    // Walk order (optional): the walk units sorted longest first by an estimate from per-body costs
    // (wcost, the previous walk's list lengths carried with the bodies through the sort).
    const float* wcost;
    uint32_t* worder;          // n + 1
    float order_thr;           // units estimated below this keep their natural (Morton) order
    WalkCounters* counters;
    // Key restarts (frames > 1). Frame f >= 1 keys of body i at fkey[(f - 1) n + i] (0 outside
    // restarted cells); cap[i] = 21 x the frames body i has (its deepest level); fseg[(f - 1) n + i]
    // = the body's restarted cell at frame f, an index into fcube (ox, oy, oz, side).
    int frames;
    const uint64_t* fkey;
    const uint32_t* cap;
    const uint32_t* fseg;
    const double4* fcube;
    int* host_over;            // pinned host words: [0] set by the level pass if a finest cell is
                               // over-full; [1] the deepest leaf level, published after the pass;
                               // [2] the node count
    int depth_hint;            // tree_moments_keep: deepest level of the tree (< 0: every level)
};

std::size_t scan_tmp_entries(int n);
// Bound on the node count for n bodies: n leaves at most, plus at most n/17 internal cells per
// level (cells of one level are disjoint and hold >= 17 bodies each).
inline long long node_capacity(long long n) { return n + (n / 17 + 1) * kLevels + 1; }
// The node arrays' first size: a quarter of the bodies (trees of smooth distributions have 0.15-0.25
// nodes per body at leaf size 16), grown to a quarter more than a build finds when one outgrows it.
inline long long node_reserve_initial(long long n) { return std::min(node_capacity(n), n / 4 + 4096); }

// Leaf levels, node enumeration, level-order layout, groups and multipoles. inv_theta as the
// reference CPU computes it. levels_done (optional) is recorded after the level pass, which sets
// host_over[0] when a deepest cell is over-full (key restarts needed, or coincident bodies) and
// host_over[1] to the deepest leaf level. With both, the host waits for the pass before enqueuing
// the multipoles: it launches only the levels the tree has, and none when host_over[0] is set (the
// caller then restarts the keys and rebuilds). Returns the deepest level built (-1: no multipoles).
int tree_build(hipStream_t s, const BuildBufs& b, float inv_theta, hipEvent_t levels_done = nullptr);
// tree_build's return when the tree has more nodes than b.cap_nodes (host_over[2] holds the count;
// only with levels_done and host_over): the node arrays must grow before the build runs again.
constexpr int kBuildGrow = -2;
// Deepest level of a tree with b.frames key frames.
inline int tree_depth(int frames) { return kMaxLevel * frames; }

// ---- key restarts (the rare path; the host drives it, src/cpu_backend.cpp:501-544)
// Level pass with frames: bl, cnt, gflag from the extended keys and caps.
void tree_levels(hipStream_t s, const BuildBufs& b);
// Over-full deepest cells: seg[i] = 1 at each one's first body, then the exclusive scan (count at
// [n], read back; the stream is synchronised). Returns the count.
int tree_restart_cells(hipStream_t s, const BuildBufs& b, uint32_t* seg);
// Node count of the tree the levels in b describe (scans b.cnt; read back).
int tree_node_count(hipStream_t s, const BuildBufs& b);
// Bounding cube of each over-full cell (side 0: its bodies coincide, no restart), from positions
// (double if f64). bbox: 6 x ncell scratch.
void tree_restart_cubes(hipStream_t s, const BuildBufs& b, const void* x, const void* y, const void* z, bool f64,
                        const uint32_t* seg, int ncell, unsigned long long* bbox, double4* cubes);
// The next frame's keys of the bodies of restarted cells (cube index cube0 + cell), their caps and
// cells; gkey[i] = the position of the first body of i's restarted cell, else i (the grouping key of
// the stable re-sort); cstart: ncell scratch. fkey/fseg of the frame are written for every body.
void tree_restart_keys(hipStream_t s, const BuildBufs& b, const void* x, const void* y, const void* z, bool f64,
                       const uint32_t* seg, const double4* cubes, int cube0, uint64_t* fkey_frame,
                       uint32_t* fseg_frame, uint32_t* cap, uint64_t* gkey, uint32_t* cstart);
// out[i] = in[perm[i]] (u64 and u32 arrays of n).
void gather_u64(hipStream_t s, const uint32_t* perm, const uint64_t* in, uint64_t* out, int n);
void gather_u32(hipStream_t s, const uint32_t* perm, const uint32_t* in, uint32_t* out, int n);
void fill_u32(hipStream_t s, uint32_t* a, uint32_t v, long long n);
void iota_u32(hipStream_t s, uint32_t* a, int n);
// In-place exclusive scan of n entries; a[n] = total. tmp: scan_tmp_entries(n) entries.
void scan_u32(hipStream_t s, uint32_t* a, int n, uint32_t* tmp);

// LET export (kernels_let.hip): the nodes of the tree in [0, count) a target box needs, in the LET's
// breadth-first order. state, ex, bcnt: count + 1 entries of scratch; tmp: scan_tmp_entries(count).
// Positions (float, or double when f64) and masses in sorted order. Writes the records to out and
// the bodies of opened leaves to bx..bm; the totals (nodes, bodies) end at ex[count], bcnt[count].
struct LetExport {
    const WalkNode* node;
    const NodeMom* mom;
    const uint32_t* levstart;
    int count;
    int depth;                // deepest level
    const void *x, *y, *z;
    bool f64;
    const float* m;
    float lo[3], hi[3];       // target box
    uint32_t *state, *ex, *bcnt, *tmp;
    LetWire* out;
    float *bx, *by, *bz, *bm;
};
void let_export(hipStream_t s, const LetExport& e);

// Mixed form, frozen topology: recompute com, mass and quadrupole (node and momd) from the
// positions in b over the tree tree_build made, keeping every decision input (node records,
// rcrit2, qrel).
void tree_moments_keep(hipStream_t s, const BuildBufs& b);

// The order in which a walk takes the units of a tree (ustart, count at ucount[n]): units estimated
// at thr or more first, longest first, then the rest in their natural order (thr = 1: all sorted).
// Estimates from b.wcost; scratch b.pkey, b.pval, b.pkey2, b.radix.
void unit_order(hipStream_t s, const BuildBufs& b, const uint32_t* ustart, const uint32_t* ucount, float thr,
                uint32_t* ranges, uint32_t* order);

struct WalkArgs {
    const float4* posm;        // plain form: sorted (position, mass)
    // mixed form (posd != null): FP64 (position, mass) records, FP64 node moments, FP64 outputs
    const double4* posd;
    const double4* posd_box;   // positions for the group boxes (decisions); posd unless replaying
    const NodeMomD* momd;
    double *axd, *ayd, *azd, *potd;
    double Gd;
    const WalkNode* node;
    const NodeMom* mom;
    const uint32_t* gstart;
    const uint32_t* gcount;    // group count = gcount[n]
    const uint32_t* order;     // optional: the order in which the waves take the groups (a permutation)
    int ranges;                // 1: order is the build's, in kWalkRanges ranges (counters->range_start),
                               // each taken first by the waves of one XCD
    int n;
    float eps2;
    float G;
    float* ax;
    float* ay;
    float* az;
    float* pot;
    float* cost;               // per-body interactions (self excluded), may be null
    WalkCounters* counters;
    int blocks;                // persistent grid size (blocks of 4 waves)
    int small_leaf;            // accepted leaves with at most this many bodies are summed directly
    float quad_tau;            // error-budgeted far field threshold (0: every quadrupole kept)
    int depth;                 // deepest level of the tree walked (at most 21 x key frames)
    // Remote trees (let_trees > 0): node records, bodies (x, y, z, m) and the root of each
    // non-empty tree; the stack threshold for their height and fan-out. Every group walks them
    // after the local tree, into the same lists. direct_blocks: the groups are blocks of kGroupMax
    // consecutive bodies in the current order that walk the remote trees only and add to the
    // direct sum's forces (the direct method's remote pass).
    const LetRec* let;
    const float4* let_body;
    const int* let_root;
    int let_trees;
    int let_thr;
    int let_fan;
    int direct_blocks;
    // Interaction lists (lpool != null, local tree only): walk_lists first traverses every group and
    // writes its list to the pool, then the walk evaluates the lists instead of traversing. A
    // group's list is one unit per kGroupMax-body chunk of it (one for every group of at most
    // kGroupMax bodies): a count entry, then that many entries (kEnt* | node index). It fills pool
    // chunks of kListChunk entries, chunk c continuing at chunk lnext[c], from chunk lhead[group].
    uint32_t* lpool;
    uint32_t* lnext;
    uint32_t* lhead;
    uint32_t list_cap;         // pool chunks; a traversal that needs more sets counters->list_chunks above it
    int list_blocks;           // persistent grid of the list traversal (blocks of 4 waves)
    // Evaluation plan per chunk size (entry nb - 1 for a chunk of nb bodies; tree_walk fills it for
    // the launch's slots per lane): walk_eval_plan's packed parts.
    uint32_t plan[kGroupMax];
};
// The evaluation plan of a walk chunk of nb bodies (1 .. kGroupMax) with at most nx target slots
// per lane: its bodies split into part A = [0, na) and part B = [na, nb) (empty when na = nb);
// part P takes ns_P slots per lane and S_P list shares, lps_P = ceil(n_P / ns_P) lanes per share
// and S_P = min(16, 64 / lps_P), so a list of L entries costs it L / S_P steps of ns_P slots. The
// plan minimises (ns_A + 1/4) / S_A + (ns_B + 1/4) / S_B per list entry: the slot-steps plus a
// quarter of one per loop step, the step's LDS read and loop overhead (the weight that timed best of
// 0, 1/8, 1/4, 3/8, 1/2 and 1 on the 1M Plummer sphere). Ties: one part, then fewer slots. Packed as
// na | ns_A << 8 | S_A << 12 | ns_B << 18 | S_B << 22 (ns_B = 0: one part).
uint32_t walk_eval_plan(int nb, int nx);
// List entries: a quadrupole cell, a monopole cell or the bodies of a leaf, and the node index.
constexpr int kListChunk = 256;
constexpr uint32_t kEntCell = 0u << 30, kEntMono = 1u << 30, kEntRange = 2u << 30, kEntIndex = (1u << 30) - 1u;
// Node indices must fit kEntIndex: lists are used only while the node capacity does.
inline bool lists_fit(long long cap_nodes) { return cap_nodes <= (long long)kEntIndex; }
void tree_walk(hipStream_t s, const WalkArgs& a, bool potential);
// Entries of a walk wave's traversal stack (LDS).
constexpr int kWalkStack = 896;
int walk_blocks_per_cu();
int walk_lists_blocks_per_cu();
constexpr int kWalkBlockWaves = 4;  // waves per walk block

// Vector-Jacobian product of the mixed-precision walk's forces with frozen topology
// (kernels_treeadj.hip). u = cotangent on the accelerations (already scaled by G and the kick
// weight), sorted order. Target-side terms are written to gt*; body source terms and the
// cell-moment terms carried down the tree are added to gs* (zeroed by the caller); gcell, gmono and
// gpriv (zeroed by the caller) and cum (9 per node) are scratch.
struct TreeVjpArgs {
    const WalkNode* node;
    const NodeMom* mom;
    const NodeMomD* momd;
    const double4* posd;
    const double *ux, *uy, *uz;
    const uint32_t* gstart;
    const uint32_t* gcount;
    const uint32_t* order;     // optional: the order in which the waves take the groups (a permutation)
    const uint32_t* ranges;    // optional: the order's kWalkRanges range starts (as the walk's)
    int n;
    float eps2;
    int small_leaf;
    float quad_tau;
    double *gtx, *gty, *gtz;
    double *gsx, *gsy, *gsz;
    double* gcell;  // per node: gC (3) and G (6) of its quadrupole-cell entries
    double* gmono;  // per node: gC (3) of its monopole-cell entries
    double* gpriv;  // blocks x kHot x 12 doubles (per-block gcell, gmono of the first kHot nodes), zeroed
    WalkCounters* counters;
    int blocks;
    int depth;      // deepest level of the tree walked (at most 21 x key frames)
    // The result: lam[id[k]] += gt_k + gs_k for every body k (sorted order) of the tree; lam holds
    // (x, y, z) per body in id order.
    const uint32_t* id;
    double* lam;
};
// Nodes at the top of the tree (levels 0-3 and part of 4 in level order) whose VJP cotangents each
// block accumulates privately.
constexpr int kHot = 512;
// node_count: the device node count of the tree (nodes = its capacity, the launch bound).
void tree_vjp(hipStream_t s, const TreeVjpArgs& a, const uint32_t* levstart, const uint32_t* node_count, int nodes,
              double* cum);
int tree_vjp_blocks_per_cu();

}  // namespace lhip
