// Device octree build and walk: buffers and launchers (kernels_build.hip, kernels_walk.hip).
#pragma once
#include "common.hpp"

namespace lhip {

// Counters written by the device during one force call, read back once at its end.
struct alignas(16) WalkCounters {
    unsigned long long interactions;  // body-body + body-node, self excluded
    unsigned long long p2p;           // body-body (self excluded)
    unsigned long long p2c;           // body-node
    unsigned long long listed;        // list entries per body summed (the reference CPU count)
    int error;                        // nonzero: walk failure
    int overfull;                     // bodies in finest-level cells with > 16 bodies
    unsigned int next_group;          // walk work counter
    int pad;
};

struct BuildBufs {
    int n;
    int cap_nodes;             // node capacity (upper bound for n bodies)
    const uint64_t* key;       // sorted keys
    const float4* posm;        // sorted bodies
    const Cube* cube;
    uint32_t* bl;              // per body: leaf level | group level << 8 | (c + 1) << 16
    uint32_t* cnt;             // n + 1: nodes starting at each body, then their exclusive scan
    uint32_t* gflag;           // n + 1: group starts, then their exclusive scan
    uint32_t* scan_tmp;        // scan block sums
    NodeLink* link;            // cap_nodes
    NodeMom* mom;              // cap_nodes
    float4* trav;              // cap_nodes: com xyz + rcrit2 (the walk's test data)
    uint64_t* prefix;          // cap_nodes
    uint32_t* gstart;          // n + 1
    WalkCounters* counters;
};

std::size_t scan_tmp_entries(int n);
// Bound on the node count for n bodies: n leaves at most, plus at most n/17 internal cells per
// level (cells of one level are disjoint and hold >= 17 bodies each).
inline long long node_capacity(long long n) { return n + (n / 17 + 1) * kLevels + 1; }

// Leaf levels, node enumeration, groups and multipoles. inv_theta as the reference CPU computes it.
void tree_build(hipStream_t s, const BuildBufs& b, float inv_theta);

struct WalkArgs {
    const float4* posm;
    const float4* trav;
    const NodeLink* link;
    const NodeMom* mom;
    const uint32_t* cnt_scan;  // node count = cnt_scan[n]
    const uint32_t* gstart;
    const uint32_t* gcount;    // group count = gcount[n]
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
};
void tree_walk(hipStream_t s, const WalkArgs& a, bool potential);
int walk_blocks_per_cu();

}  // namespace lhip
