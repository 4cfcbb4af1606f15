// Context: one backend instance (device buffers, one stream, phase events). Semantics follow the
// reference Backend contract (src/backend.hpp:89-141) and the CUDA backend's state machine
// (src/cuda_backend.cu:1289-1650): uploads synchronise, kick/drift are asynchronous, drift
// invalidates the tree, build_tree invalidates accelerations, compute_forces_with synchronises once
// at its end.
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <cstdlib>
#include <vector>

#include "../../include/lettuce_hip.h"
#include "adjoint.hpp"
#include "kernels.hpp"
#include "tree.hpp"

namespace lhip {

struct AdjSim;  // differentiable FP64-state simulation (adjoint.hip)
void adj_destroy(AdjSim* a);
struct TreeSim;  // differentiable simulation through the tree (treesim.hip)
void treesim_destroy(TreeSim* t);

// The list pool's first size for n bodies (chunks of kListChunk entries): 1 chunk per 4 bodies.
constexpr uint64_t kListChunksMax = uint64_t(INT32_MAX) / kListChunk;
inline uint32_t list_chunks_initial(long long n) {
    return uint32_t(std::min<uint64_t>(uint64_t(n) / 4 + 1024, kListChunksMax));
}

struct Context {
    AdjSim* adj = nullptr;
    TreeSim* tsim = nullptr;
    bool last_sim_tree = false;
    // Generation of the differentiable forward whose adjoint lhip_sim_backward runs: every forward
    // other than a tree replay starts a new one, so an older caller's backward can be refused.
    uint64_t sim_gen = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    std::string name, err;
    int n = 0;
    int cus = 1;
    int walk_blocks = 1;

    // bodies (backend order) and their permute alternates
    DevBuf x, y, z, vx, vy, vz, m, id, cost;
    DevBuf x2, y2, z2, vx2, vy2, vz2, m2, id2, cost2;
    DevBuf ax, ay, az, pot, posm;
    // sort
    DevBuf key, key2, idx, idx2, radix_buf;
    RadixWork radix;
    // small
    DevBuf cube, bbox_part, red_part, red_out, direct_part;
    // tree
    DevBuf bl, cnt, gflag, scan_tmp, pkey, pkey2, pval, pval2, inv, levstart, node, mom, prefix, gstart, counters;
    DevBuf worder;  // the walk's unit order, written by every build
    long long cap_nodes = 0;
    // Key restarts: key frames of the current tree (1: none), frame keys, depth caps, restarted
    // cells and their cubes; the pinned flag the level pass sets; restart scratch.
    int frames = 1;
    int* host_over = nullptr;
    // Pinned mirror of the walk counters' host-read fields (the force call's one readback).
    WalkCounters* host_counters = nullptr;
    // Copies the counters the host reads (those before the walk order's range fields) into the
    // pinned mirror, waits for the stream, and returns them in hc.
    void read_counters(WalkCounters& hc) {
        constexpr std::size_t kRead = offsetof(WalkCounters, range_start);
        LHIP_CHECK(hipMemcpyAsync(host_counters, counters.as<void>(), kRead, hipMemcpyDeviceToHost, stream));
        sync();
        std::memcpy(static_cast<void*>(&hc), host_counters, kRead);
    }
    DevBuf fkey, fseg, capb, fcube, rs_seg, rs_bbox, rs_cstart, rs_gkey, rs_ka, rs_kb, rs_va, rs_vb;

    // Mixed-precision form: FP64 positions and velocities (backend order) with permute alternates,
    // FP64 accelerations and potential, FP64 node moments and direct-sum records. Masses and ids
    // stay in m and id; the float position arrays are then only download staging.
    bool mixed = false;
    lhip_walk_options wopt{kGroupMax, kSmallLeaf, 0.0};  // the reference's walk
    int env_group = 0;  // LETTUCE_HIP_GROUP, 0 when unset
    int env_sort = 0;   // LETTUCE_HIP_SORT: 0 auto, else the body sort's top digits (3, 5, 7 or 8)
    // The body sort: the deepest level of the last tree built from these bodies (-1: none since the
    // last upload, or key restarts). With a deepest level D, every cell of a level >= D holds at most
    // a leaf's bodies, so the radix passes need only the digits above a level >= D + 2; the runs of
    // equal top digits left are short and a fix-up pass orders them.
    int sort_depth = -1;
    int sort_top_digits() const {
        if (env_sort) return env_sort;
        if (sort_depth < 0) return kRadixDigits;
        // top digits t cover the 8 t - 1 top key bits: levels 1 .. (8 t - 1) / 3
        if (sort_depth <= 5) return 3;
        if (sort_depth <= 11) return 5;
        if (sort_depth <= 16) return 7;
        return kRadixDigits;
    }
    DevBuf xd, yd, zd, vxd, vyd, vzd, xd2, yd2, zd2, vxd2, vyd2, vzd2, axd, ayd, azd, potd, momd, rec64, posd;

    // Frozen-topology replay of the differentiable tree simulation: group boxes from these recorded
    // positions instead of the current ones (null otherwise).
    const double4* posd_box = nullptr;

    // Interaction lists (the walk split into a list traversal and a list evaluation; measured slower
    // than the fused walk on both forms, so off unless LETTUCE_HIP_WALK=lists asks for it): per mode
    // whether the local walk uses them, the pool (list_cap chunks of kListChunk entries), chunk links
    // and per-group heads, and the chunks the last walk took (the next pool's size).
    bool lists_plain = false, lists_mixed = false;
    DevBuf lpool, lnext, lhead;
    uint32_t list_cap = 0, list_used = 0;
    int list_blocks = 1;
    bool lists_now() const { return mixed ? lists_mixed : lists_plain; }

    bool has_pe = false, acc_ok = false, tree_ok = false, costs_ok = false;
    bool kick_rec = false, drift_rec = false;
    lhip_phase_times t{};
    hipEvent_t ev[10] = {};  // 0 build start, 1 sorted, 2 built, 3 walk start, 4 walk end, 5-6 kick, 7-8 drift

    explicit Context(int device) : dev(device) {
        // Environment knobs, as the reference's CUDA backend reads LETTUCE_CUDA_WALK and
        // LETTUCE_CUDA_GROUP (src/cuda_backend.cu:1239-1254): LETTUCE_HIP_WALK=fused|lists picks the
        // local walk (one fused kernel, or a list traversal and a list evaluation); LETTUCE_HIP_GROUP
        // (16 .. 128, default 128) is the walk group size of both forms' default options. Other values
        // are errors (read first, before any device resource is taken).
        if (const char* w = std::getenv("LETTUCE_HIP_WALK")) {
            const std::string v = w;
            if (v != "fused" && v != "lists")
                throw Error(kInvalid, "LETTUCE_HIP_WALK must be fused or lists, not '" + v + "'");
            lists_plain = lists_mixed = v == "lists";
        }
        if (const char* q = std::getenv("LETTUCE_HIP_SORT")) {
            // The body sort's radix passes: auto (from the last tree's depth), full (all eight), or 3,
            // 5 or 7 top digits and the run fix-up (a diagnostic and test knob).
            const std::string v = q;
            if (v == "auto") env_sort = 0;
            else if (v == "full") env_sort = kRadixDigits;
            else if (v == "3" || v == "5" || v == "7") env_sort = v[0] - '0';
            else throw Error(kInvalid, "LETTUCE_HIP_SORT must be auto, full, 3, 5 or 7, not '" + v + "'");
        }
        if (const char* g = std::getenv("LETTUCE_HIP_GROUP")) {
            char* end = nullptr;
            const long v = std::strtol(g, &end, 10);
            if (!*g || *end || v < kLeafMax || v > kGroupMax)
                throw Error(kInvalid, std::string("LETTUCE_HIP_GROUP must be a group size in [16, 128], not '") + g + "'");
            env_group = int(v);
            wopt.group_max = env_group;
        }
        LHIP_CHECK(hipSetDevice(dev));
        hipDeviceProp_t prop{};
        LHIP_CHECK(hipGetDeviceProperties(&prop, dev));
        cus = prop.multiProcessorCount;
        char buf[512];
        std::snprintf(buf, sizeof buf, "hip: %s (%s, %d CUs, %.0f GB)", prop.name[0] ? prop.name : "AMD GPU",
                      prop.gcnArchName, cus,
                      double(prop.totalGlobalMem) / 1073741824.0);
        name = buf;
        // Default priority on purpose: streams of different priorities are not mixed here.
        LHIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (auto& e : ev) LHIP_CHECK(hipEventCreate(&e));
        cube.reserve(sizeof(Cube));
        bbox_part.reserve(sizeof(double) * 6 * kReduceBlocks);
        red_part.reserve(sizeof(double) * 12 * std::max<std::size_t>(kReduceBlocks, 1));
        red_out.reserve(sizeof(double) * 16);
        counters.reserve(sizeof(WalkCounters));
        walk_blocks = cus * walk_blocks_per_cu();
        list_blocks = cus * walk_lists_blocks_per_cu();
        name += lists_plain ? ", list walk" : ", fused walk";
        name += ", group " + std::to_string(wopt.group_max);
        LHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_over), 4 * sizeof(int), hipHostMallocDefault));
        LHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_counters), sizeof(WalkCounters), hipHostMallocDefault));
        host_over[0] = host_over[1] = host_over[2] = host_over[3] = 0;
    }
    ~Context() {
        adj_destroy(adj);
        treesim_destroy(tsim);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (host_over) (void)hipHostFree(host_over);
        if (host_counters) (void)hipHostFree(host_counters);
    }

    void reserve_bodies(int nn) {
        const std::size_t f = std::size_t(nn) * 4;
        for (DevBuf* b : {&x, &y, &z, &vx, &vy, &vz, &m, &id, &cost, &x2, &y2, &z2, &vx2, &vy2, &vz2, &m2, &id2, &cost2,
                          &ax, &ay, &az, &pot})
            b->reserve(f);
        posm.reserve(std::size_t(nn) * sizeof(float4));
        if (mixed) reserve_f64(nn);
    }

    void reserve_f64(int nn) {
        const std::size_t d = std::size_t(nn) * 8;
        for (DevBuf* b : {&xd, &yd, &zd, &vxd, &vyd, &vzd, &xd2, &yd2, &zd2, &vxd2, &vyd2, &vzd2, &axd, &ayd, &azd, &potd})
            b->reserve(d);
    }

    double* D(DevBuf& b) { return b.as<double>(); }
    float* F(DevBuf& b) { return b.as<float>(); }

    lhip_walk_options default_options(bool mix) const {
        const int g = env_group ? env_group : kGroupMax;
        return mix ? lhip_walk_options{g, 1, 0.05} : lhip_walk_options{g, kSmallLeaf, 0.0};
    }

    // Switches between the plain (FP32 state) and mixed-precision (FP64 state) forms, converting the
    // resident bodies, accelerations and potential; the walk options apply to later trees.
    void set_precision(int mode, const lhip_walk_options* opt) {
        sort_depth = -1;  // new bodies: the next sort takes every digit
        LHIP_CHECK(hipSetDevice(dev));
        if (mode != LHIP_MODE_PLAIN && mode != LHIP_MODE_MIXED) throw Error(kInvalid, "unknown precision mode");
        const bool mix = mode == LHIP_MODE_MIXED;
        const lhip_walk_options o = opt ? *opt : default_options(mix);
        if (o.group_max < kLeafMax || o.group_max > kGroupMax)
            throw Error(kInvalid, "walk group size must be in [16, 128], got " + std::to_string(o.group_max));
        if (o.small_leaf < 0 || o.small_leaf > kLeafMax)
            throw Error(kInvalid, "small_leaf must be in [0, 16], got " + std::to_string(o.small_leaf));
        if (!(o.quad_tau >= 0) || !(o.quad_tau < 1e30)) throw Error(kInvalid, "quad_tau must be finite and >= 0");
        if (mix != mixed && n > 0) {
            if (mix) {
                reserve_f64(n);
                const std::pair<DevBuf*, DevBuf*> cv[] = {{&x, &xd}, {&y, &yd}, {&z, &zd}, {&vx, &vxd}, {&vy, &vyd},
                                                          {&vz, &vzd}, {&ax, &axd}, {&ay, &ayd}, {&az, &azd},
                                                          {&pot, &potd}};
                for (auto& e : cv) launch_f32_to_f64(stream, F(*e.first), D(*e.second), n);
            } else {
                const std::pair<DevBuf*, DevBuf*> cv[] = {{&xd, &x}, {&yd, &y}, {&zd, &z}, {&vxd, &vx}, {&vyd, &vy},
                                                          {&vzd, &vz}, {&axd, &ax}, {&ayd, &ay}, {&azd, &az},
                                                          {&potd, &pot}};
                for (auto& e : cv) launch_f64_to_f32(stream, D(*e.first), F(*e.second), n);
            }
            sync();
        }
        mixed = mix;
        wopt = o;
        tree_ok = false;  // the next tree uses the new form and options
    }

    // FP64 state in and out (mixed form only).
    void upload_f64(uint64_t nn, const double* hx, const double* hy, const double* hz, const double* hvx,
                    const double* hvy, const double* hvz, const float* hm, const uint32_t* hid) {
        sort_depth = -1;  // new bodies: the next sort takes every digit
        if (!mixed) throw Error(kLogic, "hip upload_f64: FP64 state needs the mixed-precision mode");
        if (nn >= uint64_t(INT32_MAX)) throw Error(kInvalid, "hip backend: n = " + std::to_string(nn) +
                                                                " exceeds the 2^31-1 body limit");
        LHIP_CHECK(hipSetDevice(dev));
        reserve_bodies(int(nn));
        n = int(nn);
        reset_state();
        if (n == 0) return;
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        const std::pair<DevBuf*, const double*> src[] = {{&xd, hx}, {&yd, hy}, {&zd, hz},
                                                         {&vxd, hvx}, {&vyd, hvy}, {&vzd, hvz}};
        for (auto& e : src) LHIP_CHECK(hipMemcpyAsync(e.first->as<void>(), e.second, db, hipMemcpyHostToDevice, stream));
        LHIP_CHECK(hipMemcpyAsync(m.as<void>(), hm, fb, hipMemcpyHostToDevice, stream));
        LHIP_CHECK(hipMemcpyAsync(id.as<void>(), hid, fb, hipMemcpyHostToDevice, stream));
        for (DevBuf* b : {&axd, &ayd, &azd, &potd}) launch_fill64(stream, D(*b), 0.0, n);
        sync();
    }

    void download_f64(double* hx, double* hy, double* hz, double* hvx, double* hvy, double* hvz, float* hm,
                      uint32_t* hid) {
        if (!mixed) throw Error(kLogic, "hip download_f64: FP64 state needs the mixed-precision mode");
        LHIP_CHECK(hipSetDevice(dev));
        if (n == 0) return;
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        const std::pair<void*, DevBuf*> dst[] = {{hx, &xd}, {hy, &yd}, {hz, &zd}, {hvx, &vxd}, {hvy, &vyd}, {hvz, &vzd}};
        for (auto& e : dst)
            if (e.first) LHIP_CHECK(hipMemcpyAsync(e.first, e.second->as<void>(), db, hipMemcpyDeviceToHost, stream));
        if (hm) LHIP_CHECK(hipMemcpyAsync(hm, m.as<void>(), fb, hipMemcpyDeviceToHost, stream));
        if (hid) LHIP_CHECK(hipMemcpyAsync(hid, id.as<void>(), fb, hipMemcpyDeviceToHost, stream));
        sync();
    }

    void reset_state() {
        has_pe = false;
        acc_ok = true;  // zeroed by the upload
        tree_ok = costs_ok = false;
        kick_rec = drift_rec = false;
        t = lhip_phase_times{};
    }

    // Accelerations and potential of the last force call, in the backend's precision converted to
    // float (plain: as computed).
    void accelerations(float* hax, float* hay, float* haz, float* hpot) {
        LHIP_CHECK(hipSetDevice(dev));
        if (n == 0) return;
        if (mixed) {
            const std::pair<DevBuf*, DevBuf*> cv[] = {{&axd, &ax}, {&ayd, &ay}, {&azd, &az}, {&potd, &pot}};
            for (auto& e : cv) launch_f64_to_f32(stream, D(*e.first), F(*e.second), n);
        }
        const std::size_t fb = std::size_t(n) * 4;
        const std::pair<float*, DevBuf*> d[] = {{hax, &ax}, {hay, &ay}, {haz, &az}, {hpot, &pot}};
        for (auto& e : d)
            if (e.first) LHIP_CHECK(hipMemcpyAsync(e.first, e.second->as<void>(), fb, hipMemcpyDeviceToHost, stream));
        sync();
    }

    // FP64 direct-sum accelerations and potentials of the bodies at backend indices tgt (validation).
    void direct_f64_targets(uint64_t nt, const uint32_t* tgt, double G, double eps, double* hax, double* hay,
                            double* haz, double* hpot) {
        LHIP_CHECK(hipSetDevice(dev));
        if (nt == 0) return;
        if (nt >= uint64_t(INT32_MAX)) throw Error(kInvalid, "too many targets");
        for (uint64_t t = 0; t < nt; ++t)
            if (tgt[t] >= uint32_t(n)) throw Error(kInvalid, "target index out of range");
        const int k = int(nt);
        chk_tgt.reserve(std::size_t(k) * 4);
        chk_out.reserve(std::size_t(k) * 4 * 8);
        chk_part.reserve(std::size_t(direct_f64_split(n, k)) * 4 * std::size_t(k) * 8);
        LHIP_CHECK(hipMemcpyAsync(chk_tgt.as<void>(), tgt, std::size_t(k) * 4, hipMemcpyHostToDevice, stream));
        double* o = chk_out.as<double>();
        if (mixed)
            launch_direct_f64_targets(stream, D(xd), D(yd), D(zd), true, m.as<float>(), n, chk_tgt.as<uint32_t>(), k,
                                      eps, G, o, o + k, o + 2 * std::size_t(k), o + 3 * std::size_t(k),
                                      chk_part.as<double>());
        else
            launch_direct_f64_targets(stream, F(x), F(y), F(z), false, m.as<float>(), n, chk_tgt.as<uint32_t>(), k,
                                      eps, G, o, o + k, o + 2 * std::size_t(k), o + 3 * std::size_t(k),
                                      chk_part.as<double>());
        double* dst[4] = {hax, hay, haz, hpot};
        for (int c = 0; c < 4; ++c)
            if (dst[c])
                LHIP_CHECK(hipMemcpyAsync(dst[c], o + std::size_t(c) * k, std::size_t(k) * 8, hipMemcpyDeviceToHost,
                                          stream));
        sync();
    }
    DevBuf chk_tgt, chk_out, chk_part;

    void accelerations_f64(double* hax, double* hay, double* haz, double* hpot) {
        if (!mixed) throw Error(kLogic, "hip accelerations_f64: FP64 accelerations need the mixed-precision mode");
        LHIP_CHECK(hipSetDevice(dev));
        if (n == 0) return;
        const std::size_t db = std::size_t(n) * 8;
        const std::pair<double*, DevBuf*> d[] = {{hax, &axd}, {hay, &ayd}, {haz, &azd}, {hpot, &potd}};
        for (auto& e : d)
            if (e.first) LHIP_CHECK(hipMemcpyAsync(e.first, e.second->as<void>(), db, hipMemcpyDeviceToHost, stream));
        sync();
    }

    void reserve_tree(int nn) {
        const std::size_t N = std::size_t(nn);
        key.reserve(N * 8); key2.reserve(N * 8);
        idx.reserve(N * 4); idx2.reserve(N * 4);
        cap_nodes = std::max(cap_nodes, node_reserve_initial(nn));
        reserve_radix(std::max<long long>(nn, cap_nodes));
        bl.reserve(N * 4);
        cnt.reserve((N + 1) * 4);
        gflag.reserve((N + 1) * 4);
        gstart.reserve((N + 1) * 4);
        worder.reserve((N + 1) * 4);
        scan_tmp.reserve(2 * scan_tmp_entries(nn) * 4);  // two arrays scanned at once
        const std::size_t C = std::size_t(cap_nodes);
        pkey.reserve(C * 8); pkey2.reserve(C * 8);
        pval.reserve(C * 4); pval2.reserve(C * 4); inv.reserve(C * 4);
        levstart.reserve((kDepthLevels + 2) * 4);
        node.reserve(C * sizeof(WalkNode));
        mom.reserve(C * sizeof(NodeMom));
        prefix.reserve(C * 8);
    }

    // Node arrays for a tree of M nodes (key restarts can exceed the single-frame bound).
    void grow_nodes(long long M) {
        if (M <= cap_nodes) return;
        cap_nodes = M;
        const std::size_t C = std::size_t(cap_nodes);
        reserve_radix(std::max<long long>(n, cap_nodes));
        pkey.reserve(C * 8); pkey2.reserve(C * 8);
        pval.reserve(C * 4); pval2.reserve(C * 4); inv.reserve(C * 4);
        node.reserve(C * sizeof(WalkNode));
        mom.reserve(C * sizeof(NodeMom));
        prefix.reserve(C * 8);
        if (mixed) momd.reserve(C * sizeof(NodeMomD));
    }

    // Radix sort work memory for `elems` elements.
    void reserve_radix(long long elems) {
        radix_buf.reserve(radix_work_words(int(elems)) * 4);
        radix.words = radix_buf.as<uint32_t>();
    }

    float ms(int a, int b) const {
        float v = 0;
        LHIP_CHECK(hipEventElapsedTime(&v, ev[a], ev[b]));
        return v;
    }

    void sync() { LHIP_CHECK(hipStreamSynchronize(stream)); }

    // ---- Backend members ----

    void upload(uint64_t nn, const float* hx, const float* hy, const float* hz, const float* hvx, const float* hvy,
                const float* hvz, const float* hm, const uint32_t* hid) {
        sort_depth = -1;  // new bodies: the next sort takes every digit
        if (nn >= uint64_t(INT32_MAX)) throw Error(kInvalid, "hip backend: n = " + std::to_string(nn) +
                                                                " exceeds the 2^31-1 body limit");
        LHIP_CHECK(hipSetDevice(dev));
        reserve_bodies(int(nn));
        n = int(nn);
        reset_state();
        if (n == 0) return;
        const std::size_t fb = std::size_t(n) * 4;
        const std::pair<DevBuf*, const void*> src[] = {{&x, hx}, {&y, hy}, {&z, hz}, {&vx, hvx},
                                                       {&vy, hvy}, {&vz, hvz}, {&m, hm}, {&id, hid}};
        for (auto& s : src) LHIP_CHECK(hipMemcpyAsync(s.first->as<void>(), s.second, fb, hipMemcpyHostToDevice, stream));
        for (DevBuf* b : {&ax, &ay, &az, &pot}) launch_zero_f32(stream, b->as<float>(), n);
        if (mixed) {  // float bodies promoted to the FP64 state
            const std::pair<DevBuf*, DevBuf*> cv[] = {{&x, &xd}, {&y, &yd}, {&z, &zd},
                                                      {&vx, &vxd}, {&vy, &vyd}, {&vz, &vzd}};
            for (auto& e : cv) launch_f32_to_f64(stream, F(*e.first), D(*e.second), n);
            for (DevBuf* b : {&axd, &ayd, &azd, &potd}) launch_fill64(stream, D(*b), 0.0, n);
        }
        sync();
    }

    void download(float* hx, float* hy, float* hz, float* hvx, float* hvy, float* hvz, float* hm, uint32_t* hid) {
        LHIP_CHECK(hipSetDevice(dev));
        if (n == 0) return;
        if (mixed) {  // FP64 state rounded to float
            const std::pair<DevBuf*, DevBuf*> cv[] = {{&xd, &x}, {&yd, &y}, {&zd, &z},
                                                      {&vxd, &vx}, {&vyd, &vy}, {&vzd, &vz}};
            for (auto& e : cv) launch_f64_to_f32(stream, D(*e.first), F(*e.second), n);
        }
        const std::size_t fb = std::size_t(n) * 4;
        const std::pair<void*, DevBuf*> dst[] = {{hx, &x}, {hy, &y}, {hz, &z}, {hvx, &vx},
                                                 {hvy, &vy}, {hvz, &vz}, {hm, &m}, {hid, &id}};
        for (auto& d : dst)
            if (d.first) LHIP_CHECK(hipMemcpyAsync(d.first, d.second->as<void>(), fb, hipMemcpyDeviceToHost, stream));
        sync();
    }

    void kick(double dt) {
        LHIP_CHECK(hipSetDevice(dev));
        if (n == 0) return;
        if (!acc_ok) throw Error(kLogic, "hip kick: accelerations are stale (build_tree since the last force call)");
        LHIP_CHECK(hipEventRecord(ev[5], stream));
        if (mixed)
            launch_kick(stream, D(vxd), D(vyd), D(vzd), D(axd), D(ayd), D(azd), dt, n);
        else
            launch_kick(stream, vx.as<float>(), vy.as<float>(), vz.as<float>(), ax.as<float>(), ay.as<float>(),
                        az.as<float>(), float(dt), n);
        LHIP_CHECK(hipEventRecord(ev[6], stream));
        kick_rec = true;
    }

    void drift(double dt) {
        LHIP_CHECK(hipSetDevice(dev));
        if (n == 0) return;
        LHIP_CHECK(hipEventRecord(ev[7], stream));
        if (mixed)
            launch_drift(stream, D(xd), D(yd), D(zd), D(vxd), D(vyd), D(vzd), dt, n);
        else
            launch_drift(stream, x.as<float>(), y.as<float>(), z.as<float>(), vx.as<float>(), vy.as<float>(),
                         vz.as<float>(), float(dt), n);
        LHIP_CHECK(hipEventRecord(ev[8], stream));
        drift_rec = true;
        tree_ok = false;
        has_pe = false;
    }

    void build_tree(const lhip_force_params& p) {
        LHIP_CHECK(hipSetDevice(dev));
        tree_ok = false;
        t.sort_ms = t.build_ms = 0;
        if (p.method != LHIP_METHOD_TREE || n == 0) return;
        // The sort reorders the bodies; a, phi and the costs are stale until the next force call.
        // The costs are carried with the bodies all the same: the build orders the walk's groups by
        // them (longest first), which changes no result.
        reserve_tree(n);
        const bool keep_cost = true;
        costs_ok = false;
        has_pe = false;
        acc_ok = false;
        LHIP_CHECK(hipEventRecord(ev[0], stream));
        TreeBufs tb = tree_bufs();
        tree_sort(stream, tb, n, keep_cost, sort_top_digits());
        swap_permuted(keep_cost);
        LHIP_CHECK(hipEventRecord(ev[1], stream));
        frames = 1;
        BuildBufs bb = build_bufs();
        // inv_theta as the reference CPU forms it (src/cpu_backend.cpp:595).
        const float inv_theta = p.theta > 0 ? float(1.0 / p.theta) : FLT_MAX;
        // The level pass flags over-full finest cells; the host reads the flag while the rest of the
        // build runs, and only then takes the key-restart path (a second build).
        built_depth = build_waited(bb, inv_theta);
        if (*reinterpret_cast<volatile int*>(host_over)) {
            restart_keys(nullptr, nullptr, nullptr, true);
            rebuild(inv_theta, nullptr);
        }
        sort_depth = frames == 1 ? built_depth : -1;
        LHIP_CHECK(hipEventRecord(ev[2], stream));
        tree_ok = true;
    }

    TreeBufs tree_bufs() {
        TreeBufs tb{};
        tb.f64 = mixed;
        if (mixed) {
            tb.x = D(xd); tb.y = D(yd); tb.z = D(zd); tb.vx = D(vxd); tb.vy = D(vyd); tb.vz = D(vzd);
            tb.x2 = D(xd2); tb.y2 = D(yd2); tb.z2 = D(zd2); tb.vx2 = D(vxd2); tb.vy2 = D(vyd2); tb.vz2 = D(vzd2);
            tb.posm = nullptr;
            posd.reserve(std::size_t(n) * sizeof(double4));
            tb.posd = posd.as<double4>();
        } else {
            tb.x = F(x); tb.y = F(y); tb.z = F(z); tb.vx = F(vx); tb.vy = F(vy); tb.vz = F(vz);
            tb.x2 = F(x2); tb.y2 = F(y2); tb.z2 = F(z2); tb.vx2 = F(vx2); tb.vy2 = F(vy2); tb.vz2 = F(vz2);
            tb.posm = posm.as<float4>();
            tb.posd = nullptr;
        }
        tb.m = m.as<float>(); tb.id = id.as<uint32_t>(); tb.cost = cost.as<float>();
        tb.m2 = m2.as<float>(); tb.id2 = id2.as<uint32_t>(); tb.cost2 = cost2.as<float>();
        tb.key = key.as<uint64_t>(); tb.key2 = key2.as<uint64_t>();
        tb.idx = idx.as<uint32_t>(); tb.idx2 = idx2.as<uint32_t>();
        tb.radix = &radix;
        tb.cube = cube.as<Cube>();
        tb.bbox_part = bbox_part.as<void>();
        return tb;
    }

    // The permute wrote every body array into its alternate: swap the owners.
    void swap_permuted(bool with_cost) {
        if (mixed) {
            swap(xd, xd2); swap(yd, yd2); swap(zd, zd2); swap(vxd, vxd2); swap(vyd, vyd2); swap(vzd, vzd2);
        } else {
            swap(x, x2); swap(y, y2); swap(z, z2); swap(vx, vx2); swap(vy, vy2); swap(vz, vz2);
        }
        swap(m, m2); swap(id, id2);
        if (with_cost) swap(cost, cost2);
    }

    // Key restarts (src/cpu_backend.cpp:501-544), the rare path after a level pass that found an
    // over-full finest cell. Frame by frame: the over-full cells of the deepest frame whose bodies do
    // not coincide get keys in their own cubes, and (permute) the bodies are re-sorted stably by
    // (cell, new key). kx, ky, kz: the key positions in the current order (null: the state's own);
    // without permute the order must already be the restarted one (a replayed topology). Leaves
    // bl, cnt, gflag of the final frames; the caller rebuilds the tree (rebuild).
    void restart_keys(const void* kx, const void* ky, const void* kz, bool permute) {
        sync();
        const std::size_t N = std::size_t(n);
        const std::size_t maxcells = N / (kLeafMax + 1) + 1;
        fkey.reserve((kMaxFrames - 1) * N * 8);
        fseg.reserve((kMaxFrames - 1) * N * 4);
        capb.reserve(N * 4);
        fcube.reserve((kMaxFrames - 1) * maxcells * sizeof(double4));
        rs_seg.reserve((N + 1) * 4);
        rs_bbox.reserve(6 * maxcells * 8);
        rs_cstart.reserve(maxcells * 4);
        for (DevBuf* b : {&rs_gkey, &rs_ka, &rs_kb}) b->reserve(N * 8);
        for (DevBuf* b : {&rs_va, &rs_vb}) b->reserve(N * 4);
        fill_u32(stream, capb.as<uint32_t>(), kMaxLevel, n);
        frames = 1;
        auto bufs = [&]() {
            BuildBufs b = build_bufs();
            b.frames = frames;
            b.cap = capb.as<uint32_t>();
            b.fkey = fkey.as<uint64_t>();
            b.fseg = fseg.as<uint32_t>();
            b.fcube = fcube.as<double4>();
            b.host_over = nullptr;
            return b;
        };
        auto pos = [&](int c) -> const void* {
            if (c == 0 && kx) return kx;
            if (c == 1 && ky) return ky;
            if (c == 2 && kz) return kz;
            if (mixed) return c == 0 ? D(xd) : c == 1 ? D(yd) : D(zd);
            return c == 0 ? F(x) : c == 1 ? F(y) : F(z);
        };
        BuildBufs b = bufs();
        tree_levels(stream, b);
        int cubes_used = 0;
        for (int f = 1; f < kMaxFrames; ++f) {
            const int nc = tree_restart_cells(stream, b, rs_seg.as<uint32_t>());
            if (nc == 0) break;
            double4* cubes = fcube.as<double4>() + cubes_used;
            tree_restart_cubes(stream, b, pos(0), pos(1), pos(2), mixed, rs_seg.as<uint32_t>(), nc,
                               rs_bbox.as<unsigned long long>(), cubes);
            std::vector<double4> hc(nc);
            LHIP_CHECK(hipMemcpyAsync(hc.data(), cubes, std::size_t(nc) * sizeof(double4), hipMemcpyDeviceToHost, stream));
            sync();
            bool any = false;
            for (const double4& q : hc) any = any || q.w > 0.0;
            if (!any) break;  // every over-full cell's bodies coincide: they stay leaves
            uint64_t* fk = fkey.as<uint64_t>() + std::size_t(f - 1) * N;
            uint32_t* fs = fseg.as<uint32_t>() + std::size_t(f - 1) * N;
            tree_restart_keys(stream, b, pos(0), pos(1), pos(2), mixed, rs_seg.as<uint32_t>(), cubes, cubes_used, fk,
                              fs, capb.as<uint32_t>(), rs_gkey.as<uint64_t>(), rs_cstart.as<uint32_t>());
            cubes_used += nc;
            frames = f + 1;
            if (permute) {
                // Stable sort by the new key, then stably by the grouping key (the cell's first body,
                // or the body itself outside restarted cells): cells stay in place, sorted inside.
                uint64_t *ka = rs_ka.as<uint64_t>(), *kb = rs_kb.as<uint64_t>();
                uint32_t *va = rs_va.as<uint32_t>(), *vb = rs_vb.as<uint32_t>();
                LHIP_CHECK(hipMemcpyAsync(ka, fk, N * 8, hipMemcpyDeviceToDevice, stream));
                const RadixBufs rb{{ka, kb}, {va, vb}};
                radix_sort(stream, radix, rb, true, n, nullptr, 0, kRadixDigits);  // even: back in ka, va
                gather_u64(stream, va, rs_gkey.as<uint64_t>(), ka, n);
                int gdig = 0;  // digits of a body index
                while (gdig < 4 && (1ll << (8 * gdig)) < (long long)n) ++gdig;
                radix_sort(stream, radix, rb, false, n, nullptr, 0, gdig);
                const uint32_t* perm = rb.v[gdig & 1];
                TreeBufs tb = tree_bufs();
                tree_permute(stream, tb, perm, n, false);
                swap_permuted(false);
                gather_u64(stream, perm, key.as<uint64_t>(), key2.as<uint64_t>(), n);
                swap(key, key2);
                uint64_t* k2 = kb;  // free scratch of N keys
                for (int g = 1; g <= f; ++g) {
                    uint64_t* fg = fkey.as<uint64_t>() + std::size_t(g - 1) * N;
                    gather_u64(stream, perm, fg, k2, n);
                    LHIP_CHECK(hipMemcpyAsync(fg, k2, N * 8, hipMemcpyDeviceToDevice, stream));
                    uint32_t* sg = fseg.as<uint32_t>() + std::size_t(g - 1) * N;
                    uint32_t* t2 = perm == rs_va.as<uint32_t>() ? rs_vb.as<uint32_t>() : rs_va.as<uint32_t>();
                    gather_u32(stream, perm, sg, t2, n);
                    LHIP_CHECK(hipMemcpyAsync(sg, t2, N * 4, hipMemcpyDeviceToDevice, stream));
                }
                uint32_t* t2 = perm == rs_va.as<uint32_t>() ? rs_vb.as<uint32_t>() : rs_va.as<uint32_t>();
                gather_u32(stream, perm, capb.as<uint32_t>(), t2, n);
                LHIP_CHECK(hipMemcpyAsync(capb.as<uint32_t>(), t2, N * 4, hipMemcpyDeviceToDevice, stream));
            }
            b = bufs();
            tree_levels(stream, b);
        }
    }

    // The build after restart_keys: node arrays sized for the tree, then the whole build over the
    // key frames. box (optional) replaces the positions the moments read (a replayed topology).
    void rebuild(float inv_theta, const double* const* box) {
        BuildBufs b = build_bufs();
        grow_nodes(tree_node_count(stream, b) + 1);  // this build does not wait for the count
        b = build_bufs();
        b.host_over = nullptr;
        if (box) { b.xd = box[0]; b.yd = box[1]; b.zd = box[2]; }
        built_depth = tree_build(stream, b, inv_theta);
    }
    // Deepest level of the last build's multipoles (tree_moments_keep needs no more).
    int built_depth = -1;
    // The deepest level a walk of the current tree can reach (the traversal stacks reserve 7 entries
    // per level): the level the last build reported, or the bound for the key frames when the build
    // did not wait for the report.
    int walk_depth() const {
        return built_depth >= 0 ? std::min(built_depth, tree_depth(frames)) : tree_depth(frames);
    }

    // tree_build with the host wait on the level pass (ev[9]); a tree with more nodes than the node
    // arrays hold grows them to a quarter more than it has and is built again. bb's positions (the
    // caller's) are kept; its node arrays follow the growth. Returns tree_build's depth, or -1 when
    // host_over[0] asks for key restarts.
    int build_waited(BuildBufs bb, float inv_theta) {
        for (;;) {
            *reinterpret_cast<volatile int*>(host_over) = 0;
            const int d = tree_build(stream, bb, inv_theta, ev[9]);
            LHIP_CHECK(hipEventSynchronize(ev[9]));
            if (d != kBuildGrow) return d;
            const long long M = (long long)uint32_t(*reinterpret_cast<volatile int*>(host_over + 2));
            sync();
            grow_nodes(std::min(node_capacity(n), M + M / 4 + 1024));
            const BuildBufs g = build_bufs();
            bb.cap_nodes = g.cap_nodes;
            bb.pkey = g.pkey; bb.pkey2 = g.pkey2; bb.pval = g.pval; bb.pval2 = g.pval2; bb.inv = g.inv;
            bb.node = g.node; bb.mom = g.mom; bb.prefix = g.prefix;
            if (bb.momd) bb.momd = g.momd;
        }
    }

    BuildBufs build_bufs() {
        BuildBufs bb{};
        bb.n = n;
        bb.cap_nodes = int(cap_nodes);
        bb.gmax = wopt.group_max;
        bb.key = key.as<uint64_t>();
        bb.posm = posm.as<float4>();
        if (mixed) {
            momd.reserve(std::size_t(cap_nodes) * sizeof(NodeMomD));
            bb.xd = D(xd); bb.yd = D(yd); bb.zd = D(zd);
            bb.m = m.as<float>();
            bb.momd = momd.as<NodeMomD>();
        }
        bb.cube = cube.as<Cube>();
        bb.bl = bl.as<uint32_t>();
        bb.cnt = cnt.as<uint32_t>();
        bb.gflag = gflag.as<uint32_t>();
        bb.scan_tmp = scan_tmp.as<uint32_t>();
        bb.pkey = pkey.as<uint64_t>(); bb.pkey2 = pkey2.as<uint64_t>();
        bb.pval = pval.as<uint32_t>(); bb.pval2 = pval2.as<uint32_t>();
        bb.inv = inv.as<uint32_t>();
        bb.radix = &radix;
        bb.levstart = levstart.as<uint32_t>();
        bb.node = node.as<WalkNode>();
        bb.mom = mom.as<NodeMom>();
        bb.prefix = prefix.as<uint64_t>();
        bb.gstart = gstart.as<uint32_t>();
        bb.wcost = cost.as<float>();
        bb.worder = worder.as<uint32_t>();
        bb.order_thr = walk_order_threshold();
        bb.counters = counters.as<WalkCounters>();
        bb.frames = frames;
        if (frames > 1) {
            bb.fkey = fkey.as<uint64_t>();
            bb.cap = capb.as<uint32_t>();
            bb.fseg = fseg.as<uint32_t>();
            bb.fcube = fcube.as<double4>();
        }
        bb.host_over = host_over;
        bb.depth_hint = -1;
        return bb;
    }
    // Walk units estimated above this share of one wave's work are taken first (longest first): a
    // costly unit taken late would leave the other waves idle at the end. One wave's work is the
    // previous walk's interactions (x 1.2 for the traversal term of the estimate) over the waves of
    // the grid; with no previous walk every unit keeps its natural order.
    static constexpr double kOrderShare = 0.125;
    float walk_order_threshold() const {
        if (t.interactions == 0) return FLT_MAX;
        const double wave = 1.2 * double(t.interactions) / (double(walk_blocks) * kWalkBlockWaves);
        return float(std::max(1.0, kOrderShare * wave));
    }

    // The walk's units: the groups. The count is at [n].
    const uint32_t* walk_starts() const { return gstart.as<uint32_t>(); }
    const uint32_t* walk_count() const { return gflag.as<uint32_t>(); }

    // The walk's arguments for the current bodies: the local tree (tree method) and the outputs.
    WalkArgs walk_args(const lhip_force_params& p) {
        WalkArgs a{};
        a.posm = posm.as<float4>();
        a.node = node.as<WalkNode>();
        a.mom = mom.as<NodeMom>();
        a.gstart = walk_starts();
        a.gcount = walk_count();
        a.order = p.method == LHIP_METHOD_TREE ? worder.as<uint32_t>() : nullptr;
        a.ranges = a.order ? 1 : 0;  // the build wrote the order's ranges
        a.n = n;
        a.eps2 = float(p.eps * p.eps);
        a.G = float(p.G);
        a.ax = ax.as<float>(); a.ay = ay.as<float>(); a.az = az.as<float>(); a.pot = pot.as<float>();
        a.cost = cost.as<float>();
        if (mixed) {
            a.posd = posd.as<double4>();
            a.posd_box = posd_box ? posd_box : a.posd;
            a.momd = momd.as<NodeMomD>();
            a.axd = D(axd); a.ayd = D(ayd); a.azd = D(azd); a.potd = D(potd);
            a.Gd = p.G;
        }
        a.counters = counters.as<WalkCounters>();
        a.blocks = walk_blocks;
        a.list_blocks = list_blocks;
        a.small_leaf = wopt.small_leaf;
        a.quad_tau = float(wopt.quad_tau);
        a.depth = walk_depth();
        return a;
    }

    // ---- remote trees (LETs, src/backend.hpp:63-87)
    DevBuf let_rec, let_body, let_root;
    std::vector<LetRec> let_hrec;
    std::vector<float4> let_hbody;
    std::vector<int> let_hroot;
    int let_thr = 0, let_fan = 1;
    // Most bodies one remote node may hold: 64 of them (one walk step) stay below 2^31.
    static constexpr int32_t kLetMaxBodies = INT32_MAX / 64;

    // Validates the remote trees, converts them to device records (indices offset into the
    // concatenated arrays) and uploads them. Returns the number of non-empty trees. Children must
    // follow their parent (both reference exporters emit breadth-first): no cycles, and one forward
    // pass gives every node's depth, which with the largest child count bounds the walk's stack.
    int upload_remote(const lhip_let* remote, int nremote) {
        if (nremote < 0 || (nremote > 0 && !remote)) throw Error(kInvalid, "hip compute_forces_with: bad remote list");
        uint64_t nn = 0, nbt = 0;
        for (int t = 0; t < nremote; ++t) {
            if (remote[t].n_nodes == 0) continue;
            nn += remote[t].n_nodes;
            nbt += remote[t].n_bodies;
        }
        if (nn == 0) return 0;
        if (nn >= uint64_t(INT32_MAX) || nbt >= uint64_t(INT32_MAX))
            throw Error(kInvalid, "hip compute_forces_with: remote LETs exceed 2^31 - 1 nodes or bodies");
        let_hrec.resize(nn);
        let_hbody.resize(nbt);
        let_hroot.clear();
        std::vector<int> depth;
        int height = 0, fan = 1;
        uint64_t off = 0, boff = 0;
        for (int t = 0; t < nremote; ++t) {
            const lhip_let& L = remote[t];
            if (L.n_nodes == 0) continue;
            const uint64_t m = L.n_nodes, nb = L.n_bodies;
            if (!L.nodes || (nb > 0 && (!L.bx || !L.by || !L.bz || !L.bm)))
                throw Error(kInvalid, "hip compute_forces_with: remote LET with null arrays");
            depth.assign(m, 0);
            let_hroot.push_back(int(off));
            for (uint64_t i = 0; i < m; ++i) {
                const lhip_let_node& d = L.nodes[i];
                LetRec r;
                r.A = make_float4(d.cx, d.cy, d.cz, d.mass);
                r.B = make_float4(d.qxx, d.qxy, d.qxz, d.qyy);
                const double qf = std::sqrt(double(d.qxx) * d.qxx + double(d.qyy) * d.qyy + double(d.qzz) * d.qzz +
                                            2.0 * (double(d.qxy) * d.qxy + double(d.qxz) * d.qxz + double(d.qyz) * d.qyz));
                r.C = make_float4(d.qyz, d.qzz, d.rcrit2, d.mass > 0.f ? float(1.5 * qf / d.mass) : 0.f);
                int32_t bf = 0;
                if (d.body_count > 0) {
                    if (d.body_first < 0 || uint64_t(d.body_first) + uint64_t(d.body_count) > nb)
                        throw Error(kInvalid, "hip compute_forces_with: remote LET node body range out of bounds");
                    if (d.body_count > kLetMaxBodies)
                        throw Error(kInvalid, "hip compute_forces_with: remote LET node holds more than " +
                                                  std::to_string(kLetMaxBodies) + " bodies");
                    bf = int32_t(boff + uint64_t(d.body_first));
                }
                // first child: -1 only where the record has none (with no bodies: terminal).
                int32_t fc = d.first_child < 0 ? -1 : 0, cc = 0;
                if (d.body_count <= 0 && d.child_count > 0) {
                    if (d.first_child < 0 || uint64_t(d.first_child) <= i ||
                        uint64_t(d.first_child) + uint64_t(d.child_count) > m)
                        throw Error(kInvalid, "hip compute_forces_with: remote LET node children out of order or bounds");
                    fc = int32_t(off + uint64_t(d.first_child));
                    cc = d.child_count;
                    fan = std::max(fan, int(cc));
                    for (int32_t c = d.first_child; c < d.first_child + d.child_count; ++c)
                        depth[std::size_t(c)] = std::max(depth[std::size_t(c)], depth[i] + 1);
                }
                height = std::max(height, depth[i]);
                r.D = make_int4(fc, cc, bf, d.body_count);
                let_hrec[off + i] = r;
            }
            for (uint64_t j = 0; j < nb; ++j) let_hbody[boff + j] = make_float4(L.bx[j], L.by[j], L.bz[j], L.bm[j]);
            off += m;
            boff += nb;
        }
        // A step popping P nodes pushes at most fan P; one node at a time (above the threshold) the
        // stack grows by at most (fan - 1) per level below the top node.
        const long long thr = kWalkStack - (long long)fan * (height + 1);
        if (thr < 64)
            throw Error(kRuntime, "hip compute_forces_with: remote LETs too deep for the walk stack (height " +
                                      std::to_string(height) + ", up to " + std::to_string(fan) + " children)");
        let_thr = int(thr);
        let_fan = fan;
        let_rec.reserve(nn * sizeof(LetRec));
        let_body.reserve(std::max<uint64_t>(nbt, 1) * sizeof(float4));
        let_root.reserve(let_hroot.size() * sizeof(int));
        LHIP_CHECK(hipMemcpyAsync(let_rec.as<void>(), let_hrec.data(), nn * sizeof(LetRec), hipMemcpyHostToDevice, stream));
        if (nbt)
            LHIP_CHECK(hipMemcpyAsync(let_body.as<void>(), let_hbody.data(), nbt * sizeof(float4), hipMemcpyHostToDevice,
                                      stream));
        LHIP_CHECK(hipMemcpyAsync(let_root.as<void>(), let_hroot.data(), let_hroot.size() * sizeof(int),
                                  hipMemcpyHostToDevice, stream));
        return int(let_hroot.size());
    }

    void remote_args(WalkArgs& a, int trees, bool direct) {
        a.let = let_rec.as<LetRec>();
        a.let_body = let_body.as<float4>();
        a.let_root = let_root.as<int>();
        a.let_trees = trees;
        a.let_thr = let_thr;
        a.let_fan = let_fan;
        a.direct_blocks = direct ? 1 : 0;
    }

    void compute_forces_with(const lhip_force_params& p, bool potential, const lhip_let* remote = nullptr,
                             int nremote = 0) {
        LHIP_CHECK(hipSetDevice(dev));
        t.walk_ms = t.force_ms = 0;
        t.interactions = t.p2p = t.p2c = t.listed = t.p2c_mono = 0;
        costs_ok = false;
        if (n == 0) {
            t.sort_ms = t.build_ms = 0;
            has_pe = potential;
            acc_ok = true;
            return;
        }
        const bool tree = p.method == LHIP_METHOD_TREE;
        if (tree && !tree_ok)
            throw Error(kLogic, "hip compute_forces_with: call build_tree after the last upload/drift first");
        const float G = float(p.G), eps2 = float(p.eps * p.eps);
        // Remote trees are validated and uploaded before any force work is enqueued.
        const int trees = upload_remote(remote, nremote);
        LHIP_CHECK(hipEventRecord(ev[3], stream));
        WalkCounters hc{};
        if (!tree) {
            if (mixed) {
                direct_part.reserve(std::size_t(direct64_split(n)) * 4 * std::size_t(n) * 8);
                rec64.reserve(std::size_t(n) * sizeof(SimRec));
                launch_direct64(stream, D(xd), D(yd), D(zd), m.as<float>(), n, p.eps, p.G, potential, D(axd), D(ayd),
                                D(azd), D(potd), rec64.as<SimRec>(), direct_part.as<double>());
            } else {
                launch_pack_posm(stream, x.as<float>(), y.as<float>(), z.as<float>(), m.as<float>(), posm.as<float4>(),
                                 n);
                const int S = direct_split(n);
                if (S > 1) direct_part.reserve(std::size_t(S) * std::size_t(n) * sizeof(float4));
                launch_direct(stream, posm.as<float4>(), n, eps2, G, potential, ax.as<float>(), ay.as<float>(),
                              az.as<float>(), pot.as<float>(), direct_part.as<float4>());
            }
            if (trees > 0) {
                // Remote trees per block of kGroupMax consecutive bodies (src/cpu_backend.cpp:691-713).
                if (mixed) {
                    posd.reserve(std::size_t(n) * sizeof(double4));
                    launch_pack_posd(stream, D(xd), D(yd), D(zd), m.as<float>(), posd.as<double4>(), n);
                }
                WalkArgs a = walk_args(p);
                if (mixed) a.posd_box = a.posd;
                remote_args(a, trees, true);
                tree_walk(stream, a, potential);
            }
            LHIP_CHECK(hipEventRecord(ev[4], stream));
            if (trees > 0) read_counters(hc);
            else sync();
            const uint64_t pairs = uint64_t(n) * uint64_t(n - 1);
            hc.interactions += pairs;
            hc.p2p += pairs;
            hc.listed += uint64_t(n) * uint64_t(n);
            direct_costs_pending = trees == 0;
        } else {
            WalkArgs a = walk_args(p);
            const bool lists = trees == 0 && lists_now() && lists_fit(cap_nodes);
            if (lists) list_args(a, list_reserve());
            if (trees > 0) remote_args(a, trees, false);  // one pass: local, then remote trees
            tree_walk(stream, a, potential);
            LHIP_CHECK(hipEventRecord(ev[4], stream));
            read_counters(hc);
            if (lists && !hc.error && hc.list_chunks > a.list_cap) {
                // The pool was too small (the first call, or lists that grew): the evaluation left
                // everything alone; run the walk again with a pool of the size the traversal asked for.
                list_used = hc.list_chunks;
                list_args(a, list_reserve());
                tree_walk(stream, a, potential);
                LHIP_CHECK(hipEventRecord(ev[4], stream));
                read_counters(hc);
                if (!hc.error && hc.list_chunks > a.list_cap)
                    throw Error(kRuntime, "hip tree walk: interaction lists outgrew their pool");
            }
            if (lists) list_used = hc.list_chunks;
            direct_costs_pending = false;
        }
        if (hc.error)
            throw Error(kRuntime, "hip tree walk: traversal stack overflow; accelerations are invalid");
        t.sort_ms = tree ? ms(0, 1) : 0.0;
        t.build_ms = tree ? ms(1, 2) : 0.0;
        t.walk_ms = ms(3, 4);
        t.force_ms = t.sort_ms + t.build_ms + t.walk_ms;
        t.interactions = hc.interactions;
        t.p2p = hc.p2p;
        t.p2c = hc.p2c;
        t.listed = hc.listed;
        t.p2c_mono = hc.p2m;
        overfull = hc.overfull;
        has_pe = potential;
        acc_ok = true;
        costs_ok = true;
    }

    bool direct_costs_pending = false;
    int overfull = 0;

    // Chunks of the list pool: 1 per 4 bodies to begin with (list_chunks_initial), and a quarter
    // more than a walk asked for whenever one outgrew the pool.
    uint32_t list_reserve() {
        uint32_t c = std::max(list_cap, list_chunks_initial(n));
        if (list_used > c) c = uint32_t(std::min<uint64_t>(uint64_t(list_used) + list_used / 4 + 64, kListChunksMax));
        lpool.reserve(std::size_t(c) * kListChunk * 4);
        lnext.reserve(std::size_t(c) * 4);
        lhead.reserve((std::size_t(n) + 1) * 4);
        list_cap = c;
        return c;
    }
    void list_args(WalkArgs& a, uint32_t cap) {
        a.lpool = lpool.as<uint32_t>();
        a.lnext = lnext.as<uint32_t>();
        a.lhead = lhead.as<uint32_t>();
        a.list_cap = cap;
    }

    void compute_forces(const lhip_force_params& p, bool potential) {
        build_tree(p);
        compute_forces_with(p, potential);
    }

    lhip_diag diagnostics() {
        LHIP_CHECK(hipSetDevice(dev));
        lhip_diag d{};
        d.has_pe = has_pe ? 1 : 0;
        if (n == 0) return d;
        if (mixed)
            launch_diag(stream, D(xd), D(yd), D(zd), D(vxd), D(vyd), D(vzd), m.as<float>(), D(potd), has_pe ? 1 : 0, n,
                        red_part.as<double>(), red_out.as<double>());
        else
            launch_diag(stream, x.as<float>(), y.as<float>(), z.as<float>(), vx.as<float>(), vy.as<float>(),
                        vz.as<float>(), m.as<float>(), pot.as<float>(), has_pe ? 1 : 0, n, red_part.as<double>(),
                        red_out.as<double>());
        double v[12];
        LHIP_CHECK(hipMemcpyAsync(v, red_out.as<void>(), sizeof v, hipMemcpyDeviceToHost, stream));
        sync();
        d.mass = v[0];
        d.ke = v[1];
        d.pe = has_pe ? 0.5 * v[2] : 0.0;
        d.px = v[3]; d.py = v[4]; d.pz = v[5];
        d.lx = v[6]; d.ly = v[7]; d.lz = v[8];
        if (v[0] > 0) { d.cx = v[9] / v[0]; d.cy = v[10] / v[0]; d.cz = v[11] / v[0]; }
        return d;
    }

    double exact_pe(double G, double eps) {
        LHIP_CHECK(hipSetDevice(dev));
        if (n < 2) return 0.0;
        const std::size_t blocks = grid_for(n, kBS);
        red_part.reserve(std::max<std::size_t>(blocks, 12 * kReduceBlocks) * sizeof(double));
        if (mixed) {
            launch_exact_pe(stream, D(xd), D(yd), D(zd), m.as<float>(), n, eps * eps, red_part.as<double>(),
                            red_out.as<double>());
        } else {
            launch_pack_posm(stream, x.as<float>(), y.as<float>(), z.as<float>(), m.as<float>(), posm.as<float4>(), n);
            launch_exact_pe(stream, posm.as<float4>(), n, eps * eps, red_part.as<double>(), red_out.as<double>());
        }
        double h = 0;
        LHIP_CHECK(hipMemcpyAsync(&h, red_out.as<void>(), sizeof h, hipMemcpyDeviceToHost, stream));
        sync();
        // The tree's posm (sorted order) is rebuilt by every sort; nothing else reads it in between.
        return -0.5 * G * h;
    }

    lhip_phase_times times() {
        lhip_phase_times r = t;
        double step = 0;
        if (kick_rec) { LHIP_CHECK(hipEventSynchronize(ev[6])); step += ms(5, 6); }
        if (drift_rec) { LHIP_CHECK(hipEventSynchronize(ev[8])); step += ms(7, 8); }
        r.step_ms = step;
        return r;
    }

    void local_bounds(float lo[3], float hi[3]) {
        LHIP_CHECK(hipSetDevice(dev));
        for (int c = 0; c < 3; ++c) lo[c] = hi[c] = 0.f;
        if (n == 0) return;
        if (mixed) launch_bbox(stream, D(xd), D(yd), D(zd), n, bbox_part.as<void>(), cube.as<Cube>());
        else launch_bbox(stream, F(x), F(y), F(z), n, bbox_part.as<void>(), cube.as<Cube>());
        Cube c{};
        LHIP_CHECK(hipMemcpyAsync(&c, cube.as<void>(), sizeof c, hipMemcpyDeviceToHost, stream));
        sync();
        for (int k = 0; k < 3; ++k) { lo[k] = c.lo[k]; hi[k] = c.hi[k]; }
    }

    // ---- LET export (src/backend.hpp:125-130)
    DevBuf let_state, let_ex, let_bcnt, let_tmp, let_out, let_bods;
    bool let_ready = false, let_host = false;
    uint64_t let_nn = 0, let_nb = 0;
    std::size_t let_stride = 0;  // entries per body array in let_bods
    lhip_let_node let_hnode{};
    std::vector<float> let_hx, let_hy, let_hz, let_hm;

    void export_let(const float lo[3], const float hi[3], const lhip_force_params& p, uint64_t* nn, uint64_t* nb) {
        static_assert(sizeof(lhip_let_node) == sizeof(LetWire), "LetWire is the ABI's LetNode");
        LHIP_CHECK(hipSetDevice(dev));
        let_ready = let_host = false;
        let_nn = let_nb = 0;
        if (n > 0 && p.method != LHIP_METHOD_TREE) {
            // Direct method: one always-open root holding every body in the current order, its com
            // summed in double over that order (src/cpu_backend.cpp:727-745).
            for (auto* v : {&let_hx, &let_hy, &let_hz, &let_hm}) v->resize(std::size_t(n));
            download(let_hx.data(), let_hy.data(), let_hz.data(), nullptr, nullptr, nullptr, let_hm.data(), nullptr);
            double M = 0, sx = 0, sy = 0, sz = 0;
            for (int i = 0; i < n; ++i) {
                const double w = let_hm[i];
                M += w; sx += w * let_hx[i]; sy += w * let_hy[i]; sz += w * let_hz[i];
            }
            lhip_let_node r{};
            if (M > 0) { r.cx = float(sx / M); r.cy = float(sy / M); r.cz = float(sz / M); }
            r.mass = float(M);
            r.rcrit2 = 3.0e38f;  // always opened: the quadrupole is never read and stays zero
            r.first_child = -1; r.child_count = 0;
            r.body_first = 0; r.body_count = n;
            let_hnode = r;
            let_host = true;
            let_nn = 1;
            let_nb = uint64_t(n);
        } else if (n > 0) {
            if (!tree_ok) throw Error(kLogic, "hip export_let: call build_tree after the last upload/drift first");
            uint32_t count = 0;
            LHIP_CHECK(hipMemcpyAsync(&count, cnt.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, stream));
            sync();
            const std::size_t c1 = std::size_t(count) + 1;
            for (DevBuf* b : {&let_state, &let_ex, &let_bcnt}) b->reserve(c1 * 4);
            let_tmp.reserve(scan_tmp_entries(int(count)) * 4);
            let_out.reserve(std::size_t(count) * sizeof(LetWire));
            let_bods.reserve(std::size_t(n) * 16);
            LetExport e{};
            e.node = node.as<WalkNode>();
            e.mom = mom.as<NodeMom>();
            e.levstart = levstart.as<uint32_t>();
            e.count = int(count);
            e.depth = walk_depth();
            e.f64 = mixed;
            if (mixed) { e.x = D(xd); e.y = D(yd); e.z = D(zd); }
            else { e.x = F(x); e.y = F(y); e.z = F(z); }
            e.m = m.as<float>();
            for (int k = 0; k < 3; ++k) { e.lo[k] = lo[k]; e.hi[k] = hi[k]; }
            e.state = let_state.as<uint32_t>();
            e.ex = let_ex.as<uint32_t>();
            e.bcnt = let_bcnt.as<uint32_t>();
            e.tmp = let_tmp.as<uint32_t>();
            e.out = let_out.as<LetWire>();
            float* bods = let_bods.as<float>();
            let_stride = std::size_t(n);
            e.bx = bods; e.by = bods + let_stride; e.bz = bods + 2 * let_stride; e.bm = bods + 3 * let_stride;
            let_export(stream, e);
            uint32_t tot[2] = {0, 0};
            LHIP_CHECK(hipMemcpyAsync(&tot[0], let_ex.as<uint32_t>() + count, 4, hipMemcpyDeviceToHost, stream));
            LHIP_CHECK(hipMemcpyAsync(&tot[1], let_bcnt.as<uint32_t>() + count, 4, hipMemcpyDeviceToHost, stream));
            sync();
            let_nn = tot[0];
            let_nb = tot[1];
        }
        let_ready = true;
        *nn = let_nn;
        *nb = let_nb;
    }

    void export_let_copy(lhip_let_node* nodes, float* bx, float* by, float* bz, float* bm) {
        LHIP_CHECK(hipSetDevice(dev));
        if (!let_ready) throw Error(kLogic, "hip export_let_copy: call export_let first");
        if ((let_nn && !nodes) || (let_nb && (!bx || !by || !bz || !bm)))
            throw Error(kInvalid, "hip export_let_copy: null output array");
        if (let_host) {
            nodes[0] = let_hnode;
            const std::size_t fb = std::size_t(let_nb) * 4;
            std::memcpy(bx, let_hx.data(), fb);
            std::memcpy(by, let_hy.data(), fb);
            std::memcpy(bz, let_hz.data(), fb);
            std::memcpy(bm, let_hm.data(), fb);
            return;
        }
        if (let_nn) LHIP_CHECK(hipMemcpyAsync(nodes, let_out.as<void>(), let_nn * sizeof(LetWire), hipMemcpyDeviceToHost, stream));
        const std::size_t fb = std::size_t(let_nb) * 4;
        const float* bods = let_bods.as<float>();
        float* dst[4] = {bx, by, bz, bm};
        if (let_nb)
            for (int c = 0; c < 4; ++c)
                LHIP_CHECK(hipMemcpyAsync(dst[c], bods + c * let_stride, fb, hipMemcpyDeviceToHost, stream));
        sync();
    }

    // ---- snapshots (src/simulation.cpp:311-325 writes download() order; restart uploads it)
    void snapshot_write(const char* path, const lhip_snapshot_header& h) {
        const std::size_t k = std::size_t(n);
        std::vector<float> f[7];
        for (auto& v : f) v.resize(k);
        std::vector<uint32_t> ids(k);
        download(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), f[6].data(), ids.data());
        std::vector<double> d[6];
        if (mixed) {
            for (auto& v : d) v.resize(k);
            download_f64(d[0].data(), d[1].data(), d[2].data(), d[3].data(), d[4].data(), d[5].data(), nullptr,
                         nullptr);
        }
        const int rc = lhip_snapshot_save(path, &h, k, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(),
                                          f[5].data(), f[6].data(), ids.data(), mixed ? d[0].data() : nullptr,
                                          d[1].data(), d[2].data(), d[3].data(), d[4].data(), d[5].data());
        if (rc != LHIP_OK) throw Error(rc, lhip_snapshot_error());
    }

    void snapshot_restore(const char* path, lhip_snapshot_header* out) {
        sort_depth = -1;  // new bodies: the next sort takes every digit
        lhip_snapshot_header h{};
        int rc = lhip_snapshot_read_header(path, &h);
        if (rc != LHIP_OK) throw Error(rc, lhip_snapshot_error());
        if (h.n >= uint64_t(INT32_MAX))
            throw Error(kInvalid, "hip backend: n = " + std::to_string(h.n) + " exceeds the 2^31-1 body limit");
        const std::size_t k = std::size_t(h.n);
        std::vector<float> f[7];
        for (auto& v : f) v.resize(k);
        std::vector<uint32_t> ids(k);
        const bool want64 = mixed && h.fp64;
        std::vector<double> d[6];
        if (want64)
            for (auto& v : d) v.resize(k);
        auto p64 = [&](int c) { return want64 ? d[c].data() : nullptr; };
        rc = lhip_snapshot_load(path, &h, k, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(),
                                f[5].data(), f[6].data(), ids.data(), p64(0), p64(1), p64(2), p64(3), p64(4), p64(5));
        if (rc != LHIP_OK) throw Error(rc, lhip_snapshot_error());
        if (want64)
            upload_f64(h.n, d[0].data(), d[1].data(), d[2].data(), d[3].data(), d[4].data(), d[5].data(), f[6].data(),
                       ids.data());
        else
            upload(h.n, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), f[6].data(),
                   ids.data());
        *out = h;
    }

    bool body_costs(float* out) {
        LHIP_CHECK(hipSetDevice(dev));
        if (!costs_ok || n == 0) return false;
        if (direct_costs_pending) {
            for (int i = 0; i < n; ++i) out[i] = float(n - 1);
            return true;
        }
        LHIP_CHECK(hipMemcpyAsync(out, cost.as<void>(), std::size_t(n) * 4, hipMemcpyDeviceToHost, stream));
        sync();
        return true;
    }

    // Every buffer of the force path: the state, the sort, the tree, the walk, LETs and checks (not
    // the differentiable simulations' own, which the force path never reuses).
    std::size_t held_bytes() const {
        std::size_t held = 0;
        for (const DevBuf* b :
             {&x, &y, &z, &vx, &vy, &vz, &m, &id, &cost, &x2, &y2, &z2, &vx2, &vy2, &vz2, &m2, &id2, &cost2, &ax, &ay,
              &az, &pot, &posm, &key, &key2, &idx, &idx2, &radix_buf, &cube, &bbox_part, &red_part, &red_out,
              &direct_part, &bl, &cnt, &gflag, &scan_tmp, &pkey, &pkey2, &pval, &pval2, &inv, &levstart, &node, &mom,
              &prefix, &gstart, &worder, &counters, &fkey, &fseg, &capb, &fcube, &rs_seg, &rs_bbox,
              &rs_cstart, &rs_gkey, &rs_ka, &rs_kb, &rs_va, &rs_vb, &xd, &yd, &zd, &vxd, &vyd, &vzd, &xd2, &yd2, &zd2,
              &vxd2, &vyd2, &vzd2, &axd, &ayd, &azd, &potd, &momd, &rec64, &posd, &chk_tgt, &chk_out, &chk_part,
              &let_rec, &let_body, &let_root, &let_state, &let_ex, &let_bcnt, &let_tmp, &let_out, &let_bods, &lpool, &lnext,
              &lhead})
            held += b->capacity();
        return held;
    }

    // Free device bytes counting what this context holds (a grow-only buffer is freed before it is
    // reallocated), and the total (src/backend.hpp:138-143).
    void device_memory(uint64_t* fr, uint64_t* tot) {
        LHIP_CHECK(hipSetDevice(dev));
        std::size_t f = 0, tt = 0;
        LHIP_CHECK(hipMemGetInfo(&f, &tt));
        *fr = f + held_bytes();
        *tot = tt;
    }
};

// Device bytes a force call on n bodies reserves (Backend::memory_need, src/backend.hpp:137): the
// same sizes the reserve calls use, buffer by buffer, each at least 256 bytes, plus 64 MiB for
// allocator rounding. The node arrays are counted at their first size (n / 4 + 4096 nodes). Not
// included: node arrays grown for a tree with more nodes; the key-restart buffers (about 76 B per body
// with 4 key frames), reserved the first time a build restarts keys (clustered or coincident bodies)
// and kept; remote LETs and their exports, whose sizes depend on the other ranks; the differentiable
// simulations; and the interaction-list pool of LETTUCE_HIP_WALK=lists (about 1 KiB per 4 bodies).
// Host: nothing per body beyond the caller's arrays.
inline void memory_need(uint64_t n, int method, bool mixed, uint64_t* dev, uint64_t* host) {
    const uint64_t N = std::min<uint64_t>(n, INT32_MAX - 1);
    const int ni = int(N);
    uint64_t d = 0;
    auto buf = [&d](uint64_t bytes) { d += std::max<uint64_t>(bytes, 256); };
    for (int k = 0; k < 22; ++k) buf(N * 4);  // state, permute alternates, a, phi, costs
    buf(N * 16);                               // posm
    buf(sizeof(Cube));
    buf(sizeof(double) * 6 * kReduceBlocks);
    buf(std::max<uint64_t>(grid_for(ni, kBS), 12 * kReduceBlocks) * 8);  // reductions, exact pe
    buf(sizeof(double) * 16);
    buf(sizeof(WalkCounters));
    if (mixed)
        for (int k = 0; k < 16; ++k) buf(N * 8);  // FP64 state, alternates, a, phi
    if (method == LHIP_METHOD_TREE) {
        const uint64_t C = uint64_t(node_reserve_initial((long long)N));
        buf(N * 8); buf(N * 8); buf(N * 4); buf(N * 4);  // keys, sort index (double-buffered)
        buf(radix_work_words(int(std::max<uint64_t>(N, C))) * 4);        // radix sort work
        buf(N * 4);                                                        // levels word
        for (int k = 0; k < 4; ++k) buf((N + 1) * 4);                      // counts, groups (flags, starts), order
        buf(2 * scan_tmp_entries(ni) * 4);
        buf(C * 8); buf(C * 8); buf(C * 4); buf(C * 4); buf(C * 4);        // level sort, inverse
        buf((kDepthLevels + 2) * 4);
        buf(C * sizeof(WalkNode)); buf(C * sizeof(NodeMom)); buf(C * 8);  // records, key prefixes
        if (mixed) {
            buf(C * sizeof(NodeMomD));
            buf(N * sizeof(double4));  // posd
        }
    } else if (mixed) {
        buf(uint64_t(direct64_split(ni)) * 4 * N * 8);
        buf(N * sizeof(SimRec));
    } else {
        const int S = direct_split(ni);
        if (S > 1) buf(uint64_t(S) * N * sizeof(float4));
    }
    *dev = d + (uint64_t(64) << 20);
    *host = 0;
}

}  // namespace lhip

struct lhip_ctx {
    lhip::Context* c;
};

namespace lhip {

// Runs f(Context&) for a C entry point: exceptions become status codes and the context's message.
template <class F>
int guarded(lhip_ctx* ctx, F&& f) {
    if (!ctx || !ctx->c) return LHIP_INVALID;
    try {
        f(*ctx->c);
        ctx->c->err.clear();
        return LHIP_OK;
    } catch (const Error& e) {
        ctx->c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->c->err = e.what();
        return LHIP_RUNTIME;
    }
}

}  // namespace lhip
