// Differentiable simulation through the tree: FP64-state KDK with the mixed-precision walk, step
// checkpoints, and the adjoint with frozen topology (no reference counterpart: the reference has no
// differentiation; the forward is the reference's KDK, src/simulation.cpp:406-438).
//
// The simulation runs on the context's own state in the mixed mode (it replaces the resident
// bodies). Each force call sorts the bodies, so the body order changes every step; the cotangents
// are kept per body id.
//
// Forward: a tree per force call. RECORD also keeps, per force call, the bodies' sorted order and
// positions; REPLAY (the frozen-topology forward that finite differences need) rebuilds each
// step's tree and walk decisions from those records and evaluates the forces at the current
// positions: the state is brought into the recorded order, keys, levels, nodes, groups and the
// decision moments (rcrit2, qrel, node com) come from the recorded positions, and a second moments
// pass takes com, mass and quadrupole from the current ones; the group boxes are the recorded ones.
//
// Backward: segment by segment from the checkpoints (x, v, a, id, m at a segment start), a
// recompute that keeps each step's positions (in the order its tree was built from), then the
// merged-kick reverse KDK of the direct-sum adjoint: VJP(x_T, dt/2); for n = T-1 .. 0
// { lambda_v += dt lambda_x; VJP(x_n, n > 0 ? dt : dt/2) }. A VJP rebuilds the tree at x_n (the
// forward's tree: the same positions in the same order sort the same way), gathers
// u = G w lambda_v in sorted order and runs tree_vjp; the result is added to lambda_x by id.
#include <vector>

#include "adjoint.hpp"
#include "context.hpp"

namespace lhip {
namespace {

// u (SoA, tree order) = c lambda (x, y, z per body, id order) of the tree's bodies; the source
// sides (SoA) start at zero.
__global__ void __launch_bounds__(kBS) gather_u_kernel(const uint32_t* __restrict__ id, const double* __restrict__ lam,
                                                       double c, double* __restrict__ ux, double* __restrict__ uy,
                                                       double* __restrict__ uz, double* __restrict__ gsx,
                                                       double* __restrict__ gsy, double* __restrict__ gsz, int n) {
    const int k = blockIdx.x * kBS + threadIdx.x;
    if (k >= n) return;
    const double* l = lam + 3 * std::size_t(id[k]);
    ux[k] = c * l[0];
    uy[k] = c * l[1];
    uz[k] = c * l[2];
    gsx[k] = 0.0;
    gsy[k] = 0.0;
    gsz[k] = 0.0;
}

// Zeroes up to three double arrays in one launch (blockIdx.y picks the array).
struct ZeroSet {
    double* p[3];
    long long n[3];
};

__global__ void __launch_bounds__(kBS) zero_set_kernel(ZeroSet z) {
    double* __restrict__ p = z.p[blockIdx.y];
    const long long n = z.n[blockIdx.y];
    for (long long i = (long long)blockIdx.x * kBS + threadIdx.x; i < n; i += (long long)gridDim.x * kBS) p[i] = 0.0;
}

__global__ void __launch_bounds__(kBS) inverse_kernel(const uint32_t* __restrict__ id, uint32_t* __restrict__ inv, int n) {
    const int k = blockIdx.x * kBS + threadIdx.x;
    if (k < n) inv[id[k]] = uint32_t(k);
}

// Up to kCopySegs device-to-device copies (whole 4-byte words) in one launch: blockIdx.y picks
// the copy; 16-byte words where both ends allow them.
constexpr int kCopySegs = 12;
struct CopySet {
    const void* src[kCopySegs];
    void* dst[kCopySegs];
    long long words[kCopySegs];
};

__global__ void __launch_bounds__(kBS) copy_set_kernel(CopySet cs) {
    const int g = int(blockIdx.y);
    const long long words = cs.words[g];
    const long long t0 = (long long)blockIdx.x * kBS + threadIdx.x, step = (long long)gridDim.x * kBS;
    const uint32_t* __restrict__ src = static_cast<const uint32_t*>(cs.src[g]);
    uint32_t* __restrict__ dst = static_cast<uint32_t*>(cs.dst[g]);
    long long done = 0;
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0) {
        const long long quads = words >> 2;
        const uint4* __restrict__ s4 = reinterpret_cast<const uint4*>(src);
        uint4* __restrict__ d4 = reinterpret_cast<uint4*>(dst);
        for (long long q = t0; q < quads; q += step) d4[q] = s4[q];
        done = quads << 2;
    }
    for (long long w = done + t0; w < words; w += step) dst[w] = src[w];
}

// Collects copies and issues them as one copy_set_kernel launch per kCopySegs copies (the last
// launch at run()).
struct Copies {
    hipStream_t s;
    CopySet cs{};
    int count = 0;
    long long most = 0;
    void add(void* dst, const void* src, std::size_t bytes) {
        if (bytes == 0) return;
        if (count == kCopySegs) run();
        cs.src[count] = src;
        cs.dst[count] = dst;
        cs.words[count] = static_cast<long long>(bytes / 4);
        most = std::max(most, cs.words[count]);
        ++count;
    }
    void run() {
        if (count == 0) return;
        const unsigned gx = std::min<unsigned>(std::max(1u, grid_for((most + 3) / 4, kBS)), 1024u);
        const int segs = count;
        count = 0;
        most = 0;
        copy_set_kernel<<<dim3(gx, segs), kBS, 0, s>>>(cs);
        LHIP_LAUNCH_CHECK();
    }
};

// The six FP64 state arrays (x, y, z, vx, vy, vz) as one argument.
struct State6 {
    const double* a[6];
};

// Moves the state into a recorded order: new position k holds the body with id rec[k], currently
// at inv[rec[k]]; dst holds the six arrays at stride n.
__global__ void __launch_bounds__(kBS) reorder_kernel(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ inv,
                                                      State6 src, double* __restrict__ dst, int n) {
    const int k = blockIdx.x * kBS + threadIdx.x;
    if (k >= n) return;
    const uint32_t j = inv[rec[k]];
#pragma unroll
    for (int c = 0; c < 6; ++c) dst[std::size_t(c) * n + k] = src.a[c][j];
}

__global__ void __launch_bounds__(kBS) reorder_small_kernel(const uint32_t* __restrict__ rec,
                                                            const uint32_t* __restrict__ inv,
                                                            const float* __restrict__ m, float* __restrict__ m2,
                                                            uint32_t* __restrict__ id2, int n) {
    const int k = blockIdx.x * kBS + threadIdx.x;
    if (k >= n) return;
    m2[k] = m[inv[rec[k]]];
    id2[k] = rec[k];
}

__global__ void __launch_bounds__(kBS) pack_posd_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const float* __restrict__ m,
                                                        double4* __restrict__ p, int n) {
    const int k = blockIdx.x * kBS + threadIdx.x;
    if (k < n) p[k] = make_double4(x[k], y[k], z[k], double(m[k]));
}

// A recomputed step's tree into a store slot: the node records of the device node count (at most
// cap; flag = 1 and nothing copied beyond that).
__global__ void __launch_bounds__(kBS) store_nodes_kernel(const uint32_t* __restrict__ node_count,
                                                          const WalkNode* __restrict__ node,
                                                          const NodeMom* __restrict__ mom,
                                                          const NodeMomD* __restrict__ momd, long long cap,
                                                          WalkNode* __restrict__ snode, NodeMom* __restrict__ smom,
                                                          NodeMomD* __restrict__ smomd, uint32_t* __restrict__ flag) {
    const long long M = *node_count;
    const long long i0 = (long long)blockIdx.x * kBS + threadIdx.x;
    if (i0 == 0) *flag = M > cap ? 1u : 0u;
    if (M > cap) return;
    for (long long i = i0; i < M; i += (long long)gridDim.x * kBS) {
        snode[i] = node[i];
        smom[i] = mom[i];
        smomd[i] = momd[i];
    }
}

}  // namespace

struct TreeSim {
    int n = 0;
    lhip_sim_params p{};
    int nck = 0;
    bool have_forward = false, have_record = false;
    int rec_steps = 0;
    double fwd_ms = 0, bwd_ms = 0;
    hipEvent_t ev[4] = {};
    // checkpoints: 9 double arrays (x, y, z, vx, vy, vz, ax, ay, az) + id + m per segment start
    DevBuf ck_d, ck_id, ck_m;
    // recompute of a segment: positions (3 arrays), id, m per step
    DevBuf seg_d, seg_id, seg_m;
    // records per force call: sorted ids and sorted positions
    DevBuf rec_id, rec_x;
    // cotangents (id order) and VJP scratch (sorted order)
    DevBuf lx, lvx, u, gt, gs, gcell, gmono, gpriv, cum, inv, tmp, tmp_m, posd_rec;
    // Trees of a segment's recomputed steps (slot k = step start + k), reused by their VJPs instead
    // of a second build: node records (at most tr_cap per slot), level starts, node count, walk
    // units and their count; tr_flag[k] = 1 where a tree did not fit.
    DevBuf tr_node, tr_mom, tr_momd, tr_lev, tr_cnt, tr_starts, tr_count, tr_order, tr_flag;
    DevBuf vjp_order;  // the VJP's group order on a tree built for it
    DevBuf tr_ranges, vjp_ranges;  // the range starts of those orders (kWalkRanges per order)
    long long tr_cap = 0;
    int tr_slots = 0;
    std::vector<int> tr_frames;
    std::vector<int> tr_depth;  // deepest level of each stored tree

    TreeSim() {
        for (auto& e : ev) LHIP_CHECK(hipEventCreate(&e));
    }
    ~TreeSim() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void treesim_destroy(TreeSim* t) { delete t; }

namespace {

struct Run {
    Context& c;
    TreeSim& S;
    hipStream_t s;
    int n;
    lhip_force_params fp;

    double* D(DevBuf& b) { return b.as<double>(); }

    void force() { c.compute_forces(fp, false); }

    // State arrays of the context in checkpoint order.
    DevBuf* state(int k) {
        DevBuf* a[9] = {&c.xd, &c.yd, &c.zd, &c.vxd, &c.vyd, &c.vzd, &c.axd, &c.ayd, &c.azd};
        return a[k];
    }

    void save_checkpoint(int k) {
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        double* base = D(S.ck_d) + std::size_t(k) * 9 * n;
        Copies cp{s};
        for (int a = 0; a < 9; ++a) cp.add(base + std::size_t(a) * n, state(a)->as<void>(), db);
        cp.add(S.ck_id.as<uint32_t>() + std::size_t(k) * n, c.id.as<void>(), fb);
        cp.add(S.ck_m.as<float>() + std::size_t(k) * n, c.m.as<void>(), fb);
        cp.run();
    }

    void load_checkpoint(int k) {
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        const double* base = D(S.ck_d) + std::size_t(k) * 9 * n;
        {
            Copies cp{s};
            for (int a = 0; a < 9; ++a) cp.add(state(a)->as<void>(), base + std::size_t(a) * n, db);
            cp.add(c.id.as<void>(), S.ck_id.as<uint32_t>() + std::size_t(k) * n, fb);
            cp.add(c.m.as<void>(), S.ck_m.as<float>() + std::size_t(k) * n, fb);
            cp.run();
        }
        c.tree_ok = false;
        c.acc_ok = true;
    }

    // Positions, ids and masses of the current state into slot k of a (positions, id, m) store.
    void save_positions(DevBuf& dd, DevBuf& did, DevBuf& dm, int k) {
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        double* base = D(dd) + std::size_t(k) * 3 * n;
        Copies cp{s};
        for (int a = 0; a < 3; ++a) cp.add(base + std::size_t(a) * n, state(a)->as<void>(), db);
        cp.add(did.as<uint32_t>() + std::size_t(k) * n, c.id.as<void>(), fb);
        cp.add(dm.as<float>() + std::size_t(k) * n, c.m.as<void>(), fb);
        cp.run();
    }

    void load_positions(DevBuf& dd, DevBuf& did, DevBuf& dm, int k) {
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        const double* base = D(dd) + std::size_t(k) * 3 * n;
        {
            Copies cp{s};
            for (int a = 0; a < 3; ++a) cp.add(state(a)->as<void>(), base + std::size_t(a) * n, db);
            cp.add(c.id.as<void>(), did.as<uint32_t>() + std::size_t(k) * n, fb);
            cp.add(c.m.as<void>(), dm.as<float>() + std::size_t(k) * n, fb);
            cp.run();
        }
        c.tree_ok = false;
    }

    // The force at the current positions with the topology of recorded force call r: state into
    // the recorded order, tree and decisions from the recorded positions, moments at the current.
    void replay_force(int r) {
        const uint32_t* rid = S.rec_id.as<uint32_t>() + std::size_t(r) * n;
        const double* rx = D(S.rec_x) + std::size_t(r) * 3 * n;
        inverse_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(c.id.as<uint32_t>(), S.inv.as<uint32_t>(), n);
        LHIP_LAUNCH_CHECK();
        // 6 state arrays (x, v) into tmp in the recorded order, masses into tmp_m (ids from the
        // record), then back
        double* t = D(S.tmp);
        State6 st{};
        for (int a = 0; a < 6; ++a) st.a[a] = D(*state(a));
        reorder_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(rid, S.inv.as<uint32_t>(), st, t, n);
        LHIP_LAUNCH_CHECK();
        reorder_small_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(rid, S.inv.as<uint32_t>(), c.m.as<float>(),
                                                              S.tmp_m.as<float>(), c.id.as<uint32_t>(), n);
        LHIP_LAUNCH_CHECK();
        {
            Copies cp{s};
            for (int a = 0; a < 6; ++a) cp.add(state(a)->as<void>(), t + std::size_t(a) * n, std::size_t(n) * 8);
            cp.add(c.m.as<void>(), S.tmp_m.as<void>(), std::size_t(n) * 4);
            cp.run();
        }
        // tree from the recorded positions (already in key order)
        c.reserve_tree(n);
        c.momd.reserve(std::size_t(c.cap_nodes) * sizeof(NodeMomD));
        c.posd.reserve(std::size_t(n) * sizeof(double4));
        S.posd_rec.reserve(std::size_t(n) * sizeof(double4));
        LHIP_CHECK(hipEventRecord(c.ev[0], s));
        tree_keys_sorted(s, rx, rx + n, rx + 2 * std::size_t(n), n, c.bbox_part.as<void>(), c.cube.as<Cube>(),
                         c.key.as<uint64_t>(), c.idx.as<uint32_t>());
        LHIP_CHECK(hipEventRecord(c.ev[1], s));
        c.frames = 1;
        BuildBufs bb = c.build_bufs();
        bb.xd = rx; bb.yd = rx + n; bb.zd = rx + 2 * std::size_t(n);
        const float inv_theta = fp.theta > 0 ? float(1.0 / fp.theta) : FLT_MAX;
        c.built_depth = c.build_waited(bb, inv_theta);
        if (*reinterpret_cast<volatile int*>(c.host_over)) {
            // key restarts of the recorded call: its order is already the restarted one
            const double* box[3] = {rx, rx + n, rx + 2 * std::size_t(n)};
            c.restart_keys(box[0], box[1], box[2], false);
            c.rebuild(inv_theta, box);
        }
        bb = c.build_bufs();
        bb.xd = D(c.xd); bb.yd = D(c.yd); bb.zd = D(c.zd);
        bb.depth_hint = c.built_depth;
        tree_moments_keep(s, bb);
        pack_posd_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(rx, rx + n, rx + 2 * std::size_t(n), c.m.as<float>(),
                                                          S.posd_rec.as<double4>(), n);
        LHIP_LAUNCH_CHECK();
        pack_posd_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(D(c.xd), D(c.yd), D(c.zd), c.m.as<float>(),
                                                          c.posd.as<double4>(), n);
        LHIP_LAUNCH_CHECK();
        LHIP_CHECK(hipEventRecord(c.ev[2], s));
        c.tree_ok = true;
        c.posd_box = S.posd_rec.as<double4>();
        c.compute_forces_with(fp, false);
        c.posd_box = nullptr;
    }

    void record(int r) {
        const std::size_t db = std::size_t(n) * 8, fb = std::size_t(n) * 4;
        double* base = D(S.rec_x) + std::size_t(r) * 3 * n;
        Copies cp{s};
        for (int a = 0; a < 3; ++a) cp.add(base + std::size_t(a) * n, state(a)->as<void>(), db);
        cp.add(S.rec_id.as<uint32_t>() + std::size_t(r) * n, c.id.as<void>(), fb);
        cp.run();
    }

    // One force evaluation of the forward: r = index of the force call.
    void forward_force(int r, int mode) {
        if (mode == LHIP_TREE_REPLAY) {
            replay_force(r);
        } else {
            force();
            if (mode == LHIP_TREE_RECORD) record(r);
        }
    }
};

void host_soa(const double* aos, std::size_t n, std::vector<double>& soa) {
    soa.resize(3 * n);
    for (std::size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = aos[3 * i + c];
}

// Device SoA (3 arrays of n) in the body order given by ids -> host AoS in id order.
void to_host_by_id(hipStream_t s, const double* x, const double* y, const double* z, const uint32_t* id, int n,
                   double* out) {
    std::vector<double> h(3 * std::size_t(n));
    std::vector<uint32_t> ids(n);
    LHIP_CHECK(hipMemcpyAsync(h.data(), x, std::size_t(n) * 8, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipMemcpyAsync(h.data() + n, y, std::size_t(n) * 8, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipMemcpyAsync(h.data() + 2 * std::size_t(n), z, std::size_t(n) * 8, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipMemcpyAsync(ids.data(), id, std::size_t(n) * 4, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) out[3 * std::size_t(ids[k]) + c] = h[std::size_t(c) * n + k];
}

}  // namespace

void treesim_forward(Context& c, const lhip_sim_params& p, int n, const double* x0, const double* v0, const double* m,
                     double* xT, double* vT) {
    if (!c.tsim) c.tsim = new TreeSim();
    TreeSim& S = *c.tsim;
    const int mode = p.tree_mode;
    if (mode != LHIP_TREE_BUILD && mode != LHIP_TREE_RECORD && mode != LHIP_TREE_REPLAY)
        throw Error(kInvalid, "sim: unknown tree mode");
    const bool replay = mode == LHIP_TREE_REPLAY;
    if (replay && (!S.have_record || S.n != n || S.rec_steps != p.steps))
        throw Error(kLogic, "sim: replay needs a recorded forward of the same size and length");
    // The topology a replay freezes is the recorded forward's: same opening angle and walk options.
    if (replay && (p.theta != S.p.theta || p.walk.group_max != S.p.walk.group_max ||
                   p.walk.small_leaf != S.p.walk.small_leaf || p.walk.quad_tau != S.p.walk.quad_tau))
        throw Error(kLogic, "sim: a replay takes the recorded forward's theta and walk options");
    if (!(p.eps > 0)) throw Error(kInvalid, "sim: the tree method needs eps > 0");
    const lhip_walk_options* wo = p.walk.group_max > 0 ? &p.walk : nullptr;
    c.set_precision(LHIP_MODE_MIXED, wo);
    const int T = p.steps, K = p.checkpoint_every;
    const std::size_t N = std::size_t(n);
    // A replay leaves the recorded forward's parameters, checkpoints and adjoint in place.
    if (!replay) {
        S.n = n;
        S.p = p;
        S.have_forward = false;
        S.have_record = false;
        S.nck = T > 0 ? (T + K - 1) / K : 1;
        S.ck_d.reserve(std::size_t(S.nck) * 9 * N * 8);
        S.ck_id.reserve(std::size_t(S.nck) * N * 4);
        S.ck_m.reserve(std::size_t(S.nck) * N * 4);
    }
    if (mode == LHIP_TREE_RECORD) {
        S.rec_x.reserve(std::size_t(T + 1) * 3 * N * 8);
        S.rec_id.reserve(std::size_t(T + 1) * N * 4);
    }
    S.inv.reserve(N * 4);
    S.tmp.reserve(6 * N * 8);
    S.tmp_m.reserve(N * 4);
    // initial state: FP64 upload with ids 0..n-1 and float masses
    std::vector<double> sx, sv;
    host_soa(x0, N, sx);
    host_soa(v0, N, sv);
    std::vector<float> mf(N);
    std::vector<uint32_t> ids(N);
    for (std::size_t i = 0; i < N; ++i) { mf[i] = float(m[i]); ids[i] = uint32_t(i); }
    c.upload_f64(N, sx.data(), sx.data() + N, sx.data() + 2 * N, sv.data(), sv.data() + N, sv.data() + 2 * N,
                 mf.data(), ids.data());
    hipStream_t s = c.stream;
    lhip_force_params fp{p.G, p.eps, p.theta, LHIP_METHOD_TREE, 0};
    Run R{c, S, s, n, fp};
    const double h = 0.5 * p.dt;
    LHIP_CHECK(hipEventRecord(S.ev[0], s));
    R.forward_force(0, mode);
    if (mode != LHIP_TREE_REPLAY) R.save_checkpoint(0);
    for (int st = 0; st < T; ++st) {
        c.kick(h);
        c.drift(p.dt);
        R.forward_force(st + 1, mode);
        c.kick(h);
        if (mode != LHIP_TREE_REPLAY && (st + 1) % K == 0 && st + 1 < T) R.save_checkpoint((st + 1) / K);
    }
    LHIP_CHECK(hipEventRecord(S.ev[1], s));
    if (xT) to_host_by_id(s, R.D(c.xd), R.D(c.yd), R.D(c.zd), c.id.as<uint32_t>(), n, xT);
    if (vT) to_host_by_id(s, R.D(c.vxd), R.D(c.vyd), R.D(c.vzd), c.id.as<uint32_t>(), n, vT);
    c.sync();
    float ms = 0;
    LHIP_CHECK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
    S.fwd_ms = ms;
    if (mode == LHIP_TREE_RECORD) {
        S.have_record = true;
        S.rec_steps = T;
    }
    if (!replay) S.have_forward = true;
}

void treesim_backward(Context& c, const double* gxT, const double* gvT, double* gx0, double* gv0) {
    if (!c.tsim || !c.tsim->have_forward) throw Error(kLogic, "sim backward: call lhip_sim_forward first");
    TreeSim& S = *c.tsim;
    const int n = S.n, T = S.p.steps, K = S.p.checkpoint_every;
    if (c.n != n) throw Error(kLogic, "sim backward: the context's bodies changed since the forward");
    const std::size_t N = std::size_t(n);
    const double dt = S.p.dt;
    hipStream_t s = c.stream;
    lhip_force_params fp{S.p.G, S.p.eps, S.p.theta, LHIP_METHOD_TREE, 0};
    const lhip_walk_options* wo = S.p.walk.group_max > 0 ? &S.p.walk : nullptr;
    c.set_precision(LHIP_MODE_MIXED, wo);
    Run R{c, S, s, n, fp};
    S.lx.reserve(3 * N * 8);
    S.lvx.reserve(3 * N * 8);
    S.u.reserve(3 * N * 8);
    S.gt.reserve(3 * N * 8);
    S.gs.reserve(3 * N * 8);
    S.seg_d.reserve(std::size_t(K + 1) * 3 * N * 8);
    S.seg_id.reserve(std::size_t(K + 1) * N * 4);
    S.seg_m.reserve(std::size_t(K + 1) * N * 4);
    double* lx = R.D(S.lx);
    double* lv = R.D(S.lvx);
    {
        // lambda_x and lambda_v stay as the caller's (x, y, z) per body in id order, so that a
        // gather or scatter by id touches one record
        LHIP_CHECK(hipMemcpyAsync(lx, gxT, 3 * N * 8, hipMemcpyHostToDevice, s));
        LHIP_CHECK(hipMemcpyAsync(lv, gvT, 3 * N * 8, hipMemcpyHostToDevice, s));
        c.sync();
    }
    const int blocks = c.cus * tree_vjp_blocks_per_cu();
    // Tree store for the K - 1 inner steps of a segment, sized on the first recomputed tree (twice
    // its node count per slot) and limited to kTreeStoreBytes; off when K < 2 or over the limit.
    constexpr long long kTreeStoreBytes = 16ll << 30;
    constexpr int kLev = kDepthLevels + 2;
    S.tr_slots = 0;
    bool tr_sized = K < 2;
    auto size_store = [&]() {
        tr_sized = true;
        uint32_t M0 = 0;
        LHIP_CHECK(hipMemcpyAsync(&M0, c.cnt.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, s));
        c.sync();
        const long long cap = std::max<long long>(2ll * M0, 1024);
        const long long per = cap * (long long)(sizeof(WalkNode) + sizeof(NodeMom) + sizeof(NodeMomD)) +
                              3ll * (n + 1) * 4 + kLev * 4 + 8;
        if (per * (K - 1) > kTreeStoreBytes) return;
        S.tr_cap = cap;
        S.tr_slots = K - 1;
        const std::size_t sl = std::size_t(S.tr_slots);
        S.tr_node.reserve(sl * cap * sizeof(WalkNode));
        S.tr_mom.reserve(sl * cap * sizeof(NodeMom));
        S.tr_momd.reserve(sl * cap * sizeof(NodeMomD));
        S.tr_lev.reserve(sl * kLev * 4);
        S.tr_cnt.reserve(sl * 4);
        S.tr_starts.reserve(sl * (N + 1) * 4);
        S.tr_count.reserve(sl * (N + 1) * 4);
        S.tr_order.reserve(sl * (N + 1) * 4);
        S.tr_ranges.reserve(sl * kWalkRanges * 4);
        S.tr_flag.reserve(sl * 4);
        S.tr_frames.assign(sl, 1);
        S.tr_depth.assign(sl, tree_depth(1));
    };
    auto slot_node = [&](int j) { return S.tr_node.as<WalkNode>() + std::size_t(j) * S.tr_cap; };
    auto slot_mom = [&](int j) { return S.tr_mom.as<NodeMom>() + std::size_t(j) * S.tr_cap; };
    auto slot_momd = [&](int j) { return S.tr_momd.as<NodeMomD>() + std::size_t(j) * S.tr_cap; };
    // the tree the recompute just built at segment step k (1 <= k <= K - 1) into slot k - 1
    // The VJP's order: eight Morton ranges (one per XCD first, as the walk's), each with its units
    // above 1/64 of a wave's share of the last walk's work first, longest first.
    auto vjp_order_threshold = [&]() { return c.walk_order_threshold() * 0.125f; };
    auto store_tree = [&](int k) {
        if (!tr_sized) size_store();
        if (k < 1 || k > S.tr_slots) return;
        const int j = k - 1;
        const unsigned g = std::min<unsigned>(grid_for(S.tr_cap, kBS), 2048u);
        store_nodes_kernel<<<g, kBS, 0, s>>>(c.cnt.as<uint32_t>() + n, c.node.as<WalkNode>(), c.mom.as<NodeMom>(),
                                             c.momd.as<NodeMomD>(), S.tr_cap, slot_node(j), slot_mom(j),
                                             slot_momd(j), S.tr_flag.as<uint32_t>() + j);
        LHIP_LAUNCH_CHECK();
        {
            Copies cp{s};
            cp.add(S.tr_lev.as<uint32_t>() + std::size_t(j) * kLev, c.levstart.as<void>(), kLev * 4);
            cp.add(S.tr_cnt.as<uint32_t>() + j, c.cnt.as<uint32_t>() + n, 4);
            cp.add(S.tr_starts.as<uint32_t>() + std::size_t(j) * (N + 1), c.walk_starts(), (N + 1) * 4);
            cp.add(S.tr_count.as<uint32_t>() + std::size_t(j) * (N + 1), c.walk_count(), (N + 1) * 4);
            cp.run();
        }
        unit_order(s, c.build_bufs(), c.walk_starts(), c.walk_count(), vjp_order_threshold(),
                   S.tr_ranges.as<uint32_t>() + std::size_t(j) * kWalkRanges,
                   S.tr_order.as<uint32_t>() + std::size_t(j) * (N + 1));
        S.tr_frames[j] = c.frames;
        S.tr_depth[j] = c.walk_depth();
    };
    std::vector<uint32_t> tr_ok;  // per slot of the current segment: 1 if its tree was stored whole
    auto vjp_at = [&](int k, double w) {
        R.load_positions(S.seg_d, S.seg_id, S.seg_m, k);
        c.posd.reserve(N * sizeof(double4));
        const int j = k - 1;
        const bool stored = j >= 0 && j < int(tr_ok.size()) && tr_ok[j];
        if (stored) {
            // the recompute's tree at these positions; only the FP64 records are repacked
            pack_posd_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(R.D(c.xd), R.D(c.yd), R.D(c.zd), c.m.as<float>(),
                                                              c.posd.as<double4>(), n);
            LHIP_LAUNCH_CHECK();
        } else {
            c.build_tree(fp);  // the forward's tree at these positions
        }
        // node scratch covers the tree in use (a stored tree may exceed the last build's bound)
        const int nodes = int(stored ? std::max<long long>(c.cap_nodes, S.tr_cap) : c.cap_nodes);
        S.gcell.reserve(std::size_t(nodes) * 9 * 8);
        S.gmono.reserve(std::size_t(nodes) * 3 * 8);
        S.cum.reserve(std::size_t(nodes) * 9 * 8);
        double* u = R.D(S.u);
        double* gs = R.D(S.gs);
        double* gt = R.D(S.gt);
        gather_u_kernel<<<grid_for(n, kBS), kBS, 0, s>>>(c.id.as<uint32_t>(), lv, S.p.G * w, u, u + N, u + 2 * N, gs,
                                                         gs + N, gs + 2 * N, n);
        LHIP_LAUNCH_CHECK();
        S.gpriv.reserve(std::size_t(blocks) * kHot * 12 * 8);
        {
            const ZeroSet z{{R.D(S.gcell), R.D(S.gmono), R.D(S.gpriv)},
                            {9ll * nodes, 3ll * nodes, (long long)blocks * kHot * 12}};
            const long long most = std::max(z.n[0], std::max(z.n[1], z.n[2]));
            zero_set_kernel<<<dim3(std::min<unsigned>(std::max(1u, grid_for(most, kBS)), 1024u), 3), kBS, 0, s>>>(z);
            LHIP_LAUNCH_CHECK();
        }
        TreeVjpArgs a{};
        a.node = stored ? slot_node(j) : c.node.as<WalkNode>();
        a.mom = stored ? slot_mom(j) : c.mom.as<NodeMom>();
        a.momd = stored ? slot_momd(j) : c.momd.as<NodeMomD>();
        a.posd = c.posd.as<double4>();
        a.ux = u; a.uy = u + N; a.uz = u + 2 * N;
        a.gstart = stored ? S.tr_starts.as<uint32_t>() + std::size_t(j) * (N + 1) : c.walk_starts();
        a.gcount = stored ? S.tr_count.as<uint32_t>() + std::size_t(j) * (N + 1) : c.walk_count();
        if (!stored) {
            S.vjp_order.reserve((N + 1) * 4);
            S.vjp_ranges.reserve(kWalkRanges * 4);
            unit_order(s, c.build_bufs(), c.walk_starts(), c.walk_count(), vjp_order_threshold(),
                       S.vjp_ranges.as<uint32_t>(),
                       S.vjp_order.as<uint32_t>());
        }
        a.order = stored ? S.tr_order.as<uint32_t>() + std::size_t(j) * (N + 1) : S.vjp_order.as<uint32_t>();
        a.ranges = stored ? S.tr_ranges.as<uint32_t>() + std::size_t(j) * kWalkRanges : S.vjp_ranges.as<uint32_t>();
        a.n = n;
        a.eps2 = float(S.p.eps * S.p.eps);
        a.small_leaf = c.wopt.small_leaf;
        a.quad_tau = float(c.wopt.quad_tau);
        a.gtx = gt; a.gty = gt + N; a.gtz = gt + 2 * N;
        a.gsx = gs; a.gsy = gs + N; a.gsz = gs + 2 * N;
        a.gcell = R.D(S.gcell);
        a.gmono = R.D(S.gmono);
        a.gpriv = R.D(S.gpriv);
        a.counters = c.counters.as<WalkCounters>();
        a.blocks = blocks;
        a.depth = stored ? S.tr_depth[j] : c.walk_depth();
        a.id = c.id.as<uint32_t>();
        a.lam = lx;
        tree_vjp(s, a, stored ? S.tr_lev.as<uint32_t>() + std::size_t(j) * kLev : c.levstart.as<uint32_t>(),
                 stored ? S.tr_cnt.as<uint32_t>() + j : c.cnt.as<uint32_t>() + n, nodes, R.D(S.cum));
    };
    auto weight = [&](int q) { return (q == T || q == 0) ? 0.5 * dt : dt; };
    LHIP_CHECK(hipEventRecord(S.ev[2], s));
    for (int sg = S.nck - 1; sg >= 0 && T > 0; --sg) {
        const int start = sg * K, end = std::min(T, start + K);
        R.load_checkpoint(sg);
        R.save_positions(S.seg_d, S.seg_id, S.seg_m, 0);
        for (int t = start; t < end; ++t) {
            c.kick(0.5 * dt);
            c.drift(dt);
            if (t + 1 < end) {
                R.force();
                R.save_positions(S.seg_d, S.seg_id, S.seg_m, t + 1 - start);
                store_tree(t + 1 - start);
                c.kick(0.5 * dt);
            } else {
                R.save_positions(S.seg_d, S.seg_id, S.seg_m, t + 1 - start);
            }
        }
        // which of this segment's stored trees fit their slots (one read per segment)
        const int stored_steps = std::min(S.tr_slots, end - start - 1);
        tr_ok.assign(std::max(stored_steps, 0), 0u);
        if (stored_steps > 0) {
            std::vector<uint32_t> fl(stored_steps);
            LHIP_CHECK(hipMemcpyAsync(fl.data(), S.tr_flag.as<void>(), std::size_t(stored_steps) * 4,
                                      hipMemcpyDeviceToHost, s));
            c.sync();
            for (int q = 0; q < stored_steps; ++q) tr_ok[q] = fl[q] == 0u;
        }
        const int first = end == T ? end : end - 1;
        for (int q = first; q >= start; --q) {
            vjp_at(q - start, weight(q));
            if (q > 0) launch_axpy_f64(s, lv, lx, dt, 3 * n);
        }
    }
    LHIP_CHECK(hipEventRecord(S.ev[3], s));
    if (gx0) LHIP_CHECK(hipMemcpyAsync(gx0, lx, 3 * N * 8, hipMemcpyDeviceToHost, s));
    if (gv0) LHIP_CHECK(hipMemcpyAsync(gv0, lv, 3 * N * 8, hipMemcpyDeviceToHost, s));
    c.sync();
    float ms = 0;
    LHIP_CHECK(hipEventElapsedTime(&ms, S.ev[2], S.ev[3]));
    S.bwd_ms = ms;
}

void treesim_times(Context& c, double* f, double* b) {
    if (!c.tsim) throw Error(kLogic, "sim: no tree simulation run yet");
    if (f) *f = c.tsim->fwd_ms;
    if (b) *b = c.tsim->bwd_ms;
}

}  // namespace lhip
