// Differentiable simulation: FP64 state, KDK leapfrog forward with step checkpoints, and the adjoint
// (reverse) sweep with recompute, through the direct-sum VJP. No reference counterpart: the
// reference has no differentiation. The forward integrates exactly the reference's KDK
// (src/simulation.cpp:406-438) in FP64 state.
//
// Reverse of one KDK step n (kick at x_n, drift, kick at x_{n+1}): a kick v' = v + c a(x) leaves
// the velocity cotangent unchanged and adds c J(x)^T lambda_v to the position cotangent (and the mass
// term); a drift x' = x + dt v adds dt lambda_x to lambda_v. The two half-kicks at an interior
// position x_n therefore see the same lambda_v and merge into one VJP of weight dt:
//   VJP(x_T, dt/2); for n = T-1 .. 0 { lambda_v += dt lambda_x; VJP(x_n, n > 0 ? dt : dt/2) }.
// Positions x_n come back segment by segment: checkpoints (x, v, a) every K steps, and within a
// segment [s, e) a forward recompute from the checkpoint that stores x_s .. x_e; it needs the forces
// at x_{s+1} .. x_{e-1} only (a(x_s) is in the checkpoint, a(x_e) is never used), so the sweep costs
// (K - 1) / K forces and one VJP per step.
#include "adjoint.hpp"
#include "context.hpp"

namespace lhip {

struct AdjSim {
    int n = 0;
    lhip_sim_params p{};
    int nck = 0;
    DevBuf x, y, z, vx, vy, vz, m, ax, ay, az, rec, vrec;
    DevBuf ckpt, seg;
    DevBuf lx, ly, lz, lvx, lvy, lvz, lm, part;
    hipEvent_t ev[4] = {};
    bool have_forward = false;
    double fwd_ms = 0, bwd_ms = 0;

    AdjSim() {
        for (auto& e : ev) LHIP_CHECK(hipEventCreate(&e));
    }
    ~AdjSim() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    double* d(DevBuf& b) { return b.as<double>(); }
    SimRec* r(DevBuf& b) { return b.as<SimRec>(); }

    void reserve(int nn) {
        const std::size_t f = std::size_t(nn) * 8;
        for (DevBuf* b : {&x, &y, &z, &vx, &vy, &vz, &m, &ax, &ay, &az, &lx, &ly, &lz, &lvx, &lvy, &lvz, &lm})
            b->reserve(f);
        rec.reserve(std::size_t(nn) * sizeof(SimRec));
        vrec.reserve(std::size_t(nn) * sizeof(SimRec));
        part.reserve(std::size_t(adj_split(nn)) * 4 * f);
    }

    bool f64() const { return p.precision == LHIP_PRECISION_F64; }

    // a = force at the positions in rec; v += h a.
    void force_kick(hipStream_t s, double h) {
        launch_force_finish(s, r(rec), d(m), n, p.eps, p.G, f64(), h, d(ax), d(ay), d(az), d(vx), d(vy), d(vz),
                            d(part));
    }

    // One KDK step from (x, v, a(x)); copy (may be null) receives the new positions. With
    // need_force false the closing kick is skipped (recompute of a segment's last position).
    void step(hipStream_t s, double* copy, bool need_force = true) {
        const double h = 0.5 * p.dt;
        launch_kick_drift(s, d(x), d(y), d(z), d(vx), d(vy), d(vz), d(ax), d(ay), d(az), d(m), h, p.dt, r(rec), copy,
                          n);
        if (need_force) force_kick(s, h);
    }

    static constexpr int kCkArrays = 9;  // x, y, z, vx, vy, vz, ax, ay, az
    void copy_state(hipStream_t s, double* ck, bool save) {
        const std::size_t f = std::size_t(n) * 8;
        DevBuf* a[kCkArrays] = {&x, &y, &z, &vx, &vy, &vz, &ax, &ay, &az};
        for (int c = 0; c < kCkArrays; ++c) {
            void* dev = a[c]->as<void>();
            void* slot = ck + std::size_t(c) * n;
            LHIP_CHECK(hipMemcpyAsync(save ? slot : dev, save ? dev : slot, f, hipMemcpyDeviceToDevice, s));
        }
    }
    double* ck_at(int k) { return ckpt.as<double>() + std::size_t(k) * kCkArrays * n; }
};

void adj_destroy(AdjSim* a) { delete a; }

void treesim_forward(Context& c, const lhip_sim_params& p, int n, const double* x0, const double* v0, const double* m,
                     double* xT, double* vT);
void treesim_backward(Context& c, const double* gxT, const double* gvT, double* gx0, double* gv0);
void treesim_times(Context& c, double* f, double* b);

namespace {

void to_soa(const double* aos, std::size_t n, std::vector<double>& soa) {
    soa.resize(3 * n);
    for (std::size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = aos[3 * i + c];
}

void to_aos(const std::vector<double>& soa, std::size_t n, double* aos) {
    for (std::size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) aos[3 * i + c] = soa[c * n + i];
}

void h2d3(hipStream_t s, AdjSim& A, DevBuf& a, DevBuf& b, DevBuf& c, const std::vector<double>& soa) {
    const std::size_t f = std::size_t(A.n) * 8;
    LHIP_CHECK(hipMemcpyAsync(a.as<void>(), soa.data(), f, hipMemcpyHostToDevice, s));
    LHIP_CHECK(hipMemcpyAsync(b.as<void>(), soa.data() + A.n, f, hipMemcpyHostToDevice, s));
    LHIP_CHECK(hipMemcpyAsync(c.as<void>(), soa.data() + 2 * std::size_t(A.n), f, hipMemcpyHostToDevice, s));
}

void d2h3(hipStream_t s, AdjSim& A, DevBuf& a, DevBuf& b, DevBuf& c, std::vector<double>& soa) {
    const std::size_t f = std::size_t(A.n) * 8;
    soa.resize(3 * std::size_t(A.n));
    LHIP_CHECK(hipMemcpyAsync(soa.data(), a.as<void>(), f, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipMemcpyAsync(soa.data() + A.n, b.as<void>(), f, hipMemcpyDeviceToHost, s));
    LHIP_CHECK(hipMemcpyAsync(soa.data() + 2 * std::size_t(A.n), c.as<void>(), f, hipMemcpyDeviceToHost, s));
}

}  // namespace
}  // namespace lhip

extern "C" {

int lhip_sim_forward(lhip_ctx* ctx, const lhip_sim_params* p, uint64_t n, const double* x0, const double* v0,
                     const double* m, double* xT, double* vT) {
    if (!p || !x0 || !v0 || !m) return LHIP_INVALID;
    return lhip::guarded(ctx, [&](lhip::Context& c) {
        using namespace lhip;
        if (n == 0 || n >= uint64_t(INT32_MAX)) throw Error(kInvalid, "sim: n must be in [1, 2^31-1)");
        if (p->steps < 0 || p->checkpoint_every < 1) throw Error(kInvalid, "sim: steps >= 0, checkpoint_every >= 1");
        LHIP_CHECK(hipSetDevice(c.dev));
        if (p->method == LHIP_METHOD_TREE) {
            // A replay evaluates the recorded topology from other initial conditions and keeps the
            // recorded forward's checkpoints: the adjoint on offer stays the recorded forward's.
            if (p->tree_mode != LHIP_TREE_REPLAY) {
                ++c.sim_gen;
                c.last_sim_tree = true;
            }
            treesim_forward(c, *p, int(n), x0, v0, m, xT, vT);
            return;
        }
        if (p->method != LHIP_METHOD_DIRECT) throw Error(kInvalid, "sim: unknown method");
        if (!c.adj) c.adj = new AdjSim();
        AdjSim& A = *c.adj;
        ++c.sim_gen;
        c.last_sim_tree = false;
        A.n = int(n);
        A.p = *p;
        A.have_forward = false;
        A.reserve(A.n);
        const int T = p->steps, K = p->checkpoint_every;
        A.nck = T > 0 ? (T + K - 1) / K : 1;
        A.ckpt.reserve(std::size_t(A.nck) * AdjSim::kCkArrays * n * 8);
        A.seg.reserve(std::size_t(K + 1) * 3 * n * 8);
        hipStream_t s = c.stream;
        std::vector<double> soa;
        to_soa(x0, n, soa);
        h2d3(s, A, A.x, A.y, A.z, soa);
        c.sync();  // the host staging vector is reused below
        to_soa(v0, n, soa);
        h2d3(s, A, A.vx, A.vy, A.vz, soa);
        LHIP_CHECK(hipMemcpyAsync(A.m.as<void>(), m, n * 8, hipMemcpyHostToDevice, s));
        c.sync();
        LHIP_CHECK(hipEventRecord(A.ev[0], s));
        launch_pack(s, A.d(A.x), A.d(A.y), A.d(A.z), A.d(A.m), A.r(A.rec), A.n);
        A.force_kick(s, 0.0);
        A.copy_state(s, A.ck_at(0), true);
        for (int st = 0; st < T; ++st) {
            A.step(s, nullptr);
            if ((st + 1) % K == 0 && st + 1 < T) A.copy_state(s, A.ck_at((st + 1) / K), true);
        }
        LHIP_CHECK(hipEventRecord(A.ev[1], s));
        if (xT) {
            d2h3(s, A, A.x, A.y, A.z, soa);
            c.sync();
            to_aos(soa, n, xT);
        }
        if (vT) {
            d2h3(s, A, A.vx, A.vy, A.vz, soa);
            c.sync();
            to_aos(soa, n, vT);
        }
        c.sync();
        float ms = 0;
        LHIP_CHECK(hipEventElapsedTime(&ms, A.ev[0], A.ev[1]));
        A.fwd_ms = ms;
        A.have_forward = true;
    });
}

int lhip_sim_generation(const lhip_ctx* ctx, uint64_t* generation) {
    if (!ctx || !ctx->c || !generation) return LHIP_INVALID;
    *generation = ctx->c->sim_gen;
    return LHIP_OK;
}

int lhip_sim_backward_gen(lhip_ctx* ctx, uint64_t generation, const double* gxT, const double* gvT, double* gx0,
                          double* gv0, double* gm) {
    if (!gxT || !gvT) return LHIP_INVALID;
    return lhip::guarded(ctx, [&](lhip::Context& c) {
        using namespace lhip;
        if (generation != 0 && generation != c.sim_gen)
            throw Error(kLogic, "sim backward: another forward has run on this context since (its checkpoints are "
                                "gone); run the forward again");
        if (c.last_sim_tree) {
            if (gm) throw Error(kInvalid, "sim backward: the tree adjoint gives no mass gradient (pass NULL)");
            LHIP_CHECK(hipSetDevice(c.dev));
            treesim_backward(c, gxT, gvT, gx0, gv0);
            return;
        }
        if (!c.adj || !c.adj->have_forward) throw Error(kLogic, "sim backward: call lhip_sim_forward first");
        LHIP_CHECK(hipSetDevice(c.dev));
        AdjSim& A = *c.adj;
        hipStream_t s = c.stream;
        const int n = A.n, T = A.p.steps, K = A.p.checkpoint_every;
        const double dt = A.p.dt;
        std::vector<double> soa;
        to_soa(gxT, n, soa);
        h2d3(s, A, A.lx, A.ly, A.lz, soa);
        c.sync();
        to_soa(gvT, n, soa);
        h2d3(s, A, A.lvx, A.lvy, A.lvz, soa);
        c.sync();
        launch_fill64(s, A.d(A.lm), 0.0, n);
        LHIP_CHECK(hipEventRecord(A.ev[2], s));
        double* seg = A.seg.as<double>();
        const std::size_t stride = 3 * std::size_t(n);
        auto weight = [&](int q) { return (q == T || q == 0) ? 0.5 * dt : dt; };
        for (int sg = A.nck - 1; sg >= 0 && T > 0; --sg) {
            const int start = sg * K, end = std::min(T, start + K);
            // Recompute x_start .. x_end into seg from the checkpoint (x, v, a at x_start).
            A.copy_state(s, A.ck_at(sg), false);
            for (int cc = 0; cc < 3; ++cc)
                LHIP_CHECK(hipMemcpyAsync(seg + std::size_t(cc) * n, A.ck_at(sg) + std::size_t(cc) * n,
                                          std::size_t(n) * 8, hipMemcpyDeviceToDevice, s));
            for (int t = start; t < end; ++t) A.step(s, seg + std::size_t(t + 1 - start) * stride, t + 1 < end);
            // VJPs at x_q for q = first .. start (x_end belongs to this segment only when end == T).
            const int first = end == T ? end : end - 1;
            launch_pack_vjp(s, seg + std::size_t(first - start) * stride, A.d(A.m), A.d(A.lvx), A.d(A.lvy),
                            A.d(A.lvz), weight(first), A.r(A.vrec), n);
            for (int q = first; q >= start; --q) {
                const double* npos = q > start ? seg + std::size_t(q - 1 - start) * stride : nullptr;
                launch_vjp_update(s, A.r(A.vrec), n, A.p.eps, A.p.G, A.d(A.part), A.d(A.lx), A.d(A.ly), A.d(A.lz),
                                  A.d(A.lm), A.d(A.lvx), A.d(A.lvy), A.d(A.lvz), q > 0 ? dt : 0.0, npos, A.d(A.m),
                                  q > 0 ? weight(q - 1) : 0.0, A.r(A.vrec));
            }
        }
        LHIP_CHECK(hipEventRecord(A.ev[3], s));
        if (gx0) {
            d2h3(s, A, A.lx, A.ly, A.lz, soa);
            c.sync();
            to_aos(soa, n, gx0);
        }
        if (gv0) {
            d2h3(s, A, A.lvx, A.lvy, A.lvz, soa);
            c.sync();
            to_aos(soa, n, gv0);
        }
        if (gm) LHIP_CHECK(hipMemcpyAsync(gm, A.lm.as<void>(), std::size_t(n) * 8, hipMemcpyDeviceToHost, s));
        c.sync();
        float ms = 0;
        LHIP_CHECK(hipEventElapsedTime(&ms, A.ev[2], A.ev[3]));
        A.bwd_ms = ms;
    });
}

int lhip_sim_backward(lhip_ctx* ctx, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) {
    return lhip_sim_backward_gen(ctx, 0, gxT, gvT, gx0, gv0, gm);
}

int lhip_sim_times(lhip_ctx* ctx, double* forward_ms, double* backward_ms) {
    return lhip::guarded(ctx, [&](lhip::Context& c) {
        if (c.last_sim_tree) {
            lhip::treesim_times(c, forward_ms, backward_ms);
            return;
        }
        if (!c.adj) throw lhip::Error(LHIP_LOGIC, "sim: no simulation run yet");
        if (forward_ms) *forward_ms = c.adj->fwd_ms;
        if (backward_ms) *backward_ms = c.adj->bwd_ms;
    });
}

}  // extern "C"
