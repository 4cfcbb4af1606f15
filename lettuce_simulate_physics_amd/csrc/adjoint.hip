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
// Positions x_n come back segment by segment: checkpoints (x, v) every K steps, and within a segment
// a forward recompute that stores the K + 1 positions. Cost: one force + one VJP per step.
#include "adjoint.hpp"
#include "context.hpp"

namespace lhip {

struct AdjSim {
    int n = 0;
    lhip_sim_params p{};
    int nck = 0;
    DevBuf x, y, z, vx, vy, vz, m, ax, ay, az;
    DevBuf ckpt, seg;
    DevBuf lx, ly, lz, lvx, lvy, lvz, lm, ux, uy, uz, wx, wy, wz, gmv, part;
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

    void reserve(int nn) {
        const std::size_t f = std::size_t(nn) * 8;
        for (DevBuf* b : {&x, &y, &z, &vx, &vy, &vz, &m, &ax, &ay, &az, &lx, &ly, &lz, &lvx, &lvy, &lvz, &lm, &ux, &uy,
                          &uz, &wx, &wy, &wz, &gmv})
            b->reserve(f);
        part.reserve(std::size_t(adj_split(nn)) * 4 * f);
    }

    void force(hipStream_t s, const double* px, const double* py, const double* pz) {
        launch_force64(s, px, py, pz, d(m), n, p.eps, p.G, p.precision == LHIP_PRECISION_F64, d(ax), d(ay), d(az),
                       d(part));
    }

    void step(hipStream_t s) {
        const double h = 0.5 * p.dt;
        launch_kick64(s, d(vx), d(vy), d(vz), d(ax), d(ay), d(az), h, n);
        launch_axpy3(s, d(x), d(y), d(z), d(vx), d(vy), d(vz), p.dt, n);
        force(s, d(x), d(y), d(z));
        launch_kick64(s, d(vx), d(vy), d(vz), d(ax), d(ay), d(az), h, n);
    }

    void save_state(hipStream_t s, double* dst) {
        const std::size_t f = std::size_t(n) * 8;
        DevBuf* src[6] = {&x, &y, &z, &vx, &vy, &vz};
        for (int c = 0; c < 6; ++c)
            LHIP_CHECK(hipMemcpyAsync(dst + std::size_t(c) * n, src[c]->as<void>(), f, hipMemcpyDeviceToDevice, s));
    }
    void load_state(hipStream_t s, const double* srcp) {
        const std::size_t f = std::size_t(n) * 8;
        DevBuf* dst[6] = {&x, &y, &z, &vx, &vy, &vz};
        for (int c = 0; c < 6; ++c)
            LHIP_CHECK(hipMemcpyAsync(dst[c]->as<void>(), srcp + std::size_t(c) * n, f, hipMemcpyDeviceToDevice, s));
    }
    void save_pos(hipStream_t s, double* dst) {
        const std::size_t f = std::size_t(n) * 8;
        DevBuf* src[3] = {&x, &y, &z};
        for (int c = 0; c < 3; ++c)
            LHIP_CHECK(hipMemcpyAsync(dst + std::size_t(c) * n, src[c]->as<void>(), f, hipMemcpyDeviceToDevice, s));
    }

    // One VJP at positions pos (3 SoA arrays) with cotangent weight c on lambda_v.
    void vjp_at(hipStream_t s, const double* pos, double c) {
        launch_scale3(s, d(ux), d(uy), d(uz), d(lvx), d(lvy), d(lvz), c, n);
        launch_vjp_direct(s, pos, pos + n, pos + 2 * std::size_t(n), d(m), d(ux), d(uy), d(uz), n, p.eps, p.G, d(wx),
                          d(wy), d(wz), d(gmv), d(part));
        launch_axpy3(s, d(lx), d(ly), d(lz), d(wx), d(wy), d(wz), 1.0, n);
        launch_axpy1(s, d(lm), d(gmv), 1.0, n);
    }
};

void adj_destroy(AdjSim* a) { delete a; }

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
        if (p->method != LHIP_METHOD_DIRECT)
            throw Error(kInvalid, "sim: the differentiable path supports the direct method");
        LHIP_CHECK(hipSetDevice(c.dev));
        if (!c.adj) c.adj = new AdjSim();
        AdjSim& A = *c.adj;
        A.n = int(n);
        A.p = *p;
        A.have_forward = false;
        A.reserve(A.n);
        const int T = p->steps, K = p->checkpoint_every;
        A.nck = T > 0 ? (T + K - 1) / K : 1;
        A.ckpt.reserve(std::size_t(A.nck) * 6 * n * 8);
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
        A.save_state(s, A.ckpt.as<double>());
        A.force(s, A.d(A.x), A.d(A.y), A.d(A.z));
        for (int st = 0; st < T; ++st) {
            A.step(s);
            if ((st + 1) % K == 0 && st + 1 < T) A.save_state(s, A.ckpt.as<double>() + std::size_t((st + 1) / K) * 6 * n);
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

int lhip_sim_backward(lhip_ctx* ctx, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) {
    if (!gxT || !gvT) return LHIP_INVALID;
    return lhip::guarded(ctx, [&](lhip::Context& c) {
        using namespace lhip;
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
        for (int sg = A.nck - 1; sg >= 0 && T > 0; --sg) {
            const int start = sg * K, end = std::min(T, start + K);
            // Recompute the segment's positions from its checkpoint.
            A.load_state(s, A.ckpt.as<double>() + std::size_t(sg) * 6 * n);
            A.save_pos(s, seg);
            A.force(s, A.d(A.x), A.d(A.y), A.d(A.z));
            for (int t = start; t < end; ++t) {
                A.step(s);
                A.save_pos(s, seg + std::size_t(t + 1 - start) * stride);
            }
            for (int st = end - 1; st >= start; --st) {
                if (st == T - 1) A.vjp_at(s, seg + std::size_t(end - start) * stride, 0.5 * dt);
                launch_axpy3(s, A.d(A.lvx), A.d(A.lvy), A.d(A.lvz), A.d(A.lx), A.d(A.ly), A.d(A.lz), dt, n);
                A.vjp_at(s, seg + std::size_t(st - start) * stride, st > 0 ? dt : 0.5 * dt);
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

int lhip_sim_times(lhip_ctx* ctx, double* forward_ms, double* backward_ms) {
    return lhip::guarded(ctx, [&](lhip::Context& c) {
        if (!c.adj) throw lhip::Error(LHIP_LOGIC, "sim: no simulation run yet");
        if (forward_ms) *forward_ms = c.adj->fwd_ms;
        if (backward_ms) *backward_ms = c.adj->bwd_ms;
    });
}

}  // extern "C"
