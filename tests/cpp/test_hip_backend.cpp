// C++ drop-in test: the reference CUDA backend's suite (tests/test_cuda_backend.cpp) run through
// the Backend interface against make_hip_backend. Accelerations are read as the reference does:
// upload with v = 0, compute_forces, kick(1), download (v == a), indexed by id. The checker is the
// oracle library (FP64 brute force, Plummer generator, CPU tree restatement).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "backend.hpp"
#include "lettuce_hip.h"

extern "C" {
void orc_plummer(uint64_t n, uint64_t seed, double G, double M, double a, double cut, float* x, float* y, float* z,
                 float* vx, float* vy, float* vz, float* m, uint32_t* id);
void orc_direct_f64(uint64_t n, const double* x, const double* y, const double* z, const double* m, double G,
                    double eps, uint64_t nt, const int64_t* idx, double* ax, double* ay, double* az, double* phi);
}

namespace {
int g_fail = 0, g_pass = 0, g_skip = 0;
struct Skip { std::string why; };
#define CHECK(c) do { if (!(c)) { std::fprintf(stderr, "  %s:%d: CHECK(%s) failed\n", __FILE__, __LINE__, #c); ++fails; } } while (0)
#define CHECK_NEAR(a, b, t) do { double A = (a), B = (b); if (!(std::fabs(A - B) <= (t))) { \
    std::fprintf(stderr, "  %s:%d: CHECK_NEAR(%s, %s) failed: %.9g vs %.9g\n", __FILE__, __LINE__, #a, #b, A, B); ++fails; } } while (0)

std::vector<std::pair<const char*, std::function<void(int&)>>>& reg() {
    static std::vector<std::pair<const char*, std::function<void(int&)>>> r;
    return r;
}
struct Reg { Reg(const char* n, std::function<void(int&)> f) { reg().push_back({n, std::move(f)}); } };
#define T_CAT2(a, b) a##b
#define T_CAT(a, b) T_CAT2(a, b)
#define TEST(name) static void T_CAT(tf_, __LINE__)(int& fails); static Reg T_CAT(tr_, __LINE__)(name, T_CAT(tf_, __LINE__)); \
    static void T_CAT(tf_, __LINE__)(int& fails)

Bodies plummer(std::size_t n, uint64_t seed, double a = 3 * M_PI / 16, double cut = 20) {
    Bodies b;
    b.resize(n);
    orc_plummer(n, seed, 1.0, 1.0, a, cut, b.x.data(), b.y.data(), b.z.data(), b.vx.data(), b.vy.data(), b.vz.data(),
                b.m.data(), b.id.data());
    return b;
}

Bodies cube(std::size_t n, uint64_t seed) {
    std::srand(unsigned(seed));
    auto u = [] { return 2.0 * (std::rand() / double(RAND_MAX)) - 1.0; };
    Bodies b;
    for (std::size_t i = 0; i < n; ++i) b.push(u(), u(), u(), u(), u(), u(), (1.0 + 0.5 * u()) / double(n));
    return b;
}

struct Ref { std::vector<double> a; double pe; };
Ref brute(const Bodies& b, double eps) {
    const std::size_t n = b.size();
    std::vector<double> x(b.x.begin(), b.x.end()), y(b.y.begin(), b.y.end()), z(b.z.begin(), b.z.end()),
        m(b.m.begin(), b.m.end()), ax(n), ay(n), az(n), phi(n);
    std::vector<int64_t> idx(n);
    for (std::size_t i = 0; i < n; ++i) idx[i] = int64_t(i);
    orc_direct_f64(n, x.data(), y.data(), z.data(), m.data(), 1.0, eps, n, idx.data(), ax.data(), ay.data(), az.data(),
                   phi.data());
    Ref r;
    r.a.assign(3 * n, 0);
    r.pe = 0;
    for (std::size_t i = 0; i < n; ++i) {
        r.a[3 * b.id[i]] = ax[i]; r.a[3 * b.id[i] + 1] = ay[i]; r.a[3 * b.id[i] + 2] = az[i];
        r.pe += 0.5 * m[i] * phi[i];
    }
    return r;
}

struct Got { std::vector<double> a; Diag d; PhaseTimes t; };
Got accel(Backend& be, Bodies b, const ForceParams& p, bool pot) {
    std::fill(b.vx.begin(), b.vx.end(), 0.f);
    std::fill(b.vy.begin(), b.vy.end(), 0.f);
    std::fill(b.vz.begin(), b.vz.end(), 0.f);
    be.upload(b);
    be.compute_forces(p, pot);
    Got g;
    g.t = be.times();
    g.d = be.diagnostics();
    be.kick(1.0);
    Bodies o;
    be.download(o);
    g.a.assign(3 * o.size(), 0);
    for (std::size_t i = 0; i < o.size(); ++i) {
        g.a[3 * o.id[i]] = o.vx[i]; g.a[3 * o.id[i] + 1] = o.vy[i]; g.a[3 * o.id[i] + 2] = o.vz[i];
    }
    return g;
}

double rel_rms(const std::vector<double>& a, const std::vector<double>& r) {
    double num = 0, den = 0;
    for (std::size_t i = 0; i < a.size(); ++i) { num += (a[i] - r[i]) * (a[i] - r[i]); den += r[i] * r[i]; }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

ForceParams params(ForceMethod m, double eps, double theta = 0.6) {
    ForceParams p;
    p.eps = eps; p.theta = theta; p.method = m;
    return p;
}

std::unique_ptr<Backend> backend() {
    std::string why;
    auto be = make_hip_backend(0, why);
    if (!be) throw Skip{why};
    return be;
}
}  // namespace

TEST("hip: name, compiled flag, bad device") {
    auto be = backend();
    CHECK(hip_compiled());
    CHECK(be->name().rfind("hip: ", 0) == 0);
    std::string why;
    CHECK(make_hip_backend(hip_device_count() + 3, why) == nullptr && !why.empty());
}

TEST("hip: direct vs double brute force, 4096 Plummer") {
    auto be = backend();
    Bodies b = plummer(4096, 1);
    Ref ref = brute(b, 0.01);
    Got g = accel(*be, b, params(ForceMethod::direct, 0.01), true);
    CHECK(rel_rms(g.a, ref.a) <= 1e-5);
    CHECK(g.d.has_pe);
    CHECK_NEAR(g.d.pe / ref.pe, 1.0, 1e-5);
    CHECK_NEAR(be->exact_potential_energy(1, 0.01) / ref.pe, 1.0, 1e-9);
    CHECK(g.t.interactions == 4096ull * 4095ull);
}

TEST("hip: small and odd N, direct and tree(theta=0) vs brute force") {
    auto be = backend();
    for (std::size_t n : {1, 2, 3, 31, 32, 33, 64, 255, 257, 1000, 1023}) {
        Bodies b = cube(n, 100 + n);
        Ref ref = brute(b, 0.05);
        for (ForceMethod m : {ForceMethod::direct, ForceMethod::tree}) {
            Got g = accel(*be, b, params(m, 0.05, 0.0), true);
            if (n == 1) { CHECK(g.a[0] == 0 && g.d.pe == 0); continue; }
            CHECK(rel_rms(g.a, ref.a) <= 1e-5);
            CHECK_NEAR(g.d.pe / ref.pe, 1.0, 1e-5);
            CHECK(g.t.interactions == uint64_t(n) * (n - 1));
        }
    }
}

TEST("hip: tree vs direct, 20k Plummer, theta 0.6") {
    auto be = backend();
    Bodies b = plummer(20000, 2);
    Got d = accel(*be, b, params(ForceMethod::direct, 0.01), true);
    Got t = accel(*be, b, params(ForceMethod::tree, 0.01, 0.6), true);
    CHECK(rel_rms(t.a, d.a) <= 1e-3);
    CHECK_NEAR(t.d.pe / d.d.pe, 1.0, 1e-3);
    CHECK(t.t.interactions > 0 && t.t.interactions < 20000ull * 19999ull);
}

TEST("hip: two-body circular orbit returns after one period") {
    auto be = backend();
    for (ForceMethod m : {ForceMethod::direct, ForceMethod::tree}) {
        Bodies b;
        b.push(0.5, 0, 0, 0, 0.5, 0, 0.5);
        b.push(-0.5, 0, 0, 0, -0.5, 0, 0.5);
        const ForceParams p = params(m, 0.0);
        const double dt = 2 * M_PI / 1000;
        be->upload(b);
        be->compute_forces(p, true);
        const Diag d0 = be->diagnostics();
        for (int s = 0; s < 1000; ++s) { be->kick(dt / 2); be->drift(dt); be->compute_forces(p, true); be->kick(dt / 2); }
        const Diag d1 = be->diagnostics();
        Bodies o;
        be->download(o);
        for (std::size_t i = 0; i < 2; ++i)
            CHECK(std::hypot(o.x[i] - b.x[o.id[i]], o.y[i] - b.y[o.id[i]], o.z[i] - b.z[o.id[i]]) <= 1e-3);
        CHECK_NEAR((d1.ke + d1.pe) / (d0.ke + d0.pe), 1.0, 1e-4);
    }
}

TEST("hip: build_tree invalidates accelerations and potential") {
    auto be = backend();
    Bodies b = plummer(5000, 12);
    const ForceParams p = params(ForceMethod::tree, 0.01, 0.6);
    be->upload(b);
    be->compute_forces(p, true);
    CHECK(be->diagnostics().has_pe);
    be->build_tree(p);
    CHECK(!be->diagnostics().has_pe);
    bool threw = false;
    try { be->kick(0.01); } catch (const std::logic_error&) { threw = true; }
    CHECK(threw);
    be->compute_forces_with(p, true, {});
    CHECK(be->diagnostics().has_pe);
    be->kick(0.01);
    std::vector<float> cost;
    CHECK(be->body_costs(cost) && cost.size() == b.size());
    uint64_t fr = 0, tot = 0;
    CHECK(be->device_memory(fr, tot) && tot > 0 && fr <= tot + (uint64_t(1) << 40));
}

// ---- distributed hooks (the reference's tests/test_cuda_let.cpp): two contexts stand in for two ranks.

TEST("hip let: two halves with LETs vs one direct sum, 20k Plummer") {
    auto ref = backend();
    auto r0 = backend();
    auto r1 = backend();
    Bodies all = plummer(20000, 13, 1.0, 10.0);
    std::fill(all.vx.begin(), all.vx.end(), 0.f);
    std::fill(all.vy.begin(), all.vy.end(), 0.f);
    std::fill(all.vz.begin(), all.vz.end(), 0.f);
    const Got d = accel(*ref, all, params(ForceMethod::direct, 0.01), true);
    Bodies h[2];
    for (std::size_t i = 0; i < all.size(); ++i) {
        Bodies& t = h[all.x[i] < 0.f ? 0 : 1];
        t.push(all.x[i], all.y[i], all.z[i], 0, 0, 0, all.m[i]);
        t.id.back() = all.id[i];
    }
    for (ForceMethod m : {ForceMethod::tree, ForceMethod::direct}) {
        const bool tree = m == ForceMethod::tree;
        const ForceParams p = params(m, 0.01);
        Backend* be[2] = {r0.get(), r1.get()};
        for (int k = 0; k < 2; ++k) { be[k]->upload(h[k]); be[k]->build_tree(p); }
        const LetTree from1 = be[1]->export_let(be[0]->local_bounds(), p);
        const LetTree from0 = be[0]->export_let(be[1]->local_bounds(), p);
        be[0]->compute_forces_with(p, true, {&from1});
        be[1]->compute_forces_with(p, true, {&from0});
        std::vector<double> a(3 * all.size(), 0.0);
        double pe = 0;
        uint64_t inter = 0;
        for (int k = 0; k < 2; ++k) {
            inter += be[k]->times().interactions;
            pe += be[k]->diagnostics().pe;
            std::vector<float> c;
            CHECK(be[k]->body_costs(c) && c.size() == h[k].size());
            double cs = 0;
            for (float v : c) cs += v;
            CHECK(cs == double(be[k]->times().interactions));
            be[k]->kick(1.0);
            Bodies o;
            be[k]->download(o);
            for (std::size_t i = 0; i < o.size(); ++i) {
                a[3 * o.id[i]] = o.vx[i]; a[3 * o.id[i] + 1] = o.vy[i]; a[3 * o.id[i] + 2] = o.vz[i];
            }
        }
        const double e = rel_rms(a, d.a);
        std::fprintf(stderr, "  %s: LETs %zu/%zu and %zu/%zu, rel rms %.3g\n", tree ? "tree" : "direct",
                     from1.nodes.size(), from1.bx.size(), from0.nodes.size(), from0.bx.size(), e);
        CHECK(e <= (tree ? 1.5e-3 : 1e-5));
        CHECK_NEAR(pe / d.d.pe, 1.0, tree ? 1.5e-3 : 1e-5);
        if (!tree) CHECK(inter == uint64_t(all.size()) * (all.size() - 1));
    }
}

TEST("hip let: empty LETs change nothing; a rank without bodies exports none") {
    auto be = backend();
    const ForceParams p = params(ForceMethod::tree, 0.01);
    const Bodies b = plummer(5000, 14, 1.0, 10.0);
    const Got g0 = accel(*be, b, p, true);
    Bodies z = b;
    std::fill(z.vx.begin(), z.vx.end(), 0.f);
    std::fill(z.vy.begin(), z.vy.end(), 0.f);
    std::fill(z.vz.begin(), z.vz.end(), 0.f);
    be->upload(z);
    be->build_tree(p);
    const LetTree empty;
    be->compute_forces_with(p, true, {&empty, &empty});
    CHECK(be->diagnostics().pe == g0.d.pe);
    be->kick(1.0);
    Bodies o;
    be->download(o);
    bool same = true;
    for (std::size_t i = 0; i < o.size(); ++i)
        same = same && o.vx[i] == g0.a[3 * o.id[i]] && o.vy[i] == g0.a[3 * o.id[i] + 1] &&
               o.vz[i] == g0.a[3 * o.id[i] + 2];
    CHECK(same);
    Bodies none;
    be->upload(none);
    be->build_tree(p);
    CHECK(be->export_let(Box3{}, p).nodes.empty());
    bool threw = false;
    try {
        be->upload(b);
        (void)be->export_let(Box3{}, p);  // no tree since the upload
    } catch (const std::logic_error&) {
        threw = true;
    }
    CHECK(threw);
}

TEST("hip snapshot: a mixed-precision run restarts bit for bit from its FP64 snapshot") {
    std::string why;
    HipOptions opt;
    opt.mixed = true;
    auto a = make_hip_backend(0, why, opt);
    auto b = make_hip_backend(0, why, opt);
    if (!a || !b) throw Skip{why};
    CHECK(hip_context(*a) != nullptr);
    const ForceParams p = params(ForceMethod::tree, 0.01);
    a->upload(plummer(8192, 21, 1.0, 10.0));
    a->compute_forces(p, false);
    auto steps = [&](Backend& be, int k) {
        for (int s = 0; s < k; ++s) {
            be.kick(5e-4); be.drift(1e-3); be.compute_forces(p, false); be.kick(5e-4);
        }
    };
    steps(*a, 4);
    lhip_snapshot_header h{};
    h.step = 4;
    const char* tmp = std::getenv("TMPDIR");
    const std::string path = std::string(tmp && *tmp ? tmp : "/tmp") + "/test_hip_snapshot_fp64.snap";
    CHECK(lhip_snapshot_write(hip_context(*a), path.c_str(), &h) == LHIP_OK);
    steps(*a, 4);
    lhip_snapshot_header g{};
    CHECK(lhip_snapshot_restore(hip_context(*b), path.c_str(), &g) == LHIP_OK && g.fp64 == 1 && g.step == 4);
    b->compute_forces(p, false);
    steps(*b, 4);
    const uint64_t n = lhip_size(hip_context(*a));
    std::vector<double> xa(n), xb(n), va(n), vb(n), d(n);
    std::vector<uint32_t> ia(n), ib(n);
    lhip_download_f64(hip_context(*a), xa.data(), d.data(), d.data(), va.data(), d.data(), d.data(), nullptr, ia.data());
    lhip_download_f64(hip_context(*b), xb.data(), d.data(), d.data(), vb.data(), d.data(), d.data(), nullptr, ib.data());
    bool same = ia == ib;
    for (uint64_t i = 0; same && i < n; ++i) same = xa[i] == xb[i] && va[i] == vb[i];
    CHECK(same);
    std::remove(path.c_str());
}

int main(int argc, char** argv) {
    const std::string filter = argc > 1 ? argv[1] : "";
    for (auto& [name, fn] : reg()) {
        if (!filter.empty() && std::string(name).find(filter) == std::string::npos) continue;
        std::fprintf(stderr, "[ RUN  ] %s\n", name);
        int fails = 0;
        try {
            fn(fails);
        } catch (const Skip& s) {
            std::fprintf(stderr, "[ SKIP ] %s: %s\n", name, s.why.c_str());
            ++g_skip;
            continue;
        } catch (const std::exception& e) {
            std::fprintf(stderr, "  exception: %s\n", e.what());
            ++fails;
        }
        std::fprintf(stderr, "[ %s ] %s\n", fails ? "FAIL" : " OK ", name);
        (fails ? g_fail : g_pass) += 1;
    }
    std::fprintf(stderr, "%d passed, %d failed, %d skipped\n", g_pass, g_fail, g_skip);
    return g_fail ? 1 : (g_pass == 0 ? 77 : 0);
}
