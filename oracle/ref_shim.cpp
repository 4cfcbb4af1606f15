// C entry points over the reference lettuce CPU backend and scenario generator, compiled from the
// reference sources where they lie (see oracle/Makefile). TEST INFRASTRUCTURE ONLY: used to pin
// the oracle restatement (oracle/lettuce_oracle.cpp), to write golden fixtures, and as the
// "reference" CPU baseline in bench.py. Never linked into the HIP product library.
//
// Interfaces used (reference files): Backend, ForceParams, Diag, PhaseTimes, LetTree
// (src/backend.hpp:12-148), make_cpu_backend (src/cpu_backend.cpp:911), Config and
// build_scenario (src/config.hpp:14-70, src/scenarios.cpp:411-420), write_snapshot and
// read_snapshot (src/snapshot.hpp:22-27).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "backend.hpp"
#include "config.hpp"
#include "scenarios.hpp"
#include "snapshot.hpp"

namespace {
thread_local std::string g_err;

struct RefCtx {
    std::unique_ptr<Backend> be;
};

ForceParams make_params(double G, double eps, double theta, int method) {
    ForceParams p;
    p.G = G;
    p.eps = eps;
    p.theta = theta;
    p.method = method == 1 ? ForceMethod::tree : ForceMethod::direct;
    return p;
}

void copy_out(const Bodies& b, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* m,
              uint32_t* id) {
    const std::size_t n = b.size();
    std::memcpy(x, b.x.data(), n * 4); std::memcpy(y, b.y.data(), n * 4); std::memcpy(z, b.z.data(), n * 4);
    std::memcpy(vx, b.vx.data(), n * 4); std::memcpy(vy, b.vy.data(), n * 4); std::memcpy(vz, b.vz.data(), n * 4);
    std::memcpy(m, b.m.data(), n * 4); std::memcpy(id, b.id.data(), n * 4);
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Scenario `name` ("plummer", "disk", ...) with n bodies, the seed, G and "key=value" scenario
// parameters separated by ';'. Returns the body count (build_scenario may change n), or -1.
// Call with x == nullptr to get the count only.
int64_t ref_scenario(const char* name, uint64_t n, uint64_t seed, double G, const char* params, float* x, float* y,
                     float* z, float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    try {
        Config c;
        c.scenario = name;
        c.n = n;
        c.seed = seed;
        c.G = G;
        std::string ps = params ? params : "";
        std::size_t pos = 0;
        while (pos < ps.size()) {
            std::size_t end = ps.find(';', pos);
            if (end == std::string::npos) end = ps.size();
            std::string kv = ps.substr(pos, end - pos);
            std::size_t eq = kv.find('=');
            if (eq != std::string::npos) c.params[kv.substr(0, eq)] = kv.substr(eq + 1);
            pos = end + 1;
        }
        if (!x) return int64_t(scenario_body_count(c));
        Bodies b = build_scenario(c);
        copy_out(b, x, y, z, vx, vy, vz, m, id);
        return int64_t(b.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void* ref_cpu_create(int threads) { return new RefCtx{make_cpu_backend(threads)}; }
void ref_cpu_destroy(void* h) { delete static_cast<RefCtx*>(h); }

void ref_upload(void* h, uint64_t n, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                const float* vz, const float* m, const uint32_t* id) {
    Bodies b;
    b.x.assign(x, x + n); b.y.assign(y, y + n); b.z.assign(z, z + n);
    b.vx.assign(vx, vx + n); b.vy.assign(vy, vy + n); b.vz.assign(vz, vz + n);
    b.m.assign(m, m + n); b.id.assign(id, id + n);
    static_cast<RefCtx*>(h)->be->upload(b);
}

void ref_download(void* h, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    Bodies b;
    static_cast<RefCtx*>(h)->be->download(b);
    copy_out(b, x, y, z, vx, vy, vz, m, id);
}

int ref_compute_forces(void* h, double G, double eps, double theta, int method, int potential) {
    try {
        static_cast<RefCtx*>(h)->be->compute_forces(make_params(G, eps, theta, method), potential != 0);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

void ref_kick(void* h, double dt) { static_cast<RefCtx*>(h)->be->kick(dt); }
void ref_drift(void* h, double dt) { static_cast<RefCtx*>(h)->be->drift(dt); }

// out: mass, ke, pe, has_pe, px, py, pz, lx, ly, lz, cx, cy, cz
void ref_diagnostics(void* h, double* out) {
    Diag d = static_cast<RefCtx*>(h)->be->diagnostics();
    double v[13] = {d.mass, d.ke, d.pe, d.has_pe ? 1.0 : 0.0, d.px, d.py, d.pz, d.lx, d.ly, d.lz, d.cx, d.cy, d.cz};
    std::memcpy(out, v, sizeof v);
}

double ref_exact_pe(void* h, double G, double eps) { return static_cast<RefCtx*>(h)->be->exact_potential_energy(G, eps); }

// out: force_ms, sort_ms, build_ms, walk_ms, step_ms; *inter = interactions
void ref_times(void* h, double* out, uint64_t* inter) {
    PhaseTimes t = static_cast<RefCtx*>(h)->be->times();
    out[0] = t.force_ms; out[1] = t.sort_ms; out[2] = t.build_ms; out[3] = t.walk_ms; out[4] = t.step_ms;
    *inter = t.interactions;
}

int ref_body_costs(void* h, float* cost) {
    std::vector<float> c;
    if (!static_cast<RefCtx*>(h)->be->body_costs(c)) return 0;
    std::memcpy(cost, c.data(), c.size() * 4);
    return 1;
}

// The whole local tree, breadth first: build_tree, then export_let for a target box that holds every
// body (every node with mass is then opened, src/cpu_backend.cpp:222-271). Two-phase: call with
// nodes == nullptr to get the node and body counts, then again with buffers (60-byte LetNode records
// as 15 floats/ints each, and nb leaf-body floats per array). Keeps the LET between the calls.
struct RefTree { LetTree t; };
static thread_local RefTree g_tree;

int ref_export_tree(void* h, double G, double eps, double theta, uint64_t* nnodes, uint64_t* nbodies, void* nodes,
                    float* bx, float* by, float* bz, float* bm) {
    try {
        auto& be = *static_cast<RefCtx*>(h)->be;
        if (!nodes) {
            ForceParams p = make_params(G, eps, theta, 1);
            be.build_tree(p);
            Box3 all = be.local_bounds();
            for (int c = 0; c < 3; ++c) { all.lo[c] -= 1.0f; all.hi[c] += 1.0f; }
            g_tree.t = be.export_let(all, p);
            *nnodes = g_tree.t.nodes.size();
            *nbodies = g_tree.t.bx.size();
            return 0;
        }
        std::memcpy(nodes, g_tree.t.nodes.data(), g_tree.t.nodes.size() * sizeof(LetNode));
        const std::size_t nb = g_tree.t.bx.size();
        std::memcpy(bx, g_tree.t.bx.data(), nb * 4); std::memcpy(by, g_tree.t.by.data(), nb * 4);
        std::memcpy(bz, g_tree.t.bz.data(), nb * 4); std::memcpy(bm, g_tree.t.bm.data(), nb * 4);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// ---- distributed support: local_bounds, build_tree, export_let, compute_forces_with
// (src/backend.hpp:114-136). LETs cross as flat arrays (the layout of lhip_let in
// include/lettuce_hip.h).
struct FlatLet {
    const void* nodes;
    uint64_t n_nodes;
    const float *bx, *by, *bz, *bm;
    uint64_t n_bodies;
};
static thread_local LetTree g_let;

void ref_local_bounds(void* h, float* lo, float* hi) {
    const Box3 b = static_cast<RefCtx*>(h)->be->local_bounds();
    for (int c = 0; c < 3; ++c) { lo[c] = b.lo[c]; hi[c] = b.hi[c]; }
}

int ref_build_tree(void* h, double G, double eps, double theta, int method) {
    try {
        static_cast<RefCtx*>(h)->be->build_tree(make_params(G, eps, theta, method));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// export_let for the box [lo, hi]: returns the sizes; ref_export_let_copy copies the LET out.
int ref_export_let(void* h, const float* lo, const float* hi, double G, double eps, double theta, int method,
                   uint64_t* nn, uint64_t* nb) {
    try {
        Box3 b;
        for (int c = 0; c < 3; ++c) { b.lo[c] = lo[c]; b.hi[c] = hi[c]; }
        g_let = static_cast<RefCtx*>(h)->be->export_let(b, make_params(G, eps, theta, method));
        *nn = g_let.nodes.size();
        *nb = g_let.bx.size();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

void ref_export_let_copy(void* nodes, float* bx, float* by, float* bz, float* bm) {
    std::memcpy(nodes, g_let.nodes.data(), g_let.nodes.size() * sizeof(LetNode));
    const std::size_t b = g_let.bx.size() * 4;
    std::memcpy(bx, g_let.bx.data(), b); std::memcpy(by, g_let.by.data(), b);
    std::memcpy(bz, g_let.bz.data(), b); std::memcpy(bm, g_let.bm.data(), b);
}

int ref_forces_with(void* h, double G, double eps, double theta, int method, int potential, const FlatLet* lets,
                    int nlets) {
    try {
        std::vector<LetTree> v(std::size_t(nlets > 0 ? nlets : 0));
        std::vector<const LetTree*> r;
        for (int k = 0; k < nlets; ++k) {
            const FlatLet& f = lets[k];
            const auto* nd = static_cast<const LetNode*>(f.nodes);
            v[k].nodes.assign(nd, nd + f.n_nodes);
            v[k].bx.assign(f.bx, f.bx + f.n_bodies); v[k].by.assign(f.by, f.by + f.n_bodies);
            v[k].bz.assign(f.bz, f.bz + f.n_bodies); v[k].bm.assign(f.bm, f.bm + f.n_bodies);
            r.push_back(&v[k]);
        }
        static_cast<RefCtx*>(h)->be->compute_forces_with(make_params(G, eps, theta, method), potential != 0, r);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// ---- snapshots (src/snapshot.hpp:22-27). hdr: n, step, seed, pieces as u64 and time, G, eps, dt as
// f64 (8 slots of 8 bytes: n, step, time, G, eps, dt, seed, pieces); scenario: 32 bytes.
static thread_local Bodies g_snap;

int ref_write_snapshot(const char* path, const double* hd, const char* scenario, uint64_t n, const float* x,
                       const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                       const float* m, const uint32_t* id) {
    try {
        SnapshotHeader h;
        std::memcpy(&h.step, hd + 1, 8);
        h.time = hd[2]; h.G = hd[3]; h.eps = hd[4]; h.dt = hd[5];
        std::memcpy(&h.seed, hd + 6, 8);
        uint64_t pieces;
        std::memcpy(&pieces, hd + 7, 8);
        h.pieces = uint32_t(pieces);
        h.scenario = scenario;
        Bodies b;
        b.x.assign(x, x + n); b.y.assign(y, y + n); b.z.assign(z, z + n);
        b.vx.assign(vx, vx + n); b.vy.assign(vy, vy + n); b.vz.assign(vz, vz + n);
        b.m.assign(m, m + n); b.id.assign(id, id + n);
        write_snapshot(path, h, b);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Reads the file (keeping the bodies for ref_snapshot_copy); fills hd and scenario as above.
int ref_read_snapshot(const char* path, double* hd, char* scenario) {
    try {
        SnapshotHeader h = read_snapshot(path, g_snap);
        std::memcpy(hd, &h.n, 8);
        std::memcpy(hd + 1, &h.step, 8);
        hd[2] = h.time; hd[3] = h.G; hd[4] = h.eps; hd[5] = h.dt;
        std::memcpy(hd + 6, &h.seed, 8);
        const uint64_t pieces = h.pieces;
        std::memcpy(hd + 7, &pieces, 8);
        std::memset(scenario, 0, 32);
        std::memcpy(scenario, h.scenario.data(), std::min<std::size_t>(h.scenario.size(), 31));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

void ref_snapshot_copy(float* x, float* y, float* z, float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    copy_out(g_snap, x, y, z, vx, vy, vz, m, id);
}

}  // extern "C"
