// ORACLE — TEST INFRASTRUCTURE ONLY. A CPU restatement of the reference lettuce force path, used
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. The product
// (the HIP library) never links, loads or calls anything in this file.
//
// It restates, in plain C++ with OpenMP, the reference's CPU backend (src/cpu_backend.cpp) and the
// scenario generators the configs use (src/scenarios.cpp); every function cites what it follows.
// What pins each part (tests/test_oracle.py):
//   scenario_plummer     reference generator output, bit for bit (oracle/_ref and golden fixture)
//   morton keys / sort   bit-exact vs the reference (keys recomputed from downloaded positions);
//                        spread21 known answers; std::stable_sort on (key, index)
//   octree topology      the reference's whole-tree export (node count, child counts, leaf bodies),
//                        plus invariants (leaf = first cell with <= 16 bodies, ranges nest)
//   moments / rcrit      the reference export (float-exact in practice), plus closed forms
//                        (two-body com, traceless Q, parallel-axis vs direct sum)
//   walk forces          reference accelerations (rel rms), interaction counts and per-body costs
//                        exactly; theta = 0 equals direct; FP64 brute force bounds
//   direct forces        FP64 brute force (rel rms <= 1e-6), two-body closed form
//   export_let / remote  the reference's own export_let and compute_forces_with(remote) (oracle/_ref):
//     LETs               LETs field for field, forces, costs and counts on two-rank splits; two ranks
//                        against one direct sum; an empty LET changes nothing
//   kick/drift/diag      closed-form sums in numpy
//   scenario_disk_sech2  NOT a reference function (the reference disk is Gaussian in z): pinned by
//                        closed-form moments of the exponential / sech^2 profiles only
//   direct_f64           textbook O(N^2) sum; two-body closed form a = m/(d^2+eps^2)^1.5
//   sim_forward/backward FP64 KDK and its adjoint (no reference counterpart): pinned by central finite
//                        differences of sim_forward on small systems (tests/test_oracle.py)
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int kLeafMax = 16;   // src/cpu_backend.cpp:24
constexpr int kGroupMax = 128; // src/cpu_backend.cpp:27
constexpr int kMaxLevel = 21;  // src/cpu_backend.cpp:28
constexpr int kSmallLeaf = 4;  // src/cpu_backend.cpp:29

// ---------------------------------------------------------------- scenarios
// Rng helpers: src/scenarios.cpp:19-34 (uniform from the top 53 bits of mt19937_64, Box-Muller
// normal with one draw of each, unit vector by z-uniform + azimuth).
struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    double uni() { return double(g() >> 11) * 0x1.0p-53; }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double gauss() {
        const double u = 1.0 - uni();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * uni());
    }
    void direction(double& x, double& y, double& z) {
        z = uni(-1.0, 1.0);
        const double s = std::sqrt(std::max(0.0, 1.0 - z * z)), ph = 2.0 * M_PI * uni();
        x = s * std::cos(ph);
        y = s * std::sin(ph);
    }
};

// Centre-of-mass frame in double, then rounding to float: src/scenarios.cpp:52-73.
void to_com_frame(std::vector<double>* col, std::size_t n, float* out[7]) {
    double M = 0, c[6] = {0, 0, 0, 0, 0, 0};
    for (std::size_t i = 0; i < n; ++i) {
        const double m = col[6][i];
        M += m;
        for (int k = 0; k < 6; ++k) c[k] += m * col[k][i];
    }
    for (int k = 0; k < 6; ++k) c[k] /= M;
    for (int k = 0; k < 6; ++k)
        for (std::size_t i = 0; i < n; ++i) out[k][i] = float(col[k][i] - c[k]);
    for (std::size_t i = 0; i < n; ++i) out[6][i] = float(col[6][i]);
}

// ---------------------------------------------------------------- tree helpers
// 21-bit dilation for 63-bit Morton keys: the bit at position b moves to 3b. Written as a loop over
// bits (the reference uses the equivalent magic-mask ladder, src/cpu_backend.cpp:52-60).
inline uint64_t dilate3(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
}

struct Node {
    float cx, cy, cz, m;
    float qxx, qxy, qxz, qyy, qyz, qzz;
    float rcrit2;
    uint32_t first, nchild, lo, hi;
    int32_t level;        // octree level in its key frame (0 = root of the frame)
    int32_t depth;        // levels from the root
    uint64_t prefix;      // key prefix of the cell in its key frame (3*level bits)
    int32_t frame;        // 0 = global keys; k > 0 = k-th key restart
    float gx, gy, gz, half;  // geometric cell centre and half side
};

struct Group { float bx, by, bz, hx, hy, hz; uint32_t lo, hi; };

// Locally essential tree in the reference's wire format (LetNode, LetTree: src/backend.hpp:63-87).
struct LetNodeO {
    float cx, cy, cz, mass;
    float qxx, qxy, qxz, qyy, qyz, qzz;
    float rcrit2;
    int32_t first_child, child_count, body_first, body_count;
};
static_assert(sizeof(LetNodeO) == 60, "LetNode is a 60-byte record");
struct LetO {
    std::vector<LetNodeO> nodes;
    std::vector<float> bx, by, bz, bm;
};
struct RBody { float x, y, z, m; };

// One list entry's terms for a target at (xi, yi, zi): a softened body (inv = 0 for r2 = 0, as the
// reference's masked rsqrt) and a monopole + quadrupole cell (src/cpu_backend.cpp:816-882).
inline void pair_term(float px, float py, float pz, float pm, float xi, float yi, float zi, float e2, float& sx,
                      float& sy, float& sz, float& sp) {
    const float dx = px - xi, dy = py - yi, dz = pz - zi;
    const float r2 = dx * dx + dy * dy + dz * dz + e2;
    const float inv = r2 > 0.f ? 1.f / std::sqrt(r2) : 0.f;
    const float mi = pm * inv, f = mi * inv * inv;
    sx += f * dx; sy += f * dy; sz += f * dz; sp += mi;
}
inline void cell_term(float cx, float cy, float cz, float cm, const float* q, float xi, float yi, float zi, float e2,
                      float& sx, float& sy, float& sz, float& sp) {
    const float dx = cx - xi, dy = cy - yi, dz = cz - zi;
    const float r2 = dx * dx + dy * dy + dz * dz + e2;
    const float inv = r2 > 0.f ? 1.f / std::sqrt(r2) : 0.f;
    const float inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
    // q: xx, xy, xz, yy, yz, zz
    const float qx = q[0] * dx + q[1] * dy + q[2] * dz;
    const float qy = q[1] * dx + q[3] * dy + q[4] * dz;
    const float qz = q[2] * dx + q[4] * dy + q[5] * dz;
    const float dqd = dx * qx + dy * qy + dz * qz;
    const float f = cm * inv3 + 2.5f * dqd * inv5 * inv2;
    sx += f * dx - qx * inv5; sy += f * dy - qy * inv5; sz += f * dz - qz * inv5;
    sp += cm * inv + 0.5f * dqd * inv5;
}

// Receiver walk of a remote LET against a group box (src/cpu_backend.cpp:788-814, rules of
// src/backend.hpp:78-83): a terminal node (no children, no bodies) or one whose com is at rcrit2 or
// beyond is accepted, as its bodies if it has 1..kSmallLeaf of them, else as its multipole; an
// opened node adds its bodies if it has any, else its children (depth-first, first child first).
void let_list(const LetO& t, const Group& g, std::vector<RBody>& rb, std::vector<LetNodeO>& rc,
              std::vector<int32_t>& st) {
    if (t.nodes.empty()) return;
    st.clear();
    st.push_back(0);
    auto bodies = [&](const LetNodeO& nd) {
        for (int32_t j = nd.body_first; j < nd.body_first + nd.body_count; ++j)
            rb.push_back({t.bx[j], t.by[j], t.bz[j], t.bm[j]});
    };
    while (!st.empty()) {
        const LetNodeO& nd = t.nodes[std::size_t(st.back())];
        st.pop_back();
        const bool terminal = nd.first_child < 0 && nd.body_count == 0;
        bool accept = terminal;
        if (!terminal) {
            const float dx = std::max(std::fabs(nd.cx - g.bx) - g.hx, 0.f);
            const float dy = std::max(std::fabs(nd.cy - g.by) - g.hy, 0.f);
            const float dz = std::max(std::fabs(nd.cz - g.bz) - g.hz, 0.f);
            accept = dx * dx + dy * dy + dz * dz >= nd.rcrit2;
        }
        if (accept) {
            if (nd.body_count > 0 && nd.body_count <= kSmallLeaf) bodies(nd);
            else rc.push_back(nd);
        } else if (nd.body_count > 0) {
            bodies(nd);
        } else {
            for (int32_t c = nd.first_child + nd.child_count; c-- > nd.first_child;) st.push_back(c);
        }
    }
}

struct Oracle {
    int nt = 1;
    std::size_t n = 0;
    std::vector<float> x, y, z, vx, vy, vz, m, ax, ay, az, pot, cost;
    std::vector<uint32_t> id;
    bool has_pot = false, tree_ok = false;
    uint64_t interactions = 0, listed = 0, p2p = 0, p2c = 0;
    // tree
    double ox = 0, oy = 0, oz = 0, side = 1;
    std::vector<uint64_t> key;
    std::vector<uint32_t> perm;  // last global sort permutation (sorted position -> prior index)
    std::vector<Node> nodes;
    std::vector<Group> groups;
    std::vector<uint32_t> leaves;
    int frames = 0;
    int max_depth = 0;

    // ---- state (src/cpu_backend.cpp:146-161, 279-305) ----
    void upload(std::size_t nn, const float* px, const float* py, const float* pz, const float* pvx,
                const float* pvy, const float* pvz, const float* pm, const uint32_t* pid) {
        n = nn;
        x.assign(px, px + n); y.assign(py, py + n); z.assign(pz, pz + n);
        vx.assign(pvx, pvx + n); vy.assign(pvy, pvy + n); vz.assign(pvz, pvz + n);
        m.assign(pm, pm + n); id.assign(pid, pid + n);
        ax.assign(n, 0.f); ay.assign(n, 0.f); az.assign(n, 0.f); pot.assign(n, 0.f);
        cost.clear();
        has_pot = tree_ok = false;
    }
    void kick(double dt) {
        const float h = float(dt);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long i = 0; i < long(n); ++i) { vx[i] += ax[i] * h; vy[i] += ay[i] * h; vz[i] += az[i] * h; }
    }
    void drift(double dt) {
        const float h = float(dt);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long i = 0; i < long(n); ++i) { x[i] += vx[i] * h; y[i] += vy[i] * h; z[i] += vz[i] * h; }
        has_pot = tree_ok = false;
    }

    // ---- diagnostics (src/cpu_backend.cpp:307-329) ----
    void diagnostics(double* d) {
        double s[12] = {};
        for (std::size_t i = 0; i < n; ++i) {
            const double mi = m[i], px = x[i], py = y[i], pz = z[i], ux = vx[i], uy = vy[i], uz = vz[i];
            s[0] += mi;
            s[1] += 0.5 * mi * (ux * ux + uy * uy + uz * uz);
            s[2] += 0.5 * mi * double(pot[i]);
            s[3] += mi * ux; s[4] += mi * uy; s[5] += mi * uz;
            s[6] += mi * (py * uz - pz * uy); s[7] += mi * (pz * ux - px * uz); s[8] += mi * (px * uy - py * ux);
            s[9] += mi * px; s[10] += mi * py; s[11] += mi * pz;
        }
        d[0] = s[0]; d[1] = s[1]; d[2] = has_pot ? s[2] : 0.0; d[3] = has_pot ? 1.0 : 0.0;
        for (int k = 0; k < 6; ++k) d[4 + k] = s[3 + k];
        for (int k = 0; k < 3; ++k) d[10 + k] = s[0] > 0 ? s[9 + k] / s[0] : 0.0;
    }

    // ---- exact PE (src/cpu_backend.cpp:331-347): pairs i < j in double ----
    double exact_pe(double G, double eps) {
        const double e2 = eps * eps;
        double pe = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64) reduction(+ : pe)
        for (long i = 0; i < long(n); ++i) {
            double s = 0;
            for (std::size_t j = std::size_t(i) + 1; j < n; ++j) {
                const double dx = double(x[j]) - x[i], dy = double(y[j]) - y[i], dz = double(z[j]) - z[i];
                const double r2 = dx * dx + dy * dy + dz * dz + e2;
                if (r2 > 0) s += double(m[j]) / std::sqrt(r2);
            }
            pe -= double(m[i]) * s;
        }
        return G * pe;
    }

    // ---- direct (src/cpu_backend.cpp:353-392): float pair terms, self skipped by index ----
    void direct(float G, float e2, bool potential) {
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
        for (long i = 0; i < long(n); ++i) {
            float sx = 0, sy = 0, sz = 0, sp = 0;
            for (std::size_t j = 0; j < n; ++j) {
                const float dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                const float r2 = dx * dx + dy * dy + dz * dz + e2;
                const float inv = (r2 > 0.f && j != std::size_t(i)) ? 1.f / std::sqrt(r2) : 0.f;
                const float mi = m[j] * inv, f = mi * inv * inv;
                sx += f * dx; sy += f * dy; sz += f * dz; sp += mi;
            }
            ax[i] = G * sx; ay[i] = G * sy; az[i] = G * sz;
            if (potential) pot[i] = -G * sp;
        }
        cost.assign(n, float(n));
        interactions = uint64_t(n) * (n ? n - 1 : 0);  // self excluded, as the GPU backends count
        listed = uint64_t(n) * n;                      // the reference CPU count (self included)
        p2p = interactions; p2c = 0;
    }

    // ---- keys and sort (src/cpu_backend.cpp:395-499) ----
    void keys_in_cube(std::size_t lo, std::size_t hi, double o0, double o1, double o2, double sd) {
        const double s = double(1u << kMaxLevel) / sd, lim = double((1u << kMaxLevel) - 1);
        for (std::size_t i = lo; i < hi; ++i) {
            auto q = [&](float v, double o) { return uint64_t(std::clamp((double(v) - o) * s, 0.0, lim)); };
            key[i] = dilate3(q(x[i], o0)) << 2 | dilate3(q(y[i], o1)) << 1 | dilate3(q(z[i], o2));
        }
    }
    // Stable order of [lo, hi) by key; ties keep index order (what a stable LSD radix sort gives).
    void sort_range(std::size_t lo, std::size_t hi, std::vector<uint32_t>& order) {
        order.resize(hi - lo);
        for (std::size_t i = lo; i < hi; ++i) order[i - lo] = uint32_t(i);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    }
    void apply_order(std::size_t lo, const std::vector<uint32_t>& order) {
        const std::size_t k = order.size();
        auto g = [&](auto& v) {
            auto tmp = v;
            for (std::size_t i = 0; i < k; ++i) v[lo + i] = tmp[order[i]];
        };
        g(x); g(y); g(z); g(vx); g(vy); g(vz); g(m); g(id); g(key);
        if (cost.size() == n) g(cost);
    }
    void bbox(std::size_t lo, std::size_t hi, float* mn, float* mx) const {
        mn[0] = mx[0] = x[lo]; mn[1] = mx[1] = y[lo]; mn[2] = mx[2] = z[lo];
        for (std::size_t i = lo; i < hi; ++i) {
            mn[0] = std::min(mn[0], x[i]); mn[1] = std::min(mn[1], y[i]); mn[2] = std::min(mn[2], z[i]);
            mx[0] = std::max(mx[0], x[i]); mx[1] = std::max(mx[1], y[i]); mx[2] = std::max(mx[2], z[i]);
        }
    }
    void sort_bodies() {
        float mn[3], mx[3];
        bbox(0, n, mn, mx);
        double sd = std::max({double(mx[0]) - mn[0], double(mx[1]) - mn[1], double(mx[2]) - mn[2]});
        sd = sd > 0 ? sd * (1 + 1e-6) : 1.0;  // src/cpu_backend.cpp:404-405
        ox = mn[0]; oy = mn[1]; oz = mn[2]; side = sd;
        key.assign(n, 0);
        keys_in_cube(0, n, ox, oy, oz, sd);
        sort_range(0, n, perm);
        apply_order(0, perm);
    }

    // Key restart for an over-full finest cell (src/cpu_backend.cpp:503-530): keys in the bodies'
    // own bounding cube, stable re-sort of the range. False if the bodies coincide.
    bool restart_keys(uint32_t lo, uint32_t hi, double& gx, double& gy, double& gz, double& half) {
        float mn[3], mx[3];
        bbox(lo, hi, mn, mx);
        double sd = std::max({double(mx[0]) - mn[0], double(mx[1]) - mn[1], double(mx[2]) - mn[2]});
        if (sd == 0) return false;
        sd *= 1 + 1e-6;
        keys_in_cube(lo, hi, mn[0], mn[1], mn[2], sd);
        std::vector<uint32_t> order;
        sort_range(lo, hi, order);
        apply_order(lo, order);
        half = 0.5 * sd;
        gx = mn[0] + half; gy = mn[1] + half; gz = mn[2] + half;
        return true;
    }

    // ---- octree (src/cpu_backend.cpp:534-583): depth-first recursion made iterative. A cell is
    // a leaf iff it holds <= 16 bodies; children are the non-empty octants in order, appended as one
    // contiguous block; single-child cells are kept. Walk groups: topmost cells with <= 128
    // bodies, or leaves. ----
    struct Pending { uint32_t ni; int level, depth; double gx, gy, gz, half; bool in_group; uint64_t prefix; int frame; };
    void build_tree(double theta) {
        nodes.clear(); groups.clear(); leaves.clear();
        frames = 0; max_depth = 0;
        Node root{};
        root.lo = 0; root.hi = uint32_t(n);
        nodes.push_back(root);
        const double h0 = 0.5 * side;
        std::vector<Pending> todo;
        todo.push_back({0, 0, 0, ox + h0, oy + h0, oz + h0, h0, false, 0, 0});
        while (!todo.empty()) {
            Pending p = todo.back();
            todo.pop_back();
            Node& nd = nodes[p.ni];
            const uint32_t lo = nd.lo, hi = nd.hi;
            max_depth = std::max(max_depth, p.depth);
            bool leaf = hi - lo <= uint32_t(kLeafMax);
            if (!leaf && p.level == kMaxLevel) {
                if (restart_keys(lo, hi, p.gx, p.gy, p.gz, p.half)) { p.level = 0; p.prefix = 0; p.frame = ++frames; }
                else leaf = true;
            }
            nd.level = p.level; nd.depth = p.depth; nd.prefix = p.prefix; nd.frame = p.frame;
            nd.gx = float(p.gx); nd.gy = float(p.gy); nd.gz = float(p.gz); nd.half = float(p.half);
            bool in_group = p.in_group;
            if (!in_group && (hi - lo <= uint32_t(kGroupMax) || leaf)) {
                groups.push_back({0, 0, 0, 0, 0, 0, lo, hi});
                in_group = true;
            }
            if (leaf) {
                nd.first = lo; nd.nchild = 0;
                leaves.push_back(p.ni);
                continue;
            }
            const int shift = 3 * (kMaxLevel - 1 - p.level);
            uint32_t b[9];
            b[0] = lo; b[8] = hi;
            for (int o = 1; o < 8; ++o) {
                uint32_t j = b[o - 1];
                while (j < hi && int((key[j] >> shift) & 7) < o) ++j;
                b[o] = j;
            }
            const uint32_t first = uint32_t(nodes.size());
            uint32_t nc = 0;
            for (int o = 0; o < 8; ++o)
                if (b[o + 1] > b[o]) {
                    Node c{};
                    c.lo = b[o]; c.hi = b[o + 1];
                    nodes.push_back(c);
                    ++nc;
                }
            nodes[p.ni].first = first;
            nodes[p.ni].nchild = nc;
            // Children are visited in octant order: push them in reverse onto the LIFO.
            const double q = 0.5 * p.half;
            std::vector<Pending> kids;
            uint32_t c = first;
            for (int o = 0; o < 8; ++o) {
                if (b[o + 1] == b[o]) continue;
                kids.push_back({c++, p.level + 1, p.depth + 1, p.gx + ((o & 4) ? q : -q), p.gy + ((o & 2) ? q : -q),
                                p.gz + ((o & 1) ? q : -q), q, in_group, (p.prefix << 3) | uint64_t(o), p.frame});
            }
            for (auto it = kids.rbegin(); it != kids.rend(); ++it) todo.push_back(*it);
        }
        // Group order: the reference appends groups in depth-first order, i.e. by body range.
        std::sort(groups.begin(), groups.end(), [](const Group& a, const Group& b) { return a.lo < b.lo; });
        moments(theta);
        tree_ok = true;
    }

    // ---- moments (src/cpu_backend.cpp:585-664) ----
    void finish(Node& nd, float inv_theta) {
        const float dx = nd.cx - nd.gx, dy = nd.cy - nd.gy, dz = nd.cz - nd.gz;
        const double r = double(2 * nd.half) * inv_theta + std::sqrt(double(dx * dx + dy * dy + dz * dz));
        nd.rcrit2 = float(std::min(r * r, double(std::numeric_limits<float>::max())));
    }
    void moments(double theta) {
        const float inv_theta = theta > 0 ? float(1.0 / theta) : std::numeric_limits<float>::max();
        for (auto& g : groups) {
            float mn[3], mx[3];
            bbox(g.lo, g.hi, mn, mx);
            g.bx = 0.5f * (mn[0] + mx[0]); g.by = 0.5f * (mn[1] + mx[1]); g.bz = 0.5f * (mn[2] + mx[2]);
            g.hx = 0.5f * (mx[0] - mn[0]); g.hy = 0.5f * (mx[1] - mn[1]); g.hz = 0.5f * (mx[2] - mn[2]);
        }
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
        for (long l = 0; l < long(leaves.size()); ++l) {
            Node& nd = nodes[leaves[l]];
            double M = 0, sx = 0, sy = 0, sz = 0;
            for (uint32_t i = nd.lo; i < nd.hi; ++i) {
                const double mi = m[i];
                M += mi; sx += mi * x[i]; sy += mi * y[i]; sz += mi * z[i];
            }
            const double cx = M > 0 ? sx / M : nd.gx, cy = M > 0 ? sy / M : nd.gy, cz = M > 0 ? sz / M : nd.gz;
            double q[6] = {0, 0, 0, 0, 0, 0};
            for (uint32_t i = nd.lo; i < nd.hi; ++i) {
                const double mi = m[i], dx = x[i] - cx, dy = y[i] - cy, dz = z[i] - cz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                q[0] += mi * (3 * dx * dx - r2); q[3] += mi * (3 * dy * dy - r2); q[5] += mi * (3 * dz * dz - r2);
                q[1] += mi * 3 * dx * dy; q[2] += mi * 3 * dx * dz; q[4] += mi * 3 * dy * dz;
            }
            nd.cx = float(cx); nd.cy = float(cy); nd.cz = float(cz); nd.m = float(M);
            nd.qxx = float(q[0]); nd.qxy = float(q[1]); nd.qxz = float(q[2]);
            nd.qyy = float(q[3]); nd.qyz = float(q[4]); nd.qzz = float(q[5]);
            finish(nd, inv_theta);
        }
        // Children always have larger indices than their parent: a reverse sweep is bottom-up.
        for (long ni = long(nodes.size()) - 1; ni >= 0; --ni) {
            Node& nd = nodes[ni];
            if (nd.nchild == 0) continue;
            double M = 0, sx = 0, sy = 0, sz = 0;
            for (uint32_t c = nd.first; c < nd.first + nd.nchild; ++c) {
                const Node& ch = nodes[c];
                M += ch.m; sx += double(ch.m) * ch.cx; sy += double(ch.m) * ch.cy; sz += double(ch.m) * ch.cz;
            }
            const double cx = M > 0 ? sx / M : nd.gx, cy = M > 0 ? sy / M : nd.gy, cz = M > 0 ? sz / M : nd.gz;
            double q[6] = {0, 0, 0, 0, 0, 0};
            for (uint32_t c = nd.first; c < nd.first + nd.nchild; ++c) {
                const Node& ch = nodes[c];
                const double mc = ch.m, dx = ch.cx - cx, dy = ch.cy - cy, dz = ch.cz - cz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                q[0] += ch.qxx + mc * (3 * dx * dx - r2); q[3] += ch.qyy + mc * (3 * dy * dy - r2);
                q[5] += ch.qzz + mc * (3 * dz * dz - r2);
                q[1] += ch.qxy + mc * 3 * dx * dy; q[2] += ch.qxz + mc * 3 * dx * dz; q[4] += ch.qyz + mc * 3 * dy * dz;
            }
            nd.cx = float(cx); nd.cy = float(cy); nd.cz = float(cz); nd.m = float(M);
            nd.qxx = float(q[0]); nd.qxy = float(q[1]); nd.qxz = float(q[2]);
            nd.qyy = float(q[3]); nd.qyz = float(q[4]); nd.qzz = float(q[5]);
            finish(nd, inv_theta);
        }
    }

    // ---- walk (src/cpu_backend.cpp:667-689, 816-882) ----
    // Per group: the interaction list against the group's bounding box. A node is accepted when it
    // does not contain the group and dist^2(box, com) >= rcrit2; accepted leaves with <= 4 bodies and
    // opened leaves contribute their bodies, other accepted nodes their multipole.
    // Remote LETs (src/cpu_backend.cpp:667-689): each group's list continues with every remote
    // tree's entries, bodies after the local bodies and cells after the local cells.
    void walk(float G, float e2, bool potential, const std::vector<const LetO*>& remote = {}) {
        cost.assign(n, 0.f);
        uint64_t inter = 0, lst = 0, npp = 0, npc = 0;
#pragma omp parallel num_threads(nt) reduction(+ : inter, lst, npp, npc)
        {
            std::vector<uint32_t> st, cells;
            std::vector<uint32_t> bods;
            std::vector<RBody> rb;
            std::vector<LetNodeO> rc;
            std::vector<int32_t> lst_stack;
#pragma omp for schedule(dynamic, 4)
            for (long gi = 0; gi < long(groups.size()); ++gi) {
                const Group& g = groups[gi];
                cells.clear(); bods.clear(); st.clear();
                st.push_back(0);
                while (!st.empty()) {
                    const uint32_t ni = st.back();
                    st.pop_back();
                    const Node& nd = nodes[ni];
                    const float dx = std::max(std::fabs(nd.cx - g.bx) - g.hx, 0.f);
                    const float dy = std::max(std::fabs(nd.cy - g.by) - g.hy, 0.f);
                    const float dz = std::max(std::fabs(nd.cz - g.bz) - g.hz, 0.f);
                    const bool contains = nd.lo <= g.lo && g.hi <= nd.hi;
                    if (!contains && dx * dx + dy * dy + dz * dz >= nd.rcrit2) {
                        if (nd.nchild == 0 && nd.hi - nd.lo <= uint32_t(kSmallLeaf))
                            for (uint32_t j = nd.lo; j < nd.hi; ++j) bods.push_back(j);
                        else
                            cells.push_back(ni);
                    } else if (nd.nchild == 0) {
                        for (uint32_t j = nd.lo; j < nd.hi; ++j) bods.push_back(j);
                    } else {
                        for (uint32_t c = nd.first + nd.nchild; c-- > nd.first;) st.push_back(c);
                    }
                }
                rb.clear(); rc.clear();
                for (const LetO* t : remote) let_list(*t, g, rb, rc, lst_stack);
                for (uint32_t i = g.lo; i < g.hi; ++i) {
                    float sx = 0, sy = 0, sz = 0, sp = 0;
                    uint64_t self = 0;
                    for (uint32_t j : bods) {
                        if (j == i) { ++self; continue; }
                        pair_term(x[j], y[j], z[j], m[j], x[i], y[i], z[i], e2, sx, sy, sz, sp);
                    }
                    for (const RBody& r : rb) pair_term(r.x, r.y, r.z, r.m, x[i], y[i], z[i], e2, sx, sy, sz, sp);
                    for (uint32_t c : cells) {
                        const Node& nd = nodes[c];
                        const float q[6] = {nd.qxx, nd.qxy, nd.qxz, nd.qyy, nd.qyz, nd.qzz};
                        cell_term(nd.cx, nd.cy, nd.cz, nd.m, q, x[i], y[i], z[i], e2, sx, sy, sz, sp);
                    }
                    for (const LetNodeO& nd : rc) {
                        const float q[6] = {nd.qxx, nd.qxy, nd.qxz, nd.qyy, nd.qyz, nd.qzz};
                        cell_term(nd.cx, nd.cy, nd.cz, nd.mass, q, x[i], y[i], z[i], e2, sx, sy, sz, sp);
                    }
                    ax[i] = G * sx; ay[i] = G * sy; az[i] = G * sz;
                    if (potential) pot[i] = -G * sp;
                    const uint64_t len = bods.size() + cells.size() + rb.size() + rc.size();
                    cost[i] = float(len);
                    lst += len;
                    inter += len - self;
                    npp += bods.size() + rb.size() - self;
                    npc += cells.size() + rc.size();
                }
            }
        }
        interactions = inter; listed = lst; p2p = npp; p2c = npc;
    }

    void compute_forces(double G, double eps, double theta, int method, bool potential) {
        has_pot = potential;
        if (n == 0) { interactions = listed = p2p = p2c = 0; return; }
        const float g = float(G), e2 = float(eps * eps);
        if (method == 0) { direct(g, e2, potential); return; }
        sort_bodies();
        build_tree(theta);
        walk(g, e2, potential);
    }

    // ---- distributed support (src/backend.hpp:114-136, src/cpu_backend.cpp:165-258) ----
    void build(double theta, int method) {
        tree_ok = false;
        if (method == 0 || n == 0) return;
        sort_bodies();
        build_tree(theta);
    }

    // Direct method with remote LETs: local pairs, then the remote trees per block of kGroupMax
    // consecutive bodies in the current order, each block's box from its bodies
    // (src/cpu_backend.cpp:691-725); costs n + the remote list length.
    void forces_with(double G, double eps, int method, bool potential, const std::vector<const LetO*>& remote) {
        has_pot = potential;
        if (n == 0) { interactions = listed = p2p = p2c = 0; return; }
        const float g = float(G), e2 = float(eps * eps);
        if (method == 1) {
            if (!tree_ok) throw std::logic_error("oracle compute_forces_with: build the tree first");
            walk(g, e2, potential, remote);
            return;
        }
        direct(g, e2, potential);
        uint64_t inter = 0, npp = 0, npc = 0;
        const long nblk = long((n + kGroupMax - 1) / kGroupMax);
#pragma omp parallel num_threads(nt) reduction(+ : inter, npp, npc)
        {
            std::vector<RBody> rb;
            std::vector<LetNodeO> rc;
            std::vector<int32_t> st;
#pragma omp for schedule(dynamic, 1)
            for (long b = 0; b < nblk; ++b) {
                Group gr{};
                gr.lo = uint32_t(b * kGroupMax);
                gr.hi = uint32_t(std::min<std::size_t>(n, std::size_t(b + 1) * kGroupMax));
                float mn[3], mx[3];
                bbox(gr.lo, gr.hi, mn, mx);
                gr.bx = 0.5f * (mn[0] + mx[0]); gr.by = 0.5f * (mn[1] + mx[1]); gr.bz = 0.5f * (mn[2] + mx[2]);
                gr.hx = 0.5f * (mx[0] - mn[0]); gr.hy = 0.5f * (mx[1] - mn[1]); gr.hz = 0.5f * (mx[2] - mn[2]);
                rb.clear(); rc.clear();
                for (const LetO* t : remote) let_list(*t, gr, rb, rc, st);
                for (uint32_t i = gr.lo; i < gr.hi; ++i) {
                    float sx = 0, sy = 0, sz = 0, sp = 0;
                    for (const RBody& r : rb) pair_term(r.x, r.y, r.z, r.m, x[i], y[i], z[i], e2, sx, sy, sz, sp);
                    for (const LetNodeO& nd : rc) {
                        const float q[6] = {nd.qxx, nd.qxy, nd.qxz, nd.qyy, nd.qyz, nd.qzz};
                        cell_term(nd.cx, nd.cy, nd.cz, nd.mass, q, x[i], y[i], z[i], e2, sx, sy, sz, sp);
                    }
                    ax[i] += g * sx; ay[i] += g * sy; az[i] += g * sz;
                    if (potential) pot[i] -= g * sp;
                    cost[i] += float(rb.size() + rc.size());
                }
                const uint64_t nb = gr.hi - gr.lo;
                inter += (rb.size() + rc.size()) * nb;
                npp += rb.size() * nb;
                npc += rc.size() * nb;
            }
        }
        interactions += inter; listed += inter; p2p += npp; p2c += npc;
    }

    // export_let (src/cpu_backend.cpp:222-258): breadth first from the root, the output doubling as
    // the queue; a node is terminal when min over the box of |x - com|^2 >= rcrit2 (box as float
    // centre and half extent), an opened leaf exports its bodies, an opened internal node appends
    // its children of positive mass as one block. Direct method: export_direct
    // (src/cpu_backend.cpp:727-745), one always-open root with every body.
    LetO export_let(const float* lo, const float* hi, int method) const {
        LetO out;
        if (n == 0) return out;
        if (method == 0) {
            out.bx = x; out.by = y; out.bz = z; out.bm = m;
            double M = 0, sx = 0, sy = 0, sz = 0;
            for (std::size_t i = 0; i < n; ++i) {
                const double w = m[i];
                M += w; sx += w * x[i]; sy += w * y[i]; sz += w * z[i];
            }
            LetNodeO r{};
            if (M > 0) { r.cx = float(sx / M); r.cy = float(sy / M); r.cz = float(sz / M); }
            r.mass = float(M);
            r.rcrit2 = 3.0e38f;
            r.first_child = -1; r.child_count = 0; r.body_first = 0; r.body_count = int32_t(n);
            out.nodes.push_back(r);
            return out;
        }
        if (!tree_ok) throw std::logic_error("oracle export_let: build the tree first");
        if (!(nodes[0].m > 0)) return out;
        const float bcx = 0.5f * (lo[0] + hi[0]), bhx = 0.5f * (hi[0] - lo[0]);
        const float bcy = 0.5f * (lo[1] + hi[1]), bhy = 0.5f * (hi[1] - lo[1]);
        const float bcz = 0.5f * (lo[2] + hi[2]), bhz = 0.5f * (hi[2] - lo[2]);
        std::vector<uint32_t> src;
        auto emit = [&](uint32_t ni) {
            const Node& d = nodes[ni];
            out.nodes.push_back(LetNodeO{d.cx, d.cy, d.cz, d.m, d.qxx, d.qxy, d.qxz, d.qyy, d.qyz, d.qzz, d.rcrit2,
                                         -1, 0, -1, 0});
            src.push_back(ni);
        };
        emit(0);
        for (std::size_t k = 0; k < out.nodes.size(); ++k) {
            const Node& d = nodes[src[k]];
            const float dx = std::max(std::fabs(d.cx - bcx) - bhx, 0.f);
            const float dy = std::max(std::fabs(d.cy - bcy) - bhy, 0.f);
            const float dz = std::max(std::fabs(d.cz - bcz) - bhz, 0.f);
            if (dx * dx + dy * dy + dz * dz >= d.rcrit2) continue;
            if (d.nchild == 0) {
                out.nodes[k].body_first = int32_t(out.bx.size());
                out.nodes[k].body_count = int32_t(d.hi - d.lo);
                for (uint32_t j = d.lo; j < d.hi; ++j) {
                    out.bx.push_back(x[j]); out.by.push_back(y[j]); out.bz.push_back(z[j]); out.bm.push_back(m[j]);
                }
                continue;
            }
            const int32_t first = int32_t(out.nodes.size());
            for (uint32_t c = d.first; c < d.first + d.nchild; ++c)
                if (nodes[c].m > 0) emit(c);
            out.nodes[k].first_child = first;
            out.nodes[k].child_count = int32_t(out.nodes.size()) - first;
        }
        return out;
    }
    LetO exported;
};

// LETs as flat arrays (the layout of lhip_let in include/lettuce_hip.h).
struct FlatLet {
    const void* nodes;
    uint64_t n_nodes;
    const float *bx, *by, *bz, *bm;
    uint64_t n_bodies;
};
std::vector<LetO> unflatten(const FlatLet* lets, int nlets) {
    std::vector<LetO> v(std::size_t(std::max(nlets, 0)));
    for (int k = 0; k < nlets; ++k) {
        const FlatLet& f = lets[k];
        const auto* nd = static_cast<const LetNodeO*>(f.nodes);
        v[k].nodes.assign(nd, nd + f.n_nodes);
        v[k].bx.assign(f.bx, f.bx + f.n_bodies); v[k].by.assign(f.by, f.by + f.n_bodies);
        v[k].bz.assign(f.bz, f.bz + f.n_bodies); v[k].bm.assign(f.bm, f.bm + f.n_bodies);
    }
    return v;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Plummer sphere (src/scenarios.cpp:94-132): radius by inverse CDF with redraw beyond cut*a,
// speed fraction q by von Neumann on q^2(1-q^2)^3.5 with bound 0.1, v = q sqrt(2GM/a)(1+r^2/a^2)^-1/4,
// isotropic directions; masses M/n; centre-of-mass frame.
void orc_plummer(uint64_t n, uint64_t seed, double G, double M, double a, double cut, float* x, float* y, float* z,
                 float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    Rng rng(seed);
    std::vector<double> col[7];
    for (auto& c : col) c.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
        double r = 0;
        do {
            const double X = rng.uni();
            if (X <= 0) continue;
            r = a / std::sqrt(std::pow(X, -2.0 / 3.0) - 1);
        } while (!(r <= cut * a));
        double q;
        for (;;) {
            q = rng.uni();
            if (0.1 * rng.uni() < q * q * std::pow(1 - q * q, 3.5)) break;
        }
        const double v = q * std::sqrt(2 * G * M / a) * std::pow(1 + r * r / (a * a), -0.25);
        double ux, uy, uz;
        rng.direction(ux, uy, uz);
        col[0][i] = r * ux; col[1][i] = r * uy; col[2][i] = r * uz;
        rng.direction(ux, uy, uz);
        col[3][i] = v * ux; col[4][i] = v * uy; col[5][i] = v * uz;
        col[6][i] = M / double(n);
    }
    float* out[7] = {x, y, z, vx, vy, vz, m};
    to_com_frame(col, n, out);
    for (uint64_t i = 0; i < n; ++i) id[i] = uint32_t(i);
}

// Rotating exponential disk with a sech^2 vertical profile around a central point mass. Not a
// reference generator: the reference disk (src/scenarios.cpp:173-236) is Gaussian in z. Radii follow
// the truncated exponential surface density by bisection of 1-(1+x)e^-x (as src/scenarios.cpp:197-205),
// z = z0 atanh(2U-1) (inverse CDF of sech^2), circular speed from the enclosed mass (core + stars
// inside r) with Gaussian dispersion sigma*vc per component. The core is body 0.
// incl rotates the disk about the x axis; (cx..cvz) places it; spin = +-1.
void orc_disk_sech2(uint64_t nstars, uint64_t seed, double G, double core, double mass, double rd, double rmax,
                    double z0, double sigma, double incl, double spin, double cx, double cy, double cz, double cvx,
                    double cvy, double cvz, int com_frame, float* x, float* y, float* z, float* vx, float* vy,
                    float* vz, float* m, uint32_t* id) {
    Rng rng(seed);
    const uint64_t n = nstars + 1;
    std::vector<double> col[7];
    for (auto& c : col) c.resize(n);
    const double ms = mass / double(nstars), xmax = rmax / rd, cmax = 1 - (1 + xmax) * std::exp(-xmax);
    const double ci = std::cos(incl), si = std::sin(incl);
    auto put = [&](uint64_t i, double px, double py, double pz, double ux, double uy, double uz, double pm) {
        col[0][i] = px + cx; col[1][i] = py * ci - pz * si + cy; col[2][i] = py * si + pz * ci + cz;
        col[3][i] = ux + cvx; col[4][i] = uy * ci - uz * si + cvy; col[5][i] = uy * si + uz * ci + cvz;
        col[6][i] = pm;
    };
    put(0, 0, 0, 0, 0, 0, 0, core);
    std::vector<double> rad(nstars);
    for (auto& r : rad) {
        const double t = rng.uni() * cmax;
        double lo = 0, hi = xmax;
        for (int it = 0; it < 48; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (1 - (1 + mid) * std::exp(-mid) < t) lo = mid; else hi = mid;
        }
        r = 0.5 * (lo + hi) * rd;
    }
    std::sort(rad.begin(), rad.end());
    for (uint64_t k = 0; k < nstars; ++k) {
        const double r = std::max(rad[k], 1e-12 * rd), ph = 2 * M_PI * rng.uni();
        const double enc = core + ms * double(k + 1), vc = std::sqrt(G * enc / r), sg = sigma * vc;
        double u = rng.uni();
        u = std::min(std::max(u, 1e-16), 1 - 1e-16);
        const double zz = z0 * std::atanh(2 * u - 1);
        const double c = std::cos(ph), s = std::sin(ph);
        const double ux = -spin * vc * s + sg * rng.gauss(), uy = spin * vc * c + sg * rng.gauss(), uz = sg * rng.gauss();
        put(k + 1, r * c, r * s, zz, ux, uy, uz, ms);
    }
    float* out[7] = {x, y, z, vx, vy, vz, m};
    if (com_frame) {
        to_com_frame(col, n, out);
    } else {
        for (int k = 0; k < 7; ++k)
            for (uint64_t i = 0; i < n; ++i) out[k][i] = float(col[k][i]);
    }
    for (uint64_t i = 0; i < n; ++i) id[i] = uint32_t(i);
}

// FP64 brute force for the targets listed in `idx` (all bodies are sources, self by index skipped):
// a_i = G sum_j m_j d_ij / (|d_ij|^2 + eps^2)^1.5, phi_i = -G sum_j m_j / sqrt(|d_ij|^2 + eps^2).
void orc_direct_f64(uint64_t n, const double* x, const double* y, const double* z, const double* m, double G,
                    double eps, uint64_t nt, const int64_t* idx, double* ax, double* ay, double* az, double* phi) {
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 4)
    for (long t = 0; t < long(nt); ++t) {
        const int64_t i = idx[t];
        double sx = 0, sy = 0, sz = 0, sp = 0;
        for (uint64_t j = 0; j < n; ++j) {
            if (int64_t(j) == i) continue;
            const double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            const double r2 = dx * dx + dy * dy + dz * dz + e2;
            if (!(r2 > 0)) continue;
            const double inv = 1.0 / std::sqrt(r2);
            sx += m[j] * dx * inv * inv * inv; sy += m[j] * dy * inv * inv * inv; sz += m[j] * dz * inv * inv * inv;
            sp += m[j] * inv;
        }
        ax[t] = G * sx; ay[t] = G * sy; az[t] = G * sz; phi[t] = -G * sp;
    }
}

// ---- differentiable FP64 simulation (no reference counterpart) ----
// Forward: the reference's KDK (src/simulation.cpp:406-438) in FP64 with FP64 direct forces; every
// position is kept. Backward: the textbook reverse of each step, two half-kick VJPs per step (not
// merged), each from the explicit Jacobian blocks da_i/dx_k and da_i/dm_k.
namespace {
struct Sim {
    std::size_t n = 0;
    int steps = 0;
    double G = 1, eps = 0, dt = 0;
    std::vector<double> m;
    std::vector<std::vector<double>> pos;  // x_0 .. x_T, 3n each (AoS)
};
Sim g_sim;

void accel64(const Sim& S, const std::vector<double>& x, std::vector<double>& a) {
    const std::size_t n = S.n;
    a.assign(3 * n, 0.0);
    const double e2 = S.eps * S.eps;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
        double s[3] = {0, 0, 0};
        for (std::size_t j = 0; j < n; ++j) {
            if (long(j) == i) continue;
            double d[3], r2 = e2;
            for (int c = 0; c < 3; ++c) { d[c] = x[3 * j + c] - x[3 * i + c]; r2 += d[c] * d[c]; }
            const double inv = 1.0 / std::sqrt(r2);
            for (int c = 0; c < 3; ++c) s[c] += S.m[j] * d[c] * inv * inv * inv;
        }
        for (int c = 0; c < 3; ++c) a[3 * i + c] = S.G * s[c];
    }
}

// lx += c * sum_i (da_i/dx_k)^T u_i, lm += c * sum_i (da_i/dm_k) . u_i, with
// da_i/dx_j = G m_j T_ij (j != i), da_i/dx_i = -sum_j G m_j T_ij, da_i/dm_j = G d_ij / s^3.
void vjp64(const Sim& S, const std::vector<double>& x, const std::vector<double>& u, double c,
           std::vector<double>& lx, std::vector<double>& lm) {
    const std::size_t n = S.n;
    const double e2 = S.eps * S.eps;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < long(n); ++k) {
        double w[3] = {0, 0, 0}, gm = 0;
        for (std::size_t i = 0; i < n; ++i) {
            if (long(i) == k) continue;
            double d[3], r2 = e2;  // d = x_k - x_i (from target i to source k)
            for (int q = 0; q < 3; ++q) { d[q] = x[3 * k + q] - x[3 * i + q]; r2 += d[q] * d[q]; }
            const double inv = 1.0 / std::sqrt(r2), s3 = inv * inv * inv, s5 = s3 * inv * inv;
            double T[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) T[a][b] = (a == b ? s3 : 0.0) - 3 * s5 * d[a] * d[b];
            // i != k: da_i/dx_k = G m_k T_ik applied to u_i; self block of a_k: -G m_i T_ki applied to u_k.
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) w[a] += S.G * (S.m[k] * T[a][b] * u[3 * i + b] - S.m[i] * T[a][b] * u[3 * k + b]);
            for (int q = 0; q < 3; ++q) gm += S.G * d[q] * s3 * u[3 * i + q];
        }
        for (int q = 0; q < 3; ++q) lx[3 * k + q] += c * w[q];
        lm[k] += c * gm;
    }
}
}  // namespace

void orc_sim_forward(uint64_t n, const double* x0, const double* v0, const double* m, double G, double eps, double dt,
                     int steps, double* xT, double* vT) {
    Sim& S = g_sim;
    S.n = n; S.steps = steps; S.G = G; S.eps = eps; S.dt = dt;
    S.m.assign(m, m + n);
    std::vector<double> x(x0, x0 + 3 * n), v(v0, v0 + 3 * n), a;
    S.pos.assign(1, x);
    accel64(S, x, a);
    for (int s = 0; s < steps; ++s) {
        for (std::size_t k = 0; k < 3 * n; ++k) v[k] += 0.5 * dt * a[k];
        for (std::size_t k = 0; k < 3 * n; ++k) x[k] += dt * v[k];
        accel64(S, x, a);
        for (std::size_t k = 0; k < 3 * n; ++k) v[k] += 0.5 * dt * a[k];
        S.pos.push_back(x);
    }
    std::memcpy(xT, x.data(), 3 * n * 8);
    std::memcpy(vT, v.data(), 3 * n * 8);
}

void orc_sim_backward(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) {
    Sim& S = g_sim;
    const std::size_t n = S.n;
    std::vector<double> lx(gxT, gxT + 3 * n), lv(gvT, gvT + 3 * n), lm(n, 0.0);
    for (int s = S.steps - 1; s >= 0; --s) {
        vjp64(S, S.pos[s + 1], lv, 0.5 * S.dt, lx, lm);                 // second half kick
        for (std::size_t k = 0; k < 3 * n; ++k) lv[k] += S.dt * lx[k];  // drift
        vjp64(S, S.pos[s], lv, 0.5 * S.dt, lx, lm);                     // first half kick
    }
    std::memcpy(gx0, lx.data(), 3 * n * 8);
    std::memcpy(gv0, lv.data(), 3 * n * 8);
    std::memcpy(gm, lm.data(), n * 8);
}

void* orc_create(int threads) {
    auto* o = new Oracle;
    o->nt = threads > 0 ? threads : omp_get_max_threads();
    return o;
}
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }
uint64_t orc_size(void* h) { return static_cast<Oracle*>(h)->n; }

void orc_upload(void* h, uint64_t n, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                const float* vz, const float* m, const uint32_t* id) {
    static_cast<Oracle*>(h)->upload(n, x, y, z, vx, vy, vz, m, id);
}

void orc_download(void* h, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    const Oracle& o = *static_cast<Oracle*>(h);
    const std::size_t b = o.n * 4;
    std::memcpy(x, o.x.data(), b); std::memcpy(y, o.y.data(), b); std::memcpy(z, o.z.data(), b);
    std::memcpy(vx, o.vx.data(), b); std::memcpy(vy, o.vy.data(), b); std::memcpy(vz, o.vz.data(), b);
    std::memcpy(m, o.m.data(), b); std::memcpy(id, o.id.data(), b);
}

int orc_compute_forces(void* h, double G, double eps, double theta, int method, int potential) {
    try {
        static_cast<Oracle*>(h)->compute_forces(G, eps, theta, method, potential != 0);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

void orc_kick(void* h, double dt) { static_cast<Oracle*>(h)->kick(dt); }
void orc_drift(void* h, double dt) { static_cast<Oracle*>(h)->drift(dt); }
void orc_diagnostics(void* h, double* out) { static_cast<Oracle*>(h)->diagnostics(out); }
double orc_exact_pe(void* h, double G, double eps) { return static_cast<Oracle*>(h)->exact_pe(G, eps); }

// counts: interactions (self excluded), listed (the reference CPU count), p2p, p2c
void orc_counts(void* h, uint64_t* out) {
    const Oracle& o = *static_cast<Oracle*>(h);
    out[0] = o.interactions; out[1] = o.listed; out[2] = o.p2p; out[3] = o.p2c;
}

void orc_accel(void* h, float* ax, float* ay, float* az, float* pot) {
    const Oracle& o = *static_cast<Oracle*>(h);
    const std::size_t b = o.n * 4;
    std::memcpy(ax, o.ax.data(), b); std::memcpy(ay, o.ay.data(), b); std::memcpy(az, o.az.data(), b);
    std::memcpy(pot, o.pot.data(), b);
}

int orc_body_costs(void* h, float* cost) {
    const Oracle& o = *static_cast<Oracle*>(h);
    if (o.cost.size() != o.n || o.n == 0) return 0;
    std::memcpy(cost, o.cost.data(), o.n * 4);
    return 1;
}

// Sorted keys (final frame per body: bodies in a restarted cell carry their restart keys) and the
// global sort permutation of the last build.
void orc_keys(void* h, uint64_t* keys, uint32_t* perm) {
    const Oracle& o = *static_cast<Oracle*>(h);
    std::memcpy(keys, o.key.data(), o.n * 8);
    if (perm) std::memcpy(perm, o.perm.data(), o.n * 4);
}

// Cube of the last build: origin xyz and side.
void orc_cube(void* h, double* out) {
    const Oracle& o = *static_cast<Oracle*>(h);
    out[0] = o.ox; out[1] = o.oy; out[2] = o.oz; out[3] = o.side;
}

// ---- distributed support ----
void orc_local_bounds(void* h, float* lo, float* hi) {
    const Oracle& o = *static_cast<Oracle*>(h);
    for (int c = 0; c < 3; ++c) lo[c] = hi[c] = 0.f;
    if (o.n) o.bbox(0, o.n, lo, hi);
}

int orc_build_tree(void* h, double theta, int method) {
    static_cast<Oracle*>(h)->build(theta, method);
    return 0;
}

int orc_forces_with(void* h, double G, double eps, int method, int potential, const FlatLet* lets, int nlets) {
    try {
        const std::vector<LetO> v = unflatten(lets, nlets);
        std::vector<const LetO*> r;
        for (const LetO& t : v) r.push_back(&t);
        static_cast<Oracle*>(h)->forces_with(G, eps, method, potential != 0, r);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// export_let for the box [lo, hi]: returns the sizes; orc_export_let_copy copies the LET out.
int orc_export_let(void* h, const float* lo, const float* hi, int method, uint64_t* nn, uint64_t* nb) {
    try {
        Oracle& o = *static_cast<Oracle*>(h);
        o.exported = o.export_let(lo, hi, method);
        *nn = o.exported.nodes.size();
        *nb = o.exported.bx.size();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

void orc_export_let_copy(void* h, void* nodes, float* bx, float* by, float* bz, float* bm) {
    const LetO& t = static_cast<Oracle*>(h)->exported;
    std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(LetNodeO));
    const std::size_t b = t.bx.size() * 4;
    std::memcpy(bx, t.bx.data(), b); std::memcpy(by, t.by.data(), b);
    std::memcpy(bz, t.bz.data(), b); std::memcpy(bm, t.bm.data(), b);
}

uint64_t orc_node_count(void* h) { return static_cast<Oracle*>(h)->nodes.size(); }
uint64_t orc_group_count(void* h) { return static_cast<Oracle*>(h)->groups.size(); }

// Node records in the reference's own node order (root first, each node's children contiguous).
// ints (8 per node): level, depth, frame, lo, hi, nchild, first, pad; prefix: u64 per node;
// flt (16 per node): cx, cy, cz, m, qxx, qxy, qxz, qyy, qyz, qzz, rcrit2, gx, gy, gz, half, 0.
void orc_nodes(void* h, int32_t* ints, uint64_t* prefix, float* flt) {
    const Oracle& o = *static_cast<Oracle*>(h);
    for (std::size_t k = 0; k < o.nodes.size(); ++k) {
        const Node& d = o.nodes[k];
        int32_t* I = ints + 8 * k;
        I[0] = d.level; I[1] = d.depth; I[2] = d.frame; I[3] = int32_t(d.lo); I[4] = int32_t(d.hi);
        I[5] = int32_t(d.nchild); I[6] = int32_t(d.first); I[7] = 0;
        prefix[k] = d.prefix;
        float* F = flt + 16 * k;
        const float v[16] = {d.cx, d.cy, d.cz, d.m, d.qxx, d.qxy, d.qxz, d.qyy, d.qyz, d.qzz, d.rcrit2,
                             d.gx, d.gy, d.gz, d.half, 0.f};
        std::memcpy(F, v, sizeof v);
    }
}

// Walk groups: body range [lo, hi) per group, in body order.
void orc_groups(void* h, uint32_t* lohi) {
    const Oracle& o = *static_cast<Oracle*>(h);
    for (std::size_t k = 0; k < o.groups.size(); ++k) { lohi[2 * k] = o.groups[k].lo; lohi[2 * k + 1] = o.groups[k].hi; }
}

}  // extern "C"
