// Initial-condition generators of the library (host code): the reference's Plummer sphere, and the
// rotating sech^2 disk and two-disk encounter that the mixed-precision and adjoint workloads use.
// Outputs are float (the reference's Bodies, rounded from double after the centre-of-mass shift)
// or double (the same values before rounding, for FP64 initial conditions).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/lettuce_hip.h"

namespace {

// The reference's random stream (src/scenarios.cpp:19-47): mt19937_64, 53-bit uniforms in [0, 1),
// Box-Muller normals from (1 - u, u'), isotropic unit vectors from (z, phi).
class Stream {
public:
    explicit Stream(uint64_t seed) : g_(seed) {}
    double uniform() { return double(g_() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        const double u = 1.0 - uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * uniform());
    }
    void unit_vector(double& x, double& y, double& z) {
        z = uniform(-1.0, 1.0);
        const double s = std::sqrt(std::max(0.0, 1.0 - z * z)), ph = 2.0 * M_PI * uniform();
        x = s * std::cos(ph);
        y = s * std::sin(ph);
    }

private:
    std::mt19937_64 g_;
};

// Seven double columns (x, y, z, vx, vy, vz, m).
struct Columns {
    std::vector<double> c[7];
    explicit Columns(std::size_t n) {
        for (auto& v : c) v.assign(n, 0.0);
    }
    std::size_t size() const { return c[0].size(); }
};

// Centre-of-mass frame in double (src/scenarios.cpp:52-73).
void com_frame(Columns& b) {
    double M = 0, s[6] = {0, 0, 0, 0, 0, 0};
    for (std::size_t i = 0; i < b.size(); ++i) {
        M += b.c[6][i];
        for (int k = 0; k < 6; ++k) s[k] += b.c[6][i] * b.c[k][i];
    }
    if (!(M > 0)) return;
    for (int k = 0; k < 6; ++k) {
        const double mean = s[k] / M;
        for (double& v : b.c[k]) v -= mean;
    }
}

template <class T>
void emit(const Columns& b, T* const out[7], uint32_t* id, uint32_t id0) {
    for (int k = 0; k < 7; ++k)
        if (out[k])
            for (std::size_t i = 0; i < b.size(); ++i) out[k][i] = T(b.c[k][i]);
    if (id)
        for (std::size_t i = 0; i < b.size(); ++i) id[i] = id0 + uint32_t(i);
}

// Plummer sphere (src/scenarios.cpp:94-132): r = a / sqrt(U^(-2/3) - 1) redrawn beyond cut * a;
// q by von Neumann on q^2 (1 - q^2)^3.5 under 0.1; speed q v_esc(r) with
// v_esc = sqrt(2GM/a) (1 + r^2/a^2)^(-1/4); isotropic directions; masses M / n.
Columns plummer(uint64_t n, uint64_t seed, double G, double M, double a, double cut) {
    Stream rng(seed);
    Columns b(n);
    for (uint64_t i = 0; i < n; ++i) {
        double r = 0;
        do {
            const double u = rng.uniform();
            if (u <= 0) continue;
            r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1);
        } while (!(r <= cut * a));
        double q;
        for (;;) {
            q = rng.uniform();
            if (0.1 * rng.uniform() < q * q * std::pow(1 - q * q, 3.5)) break;
        }
        const double v = q * std::sqrt(2 * G * M / a) * std::pow(1 + r * r / (a * a), -0.25);
        double ux, uy, uz;
        rng.unit_vector(ux, uy, uz);
        b.c[0][i] = r * ux; b.c[1][i] = r * uy; b.c[2][i] = r * uz;
        rng.unit_vector(ux, uy, uz);
        b.c[3][i] = v * ux; b.c[4][i] = v * uy; b.c[5][i] = v * uz;
        b.c[6][i] = M / double(n);
    }
    com_frame(b);
    return b;
}

// A rotating exponential disk with a sech^2 vertical profile around a central point mass, in its
// own frame (disk plane z = 0, spin about +z), core first. Radii by bisection of the truncated
// exponential mass fraction 1 - (1 + x) e^-x (as src/scenarios.cpp:197-205), z = z0 atanh(2U - 1)
// (the inverse CDF of sech^2), circular speed sqrt(G M(<r) / r) from the core plus the stars
// inside r, Gaussian dispersion sigma * vc per component.
Columns disk_local(uint64_t nstars, uint64_t seed, double G, double core, double mass, double rd, double rmax,
                   double z0, double sigma, double spin) {
    Stream rng(seed);
    Columns b(nstars + 1);
    b.c[6][0] = core;
    const double ms = mass / double(nstars), xmax = rmax / rd, cmax = 1 - (1 + xmax) * std::exp(-xmax);
    std::vector<double> rad(nstars);
    for (double& r : rad) {
        const double t = rng.uniform() * cmax;
        double lo = 0, hi = xmax;
        for (int it = 0; it < 48; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (1 - (1 + mid) * std::exp(-mid) < t) lo = mid;
            else hi = mid;
        }
        r = 0.5 * (lo + hi) * rd;
    }
    std::sort(rad.begin(), rad.end());
    for (uint64_t k = 0; k < nstars; ++k) {
        const double r = std::max(rad[k], 1e-12 * rd), ph = 2 * M_PI * rng.uniform();
        const double vc = std::sqrt(G * (core + ms * double(k + 1)) / r), sg = sigma * vc;
        const double u = std::min(std::max(rng.uniform(), 1e-16), 1 - 1e-16);
        const double zz = z0 * std::atanh(2 * u - 1);
        const double c = std::cos(ph), s = std::sin(ph);
        const double ux = -spin * vc * s + sg * rng.normal(), uy = spin * vc * c + sg * rng.normal();
        const double uz = sg * rng.normal();
        const std::size_t i = k + 1;
        b.c[0][i] = r * c; b.c[1][i] = r * s; b.c[2][i] = zz;
        b.c[3][i] = ux; b.c[4][i] = uy; b.c[5][i] = uz;
        b.c[6][i] = ms;
    }
    return b;
}

// Rotate about the x axis by incl, then translate by (p, v).
void place(Columns& b, double incl, const double p[3], const double v[3]) {
    const double ci = std::cos(incl), si = std::sin(incl);
    for (std::size_t i = 0; i < b.size(); ++i) {
        const double y = b.c[1][i], z = b.c[2][i], vy = b.c[4][i], vz = b.c[5][i];
        b.c[0][i] += p[0];
        b.c[1][i] = y * ci - z * si + p[1];
        b.c[2][i] = y * si + z * ci + p[2];
        b.c[3][i] += v[0];
        b.c[4][i] = vy * ci - vz * si + v[1];
        b.c[5][i] = vy * si + vz * ci + v[2];
    }
}

int checked(uint64_t n) { return n >= uint64_t(INT32_MAX) ? LHIP_INVALID : LHIP_OK; }

// Two-disk encounter. Each disk (disk_local, spin +1) is tilted by its inclination about the x axis
// and its core placed on a parabolic relative orbit in the xy plane: cores r0 apart, pericentre q,
// approaching, in the frame of the two cores' centre of mass. True anomaly at the start:
// cos f0 = 2q/r0 - 1, f0 < 0; relative position r0 (cos f0, sin f0, 0); relative velocity
// sqrt(G Mt / 2q) (-sin f0, 1 + cos f0, 0) (parabola, semi-latus rectum 2q). Disk 1 (bodies
// [0, n1)) sits at -(M2/Mt) r, disk 2 at +(M1/Mt) r, M_k = core + disk mass.
struct Encounter {
    Columns d1, d2;  // disk-local bodies (core first)
    double M1, M2, Mt;
    double P1[3], P2[3], V1[3], V2[3];        // core positions and velocities
    double dP1[3], dP2[3], dV1[3], dV2[3];    // their derivatives in q
};

Encounter encounter(const lhip_collision_params* p) {
    Encounter e{disk_local(p->nstars, p->seed, p->G, p->core, p->mass, p->rd, p->rmax, p->z0, p->sigma, 1.0),
                disk_local(p->nstars, p->seed + 1, p->G, p->core, p->mass, p->rd, p->rmax, p->z0, p->sigma, 1.0),
                0, 0, 0, {}, {}, {}, {}, {}, {}, {}, {}};
    e.M1 = e.M2 = p->core + p->mass;
    e.Mt = e.M1 + e.M2;
    const double q = p->pericentre, r0 = p->separation;
    const double cf = 2 * q / r0 - 1, f = -std::acos(cf), sf = std::sin(f);
    const double dfdq = -2.0 / (r0 * sf);
    const double k = std::sqrt(p->G * e.Mt / (2 * q)), dkdq = -k / (2 * q);
    const double r[3] = {r0 * cf, r0 * sf, 0}, dr[3] = {-r0 * sf * dfdq, r0 * cf * dfdq, 0};
    const double v[3] = {-k * sf, k * (1 + cf), 0};
    const double dv[3] = {-dkdq * sf - k * cf * dfdq, dkdq * (1 + cf) - k * sf * dfdq, 0};
    const double w1 = -e.M2 / e.Mt, w2 = e.M1 / e.Mt;
    for (int c = 0; c < 3; ++c) {
        e.P1[c] = w1 * r[c]; e.P2[c] = w2 * r[c]; e.V1[c] = w1 * v[c]; e.V2[c] = w2 * v[c];
        e.dP1[c] = w1 * dr[c]; e.dP2[c] = w2 * dr[c]; e.dV1[c] = w1 * dv[c]; e.dV2[c] = w2 * dv[c];
    }
    return e;
}

int encounter_checked(const lhip_collision_params* p) {
    if (!p || p->nstars < 1 || checked(2 * (p->nstars + 1)) != LHIP_OK || !(p->rd > 0) || !(p->z0 > 0) ||
        !(p->pericentre > 0) || !(p->separation > 2 * p->pericentre))
        return LHIP_INVALID;
    return LHIP_OK;
}

}  // namespace

extern "C" {

int lhip_scenario_plummer(uint64_t n, uint64_t seed, double G, double M, double a, double cut, float* x, float* y,
                          float* z, float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    if (checked(n) != LHIP_OK || !(a > 0) || !(cut > 0)) return LHIP_INVALID;
    Columns b = plummer(n, seed, G, M, a, cut);
    float* out[7] = {x, y, z, vx, vy, vz, m};
    emit(b, out, id, 0);
    return LHIP_OK;
}

int lhip_scenario_plummer_f64(uint64_t n, uint64_t seed, double G, double M, double a, double cut, double* x,
                              double* y, double* z, double* vx, double* vy, double* vz, double* m, uint32_t* id) {
    if (checked(n) != LHIP_OK || !(a > 0) || !(cut > 0)) return LHIP_INVALID;
    Columns b = plummer(n, seed, G, M, a, cut);
    double* out[7] = {x, y, z, vx, vy, vz, m};
    emit(b, out, id, 0);
    return LHIP_OK;
}

int lhip_scenario_disk_sech2(const lhip_disk_params* p, int32_t com_frame_flag, float* x, float* y, float* z,
                             float* vx, float* vy, float* vz, float* m, uint32_t* id) {
    if (!p || p->nstars < 1 || checked(p->nstars + 1) != LHIP_OK || !(p->rd > 0) || !(p->rmax > 0) || !(p->z0 > 0))
        return LHIP_INVALID;
    Columns b = disk_local(p->nstars, p->seed, p->G, p->core, p->mass, p->rd, p->rmax, p->z0, p->sigma, p->spin);
    place(b, p->incl, p->pos, p->vel);
    if (com_frame_flag) com_frame(b);
    float* out[7] = {x, y, z, vx, vy, vz, m};
    emit(b, out, id, 0);
    return LHIP_OK;
}

int lhip_scenario_disk_sech2_f64(const lhip_disk_params* p, int32_t com_frame_flag, double* x, double* y, double* z,
                                 double* vx, double* vy, double* vz, double* m, uint32_t* id) {
    if (!p || p->nstars < 1 || checked(p->nstars + 1) != LHIP_OK || !(p->rd > 0) || !(p->rmax > 0) || !(p->z0 > 0))
        return LHIP_INVALID;
    Columns b = disk_local(p->nstars, p->seed, p->G, p->core, p->mass, p->rd, p->rmax, p->z0, p->sigma, p->spin);
    place(b, p->incl, p->pos, p->vel);
    if (com_frame_flag) com_frame(b);
    double* out[7] = {x, y, z, vx, vy, vz, m};
    emit(b, out, id, 0);
    return LHIP_OK;
}

int lhip_scenario_collision(const lhip_collision_params* p, double* x, double* v, double* m, uint32_t* id) {
    if (encounter_checked(p) != LHIP_OK) return LHIP_INVALID;
    const Encounter e = encounter(p);
    const std::size_t n1 = e.d1.size();
    const Columns* d[2] = {&e.d1, &e.d2};
    const double inc[2] = {p->incl1, p->incl2};
    const double* P[2] = {e.P1, e.P2};
    const double* V[2] = {e.V1, e.V2};
    for (int k = 0; k < 2; ++k) {
        const double ci = std::cos(inc[k]), si = std::sin(inc[k]);
        for (std::size_t i = 0; i < d[k]->size(); ++i) {
            const std::size_t o = k * n1 + i;
            const auto& c = d[k]->c;
            if (x) {
                x[3 * o + 0] = c[0][i] + P[k][0];
                x[3 * o + 1] = c[1][i] * ci - c[2][i] * si + P[k][1];
                x[3 * o + 2] = c[1][i] * si + c[2][i] * ci + P[k][2];
            }
            if (v) {
                v[3 * o + 0] = c[3][i] + V[k][0];
                v[3 * o + 1] = c[4][i] * ci - c[5][i] * si + V[k][1];
                v[3 * o + 2] = c[4][i] * si + c[5][i] * ci + V[k][2];
            }
            if (m) m[o] = c[6][i];
            if (id) id[o] = uint32_t(o);
        }
    }
    return LHIP_OK;
}

int lhip_scenario_collision_vjp(const lhip_collision_params* p, const double* gx, const double* gv, double out[3]) {
    if (encounter_checked(p) != LHIP_OK || !gx || !gv || !out) return LHIP_INVALID;
    const Encounter e = encounter(p);
    const std::size_t n1 = e.d1.size();
    const Columns* d[2] = {&e.d1, &e.d2};
    const double inc[2] = {p->incl1, p->incl2};
    const double* dP[2] = {e.dP1, e.dP2};
    const double* dV[2] = {e.dV1, e.dV2};
    double gq = 0, gi[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
        // d/di of (y ci - z si, y si + z ci) = (-y si - z ci, y ci - z si)
        const double ci = std::cos(inc[k]), si = std::sin(inc[k]);
        for (std::size_t i = 0; i < d[k]->size(); ++i) {
            const std::size_t o = k * n1 + i;
            const auto& c = d[k]->c;
            const double* ax = gx + 3 * o;
            const double* av = gv + 3 * o;
            gq += ax[0] * dP[k][0] + ax[1] * dP[k][1] + ax[2] * dP[k][2];
            gq += av[0] * dV[k][0] + av[1] * dV[k][1] + av[2] * dV[k][2];
            gi[k] += ax[1] * (-c[1][i] * si - c[2][i] * ci) + ax[2] * (c[1][i] * ci - c[2][i] * si);
            gi[k] += av[1] * (-c[4][i] * si - c[5][i] * ci) + av[2] * (c[4][i] * ci - c[5][i] * si);
        }
    }
    out[0] = gq;
    out[1] = gi[0];
    out[2] = gi[1];
    return LHIP_OK;
}

}  // extern "C"
