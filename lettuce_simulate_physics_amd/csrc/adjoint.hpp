// Differentiable FP64-state simulation: kernel launchers (kernels_adjoint.hip).
#pragma once
#include "common.hpp"

namespace lhip {

// Per-body source record of the pair kernels: FP64 position, FP32 mass and FP32 cotangent u on the
// body's acceleration (VJP sources only). 48 bytes, read with scalar loads.
struct SimRec {
    double x, y, z;
    float m, ux, uy, uz;
    float pad[2];
};
static_assert(sizeof(SimRec) == 48, "SimRec layout");

// Largest source split the adjoint kernels use for n bodies (partial buffers: split * 4 * n doubles).
int adj_split(int n);

// rec <- (x, m).
void launch_pack(hipStream_t s, const double* x, const double* y, const double* z, const double* m, SimRec* rec,
                 int n);
// v += h a; x += dt v; rec <- (x, m); copy (3 SoA arrays of n, may be null) <- x.
void launch_kick_drift(hipStream_t s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                       const double* ax, const double* ay, const double* az, const double* m, double h, double dt,
                       SimRec* rec, double* copy, int n);
// a = direct force at the positions in rec (mixed precision, or all-FP64 with FP64 masses m when
// pure_f64); then v += h a (skipped when h == 0).
void launch_force_finish(hipStream_t s, const SimRec* rec, const double* m, int n, double eps, double G, bool pure_f64,
                         double h, double* ax, double* ay, double* az, double* vx, double* vy, double* vz,
                         double* part);
// rec <- (pos, m, c * lambda_v), pos = 3 SoA arrays of n.
void launch_pack_vjp(hipStream_t s, const double* pos, const double* m, const double* lvx, const double* lvy,
                     const double* lvz, double c, SimRec* rec, int n);
// One VJP at the sources in rec: lambda_x += G sum_j T_kj (m_k u_j - m_j u_k),
// lambda_m += -G sum_j s^-3 (u_j . d_kj); then lambda_v += drift * lambda_x (drift != 0) and, when
// npos is not null, next_rec <- (npos, m, c_next * lambda_v). next_rec may alias rec.
void launch_vjp_update(hipStream_t s, const SimRec* rec, int n, double eps, double G, double* part, double* lx,
                       double* ly, double* lz, double* lm, double* lvx, double* lvy, double* lvz, double drift,
                       const double* npos, const double* m, double c_next, SimRec* next_rec);
void launch_fill64(hipStream_t s, double* y, double v, int n);
// y += c * x.
void launch_axpy_f64(hipStream_t s, double* y, const double* x, double c, int n);

// Direct summation on FP64 positions with float masses (the mixed-precision backend form): FP32
// pair arithmetic on FP64 separations, FP64 carry every 64 sources; a = G sum, phi = -G sum m/s.
// rec: n records of scratch; part: direct64_split(n) * 4 * n doubles.
int direct64_split(int n);
void launch_direct64(hipStream_t s, const double* x, const double* y, const double* z, const float* m, int n,
                     double eps, double G, bool potential, double* ax, double* ay, double* az, double* pot,
                     SimRec* rec, double* part);

}  // namespace lhip
