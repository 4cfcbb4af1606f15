// Differentiable FP64-state simulation: kernel launchers (kernels_adjoint.hip).
#pragma once
#include "common.hpp"

namespace lhip {

// Largest source split the adjoint kernels use for n bodies (partial buffers: split * 4 * n doubles).
int adj_split(int n);
void launch_kick64(hipStream_t s, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                   const double* az, double dt, int n);
void launch_axpy3(hipStream_t s, double* y0, double* y1, double* y2, const double* x0, const double* x1,
                  const double* x2, double c, int n);
void launch_axpy1(hipStream_t s, double* y, const double* x, double c, int n);
void launch_fill64(hipStream_t s, double* y, double v, int n);
void launch_scale3(hipStream_t s, double* y0, double* y1, double* y2, const double* x0, const double* x1,
                   const double* x2, double c, int n);
// Direct force on FP64 positions: mixed precision (FP32 pair arithmetic on FP64 separations), or
// all-FP64 when pure_f64.
void launch_force64(hipStream_t s, const double* x, const double* y, const double* z, const double* m, int n,
                    double eps, double G, bool pure_f64, double* ax, double* ay, double* az, double* part);
// w = G sum_j T_kj (m_k u_j - m_j u_k), gm = -G sum_j s^-3 (u_j . d_kj).
void launch_vjp_direct(hipStream_t s, const double* x, const double* y, const double* z, const double* m,
                       const double* ux, const double* uy, const double* uz, int n, double eps, double G, double* wx,
                       double* wy, double* wz, double* gm, double* part);

}  // namespace lhip
