// Discrete-adjoint element kernels of the full-potential residual
// R_i = sum_e rho_tilde_e V_e (g_e . B_e,i) at a FROZEN upstream selection
// and branch (pyfp3d/kernels/upwind.py:544-643 semantics):
//   * d rho_tilde_e / d M_inf (branch-wise; feeds dR/dM_inf through the
//     coloured residual scatter),
//   * the reverse-mode product lambda^T dR/dx_nodes: per element, with
//     Lambda_e = sum_i lambda_i B_e,i, gL = g . Lambda and
//     sigma_e = V_e gL s_e + sum_{e' : u(e') = e} V_e' gL_e' s_u,e'
//     (the upstream chain), node a of e receives
//       sigma_e (-2 (g.B_a) g) + rho_tilde V gL B_a
//       - rho_tilde V ((B_a.Lambda) g + (g.B_a) Lambda)
//     from dB_j = -B_a (B_j . dx_a), dV = V (B_a . dx_a).
// Node sums use the node -> (4k+i) incidence lists (fixed order, no atomics).
#include <cmath>

#include "fp3d_common.h"

namespace fp3d {

__device__ __forceinline__ double a_mach2(double q2, double m, double g) {
  double num = q2 * (m * m);
  double den = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2);
  return num / den;
}
__device__ __forceinline__ double a_base(double q2, double m, double g) {
  double b = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2);
  return b < 1e-10 ? 1e-10 : b;
}
// d rho / d M at fixed q2: rho M (1 - q2) / base (0 on the vacuum floor)
__device__ __forceinline__ double a_drho_dm(double q2, double rho, double m,
                                            double g) {
  double b = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2);
  if (b < 1e-10) return 0.0;
  return rho * m * (1.0 - q2) / b;
}
// d M^2_local / d M at fixed q2: 2 q2 M / base^2
__device__ __forceinline__ double a_dm2_dm(double q2, double m, double g) {
  double b = a_base(q2, m, g);
  return 2.0 * q2 * m / (b * b);
}
__device__ __forceinline__ double a_nu(double m2, double mc2, double C) {
  double mm = m2 > mc2 ? m2 : mc2;
  double v = 1.0 - mc2 / mm;
  return C * (v > 0.0 ? v : 0.0);
}
__device__ __forceinline__ double a_dnu_dm(double q2, double m2, double m,
                                           double mc2, double C, double g) {
  if (!(m2 > mc2)) return 0.0;
  return C * (mc2 / (m2 * m2)) * a_dm2_dm(q2, m, g);
}

__global__ __launch_bounds__(kBlock) void k_drt_dmach(
    int64_t n, const double *__restrict__ q2, const double *__restrict__ rho,
    const int32_t *__restrict__ up, const int8_t *__restrict__ br, double m,
    double mc2, double C, double g, double *__restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int b = br[e];
  if (b == 3) {
    out[e] = 0.0;
    return;
  }
  double dre = a_drho_dm(q2[e], rho[e], m, g);
  if (b == 0) {
    out[e] = dre;
    return;
  }
  int64_t u = up[e];
  double dru = a_drho_dm(q2[u], rho[u], m, g);
  int64_t s = (b == 1) ? e : u;     // whose switch drives nu
  double m2 = a_mach2(q2[s], m, g);
  double nu = a_nu(m2, mc2, C);
  double dnu = a_dnu_dm(q2[s], m2, m, mc2, C, g);
  out[e] = dre * (1.0 - nu) + nu * dru - (rho[e] - rho[u]) * dnu;
}

// per element: w_e = V_e (g_e . Lambda_e), Lambda_e = sum_i lam_i B_e,i
__global__ __launch_bounds__(kBlock) void k_adj_weight(
    int64_t n, const int32_t *__restrict__ el, const double *__restrict__ B,
    const double *__restrict__ V, const double *__restrict__ grad,
    const double *__restrict__ lam, double *__restrict__ w,
    double *__restrict__ Lam) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double Lx = 0.0, Ly = 0.0, Lz = 0.0;
  for (int i = 0; i < 4; ++i) {
    double l = lam[el[4 * e + i]];
    Lx += l * B[12 * e + 3 * i];
    Ly += l * B[12 * e + 3 * i + 1];
    Lz += l * B[12 * e + 3 * i + 2];
  }
  Lam[3 * e] = Lx;
  Lam[3 * e + 1] = Ly;
  Lam[3 * e + 2] = Lz;
  w[e] = V[e] * (grad[3 * e] * Lx + grad[3 * e + 1] * Ly + grad[3 * e + 2] * Lz);
}

// sigma_e = w_e s_e + sum over downstream e' (u(e') = e, e' != e) of
// w_e' s_u,e' ; the downstream lists dptr/dlist are sorted by element
__global__ __launch_bounds__(kBlock) void k_adj_sigma(
    int64_t n, const double *__restrict__ w, const double *__restrict__ se,
    const double *__restrict__ su, const int32_t *__restrict__ dptr,
    const int32_t *__restrict__ dlist, double *__restrict__ sigma) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = w[e] * se[e];
  for (int32_t q = dptr[e]; q < dptr[e + 1]; ++q) {
    int32_t d = dlist[q];
    s += w[d] * su[d];
  }
  sigma[e] = s;
}

// per element and local node a: the 3-vector contribution to
// d(lambda^T R)/d x_a, stored at contrib[(4e + a) * 3 + :]
__global__ __launch_bounds__(kBlock) void k_adj_contrib(
    int64_t n, const double *__restrict__ B, const double *__restrict__ V,
    const double *__restrict__ grad, const double *__restrict__ rt,
    const double *__restrict__ Lam, const double *__restrict__ sigma,
    double *__restrict__ contrib) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double gx = grad[3 * e], gy = grad[3 * e + 1], gz = grad[3 * e + 2];
  double Lx = Lam[3 * e], Ly = Lam[3 * e + 1], Lz = Lam[3 * e + 2];
  double gL = gx * Lx + gy * Ly + gz * Lz;
  double rV = rt[e] * V[e];
  double sg = sigma[e];
  for (int a = 0; a < 4; ++a) {
    double bx = B[12 * e + 3 * a], by = B[12 * e + 3 * a + 1],
           bz = B[12 * e + 3 * a + 2];
    double gB = gx * bx + gy * by + gz * bz;
    double LB = Lx * bx + Ly * by + Lz * bz;
    double cg = -2.0 * sg * gB - rV * LB;
    double cb = rV * gL;
    double cl = -rV * gB;
    double *o = contrib + (4 * e + a) * 3;
    o[0] = cg * gx + cb * bx + cl * Lx;
    o[1] = cg * gy + cb * by + cl * Ly;
    o[2] = cg * gz + cb * bz + cl * Lz;
  }
}

// node gather of the per-(element, local) 3-vectors through the incidence
__global__ __launch_bounds__(kBlock) void k_node_gather3(
    int64_t nn, const int32_t *__restrict__ iptr,
    const int32_t *__restrict__ ient, const double *__restrict__ contrib,
    double *__restrict__ out) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  double x = 0.0, y = 0.0, z = 0.0;
  for (int32_t q = iptr[v]; q < iptr[v + 1]; ++q) {
    const double *c = contrib + 3 * (int64_t)ient[q];
    x += c[0];
    y += c[1];
    z += c[2];
  }
  out[3 * v] = x;
  out[3 * v + 1] = y;
  out[3 * v + 2] = z;
}

}  // namespace fp3d

using namespace fp3d;

extern "C" {

int fp3d_drho_tilde_dmach(int64_t n, const double *q2, const double *rho,
                          const int32_t *upstream, const int8_t *branch,
                          double m_inf, double m_crit, double C, double gamma,
                          double *out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && rho && upstream && branch && out),
               "drho_tilde_dmach: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_drt_dmach, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, q2, rho, upstream, branch, m_inf,
                     m_crit * m_crit, C, gamma, out);
  return check_launch("drho_tilde_dmach");
}

int fp3d_adjoint_dx(int64_t n_tets, int64_t n_nodes, const int32_t *elements,
                    const double *B, const double *V, const double *grad,
                    const double *rho_tilde, const double *s_e,
                    const double *s_u, const int32_t *down_ptr,
                    const int32_t *down_list, const int32_t *inc_ptr,
                    const int32_t *inc_ent, const double *lam, double *work,
                    double *out, void *stream) {
  FP3D_REQUIRE(n_tets > 0 && n_nodes > 0 && elements && B && V && grad &&
                   rho_tilde && s_e && s_u && down_ptr && down_list &&
                   inc_ptr && inc_ent && lam && work && out,
               "adjoint_dx: NULL pointer");
  hipStream_t s = FP3D_STREAM(stream);
  double *w = work;
  double *Lam = w + n_tets;
  double *sigma = Lam + 3 * n_tets;
  double *contrib = sigma + n_tets;
  unsigned g = grid_for(n_tets);
  hipLaunchKernelGGL(k_adj_weight, dim3(g), dim3(kBlock), 0, s, n_tets,
                     elements, B, V, grad, lam, w, Lam);
  hipLaunchKernelGGL(k_adj_sigma, dim3(g), dim3(kBlock), 0, s, n_tets, w, s_e,
                     s_u, down_ptr, down_list, sigma);
  hipLaunchKernelGGL(k_adj_contrib, dim3(g), dim3(kBlock), 0, s, n_tets, B, V,
                     grad, rho_tilde, Lam, sigma, contrib);
  hipLaunchKernelGGL(k_node_gather3, dim3(grid_for(n_nodes)), dim3(kBlock), 0,
                     s, n_nodes, inc_ptr, inc_ent, contrib, out);
  return check_launch("adjoint_dx");
}

int64_t fp3d_adjoint_dx_work_doubles(int64_t n_tets) { return 17 * n_tets; }

}  // extern "C"
