// Element kernels of the full-potential solve on P1 tetrahedra (gfx950).
//
// Every kernel here is HBM-bound integer/FP64 gather-scatter work: one
// work-item per element, wave64, no MFMA.  Arithmetic is written in the
// reference's evaluation order and the library is compiled with
// -ffp-contract=off, so -- apart from pow() -- each element's numbers are
// bit-identical to an IEEE restatement of the same formulas.  The coloured
// scatters are launched colour by colour with plain read-modify-writes
// (colour classes share no node), which fixes the accumulation order and
// makes every assembly bitwise repeatable, with no atomics.
#include <cmath>
#include <mutex>

#include "fp3d_common.h"

namespace fp3d {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return FP3D_EHIP;
  }
  return FP3D_OK;
}

// ---------------------------------------------------------------- physics
// pyfp3d/physics/isentropic.py:29-58 density_isentropic (vacuum floor 1e-10)
__device__ __forceinline__ double d_density(double q2, double m, double g) {
  double ex = 1.0 / (g - 1.0);
  double base = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2);
  if (base < 1e-10) base = 1e-10;
  return pow(base, ex);
}
// pyfp3d/physics/isentropic.py:79-97 mach_number_squared
__device__ __forceinline__ double d_mach2(double q2, double m, double g) {
  double num = q2 * (m * m);
  double den = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2);
  return num / den;
}
// pyfp3d/physics/isentropic.py:163-180 density_derivative_wrt_q_sq
__device__ __forceinline__ double d_drho(double q2, double m, double g) {
  double rho = d_density(q2, m, g);
  return -0.5 * (m * m) * pow(rho, 2.0 - g);
}
// pyfp3d/physics/isentropic.py:183-206 mach_squared_derivative_wrt_q_sq
__device__ __forceinline__ double d_dmach2(double q2, double m, double g) {
  double D = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2);
  return (m * m) * (1.0 + 0.5 * (g - 1.0) * (m * m)) / (D * D);
}
// nu switch C * max(0, 1 - Mc^2 / max(M^2, Mc^2))
__device__ __forceinline__ double d_nu(double m2, double mc2, double C) {
  double mm = m2 > mc2 ? m2 : mc2;
  double v = 1.0 - mc2 / mm;
  return C * (v > 0.0 ? v : 0.0);
}

__global__ void k_density_field(int64_t n, const double *q2, double m,
                                double g, double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = d_density(q2[i], m, g);
}
__global__ void k_mach2_field(int64_t n, const double *q2, double m, double g,
                              double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = d_mach2(q2[i], m, g);
}
__global__ void k_limit_q2(int64_t n, const double *q2, double cap,
                           double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double v = q2[i];
    out[i] = v < cap ? v : cap;
  }
}
__global__ void k_drho_field(int64_t n, const double *q2, double m, double g,
                             double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = d_drho(q2[i], m, g);
}

// ---------------------------------------------------------- element sweeps
__device__ __forceinline__ void elem_grad(int64_t e, const int32_t *elements,
                                          const double *B, const double *phi,
                                          double &gx, double &gy, double &gz) {
  const int4 nd = *reinterpret_cast<const int4 *>(elements + 4 * e);
  const double *b = B + 12 * e;
  double p0 = phi[nd.x], p1 = phi[nd.y], p2 = phi[nd.z], p3 = phi[nd.w];
  gx = 0.0; gy = 0.0; gz = 0.0;
  gx += p0 * b[0]; gy += p0 * b[1]; gz += p0 * b[2];
  gx += p1 * b[3]; gy += p1 * b[4]; gz += p1 * b[5];
  gx += p2 * b[6]; gy += p2 * b[7]; gz += p2 * b[8];
  gx += p3 * b[9]; gy += p3 * b[10]; gz += p3 * b[11];
}

__global__ __launch_bounds__(kBlock) void k_velocity_q2(
    int64_t n, const int32_t *__restrict__ elements,
    const double *__restrict__ B, const double *__restrict__ phi,
    double *__restrict__ grad, double *__restrict__ q2) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double gx, gy, gz;
  elem_grad(e, elements, B, phi, gx, gy, gz);
  grad[3 * e] = gx;
  grad[3 * e + 1] = gy;
  grad[3 * e + 2] = gz;
  q2[e] = gx * gx + gy * gy + gz * gz;
}

__device__ __forceinline__ void k_element_state_body(int64_t bx,
    int64_t n, const int32_t *__restrict__ elements,
    const double *__restrict__ B, const double *__restrict__ phi, double u2,
    double m, double g, double cap, double *__restrict__ grad,
    double *__restrict__ q2n_out, double *__restrict__ q2l_out,
    double *__restrict__ rho_out, uint8_t *__restrict__ lim_out,
    double *__restrict__ gdotb) {
  int64_t e = (int64_t)bx * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double gx, gy, gz;
  elem_grad(e, elements, B, phi, gx, gy, gz);
  grad[3 * e] = gx;
  grad[3 * e + 1] = gy;
  grad[3 * e + 2] = gz;
  if (gdotb) {
    // grad . B_a, the factor every later pass applies to node a: the same
    // expression (and so the same bits) the residual and Term-3 passes
    // evaluate when they are not handed it
    const double *b = B + 12 * e;
    double4 o;
    o.x = gx * b[0] + gy * b[1] + gz * b[2];
    o.y = gx * b[3] + gy * b[4] + gz * b[5];
    o.z = gx * b[6] + gy * b[7] + gz * b[8];
    o.w = gx * b[9] + gy * b[10] + gz * b[11];
    *reinterpret_cast<double4 *>(gdotb + 4 * e) = o;
  }
  double q2 = gx * gx + gy * gy + gz * gz;
  double q2n = q2 / u2;
  double q2l = q2n < cap ? q2n : cap;
  q2n_out[e] = q2n;
  q2l_out[e] = q2l;
  rho_out[e] = d_density(q2l, m, g);
  if (lim_out) lim_out[e] = (q2l == q2n) ? 1 : 0;
}
__global__ __launch_bounds__(kBlock) void k_element_state(
    int64_t n, const int32_t *__restrict__ elements,
    const double *__restrict__ B, const double *__restrict__ phi, double u2,
    double m, double g, double cap, double *__restrict__ grad,
    double *__restrict__ q2n_out, double *__restrict__ q2l_out,
    double *__restrict__ rho_out, uint8_t *__restrict__ lim_out,
    double *__restrict__ gdotb) {
  k_element_state_body(blockIdx.x, n, elements, B, phi, u2, m, g, cap, grad, q2n_out,
                       q2l_out, rho_out, lim_out, gdotb);
}

// ---------------------------------------------------------------- upwinding
// pyfp3d/kernels/upwind.py:68-125: step to the face neighbour with the most
// negative centroid displacement along grad[e] (strict <, first face wins a
// tie) until the accumulated reach covers 0.8 of e's streamwise extent or
// max_hops steps were taken; u(e) = e when no neighbour lies upstream.
__device__ __forceinline__ int32_t walk_upstream(
    int64_t e, const int32_t *__restrict__ fn, const double *__restrict__ cen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    double gx, double gy, double gz, int max_hops) {
  double pmin = 1e300, pmax = -1e300;
  for (int k = 0; k < 4; ++k) {
    int32_t nd = elements[4 * e + k];
    double p = nodes[3 * nd] * gx + nodes[3 * nd + 1] * gy + nodes[3 * nd + 2] * gz;
    if (p < pmin) pmin = p;
    if (p > pmax) pmax = p;
  }
  double target = 0.8 * (pmax - pmin);
  int32_t cur = (int32_t)e;
  double reach = 0.0;
  int hops = 0;
  while (hops < max_hops && reach < target) {
    int32_t best = -1;
    double best_disp = 0.0;
    double cx = cen[3 * (int64_t)cur], cy = cen[3 * (int64_t)cur + 1],
           cz = cen[3 * (int64_t)cur + 2];
    for (int f = 0; f < 4; ++f) {
      int32_t nb = fn[4 * (int64_t)cur + f];
      if (nb >= 0) {
        double d = (cen[3 * (int64_t)nb] - cx) * gx +
                   (cen[3 * (int64_t)nb + 1] - cy) * gy +
                   (cen[3 * (int64_t)nb + 2] - cz) * gz;
        if (d < best_disp) {
          best_disp = d;
          best = nb;
        }
      }
    }
    if (best < 0) break;
    cur = best;
    reach += -best_disp;
    hops += 1;
  }
  return cur;
}

// Same walk reading the per-face displacement table dcen[k][f] =
// centroid(nb) - centroid(k) (0 on the boundary, precomputed once in
// float64 with the same subtraction): one 16 B neighbour row and one 96 B
// displacement row per hop instead of five scattered centroid gathers.
// The dot products are the same IEEE operations, so u(e) is identical.
__device__ __forceinline__ int32_t walk_upstream_d(
    int64_t e, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    double gx, double gy, double gz, int max_hops) {
  double pmin = 1e300, pmax = -1e300;
  const int4 nd4 = *reinterpret_cast<const int4 *>(elements + 4 * e);
  const int32_t nds[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int64_t nd = nds[k];
    double p = nodes[3 * nd] * gx + nodes[3 * nd + 1] * gy + nodes[3 * nd + 2] * gz;
    if (p < pmin) pmin = p;
    if (p > pmax) pmax = p;
  }
  double target = 0.8 * (pmax - pmin);
  int32_t cur = (int32_t)e;
  double reach = 0.0;
  int hops = 0;
  while (hops < max_hops && reach < target) {
    const int4 nb4 = *reinterpret_cast<const int4 *>(fn + 4 * (int64_t)cur);
    const double2 *dp = reinterpret_cast<const double2 *>(dcen + 12 * (int64_t)cur);
    double2 d01 = dp[0], d23 = dp[1], d45 = dp[2], d67 = dp[3], d89 = dp[4],
            dab = dp[5];
    const int32_t nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const double dx[4] = {d01.x, d23.y, d67.x, d89.y};
    const double dy[4] = {d01.y, d45.x, d67.y, dab.x};
    const double dz[4] = {d23.x, d45.y, d89.x, dab.y};
    int32_t best = -1;
    double best_disp = 0.0;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (nbs[f] >= 0) {
        double d = dx[f] * gx + dy[f] * gy + dz[f] * gz;
        if (d < best_disp) {
          best_disp = d;
          best = nbs[f];
        }
      }
    }
    if (best < 0) break;
    cur = best;
    reach += -best_disp;
    hops += 1;
  }
  return cur;
}

__global__ __launch_bounds__(kBlock) void k_upstream_d(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, int max_hops, int32_t *__restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  out[e] = walk_upstream_d(e, fn, dcen, nodes, elements, grad[3 * e],
                           grad[3 * e + 1], grad[3 * e + 2], max_hops);
}

__global__ __launch_bounds__(kBlock) void k_upstream(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ cen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, int max_hops, int32_t *__restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  out[e] = walk_upstream(e, fn, cen, nodes, elements, grad[3 * e],
                         grad[3 * e + 1], grad[3 * e + 2], max_hops);
}

// pyfp3d/kernels/upwind.py:483-500
__device__ __forceinline__ void sweep_one(int64_t e, int32_t u, const double *q2,
                                          const double *rho, double m,
                                          double mc2, double C, double g,
                                          double floor_, double &nu_o,
                                          double &rt_o) {
  double m2 = d_mach2(q2[e], m, g);
  double m2u = d_mach2(q2[u], m, g);
  double nu_e = d_nu(m2, mc2, C);
  double nu_u = d_nu(m2u, mc2, C);
  double nu = nu_e >= nu_u ? nu_e : nu_u;
  nu_o = nu;
  double rt = rho[e] - nu * (rho[e] - rho[u]);
  rt_o = rt > floor_ ? rt : floor_;
}

__global__ __launch_bounds__(kBlock) void k_rho_tilde(
    int64_t n, const double *__restrict__ q2, const double *__restrict__ rho,
    const int32_t *__restrict__ up, double m, double mc2, double C, double g,
    double floor_, double *__restrict__ nu_out, double *__restrict__ rt_out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double nu, rt;
  sweep_one(e, up[e], q2, rho, m, mc2, C, g, floor_, nu, rt);
  nu_out[e] = nu;
  rt_out[e] = rt;
}

__global__ __launch_bounds__(kBlock) void k_upwind_fused(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ cen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, const double *__restrict__ q2,
    const double *__restrict__ rho, int max_hops, double m, double mc2,
    double C, double g, double floor_, int32_t *__restrict__ up_out,
    double *__restrict__ nu_out, double *__restrict__ rt_out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int32_t u = walk_upstream(e, fn, cen, nodes, elements, grad[3 * e],
                            grad[3 * e + 1], grad[3 * e + 2], max_hops);
  if (up_out) up_out[e] = u;
  double nu, rt;
  sweep_one(e, u, q2, rho, m, mc2, C, g, floor_, nu, rt);
  nu_out[e] = nu;
  rt_out[e] = rt;
}

__global__ __launch_bounds__(kBlock) void k_upwind_fused_d(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, const double *__restrict__ q2,
    const double *__restrict__ rho, int max_hops, double m, double mc2,
    double C, double g, double floor_, int32_t *__restrict__ up_out,
    double *__restrict__ nu_out, double *__restrict__ rt_out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int32_t u = walk_upstream_d(e, fn, dcen, nodes, elements, grad[3 * e],
                              grad[3 * e + 1], grad[3 * e + 2], max_hops);
  if (up_out) up_out[e] = u;
  double nu, rt;
  sweep_one(e, u, q2, rho, m, mc2, C, g, floor_, nu, rt);
  nu_out[e] = nu;
  rt_out[e] = rt;
}

// pyfp3d/kernels/upwind.py:503-541
__global__ __launch_bounds__(kBlock) void k_classify(
    int64_t n, const double *__restrict__ q2, const double *__restrict__ rho,
    const int32_t *__restrict__ up, double m, double mc2, double C, double g,
    double floor_, int8_t *__restrict__ br) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int32_t u = up[e];
  if (u == e) {
    br[e] = rho[e] <= floor_ ? 3 : 0;
    return;
  }
  double nu_e = d_nu(d_mach2(q2[e], m, g), mc2, C);
  double nu_u = d_nu(d_mach2(q2[u], m, g), mc2, C);
  double nu = nu_e >= nu_u ? nu_e : nu_u;
  double rt = rho[e] - nu * (rho[e] - rho[u]);
  if (rt <= floor_) br[e] = 3;
  else if (nu == 0.0) br[e] = 0;
  else if (nu_e >= nu_u) br[e] = 1;
  else br[e] = 2;
}

// pyfp3d/kernels/upwind.py:544-593
__global__ __launch_bounds__(kBlock) void k_frozen(
    int64_t n, const double *__restrict__ q2, const double *__restrict__ rho,
    const int32_t *__restrict__ up, const int8_t *__restrict__ br, double m,
    double mc2, double C, double g, double floor_, double *__restrict__ nu_out,
    double *__restrict__ rt_out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int b = br[e];
  if (b == 3) {
    nu_out[e] = 0.0;
    rt_out[e] = floor_;
    return;
  }
  if (b == 0) {
    nu_out[e] = 0.0;
    rt_out[e] = rho[e];
    return;
  }
  int32_t u = up[e];
  double nu = (b == 1) ? d_nu(d_mach2(q2[e], m, g), mc2, C)
                       : d_nu(d_mach2(q2[u], m, g), mc2, C);
  nu_out[e] = nu;
  rt_out[e] = rho[e] - nu * (rho[e] - rho[u]);
}

// dnu/dq2 on the active switch branch: C (Mc^2 / M^4) dM^2/dq2
__device__ __forceinline__ double d_dnu(double q2, double m2, double m,
                                        double mc2, double C, double g) {
  return C * (mc2 / (m2 * m2)) * d_dmach2(q2, m, g);
}

// Walk + switch + rho_tilde + the exact sensitivities (s_e, s_u) of the
// same selection in one sweep: the Newton residual evaluation keeps them for
// the Jacobian of an accepted state, so the linearisation skips its own
// pass over q2 / rho of e and of u.  The same expressions as sweep_one and
// k_sens, so the same bits.
// nu, rho_tilde and (SENS) the exact (s_e, s_u) of element e with upstream u
template <bool SENS>
__device__ __forceinline__ void sens_of_selection(
    int64_t e, int32_t u, const double *__restrict__ q2,
    const double *__restrict__ rho, double m, double mc2, double C, double g,
    double floor_, double *__restrict__ nu_out, double *__restrict__ rt_out,
    double *__restrict__ se, double *__restrict__ su) {
  const double q2e = q2[e], q2u = q2[u];
  const double rho_e = rho[e], rho_u = rho[u];
  const double m2 = d_mach2(q2e, m, g);
  const double m2u = d_mach2(q2u, m, g);
  const double nu_e = d_nu(m2, mc2, C);
  const double nu_u = d_nu(m2u, mc2, C);
  const double nu = nu_e >= nu_u ? nu_e : nu_u;
  const double rt = rho_e - nu * (rho_e - rho_u);
  nu_out[e] = nu;
  rt_out[e] = rt > floor_ ? rt : floor_;
  if (!SENS) return;
  const double drho_e = d_drho(q2e, m, g);
  if (u == e) {
    se[e] = rho_e <= floor_ ? 0.0 : drho_e;
    su[e] = 0.0;
    return;
  }
  if (rt <= floor_) {
    se[e] = 0.0;
    su[e] = 0.0;
    return;
  }
  if (nu == 0.0) {
    se[e] = drho_e;
    su[e] = 0.0;
    return;
  }
  const double drho_u = d_drho(q2u, m, g);
  const double djump = rho_e - rho_u;
  if (nu_e >= nu_u) {
    const double dnu = m2 > mc2 ? d_dnu(q2e, m2, m, mc2, C, g) : 0.0;
    se[e] = drho_e * (1.0 - nu) - djump * dnu;
    su[e] = nu * drho_u;
  } else {
    const double dnu = d_dnu(q2u, m2u, m, mc2, C, g);
    se[e] = drho_e * (1.0 - nu);
    su[e] = nu * drho_u - djump * dnu;
  }
}

__global__ __launch_bounds__(kBlock) void k_upwind_sens_d(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, const double *__restrict__ q2,
    const double *__restrict__ rho, int max_hops, double m, double mc2,
    double C, double g, double floor_, int32_t *__restrict__ up_out,
    double *__restrict__ nu_out, double *__restrict__ rt_out,
    double *__restrict__ se, double *__restrict__ su) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int32_t u = walk_upstream_d(e, fn, dcen, nodes, elements, grad[3 * e],
                                    grad[3 * e + 1], grad[3 * e + 2], max_hops);
  if (up_out) up_out[e] = u;
  sens_of_selection<true>(e, u, q2, rho, m, mc2, C, g, floor_, nu_out, rt_out, se,
                          su);
}

// One walk's state, for the two-walk lanes below.
struct WalkState {
  int32_t cur;
  int hops;
  double reach, target, gx, gy, gz;
  bool active;
};

__device__ __forceinline__ void walk_begin(WalkState &w, int64_t e, bool valid,
                                           const double *__restrict__ nodes,
                                           const int32_t *__restrict__ elements,
                                           const double *__restrict__ grad,
                                           int max_hops) {
  w.cur = (int32_t)e;
  w.hops = 0;
  w.reach = 0.0;
  w.active = false;
  if (!valid) return;
  w.gx = grad[3 * e];
  w.gy = grad[3 * e + 1];
  w.gz = grad[3 * e + 2];
  double pmin = 1e300, pmax = -1e300;
  const int4 nd4 = *reinterpret_cast<const int4 *>(elements + 4 * e);
  const int32_t nds[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int64_t nd = nds[k];
    double p = nodes[3 * nd] * w.gx + nodes[3 * nd + 1] * w.gy + nodes[3 * nd + 2] * w.gz;
    if (p < pmin) pmin = p;
    if (p > pmax) pmax = p;
  }
  w.target = 0.8 * (pmax - pmin);
  w.active = w.hops < max_hops && w.reach < w.target;
}

// One hop from rows already loaded: the same comparisons and updates as
// walk_upstream_d.
__device__ __forceinline__ void walk_hop(WalkState &w, const int4 nb4,
                                         const double2 *dp, int max_hops) {
  const int32_t nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
  const double dx[4] = {dp[0].x, dp[1].y, dp[3].x, dp[4].y};
  const double dy[4] = {dp[0].y, dp[2].x, dp[3].y, dp[5].x};
  const double dz[4] = {dp[1].x, dp[2].y, dp[4].x, dp[5].y};
  int32_t best = -1;
  double best_disp = 0.0;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (nbs[f] >= 0) {
      double d = dx[f] * w.gx + dy[f] * w.gy + dz[f] * w.gz;
      if (d < best_disp) {
        best_disp = d;
        best = nbs[f];
      }
    }
  }
  if (best < 0) {
    w.active = false;
    return;
  }
  w.cur = best;
  w.reach += -best_disp;
  w.hops += 1;
  w.active = w.hops < max_hops && w.reach < w.target;
}

// W elements per lane (e + k blockDim.x, k < W, of the block's W blockDim.x
// range, each coalesced): their walks advance together, every active walk's
// rows loaded before any is used, so W dependent chains of random reads are
// in flight per lane.  Per element the same u(e), nu, rho_tilde and
// (s_e, s_u) as k_upwind_sens_d (SENS; without, the same u, nu and
// rho_tilde as k_upwind_fused_d); W trades chains per lane against
// registers (occupancy), chosen by fp3d_set_walk_width.
template <int W, bool SENS>
__device__ __forceinline__ void k_upwind_sens_dw_body(int64_t bx,
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, const double *__restrict__ q2,
    const double *__restrict__ rho, int max_hops, double m, double mc2,
    double C, double g, double floor_, int32_t *__restrict__ up_out,
    double *__restrict__ nu_out, double *__restrict__ rt_out,
    double *__restrict__ se, double *__restrict__ su) {
  const int64_t e0 = (int64_t)bx * W * blockDim.x + threadIdx.x;
  if (e0 >= n) return;
  WalkState w[W];
#pragma unroll
  for (int k = 0; k < W; ++k)
    walk_begin(w[k], e0 + (int64_t)k * blockDim.x, e0 + (int64_t)k * blockDim.x < n,
               nodes, elements, grad, max_hops);
  bool any = true;
  while (any) {
    int4 nb[W];
    double2 d[W][6];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      if (w[k].active) {
        nb[k] = *reinterpret_cast<const int4 *>(fn + 4 * (int64_t)w[k].cur);
        const double2 *p = reinterpret_cast<const double2 *>(dcen + 12 * (int64_t)w[k].cur);
#pragma unroll
        for (int j = 0; j < 6; ++j) d[k][j] = p[j];
      }
    }
    any = false;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      if (w[k].active) walk_hop(w[k], nb[k], d[k], max_hops);
      any = any || w[k].active;
    }
  }
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int64_t e = e0 + (int64_t)k * blockDim.x;
    if (e < n) {
      if (up_out) up_out[e] = w[k].cur;
      sens_of_selection<SENS>(e, w[k].cur, q2, rho, m, mc2, C, g, floor_, nu_out,
                              rt_out, se, su);
    }
  }
}
template <int W, bool SENS>
__global__ __launch_bounds__(kBlock) void k_upwind_sens_dw(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ elements,
    const double *__restrict__ grad, const double *__restrict__ q2,
    const double *__restrict__ rho, int max_hops, double m, double mc2,
    double C, double g, double floor_, int32_t *__restrict__ up_out,
    double *__restrict__ nu_out, double *__restrict__ rt_out,
    double *__restrict__ se, double *__restrict__ su) {
  k_upwind_sens_dw_body<W, SENS>(blockIdx.x, n, fn, dcen, nodes, elements, grad, q2,
                                 rho, max_hops, m, mc2, C, g, floor_, up_out, nu_out,
                                 rt_out, se, su);
}

// pyfp3d/kernels/upwind.py:596-643
__global__ __launch_bounds__(kBlock) void k_frozen_sens(
    int64_t n, const double *__restrict__ q2, const double *__restrict__ rho,
    const int32_t *__restrict__ up, const int8_t *__restrict__ br, double m,
    double mc2, double C, double g, double *__restrict__ se,
    double *__restrict__ su) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int b = br[e];
  if (b == 3) {
    se[e] = 0.0;
    su[e] = 0.0;
    return;
  }
  double drho_e = d_drho(q2[e], m, g);
  if (b == 0) {
    se[e] = drho_e;
    su[e] = 0.0;
    return;
  }
  int32_t u = up[e];
  double drho_u = d_drho(q2[u], m, g);
  double djump = rho[e] - rho[u];
  if (b == 1) {
    double m2 = d_mach2(q2[e], m, g);
    double nu = d_nu(m2, mc2, C);
    double dnu = m2 > mc2 ? d_dnu(q2[e], m2, m, mc2, C, g) : 0.0;
    se[e] = drho_e * (1.0 - nu) - djump * dnu;
    su[e] = nu * drho_u;
  } else {
    double m2u = d_mach2(q2[u], m, g);
    double nu = d_nu(m2u, mc2, C);
    double dnu = m2u > mc2 ? d_dnu(q2[u], m2u, m, mc2, C, g) : 0.0;
    se[e] = drho_e * (1.0 - nu);
    su[e] = nu * drho_u - djump * dnu;
  }
}

// pyfp3d/kernels/upwind.py:646-730
__global__ __launch_bounds__(kBlock) void k_sens(
    int64_t n, const double *__restrict__ q2, const double *__restrict__ rho,
    const int32_t *__restrict__ up, double m, double mc2, double C, double g,
    double floor_, double *__restrict__ se, double *__restrict__ su) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int32_t u = up[e];
  double rho_e = rho[e];
  double drho_e = d_drho(q2[e], m, g);
  if (u == e) {
    if (rho_e <= floor_) {
      se[e] = 0.0;
      su[e] = 0.0;
    } else {
      se[e] = drho_e;
      su[e] = 0.0;
    }
    return;
  }
  double m2 = d_mach2(q2[e], m, g);
  double m2u = d_mach2(q2[u], m, g);
  double nu_e = d_nu(m2, mc2, C);
  double nu_u = d_nu(m2u, mc2, C);
  double nu = nu_e >= nu_u ? nu_e : nu_u;
  double rho_u = rho[u];
  double rt = rho_e - nu * (rho_e - rho_u);
  if (rt <= floor_) {
    se[e] = 0.0;
    su[e] = 0.0;
    return;
  }
  if (nu == 0.0) {
    se[e] = drho_e;
    su[e] = 0.0;
    return;
  }
  double drho_u = d_drho(q2[u], m, g);
  double djump = rho_e - rho_u;
  if (nu_e >= nu_u) {
    double dnu = m2 > mc2 ? d_dnu(q2[e], m2, m, mc2, C, g) : 0.0;
    se[e] = drho_e * (1.0 - nu) - djump * dnu;
    su[e] = nu * drho_u;
  } else {
    double dnu = d_dnu(q2[u], m2u, m, mc2, C, g);
    se[e] = drho_e * (1.0 - nu);
    su[e] = nu * drho_u - djump * dnu;
  }
}

// ------------------------------------------------- deterministic reductions
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// pass 1: per-block {max nu, count nu>0, count rt==floor} -> work[3*b+k]
__device__ __forceinline__ void k_monitors_p1_body(int64_t bx, int64_t nbx,
    int64_t n, const double *__restrict__ nu, const double *__restrict__ rt,
    double floor_, double *__restrict__ work) {
  __shared__ double s[3][kBlock / 64];
  double vmax = -INFINITY, cpos = 0.0, cflr = 0.0;
  for (int64_t i = (int64_t)bx * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)nbx * blockDim.x) {
    double v = nu[i];
    vmax = v > vmax ? v : vmax;
    cpos += v > 0.0 ? 1.0 : 0.0;
    cflr += rt[i] == floor_ ? 1.0 : 0.0;
  }
  vmax = wave_max(vmax);
  cpos = wave_sum(cpos);
  cflr = wave_sum(cflr);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s[0][w] = vmax;
    s[1][w] = cpos;
    s[2][w] = cflr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = s[0][0], b = s[1][0], c = s[2][0];
    for (int k = 1; k < kBlock / 64; ++k) {
      a = s[0][k] > a ? s[0][k] : a;
      b += s[1][k];
      c += s[2][k];
    }
    work[3 * bx] = a;
    work[3 * bx + 1] = b;
    work[3 * bx + 2] = c;
  }
}
__global__ __launch_bounds__(kBlock) void k_monitors_p1(
    int64_t n, const double *__restrict__ nu, const double *__restrict__ rt,
    double floor_, double *__restrict__ work) {
  k_monitors_p1_body(blockIdx.x, gridDim.x, n, nu, rt, floor_, work);
}
// second pass: 256 lanes each fold a fixed strided subset, then a fixed
// shuffle tree and a fixed-order combine of the 4 wavefront results
__device__ __forceinline__ void k_monitors_p2_body(int64_t bx,
    int nb, const double *__restrict__ work, double *__restrict__ out) {
  __shared__ double s[3][kBlock / 64];
  double a = -INFINITY, b = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    a = work[3 * k] > a ? work[3 * k] : a;
    b += work[3 * k + 1];
    c += work[3 * k + 2];
  }
  a = wave_max(a);
  b = wave_sum(b);
  c = wave_sum(c);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s[0][w] = a;
    s[1][w] = b;
    s[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / 64; ++k) {
      a = s[0][0] > s[0][k] ? s[0][0] : s[0][k];
      s[0][0] = a;
      s[1][0] += s[1][k];
      s[2][0] += s[2][k];
    }
    out[0] = s[0][0];
    out[1] = s[1][0];
    out[2] = s[2][0];
  }
}
__global__ __launch_bounds__(kBlock) void k_monitors_p2(
    int nb, const double *__restrict__ work, double *__restrict__ out) {
  k_monitors_p2_body(blockIdx.x, nb, work, out);
}
__device__ __forceinline__ void k_count_zero_p1_body(int64_t bx, int64_t nbx,
    int64_t n, const uint8_t *__restrict__ m, double *__restrict__ work) {
  __shared__ double s[kBlock / 64];
  double c = 0.0;
  for (int64_t i = (int64_t)bx * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)nbx * blockDim.x)
    c += m[i] == 0 ? 1.0 : 0.0;
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) t += s[k];
    work[bx] = t;
  }
}
__global__ __launch_bounds__(kBlock) void k_count_zero_p1(
    int64_t n, const uint8_t *__restrict__ m, double *__restrict__ work) {
  k_count_zero_p1_body(blockIdx.x, gridDim.x, n, m, work);
}
__device__ __forceinline__ void k_sum_p2_body(int64_t bx,
    int nb, const double *__restrict__ work, double *__restrict__ out) {
  __shared__ double s[kBlock / 64];
  double t = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) t += work[k];
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) r += s[k];
    out[0] = r;
  }
}
__global__ __launch_bounds__(kBlock) void k_sum_p2(
    int nb, const double *__restrict__ work, double *__restrict__ out) {
  k_sum_p2_body(blockIdx.x, nb, work, out);
}

inline int reduce_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b > kReduceBlocks) b = kReduceBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

// ------------------------------------------------------ coloured assembly
// pyfp3d/kernels/residual.py:64-76 (one launch per colour class)
__global__ __launch_bounds__(kBlock) void k_residual_color(
    int64_t k0, int64_t k1, const int32_t *__restrict__ eidx,
    const int32_t *__restrict__ elements, const double *__restrict__ B,
    const double *__restrict__ V, const double *__restrict__ rt,
    const double *__restrict__ grad, double *__restrict__ R) {
  int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= k1) return;
  int64_t e = eidx ? (int64_t)eidx[k] : k;
  double w = rt[e] * V[e];
  double gx = grad[3 * e], gy = grad[3 * e + 1], gz = grad[3 * e + 2];
  const int4 nd = *reinterpret_cast<const int4 *>(elements + 4 * e);
  const double *b = B + 12 * e;
  // the 4 nodes of one element are distinct: issue all loads, then stores
  double r0 = R[nd.x], r1 = R[nd.y], r2 = R[nd.z], r3 = R[nd.w];
  r0 += w * (gx * b[0] + gy * b[1] + gz * b[2]);
  r1 += w * (gx * b[3] + gy * b[4] + gz * b[5]);
  r2 += w * (gx * b[6] + gy * b[7] + gz * b[8]);
  r3 += w * (gx * b[9] + gy * b[10] + gz * b[11]);
  R[nd.x] = r0;
  R[nd.y] = r1;
  R[nd.z] = r2;
  R[nd.w] = r3;
}

// pyfp3d/kernels/jacobian.py:117-129
__global__ __launch_bounds__(kBlock) void k_matrix_color(
    int64_t k0, int64_t k1, const int32_t *__restrict__ eidx,
    const int32_t *__restrict__ e2c, const double *__restrict__ B,
    const double *__restrict__ V, const double *__restrict__ rt,
    double *__restrict__ data) {
  int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= k1) return;
  int64_t e = eidx ? (int64_t)eidx[k] : k;
  double w = rt[e] * V[e];
  double b[12];
#pragma unroll
  for (int t = 0; t < 12; ++t) b[t] = B[12 * e + t];
  // the 16 CSR slots of one element are distinct: load all, add, store all
  int32_t q[16];
  const int4 *s4 = reinterpret_cast<const int4 *>(e2c + 16 * e);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int4 v = s4[t];
    q[4 * t] = v.x; q[4 * t + 1] = v.y; q[4 * t + 2] = v.z; q[4 * t + 3] = v.w;
  }
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = data[q[t]];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double bi0 = b[3 * i], bi1 = b[3 * i + 1], bi2 = b[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[4 * i + j] +=
          w * (bi0 * b[3 * j] + bi1 * b[3 * j + 1] + bi2 * b[3 * j + 2]);
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) data[q[t]] = acc[t];
}

// pyfp3d/kernels/jacobian.py:146-180 (Term 1 kept as its own expression so
// J == A bitwise where s_e == 0)
__global__ __launch_bounds__(kBlock) void k_jacobian_color(
    int64_t k0, int64_t k1, const int32_t *__restrict__ eidx,
    const int32_t *__restrict__ e2c, const double *__restrict__ B,
    const double *__restrict__ V, const double *__restrict__ rt,
    const double *__restrict__ grad, const double *__restrict__ s_e,
    const uint8_t *__restrict__ lim, double inv_u2, double clamp,
    double *__restrict__ data) {
  int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= k1) return;
  int64_t e = eidx ? (int64_t)eidx[k] : k;
  double w1 = rt[e] * V[e];
  double g0 = grad[3 * e], g1 = grad[3 * e + 1], g2 = grad[3 * e + 2];
  double se = s_e[e];
  if (lim) se = se * (lim[e] ? 1.0 : 0.0);
  double w2 = 2.0 * inv_u2 * se * V[e];
  if (clamp > 0.0) {
    // SPD preconditioner variant: the streamwise softening never removes
    // more than a fraction `clamp` of the element's stiffness along g
    double q2 = g0 * g0 + g1 * g1 + g2 * g2;
    double lo = q2 > 0.0 ? -clamp * w1 / q2 : 0.0;
    w2 = w2 > lo ? w2 : lo;
  }
  double b[12];
#pragma unroll
  for (int t = 0; t < 12; ++t) b[t] = B[12 * e + t];
  int32_t q[16];
  const int4 *s4 = reinterpret_cast<const int4 *>(e2c + 16 * e);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int4 v = s4[t];
    q[4 * t] = v.x; q[4 * t + 1] = v.y; q[4 * t + 2] = v.z; q[4 * t + 3] = v.w;
  }
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = data[q[t]];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double bi0 = b[3 * i], bi1 = b[3 * i + 1], bi2 = b[3 * i + 2];
    double gi = g0 * bi0 + g1 * bi1 + g2 * bi2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double bj0 = b[3 * j], bj1 = b[3 * j + 1], bj2 = b[3 * j + 2];
      double val = w1 * (bi0 * bj0 + bi1 * bj1 + bi2 * bj2);
      if (w2 != 0.0) val += w2 * gi * (g0 * bj0 + g1 * bj1 + g2 * bj2);
      acc[4 * i + j] += val;
    }
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) data[q[t]] = acc[t];
}

// ---------------------------------------------------- node-gather assembly
// The same sums as the coloured scatters, computed row by row: node v
// gathers the contributions of its incident (element, local index) entries
// in ascending element order -- the accumulation order of the reference's
// serial assembly (pyfp3d/kernels/residual.py:144-189,
// pyfp3d/kernels/residual.py:106-129 + COO->CSR) -- with no colour passes,
// no atomics, and each output written exactly once.
// Blocks are mapped to XCDs by xcd_block().

// Two stages: an element pass writes each element's contributions once,
// coalesced (4 residual terms, or the 16 matrix values row-major [4a + b]);
// the node pass then reads one short record per incidence entry instead of
// the element's geometry and state (reading those per entry costs about
// 1.6x more on the 4.7M-tet channel).
__device__ __forceinline__ void k_elem_residual_terms_body(int64_t bx,
    int64_t n, const double *__restrict__ B, const double *__restrict__ V,
    const double *__restrict__ rt, const double *__restrict__ grad,
    const double *__restrict__ gdotb, const int32_t *__restrict__ pos,
    double *__restrict__ c) {
  int64_t e = (int64_t)bx * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double w = rt[e] * V[e];
  double4 o;
  if (gdotb) {
    const double4 gb = *reinterpret_cast<const double4 *>(gdotb + 4 * e);
    o.x = w * gb.x;
    o.y = w * gb.y;
    o.z = w * gb.z;
    o.w = w * gb.w;
  } else {
    double gx = grad[3 * e], gy = grad[3 * e + 1], gz = grad[3 * e + 2];
    const double *b = B + 12 * e;
    o.x = w * (gx * b[0] + gy * b[1] + gz * b[2]);
    o.y = w * (gx * b[3] + gy * b[4] + gz * b[5]);
    o.z = w * (gx * b[6] + gy * b[7] + gz * b[8]);
    o.w = w * (gx * b[9] + gy * b[10] + gz * b[11]);
  }
  if (pos) {
    // incidence order: term a of e goes to the slot of (e, a) in its node's
    // incidence list, so the node pass reads its terms contiguously
    const int4 p = *reinterpret_cast<const int4 *>(pos + 4 * e);
    c[p.x] = o.x;
    c[p.y] = o.y;
    c[p.z] = o.z;
    c[p.w] = o.w;
  } else {
    *reinterpret_cast<double4 *>(c + 4 * e) = o;
  }
}
__global__ __launch_bounds__(kBlock) void k_elem_residual_terms(
    int64_t n, const double *__restrict__ B, const double *__restrict__ V,
    const double *__restrict__ rt, const double *__restrict__ grad,
    const double *__restrict__ gdotb, const int32_t *__restrict__ pos,
    double *__restrict__ c) {
  k_elem_residual_terms_body(blockIdx.x, n, B, V, rt, grad, gdotb, pos, c);
}

// The loads of kUnroll consecutive entries are issued together (they are
// independent), then added one by one in entry order, so the sum keeps the
// ascending-element order while kUnroll gathers are in flight per lane.
constexpr int kUnroll = 8;

__global__ __launch_bounds__(kBlock) void k_gather_terms(
    int64_t nn, int64_t nb_pad, const int32_t *__restrict__ iptr,
    const int32_t *__restrict__ ient, const double *__restrict__ c,
    double *__restrict__ R) {
  int64_t v = xcd_block(blockIdx.x, nb_pad) * kBlock + threadIdx.x;
  if (v >= nn) return;
  double r = 0.0;
  const int32_t q1 = iptr[v + 1];
  for (int32_t q = iptr[v]; q < q1; q += kUnroll) {
    double t[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k)
      t[k] = (q + k < q1) ? c[ient ? ient[q + k] : q + k] : 0.0;
#pragma unroll
    for (int k = 0; k < kUnroll; ++k)
      if (q + k < q1) r += t[k];
  }
  R[v] = r;
}

// Incidence-ordered terms: the terms of a block's kBlock nodes are one
// contiguous range [iptr[v0], iptr[v1]).  The block streams it through LDS
// in coalesced chunks (the next chunk's loads are issued before the current
// one is summed), and each lane adds its own node's terms in entry order:
// the same sums, in the same order, as k_gather_terms.
constexpr int kTermChunk = 2048;               // doubles per chunk (16 KiB)
__device__ __forceinline__ void k_gather_terms_staged_body(int64_t bx,
    int64_t nn, int64_t nb_pad, const int32_t *__restrict__ iptr,
    const double *__restrict__ c, double *__restrict__ R) {
  __shared__ double buf[kTermChunk];
  constexpr int P = kTermChunk / kBlock;
  const int64_t v0 = xcd_block(bx, nb_pad) * kBlock;
  if (v0 >= nn) return;                        // uniform over the block
  const int64_t v1 = v0 + kBlock < nn ? v0 + kBlock : nn;
  const int64_t v = v0 + threadIdx.x;
  const int64_t qb = iptr[v0], qe = iptr[v1];
  int64_t q = 0, q1 = 0;
  if (v < v1) {
    q = iptr[v];
    q1 = iptr[v + 1];
  }
  double nxt[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t i = qb + threadIdx.x + k * kBlock;
    nxt[k] = i < qe ? c[i] : 0.0;
  }
  double r = 0.0;
  for (int64_t c0 = qb; c0 < qe; c0 += kTermChunk) {
    __syncthreads();                           // previous chunk consumed
#pragma unroll
    for (int k = 0; k < P; ++k) buf[threadIdx.x + k * kBlock] = nxt[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int64_t i = c0 + kTermChunk + threadIdx.x + k * kBlock;
      nxt[k] = i < qe ? c[i] : 0.0;
    }
    const int64_t hi = q1 < c0 + kTermChunk ? q1 : c0 + kTermChunk;
    for (; q < hi; ++q) r += buf[q - c0];
  }
  if (v < v1) R[v] = r;
}
__global__ __launch_bounds__(kBlock) void k_gather_terms_staged(
    int64_t nn, int64_t nb_pad, const int32_t *__restrict__ iptr,
    const double *__restrict__ c, double *__restrict__ R) {
  k_gather_terms_staged_body(blockIdx.x, nn, nb_pad, iptr, c, R);
}

__global__ __launch_bounds__(kBlock) void k_elem_matrix_terms(
    int64_t n, const double *__restrict__ B, const double *__restrict__ V,
    const double *__restrict__ rt, const double *__restrict__ grad,
    const double *__restrict__ s_e, const uint8_t *__restrict__ lim,
    double inv_u2, double clamp, const int32_t *__restrict__ pos,
    double *__restrict__ K) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double w1 = rt[e] * V[e];
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, w2 = 0.0;
  if (s_e) {
    g0 = grad[3 * e];
    g1 = grad[3 * e + 1];
    g2 = grad[3 * e + 2];
    double se = s_e[e];
    if (lim) se = se * (lim[e] ? 1.0 : 0.0);
    w2 = 2.0 * inv_u2 * se * V[e];
    if (clamp > 0.0) {
      double q2 = g0 * g0 + g1 * g1 + g2 * g2;
      double lo = q2 > 0.0 ? -clamp * w1 / q2 : 0.0;
      w2 = w2 > lo ? w2 : lo;
    }
  }
  double b[12];
#pragma unroll
  for (int t = 0; t < 12; ++t) b[t] = B[12 * e + t];
  double k[16];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double bi0 = b[3 * i], bi1 = b[3 * i + 1], bi2 = b[3 * i + 2];
    double gi = g0 * bi0 + g1 * bi1 + g2 * bi2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double bj0 = b[3 * j], bj1 = b[3 * j + 1], bj2 = b[3 * j + 2];
      double val = w1 * (bi0 * bj0 + bi1 * bj1 + bi2 * bj2);
      if (w2 != 0.0) val += w2 * gi * (g0 * bj0 + g1 * bj1 + g2 * bj2);
      k[4 * i + j] = val;
    }
  }
  if (pos) {
    // incidence order: row a of e goes to the slot of (e, a)
    const int4 p = *reinterpret_cast<const int4 *>(pos + 4 * e);
    const int32_t pa[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
      *reinterpret_cast<double4 *>(K + 4 * (int64_t)pa[i]) =
          make_double4(k[4 * i], k[4 * i + 1], k[4 * i + 2], k[4 * i + 3]);
  } else {
    double4 *o = reinterpret_cast<double4 *>(K + 16 * e);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      o[i] = make_double4(k[4 * i], k[4 * i + 1], k[4 * i + 2], k[4 * i + 3]);
  }
}

// rel8[4q + b] = elem_to_csr[e][a][b] - indptr[elements[e][a]] for the
// incidence entry q = inc_ent index of (e, a): the slot's offset inside its
// row (rows are at most 256 long), one 32-bit word per incidence, stored in
// incidence order so each row reads its words contiguously.
__global__ __launch_bounds__(kGatherRows) void k_gather_csr(
    int64_t nn, int64_t nb_pad, int max_seg, const int32_t *__restrict__ indptr,
    const int32_t *__restrict__ iptr, const int32_t *__restrict__ ient,
    const uint32_t *__restrict__ rel8, const double *__restrict__ K,
    double *__restrict__ data) {
  extern __shared__ double seg[];
  const int64_t r0 = xcd_block(blockIdx.x, nb_pad) * kGatherRows;
  if (r0 >= nn) return;
  const int64_t r1 = r0 + kGatherRows < nn ? r0 + kGatherRows : nn;
  const int32_t base = indptr[r0];
  const int32_t len = indptr[r1] - base;
  if (len > max_seg) return;
  for (int t = threadIdx.x; t < len; t += kGatherRows) seg[t] = 0.0;
  __syncthreads();
  const int64_t v = r0 + threadIdx.x;
  if (v < r1) {
    constexpr int U = 4;
    const int32_t rb = indptr[v] - base;
    const int32_t q1 = iptr[v + 1];
    for (int32_t q = iptr[v]; q < q1; q += U) {
      uint32_t of[U];
      double4 kv[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (q + k < q1) {
          // row a of element e, ient[q] = 4 e + a (or the row's own slot q
          // when the element pass wrote in incidence order)
          const int64_t ent = ient ? (int64_t)ient[q + k] : (int64_t)(q + k);
          of[k] = rel8[q + k];
          kv[k] = *reinterpret_cast<const double4 *>(K + 4 * ent);
        }
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (q + k < q1) {
          // four distinct slots (one element's four nodes): read, then write
          const int o0 = rb + (int)(of[k] & 0xffu);
          const int o1 = rb + (int)((of[k] >> 8) & 0xffu);
          const int o2 = rb + (int)((of[k] >> 16) & 0xffu);
          const int o3 = rb + (int)(of[k] >> 24);
          const double s0 = seg[o0], s1 = seg[o1], s2 = seg[o2], s3 = seg[o3];
          seg[o0] = s0 + kv[k].x;
          seg[o1] = s1 + kv[k].y;
          seg[o2] = s2 + kv[k].z;
          seg[o3] = s3 + kv[k].w;
        }
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < len; t += kGatherRows) data[base + t] = seg[t];
}

// Incidence-ordered matrix rows: the rows and slot offsets of a block's
// kGatherRows CSR rows are one contiguous range [iptr[r0], iptr[r1]),
// streamed through LDS in coalesced chunks of kRowChunk entries (the next
// chunk in flight while the current one is added; 1024 by default, 256,
// 512 or 2048 by fp3d_set_gather_chunk).  Four lanes per CSR row: lane c
// adds component c of each entry (the value for the element's c-th node)
// into that column's slot of the LDS-staged CSR block.  The four slots of
// one entry are distinct, and the four lanes of a row sit in one wavefront
// and step through the row's entries together (one LDS read and one LDS
// write per entry, issued in program order, which the LDS executes in
// order for a wavefront), so every slot still receives its additions in
// entry order -- the sums of k_gather_csr, bit for bit.  The chunk size
// changes only the staging, never the sums or their order.
template <int kRowChunk>
__global__ __launch_bounds__(kGatherLanes) void k_gather_csr_staged(
    int64_t nn, int64_t nb_pad, int max_seg, const int32_t *__restrict__ indptr,
    const int32_t *__restrict__ iptr, const uint32_t *__restrict__ rel8,
    const double4 *__restrict__ K4, double *__restrict__ data) {
  extern __shared__ double seg[];
  __shared__ double4 kb[kRowChunk];
  __shared__ uint32_t ob[kRowChunk];
  constexpr int NT = kGatherLanes;
  constexpr int P = kRowChunk / NT;
  static_assert(P >= 1 && P * NT == kRowChunk, "chunk must be a multiple of the lanes");
  static_assert(NT == 4 * kGatherRows, "four lanes per row");
  const int64_t r0 = xcd_block(blockIdx.x, nb_pad) * kGatherRows;
  if (r0 >= nn) return;                        // uniform over the block
  const int64_t r1 = r0 + kGatherRows < nn ? r0 + kGatherRows : nn;
  const int32_t base = indptr[r0];
  const int32_t len = indptr[r1] - base;
  if (len > max_seg) return;
  for (int t = threadIdx.x; t < len; t += NT) seg[t] = 0.0;
  const int c = threadIdx.x & 3;
  const int64_t v = r0 + (threadIdx.x >> 2);
  const int64_t qb = iptr[r0], qe = iptr[r1];
  int64_t q = 0, q1 = 0;
  int32_t rb = 0;
  if (v < r1) {
    rb = indptr[v] - base;
    q = iptr[v];
    q1 = iptr[v + 1];
  }
  double4 nk[P];
  uint32_t no[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t i = qb + threadIdx.x + k * NT;
    if (i < qe) {
      nk[k] = K4[i];
      no[k] = rel8[i];
    } else {
      nk[k] = make_double4(0.0, 0.0, 0.0, 0.0);
      no[k] = 0u;
    }
  }
  const double *kbc = reinterpret_cast<const double *>(kb) + c;
  const int sh = 8 * c;
  for (int64_t c0 = qb; c0 < qe; c0 += kRowChunk) {
    __syncthreads();                 // previous chunk consumed, seg zeroed
#pragma unroll
    for (int k = 0; k < P; ++k) {
      kb[threadIdx.x + k * NT] = nk[k];
      ob[threadIdx.x + k * NT] = no[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int64_t i = c0 + kRowChunk + threadIdx.x + k * NT;
      if (i < qe) {
        nk[k] = K4[i];
        no[k] = rel8[i];
      }
    }
    const int64_t hi = q1 < c0 + kRowChunk ? q1 : c0 + kRowChunk;
    for (; q < hi; ++q) {
      const int j = (int)(q - c0);
      const double kv = kbc[4 * j];
      const int o = rb + (int)((ob[j] >> sh) & 0xffu);
      seg[o] = seg[o] + kv;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < len; t += NT) data[base + t] = seg[t];
}

// pyfp3d/kernels/jacobian.py:197-228 (fixed slot slice [16a, 16a+16))
__global__ __launch_bounds__(kBlock) void k_term3_coo(
    int64_t n_active, const int32_t *__restrict__ active,
    const int32_t *__restrict__ elements, const double *__restrict__ B,
    const double *__restrict__ V, const double *__restrict__ grad,
    const int32_t *__restrict__ up, const double *__restrict__ s_u,
    double inv_u2, int32_t *__restrict__ rows, int32_t *__restrict__ cols,
    double *__restrict__ vals) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_active) return;
  int64_t e = active[a];
  int64_t u = up[e];
  double ge0 = grad[3 * e], ge1 = grad[3 * e + 1], ge2 = grad[3 * e + 2];
  double gu0 = grad[3 * u], gu1 = grad[3 * u + 1], gu2 = grad[3 * u + 2];
  double w = 2.0 * inv_u2 * s_u[e] * V[e];
  int64_t base = 16 * a;
  for (int i = 0; i < 4; ++i) {
    double gi = ge0 * B[12 * e + 3 * i] + ge1 * B[12 * e + 3 * i + 1] +
                ge2 * B[12 * e + 3 * i + 2];
    int32_t ri = elements[4 * e + i];
    double wi = w * gi;
    for (int k = 0; k < 4; ++k) {
      int64_t idx = base + 4 * i + k;
      rows[idx] = ri;
      cols[idx] = elements[4 * u + k];
      vals[idx] = wi * (gu0 * B[12 * u + 3 * k] + gu1 * B[12 * u + 3 * k + 1] +
                        gu2 * B[12 * u + 3 * k + 2]);
    }
  }
}

__global__ void k_fill(int64_t n, double v, double *x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// ------------------------------------------ batched residual evaluation
// The live residual evaluation of nb solves on one mesh, one launch per
// kernel for all of them (blockIdx.y = solve, its buffers from an EvalJob):
// every block runs the single-solve kernel's body on the same grid, so each
// solve gets the bits of its own evaluation.
__global__ __launch_bounds__(kBlock) void k_be_state(int64_t n,
                                                    const int32_t *__restrict__ el,
                                                    const double *__restrict__ B, double u2,
                                                    double g,
                                                    const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_element_state_body(blockIdx.x, n, el, B, j.phi, u2, j.m_inf, g, j.q2cap, j.grad, j.q2n,
                       j.q2l, j.rho, j.lim, j.gdotb);
}
__global__ __launch_bounds__(kBlock) void k_be_count_p1(int64_t n,
                                                       const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_count_zero_p1_body(blockIdx.x, gridDim.x, n, j.lim, j.work);
}
__global__ __launch_bounds__(kBlock) void k_be_count_p2(int nb,
                                                       const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_sum_p2_body(0, nb, j.work, j.out);
}
template <int W>
__global__ __launch_bounds__(kBlock) void k_be_walk(
    int64_t n, const int32_t *__restrict__ fn, const double *__restrict__ dcen,
    const double *__restrict__ nodes, const int32_t *__restrict__ el, int max_hops,
    double mc2, double C, double g, double floor_, const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_upwind_sens_dw_body<W, true>(blockIdx.x, n, fn, dcen, nodes, el, j.grad, j.q2l, j.rho,
                                 max_hops, j.m_inf, mc2, C, g, floor_, j.up, j.nu, j.rt,
                                 j.se, j.su);
}
__global__ __launch_bounds__(kBlock) void k_be_mon_p1(int64_t n, double floor_,
                                                     const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_monitors_p1_body(blockIdx.x, gridDim.x, n, j.nu, j.rt, floor_, j.work);
}
__global__ __launch_bounds__(kBlock) void k_be_mon_p2(int nb,
                                                     const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_monitors_p2_body(0, nb, j.work, j.out + 1);
}
__global__ __launch_bounds__(kBlock) void k_be_terms(int64_t n, const double *__restrict__ B,
                                                    const double *__restrict__ V,
                                                    const int32_t *__restrict__ pos,
                                                    const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_elem_residual_terms_body(blockIdx.x, n, B, V, j.rt, j.grad, j.gdotb, pos, j.terms);
}
__global__ __launch_bounds__(kBlock) void k_be_gather(int64_t nn, int64_t nb_pad,
                                                     const int32_t *__restrict__ iptr,
                                                     const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  k_gather_terms_staged_body(blockIdx.x, nn, nb_pad, iptr, j.terms, j.R);
}

}  // namespace fp3d

using namespace fp3d;

// ======================================================================
// C-ABI
// ======================================================================
extern "C" {

const char *fp3d_last_error(void) { return fp3d::g_err.c_str(); }
int fp3d_version(void) { return 1; }

int fp3d_graph_launch(void *graph_exec, void *stream) {
  FP3D_REQUIRE(graph_exec, "graph_launch: NULL graph");
  FP3D_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, FP3D_STREAM(stream)));
  return FP3D_OK;
}

int fp3d_device_count(int *count) {
  FP3D_REQUIRE(count, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return FP3D_OK;
}

int fp3d_stream_sync(void *stream) {
  FP3D_HIP(hipStreamSynchronize(FP3D_STREAM(stream)));
  return FP3D_OK;
}

int fp3d_reduce_work_doubles(void) { return 3 * kReduceBlocks; }

// Threads per block of the element-wise kernels (every one of them indexes
// with blockDim.x and uses no LDS): 64, 128, 192 or 256, process-wide, set
// by the autotuner through fp3d_set_element_block.
static int g_elem_block = kBlock;
// rows per staged chunk of the incidence-ordered CSR gather (autotuned)
static int g_gather_chunk = 1024;
static int g_walk_width = 2;

// The walk sweep at the process-wide width (1: one walk per lane, the
// single-walk kernels, which need the fewest registers).
extern "C++" template <bool SENS>
void launch_walk_width(int64_t n, const int32_t *fn, const double *dcen,
                       const double *nodes, const int32_t *elements,
                       const double *grad, const double *q2, const double *rho,
                       int max_hops, double m, double mc2, double C, double g,
                       double floor_, int32_t *up, double *nu, double *rt,
                       double *se, double *su, hipStream_t st) {
#define FP3D_WALKW(W)                                                          \
  hipLaunchKernelGGL((k_upwind_sens_dw<W, SENS>), dim3(grid_for(n, W * g_elem_block)), \
                     dim3(g_elem_block), 0, st, n, fn, dcen, nodes, elements,    \
                     grad, q2, rho, max_hops, m, mc2, C, g, floor_, up, nu, rt,  \
                     se, su)
  if (g_walk_width == 1) {
    if (SENS)
      hipLaunchKernelGGL(k_upwind_sens_d, dim3(grid_for(n, g_elem_block)),
                         dim3(g_elem_block), 0, st, n, fn, dcen, nodes, elements,
                         grad, q2, rho, max_hops, m, mc2, C, g, floor_, up, nu, rt,
                         se, su);
    else
      hipLaunchKernelGGL(k_upwind_fused_d, dim3(grid_for(n, g_elem_block)),
                         dim3(g_elem_block), 0, st, n, fn, dcen, nodes, elements,
                         grad, q2, rho, max_hops, m, mc2, C, g, floor_, up, nu, rt);
  } else if (g_walk_width == 3) {
    FP3D_WALKW(3);
  } else if (g_walk_width == 4) {
    FP3D_WALKW(4);
  } else {
    FP3D_WALKW(2);
  }
#undef FP3D_WALKW
}

#define ELEMWISE(kernel, n, ...)                                       \
  do {                                                                 \
    FP3D_REQUIRE(n >= 0, #kernel ": negative size");                   \
    if (n == 0) return FP3D_OK;                                        \
    hipLaunchKernelGGL(kernel, dim3(grid_for(n, g_elem_block)),        \
                       dim3(g_elem_block), 0,                          \
                       FP3D_STREAM(stream), n, __VA_ARGS__);           \
    return check_launch(#kernel);                                      \
  } while (0)

int fp3d_density_field(int64_t n, const double *q2, double m_inf, double gamma,
                       double *out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && out), "density_field: NULL pointer");
  ELEMWISE(k_density_field, n, q2, m_inf, gamma, out);
}
int fp3d_mach_squared_field(int64_t n, const double *q2, double m_inf,
                            double gamma, double *out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && out), "mach_squared_field: NULL pointer");
  ELEMWISE(k_mach2_field, n, q2, m_inf, gamma, out);
}
int fp3d_limit_q2_field(int64_t n, const double *q2, double q2cap, double *out,
                        void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && out), "limit_q2_field: NULL pointer");
  ELEMWISE(k_limit_q2, n, q2, q2cap, out);
}
int fp3d_density_derivative_field(int64_t n, const double *q2, double m_inf,
                                  double gamma, double *out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && out), "density_derivative_field: NULL");
  ELEMWISE(k_drho_field, n, q2, m_inf, gamma, out);
}

int fp3d_element_velocity_q2(int64_t n_tets, const int32_t *elements,
                             const double *B, const double *phi,
                             double *grad_out, double *q2_out, void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (elements && B && phi && grad_out && q2_out),
               "element_velocity_q2: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements), "elements must be 16-byte aligned");
  ELEMWISE(k_velocity_q2, n_tets, elements, B, phi, grad_out, q2_out);
}

int fp3d_element_state(int64_t n_tets, const int32_t *elements, const double *B,
                       const double *phi, double u2, double m_inf, double gamma,
                       double q2cap, double *grad_out, double *q2n_out,
                       double *q2l_out, double *rho_out, uint8_t *lim_out,
                       double *gdotb_out, void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (elements && B && phi && grad_out && q2n_out &&
                               q2l_out && rho_out),
               "element_state: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements), "elements must be 16-byte aligned");
  FP3D_REQUIRE(((uintptr_t)gdotb_out & 31) == 0,
               "gdotb must be 32-byte aligned");
  FP3D_REQUIRE(u2 > 0.0, "element_state: u_inf^2 must be > 0");
  ELEMWISE(k_element_state, n_tets, elements, B, phi, u2, m_inf, gamma, q2cap,
           grad_out, q2n_out, q2l_out, rho_out, lim_out, gdotb_out);
}

int fp3d_upstream_elements(int64_t n_tets, const int32_t *face_neighbors,
                           const double *centroids, const double *nodes,
                           const int32_t *elements, const double *grad,
                           int max_hops, int32_t *out, void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (face_neighbors && centroids && nodes &&
                               elements && grad && out),
               "upstream_elements: NULL pointer");
  FP3D_REQUIRE(max_hops >= 0, "upstream_elements: max_hops < 0");
  ELEMWISE(k_upstream, n_tets, face_neighbors, centroids, nodes, elements, grad,
           max_hops, out);
}

int fp3d_upstream_elements_d(int64_t n_tets, const int32_t *face_neighbors,
                             const double *face_disp, const double *nodes,
                             const int32_t *elements, const double *grad,
                             int max_hops, int32_t *out, void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (face_neighbors && face_disp && nodes &&
                               elements && grad && out),
               "upstream_elements_d: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements) && FP3D_ALIGNED16(face_neighbors) &&
                   FP3D_ALIGNED16(face_disp),
               "upstream_elements_d: arrays must be 16-byte aligned");
  ELEMWISE(k_upstream_d, n_tets, face_neighbors, face_disp, nodes, elements,
           grad, max_hops, out);
}

int fp3d_upwind_rho_tilde_d(int64_t n_tets, const int32_t *face_neighbors,
                            const double *face_disp, const double *nodes,
                            const int32_t *elements, const double *grad,
                            const double *q2, const double *rho, int max_hops,
                            double m_inf, double m_crit, double C, double gamma,
                            double rho_floor, int32_t *upstream_out,
                            double *nu_out, double *rho_tilde_out,
                            void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (face_neighbors && face_disp && nodes &&
                               elements && grad && q2 && rho && nu_out &&
                               rho_tilde_out),
               "upwind_rho_tilde_d: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements) && FP3D_ALIGNED16(face_neighbors) &&
                   FP3D_ALIGNED16(face_disp),
               "upwind_rho_tilde_d: arrays must be 16-byte aligned");
  FP3D_REQUIRE(n_tets >= 0, "upwind_rho_tilde_d: n_tets < 0");
  if (n_tets == 0) return FP3D_OK;
  launch_walk_width<false>(n_tets, face_neighbors, face_disp, nodes, elements, grad, q2,
                           rho, max_hops, m_inf, m_crit * m_crit, C, gamma, rho_floor,
                           upstream_out, nu_out, rho_tilde_out, nullptr, nullptr,
                           FP3D_STREAM(stream));
  return check_launch("upwind_rho_tilde_d");
}

int fp3d_upwind_rho_tilde_sens_d(
    int64_t n_tets, const int32_t *face_neighbors, const double *face_disp,
    const double *nodes, const int32_t *elements, const double *grad,
    const double *q2, const double *rho, int max_hops, double m_inf,
    double m_crit, double C, double gamma, double rho_floor,
    int32_t *upstream_out, double *nu_out, double *rho_tilde_out,
    double *s_e_out, double *s_u_out, void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (face_neighbors && face_disp && nodes &&
                               elements && grad && q2 && rho && nu_out &&
                               rho_tilde_out && s_e_out && s_u_out),
               "upwind_rho_tilde_sens_d: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements) && FP3D_ALIGNED16(face_neighbors) &&
                   FP3D_ALIGNED16(face_disp),
               "upwind_rho_tilde_sens_d: arrays must be 16-byte aligned");
  FP3D_REQUIRE(n_tets >= 0, "upwind_rho_tilde_sens_d: n_tets < 0");
  if (n_tets == 0) return FP3D_OK;
  launch_walk_width<true>(n_tets, face_neighbors, face_disp, nodes, elements, grad, q2,
                          rho, max_hops, m_inf, m_crit * m_crit, C, gamma, rho_floor,
                          upstream_out, nu_out, rho_tilde_out, s_e_out, s_u_out,
                          FP3D_STREAM(stream));
  return check_launch("upwind_rho_tilde_sens_d");
}

int fp3d_rho_tilde_sweep(int64_t n, const double *q2, const double *rho,
                         const int32_t *upstream, double m_inf, double m_crit,
                         double C, double gamma, double rho_floor,
                         double *nu_out, double *rho_tilde_out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && rho && upstream && nu_out && rho_tilde_out),
               "rho_tilde_sweep: NULL pointer");
  ELEMWISE(k_rho_tilde, n, q2, rho, upstream, m_inf, m_crit * m_crit, C, gamma,
           rho_floor, nu_out, rho_tilde_out);
}

int fp3d_upwind_rho_tilde(int64_t n_tets, const int32_t *face_neighbors,
                          const double *centroids, const double *nodes,
                          const int32_t *elements, const double *grad,
                          const double *q2, const double *rho, int max_hops,
                          double m_inf, double m_crit, double C, double gamma,
                          double rho_floor, int32_t *upstream_out,
                          double *nu_out, double *rho_tilde_out, void *stream) {
  FP3D_REQUIRE(n_tets == 0 || (face_neighbors && centroids && nodes &&
                               elements && grad && q2 && rho && nu_out &&
                               rho_tilde_out),
               "upwind_rho_tilde: NULL pointer");
  ELEMWISE(k_upwind_fused, n_tets, face_neighbors, centroids, nodes, elements,
           grad, q2, rho, max_hops, m_inf, m_crit * m_crit, C, gamma,
           rho_floor, upstream_out, nu_out, rho_tilde_out);
}

int fp3d_classify_upwind_branches(int64_t n, const double *q2, const double *rho,
                                  const int32_t *upstream, double m_inf,
                                  double m_crit, double C, double gamma,
                                  double rho_floor, int8_t *branch_out,
                                  void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && rho && upstream && branch_out),
               "classify_upwind_branches: NULL pointer");
  ELEMWISE(k_classify, n, q2, rho, upstream, m_inf, m_crit * m_crit, C, gamma,
           rho_floor, branch_out);
}

int fp3d_rho_tilde_frozen_sweep(int64_t n, const double *q2, const double *rho,
                                const int32_t *upstream, const int8_t *branch,
                                double m_inf, double m_crit, double C,
                                double gamma, double rho_floor, double *nu_out,
                                double *rho_tilde_out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && rho && upstream && branch && nu_out &&
                          rho_tilde_out),
               "rho_tilde_frozen_sweep: NULL pointer");
  ELEMWISE(k_frozen, n, q2, rho, upstream, branch, m_inf, m_crit * m_crit, C,
           gamma, rho_floor, nu_out, rho_tilde_out);
}

int fp3d_rho_tilde_frozen_sensitivities(int64_t n, const double *q2,
                                        const double *rho,
                                        const int32_t *upstream,
                                        const int8_t *branch, double m_inf,
                                        double m_crit, double C, double gamma,
                                        double rho_floor, double *se_out,
                                        double *su_out, void *stream) {
  (void)rho_floor;
  FP3D_REQUIRE(n == 0 || (q2 && rho && upstream && branch && se_out && su_out),
               "rho_tilde_frozen_sensitivities: NULL pointer");
  ELEMWISE(k_frozen_sens, n, q2, rho, upstream, branch, m_inf,
           m_crit * m_crit, C, gamma, se_out, su_out);
}

int fp3d_rho_tilde_sensitivities(int64_t n, const double *q2, const double *rho,
                                 const int32_t *upstream, double m_inf,
                                 double m_crit, double C, double gamma,
                                 double rho_floor, double *se_out,
                                 double *su_out, void *stream) {
  FP3D_REQUIRE(n == 0 || (q2 && rho && upstream && se_out && su_out),
               "rho_tilde_sensitivities: NULL pointer");
  ELEMWISE(k_sens, n, q2, rho, upstream, m_inf, m_crit * m_crit, C, gamma,
           rho_floor, se_out, su_out);
}

int fp3d_batch_eval_job_bytes(void) { return (int)sizeof(EvalJob); }

int fp3d_batch_residual_eval(int nb, const void *jobs, int64_t n_tets, int64_t n_nodes,
                             const int32_t *elements, const double *B, const double *V,
                             const int32_t *face_neighbors, const double *face_disp,
                             const double *nodes, const int32_t *inc_ptr,
                             const int32_t *inc_pos, double u2, double gamma,
                             double m_crit, double C, double rho_floor, int max_hops,
                             void *stream) {
  FP3D_REQUIRE(nb >= 1 && nb <= 1024, "batch_residual_eval: nb in [1, 1024]");
  FP3D_REQUIRE(jobs && n_tets > 0 && n_nodes > 0 && elements && B && V && face_neighbors &&
                   face_disp && nodes && inc_ptr && inc_pos,
               "batch_residual_eval: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements) && FP3D_ALIGNED16(face_neighbors) &&
                   FP3D_ALIGNED16(face_disp) && FP3D_ALIGNED16(inc_pos),
               "batch_residual_eval: arrays must be 16-byte aligned");
  FP3D_REQUIRE(u2 > 0.0 && max_hops >= 0, "batch_residual_eval: u2 > 0, max_hops >= 0");
  hipStream_t s = FP3D_STREAM(stream);
  const EvalJob *J = (const EvalJob *)jobs;
  const int eb = g_elem_block;
  hipLaunchKernelGGL(k_be_state, dim3(grid_for(n_tets, eb), nb), dim3(eb), 0, s, n_tets,
                     elements, B, u2, gamma, J);
  const int rb = reduce_blocks(n_tets);
  hipLaunchKernelGGL(k_be_count_p1, dim3(rb, nb), dim3(kBlock), 0, s, n_tets, J);
  hipLaunchKernelGGL(k_be_count_p2, dim3(1, nb), dim3(kBlock), 0, s, rb, J);
  hipLaunchKernelGGL(k_be_walk<2>, dim3(grid_for(n_tets, 2 * eb), nb), dim3(eb), 0, s,
                     n_tets, face_neighbors, face_disp, nodes, elements, max_hops,
                     m_crit * m_crit, C, gamma, rho_floor, J);
  hipLaunchKernelGGL(k_be_mon_p1, dim3(rb, nb), dim3(kBlock), 0, s, n_tets, rho_floor, J);
  hipLaunchKernelGGL(k_be_mon_p2, dim3(1, nb), dim3(kBlock), 0, s, rb, J);
  hipLaunchKernelGGL(k_be_terms, dim3(grid_for(n_tets, eb), nb), dim3(eb), 0, s, n_tets, B,
                     V, inc_pos, J);
  const unsigned g = xcd_grid((n_nodes + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_be_gather, dim3(g, nb), dim3(kBlock), 0, s, n_nodes, (int64_t)g,
                     inc_ptr, J);
  int rc = check_launch("batch_residual_eval");
  if (rc) return rc;
  return batch_masked_norms(nb, J, n_nodes, s);
}

int fp3d_upwind_monitors(int64_t n, const double *nu, const double *rho_tilde,
                         double rho_floor, double *work, double *out3,
                         void *stream) {
  FP3D_REQUIRE(nu && rho_tilde && work && out3, "upwind_monitors: NULL");
  int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_monitors_p1, dim3(nb), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, nu, rho_tilde, rho_floor, work);
  hipLaunchKernelGGL(k_monitors_p2, dim3(1), dim3(kBlock), 0,
                     FP3D_STREAM(stream), nb, work, out3);
  return check_launch("upwind_monitors");
}

int fp3d_count_limited(int64_t n, const uint8_t *lim, double *work,
                       double *out1, void *stream) {
  FP3D_REQUIRE(lim && work && out1, "count_limited: NULL");
  int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_count_zero_p1, dim3(nb), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, lim, work);
  hipLaunchKernelGGL(k_sum_p2, dim3(1), dim3(kBlock), 0, FP3D_STREAM(stream),
                     nb, work, out1);
  return check_launch("count_limited");
}

static int check_offsets(int n_colors, const int64_t *off) {
  FP3D_REQUIRE(n_colors >= 0 && (n_colors == 0 || off),
               "colour offsets missing");
  for (int c = 0; c < n_colors; ++c)
    FP3D_REQUIRE(off[c] >= 0 && off[c + 1] >= off[c],
                 "colour offsets not non-decreasing");
  return FP3D_OK;
}

int fp3d_assemble_residual_colored(int n_colors, const int64_t *color_offsets,
                                   const int32_t *elem_index,
                                   const int32_t *elements, const double *B,
                                   const double *V, const double *rho_tilde,
                                   const double *grad, int64_t n_nodes,
                                   double *R, void *stream) {
  int rc = check_offsets(n_colors, color_offsets);
  if (rc) return rc;
  FP3D_REQUIRE(R && n_nodes >= 0, "assemble_residual: R missing");
  FP3D_REQUIRE(elements && B && V && rho_tilde && grad,
               "assemble_residual: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elements), "elements must be 16-byte aligned");
  if (n_nodes > 0)
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n_nodes)), dim3(kBlock), 0,
                       FP3D_STREAM(stream), n_nodes, 0.0, R);
  for (int c = 0; c < n_colors; ++c) {
    int64_t k0 = color_offsets[c], k1 = color_offsets[c + 1];
    if (k1 == k0) continue;
    hipLaunchKernelGGL(k_residual_color, dim3(grid_for(k1 - k0)), dim3(kBlock),
                       0, FP3D_STREAM(stream), k0, k1, elem_index, elements, B,
                       V, rho_tilde, grad, R);
  }
  return check_launch("assemble_residual_colored");
}

int fp3d_assemble_matrix_colored(int n_colors, const int64_t *color_offsets,
                                 const int32_t *elem_index,
                                 const int32_t *elem_to_csr, const double *B,
                                 const double *V, const double *rho_tilde,
                                 double *data, void *stream) {
  int rc = check_offsets(n_colors, color_offsets);
  if (rc) return rc;
  FP3D_REQUIRE(elem_to_csr && B && V && rho_tilde && data,
               "assemble_matrix: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elem_to_csr), "elem_to_csr must be 16-byte aligned");
  for (int c = 0; c < n_colors; ++c) {
    int64_t k0 = color_offsets[c], k1 = color_offsets[c + 1];
    if (k1 == k0) continue;
    hipLaunchKernelGGL(k_matrix_color, dim3(grid_for(k1 - k0)), dim3(kBlock),
                       0, FP3D_STREAM(stream), k0, k1, elem_index, elem_to_csr,
                       B, V, rho_tilde, data);
  }
  return check_launch("assemble_matrix_colored");
}

int fp3d_assemble_jacobian_colored(int n_colors, const int64_t *color_offsets,
                                   const int32_t *elem_index,
                                   const int32_t *elem_to_csr, const double *B,
                                   const double *V, const double *rho_tilde,
                                   const double *grad, const double *s_e,
                                   const uint8_t *lim, double inv_u2,
                                   double *data, void *stream) {
  int rc = check_offsets(n_colors, color_offsets);
  if (rc) return rc;
  FP3D_REQUIRE(elem_to_csr && B && V && rho_tilde && grad && s_e && data,
               "assemble_jacobian: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elem_to_csr), "elem_to_csr must be 16-byte aligned");
  for (int c = 0; c < n_colors; ++c) {
    int64_t k0 = color_offsets[c], k1 = color_offsets[c + 1];
    if (k1 == k0) continue;
    hipLaunchKernelGGL(k_jacobian_color, dim3(grid_for(k1 - k0)), dim3(kBlock),
                       0, FP3D_STREAM(stream), k0, k1, elem_index, elem_to_csr,
                       B, V, rho_tilde, grad, s_e, lim, inv_u2, 0.0, data);
  }
  return check_launch("assemble_jacobian_colored");
}

int fp3d_assemble_jacobian_spd_colored(int n_colors,
                                       const int64_t *color_offsets,
                                       const int32_t *elem_index,
                                       const int32_t *elem_to_csr,
                                       const double *B, const double *V,
                                       const double *rho_tilde,
                                       const double *grad, const double *s_e,
                                       const uint8_t *lim, double inv_u2,
                                       double clamp, double *data,
                                       void *stream) {
  int rc = check_offsets(n_colors, color_offsets);
  if (rc) return rc;
  FP3D_REQUIRE(elem_to_csr && B && V && rho_tilde && grad && s_e && data,
               "assemble_jacobian_spd: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(elem_to_csr), "elem_to_csr must be 16-byte aligned");
  FP3D_REQUIRE(clamp > 0.0 && clamp < 1.0, "clamp must lie in (0, 1)");
  for (int c = 0; c < n_colors; ++c) {
    int64_t k0 = color_offsets[c], k1 = color_offsets[c + 1];
    if (k1 == k0) continue;
    hipLaunchKernelGGL(k_jacobian_color, dim3(grid_for(k1 - k0)), dim3(kBlock),
                       0, FP3D_STREAM(stream), k0, k1, elem_index, elem_to_csr,
                       B, V, rho_tilde, grad, s_e, lim, inv_u2, clamp, data);
  }
  return check_launch("assemble_jacobian_spd_colored");
}

int fp3d_gather_rows_per_block(void) { return kGatherRows; }

int fp3d_element_residual_terms(int64_t n_tets, const double *B,
                                const double *V, const double *rho_tilde,
                                const double *grad, const double *gdotb,
                                const int32_t *inc_pos, double *terms,
                                void *stream) {
  FP3D_REQUIRE(n_tets >= 0, "element_residual_terms: n_tets < 0");
  if (n_tets == 0) return FP3D_OK;
  FP3D_REQUIRE(V && rho_tilde && terms && (gdotb || (B && grad)),
               "element_residual_terms: NULL pointer");
  FP3D_REQUIRE(((uintptr_t)gdotb & 31) == 0, "gdotb must be 32-byte aligned");
  FP3D_REQUIRE(((uintptr_t)terms & 31) == 0, "terms must be 32-byte aligned");
  FP3D_REQUIRE(!inc_pos || ((uintptr_t)inc_pos & 15) == 0,
               "inc_pos must be 16-byte aligned");
  hipLaunchKernelGGL(k_elem_residual_terms, dim3(grid_for(n_tets, g_elem_block)),
                     dim3(g_elem_block), 0, FP3D_STREAM(stream), n_tets, B, V,
                     rho_tilde, grad, gdotb, inc_pos, terms);
  return check_launch("element_residual_terms");
}

int fp3d_gather_node_terms(int64_t n_nodes, const int32_t *inc_ptr,
                           const int32_t *inc_ent, const double *terms,
                           double *R, void *stream) {
  FP3D_REQUIRE(n_nodes >= 0, "gather_node_terms: n_nodes < 0");
  if (n_nodes == 0) return FP3D_OK;
  FP3D_REQUIRE(inc_ptr && terms && R, "gather_node_terms: NULL pointer");
  int64_t nb = (n_nodes + kBlock - 1) / kBlock;
  unsigned g = xcd_grid(nb);
  if (inc_ent)
    hipLaunchKernelGGL(k_gather_terms, dim3(g), dim3(kBlock), 0,
                       FP3D_STREAM(stream), n_nodes, (int64_t)g, inc_ptr,
                       inc_ent, terms, R);
  else
    hipLaunchKernelGGL(k_gather_terms_staged, dim3(g), dim3(kBlock), 0,
                       FP3D_STREAM(stream), n_nodes, (int64_t)g, inc_ptr,
                       terms, R);
  return check_launch("gather_node_terms");
}

int fp3d_element_matrix_terms(int64_t n_tets, const double *B, const double *V,
                              const double *rho_tilde, const double *grad,
                              const double *s_e, const uint8_t *lim,
                              double inv_u2, double clamp,
                              const int32_t *inc_pos, double *K, void *stream) {
  FP3D_REQUIRE(n_tets >= 0, "element_matrix_terms: n_tets < 0");
  if (n_tets == 0) return FP3D_OK;
  FP3D_REQUIRE(B && V && rho_tilde && K, "element_matrix_terms: NULL pointer");
  FP3D_REQUIRE(!s_e || grad, "element_matrix_terms: s_e needs grad");
  FP3D_REQUIRE(((uintptr_t)K & 31) == 0, "K must be 32-byte aligned");
  FP3D_REQUIRE(clamp == 0.0 || (clamp > 0.0 && clamp < 1.0 && s_e),
               "element_matrix_terms: clamp must be 0 or in (0, 1) with s_e");
  hipLaunchKernelGGL(k_elem_matrix_terms, dim3(grid_for(n_tets, g_elem_block)),
                     dim3(g_elem_block),
                     0, FP3D_STREAM(stream), n_tets, B, V, rho_tilde,
                     s_e ? grad : nullptr, s_e, lim, inv_u2, clamp, inc_pos, K);
  return check_launch("element_matrix_terms");
}

int fp3d_gather_csr_terms(int64_t n_nodes, int max_seg, const int32_t *indptr,
                          const int32_t *inc_ptr, const int32_t *inc_ent,
                          const uint8_t *rel8, const double *K, double *data,
                          void *stream) {
  FP3D_REQUIRE(n_nodes >= 0, "gather_csr_terms: n_nodes < 0");
  if (n_nodes == 0) return FP3D_OK;
  FP3D_REQUIRE(indptr && inc_ptr && rel8 && K && data,
               "gather_csr_terms: NULL pointer");
  FP3D_REQUIRE(((uintptr_t)rel8 & 3) == 0, "rel8 must be 4-byte aligned");
  FP3D_REQUIRE(((uintptr_t)K & 31) == 0, "K must be 32-byte aligned");
  FP3D_REQUIRE(max_seg > 0 && (int64_t)max_seg * 8 <= 65536,
               "gather_csr_terms: max_seg must lie in (0, 8192]");
  int64_t nb = (n_nodes + kGatherRows - 1) / kGatherRows;
  unsigned g = xcd_grid(nb);
  if (inc_ent) {
    hipLaunchKernelGGL(k_gather_csr, dim3(g), dim3(kGatherRows),
                       (size_t)max_seg * sizeof(double), FP3D_STREAM(stream),
                       n_nodes, (int64_t)g, max_seg, indptr, inc_ptr, inc_ent,
                       (const uint32_t *)rel8, K, data);
  } else {
#define FP3D_STAGED(CH)                                                      \
  hipLaunchKernelGGL(k_gather_csr_staged<CH>, dim3(g), dim3(kGatherLanes),   \
                     (size_t)max_seg * sizeof(double), FP3D_STREAM(stream),  \
                     n_nodes, (int64_t)g, max_seg, indptr, inc_ptr,          \
                     (const uint32_t *)rel8, (const double4 *)K, data)
    if (g_gather_chunk == 256)
      FP3D_STAGED(256);
    else if (g_gather_chunk == 2048)
      FP3D_STAGED(2048);
    else if (g_gather_chunk == 512)
      FP3D_STAGED(512);
    else
      FP3D_STAGED(1024);
#undef FP3D_STAGED
  }
  return check_launch("gather_csr_terms");
}

int fp3d_fill_term3_coo(int64_t n_active, const int32_t *active,
                        const int32_t *elements, const double *B,
                        const double *V, const double *grad,
                        const int32_t *upstream, const double *s_u,
                        double inv_u2, int32_t *rows, int32_t *cols,
                        double *vals, void *stream) {
  FP3D_REQUIRE(n_active == 0 || (active && elements && B && V && grad &&
                                 upstream && s_u && rows && cols && vals),
               "fill_term3_coo: NULL pointer");
  ELEMWISE(k_term3_coo, n_active, active, elements, B, V, grad, upstream, s_u,
           inv_u2, rows, cols, vals);
}

int fp3d_fill(int64_t n, double value, double *x, void *stream) {
  FP3D_REQUIRE(n == 0 || x, "fill: NULL");
  ELEMWISE(k_fill, n, value, x);
}

int fp3d_set_element_block(int threads) {
  FP3D_REQUIRE(threads == 64 || threads == 128 || threads == 192 || threads == 256,
               "set_element_block: threads must be 64, 128, 192 or 256");
  g_elem_block = threads;
  return FP3D_OK;
}

int fp3d_get_element_block(void) { return g_elem_block; }

int fp3d_set_gather_chunk(int rows) {
  FP3D_REQUIRE(rows == 256 || rows == 512 || rows == 1024 || rows == 2048,
               "set_gather_chunk: rows must be 256, 512, 1024 or 2048");
  g_gather_chunk = rows;
  return FP3D_OK;
}

int fp3d_get_gather_chunk(void) { return g_gather_chunk; }

int fp3d_set_walk_width(int walks) {
  FP3D_REQUIRE(walks >= 1 && walks <= 4, "set_walk_width: walks must be 1, 2, 3 or 4");
  g_walk_width = walks;
  return FP3D_OK;
}

int fp3d_get_walk_width(void) { return g_walk_width; }

}  // extern "C"
