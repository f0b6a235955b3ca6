/*
 * ORACLE -- test infrastructure only.  A plain-C restatement of the
 * reference's numba kernels (swayli94/UP3D, pyfp3d/) used as the parity
 * checker for the HIP path and as the timed CPU baseline.  Nothing in the
 * product (up3d_amd/) may link or call this file.
 *
 * Each function follows the cited reference function statement by
 * statement, in float64, with the reference's loop structure: colours run
 * serially and elements within a colour run in parallel (OpenMP here,
 * numba prange there).  Compiled with -ffp-contract=off so the arithmetic
 * is the written IEEE sequence.
 *
 * Layout: elements int32 (n,4); B float64 (n,4,3); grad (n,3); upstream
 * int32; elem_to_csr int32 (n,4,4); face_neighbors int32 (n,4).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* OpenMP team size of the parallel loops below (the timed CPU baseline runs
 * at several thread counts in one process). */
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_get_threads(void) { return omp_get_max_threads(); }

/* ---- pyfp3d/physics/isentropic.py ------------------------------------- */
/* :29-58 density_isentropic */
static double o_density(double q2, double M, double g) {
  double exponent = 1.0 / (g - 1.0);
  double base = 1.0 + 0.5 * (g - 1.0) * (M * M) * (1.0 - q2);
  if (base < 1e-10) base = 1e-10;
  return pow(base, exponent);
}
/* :79-97 mach_number_squared */
static double o_mach2(double q2, double M, double g) {
  double numerator = q2 * (M * M);
  double denominator = 1.0 + 0.5 * (g - 1.0) * (M * M) * (1.0 - q2);
  return numerator / denominator;
}
/* :163-180 density_derivative_wrt_q_sq */
static double o_drho(double q2, double M, double g) {
  double rho = o_density(q2, M, g);
  return -0.5 * (M * M) * pow(rho, 2.0 - g);
}
/* :183-206 mach_squared_derivative_wrt_q_sq */
static double o_dmach2(double q2, double M, double g) {
  double D = 1.0 + 0.5 * (g - 1.0) * (M * M) * (1.0 - q2);
  return (M * M) * (1.0 + 0.5 * (g - 1.0) * (M * M)) / (D * D);
}
static double o_max(double a, double b) { return a >= b ? a : b; }
/* pyfp3d/kernels/upwind.py:487-488: C * max(0, 1 - mc2 / max(m2, mc2)) */
static double o_nu(double m2, double mc2, double C) {
  return C * o_max(0.0, 1.0 - mc2 / o_max(m2, mc2));
}

void or_density_field(int64_t n, const double *q2, double M, double g,
                      double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = o_density(q2[i], M, g);
}
void or_mach2_field(int64_t n, const double *q2, double M, double g,
                    double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = o_mach2(q2[i], M, g);
}
void or_drho_field(int64_t n, const double *q2, double M, double g,
                   double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = o_drho(q2[i], M, g);
}
void or_dmach2_field(int64_t n, const double *q2, double M, double g,
                     double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = o_dmach2(q2[i], M, g);
}

/* ---- pyfp3d/kernels/gradient.py:32-64 element_velocity_q2 ------------- */
void or_element_velocity_q2(int64_t n, const int32_t *el, const double *B,
                            const double *phi, double *grad, double *q2) {
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int i = 0; i < 4; ++i) {
      double p = phi[el[4 * e + i]];
      gx += p * B[12 * e + 3 * i];
      gy += p * B[12 * e + 3 * i + 1];
      gz += p * B[12 * e + 3 * i + 2];
    }
    grad[3 * e] = gx;
    grad[3 * e + 1] = gy;
    grad[3 * e + 2] = gz;
    q2[e] = gx * gx + gy * gy + gz * gz;
  }
}

/* ---- pyfp3d/kernels/residual.py:43-76 assemble_residual_colored ------- */
void or_assemble_residual_colored(int n_colors, const int64_t *off,
                                  const int64_t *color_elems, const int32_t *el,
                                  const double *B, const double *V,
                                  const double *rt, const double *grad,
                                  int64_t n_nodes, double *R) {
  for (int64_t i = 0; i < n_nodes; ++i) R[i] = 0.0;
  for (int c = 0; c < n_colors; ++c) {
#pragma omp parallel for schedule(static)
    for (int64_t k = off[c]; k < off[c + 1]; ++k) {
      int64_t e = color_elems[k];
      double w = rt[e] * V[e];
      double gx = grad[3 * e], gy = grad[3 * e + 1], gz = grad[3 * e + 2];
      for (int i = 0; i < 4; ++i)
        R[el[4 * e + i]] += w * (gx * B[12 * e + 3 * i] + gy * B[12 * e + 3 * i + 1] +
                                 gz * B[12 * e + 3 * i + 2]);
    }
  }
}

/* ---- pyfp3d/kernels/residual.py:144-189 assemble_residual (serial) ----
 * The coloured residual's per-element arithmetic, accumulated in ascending
 * element order as the reference's serial assembly does. */
void or_assemble_residual_serial(int64_t n, const int32_t *el, const double *B,
                                 const double *V, const double *rt,
                                 const double *grad, int64_t n_nodes,
                                 double *R) {
  for (int64_t i = 0; i < n_nodes; ++i) R[i] = 0.0;
  for (int64_t e = 0; e < n; ++e) {
    double w = rt[e] * V[e];
    double gx = grad[3 * e], gy = grad[3 * e + 1], gz = grad[3 * e + 2];
    for (int i = 0; i < 4; ++i)
      R[el[4 * e + i]] += w * (gx * B[12 * e + 3 * i] + gy * B[12 * e + 3 * i + 1] +
                               gz * B[12 * e + 3 * i + 2]);
  }
}

/* ---- Picard matrix / Newton Terms 1+2 in serial element order ----------
 * pyfp3d/kernels/jacobian.py:105-180 per-element values, summed per CSR slot
 * in ascending element order (the serial triplet assembly of
 * pyfp3d/kernels/residual.py:106-129 followed by COO->CSR).  s_e == NULL
 * gives the Picard matrix; clamp > 0 the clamped (SPD) variant. */
void or_assemble_jacobian_serial(int64_t n, const int32_t *emap,
                                 const double *B, const double *V,
                                 const double *rt, const double *grad,
                                 const double *s_e, const uint8_t *lim,
                                 double inv_u2, double clamp, int64_t nnz,
                                 double *data) {
  for (int64_t k = 0; k < nnz; ++k) data[k] = 0.0;
  for (int64_t e = 0; e < n; ++e) {
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
    for (int i = 0; i < 4; ++i) {
      double bi0 = B[12 * e + 3 * i], bi1 = B[12 * e + 3 * i + 1],
             bi2 = B[12 * e + 3 * i + 2];
      double gi = g0 * bi0 + g1 * bi1 + g2 * bi2;
      for (int j = 0; j < 4; ++j) {
        double bj0 = B[12 * e + 3 * j], bj1 = B[12 * e + 3 * j + 1],
               bj2 = B[12 * e + 3 * j + 2];
        double val = w1 * (bi0 * bj0 + bi1 * bj1 + bi2 * bj2);
        if (w2 != 0.0) val += w2 * gi * (g0 * bj0 + g1 * bj1 + g2 * bj2);
        data[emap[16 * e + 4 * i + j]] += val;
      }
    }
  }
}

/* ---- pyfp3d/kernels/jacobian.py:77-102 build_elem_to_csr -------------- */
void or_build_elem_to_csr(int64_t n, const int32_t *el, const int32_t *indptr,
                          const int32_t *indices, int32_t *emap) {
  for (int64_t e = 0; e < n; ++e)
    for (int i = 0; i < 4; ++i) {
      int32_t row = el[4 * e + i];
      int64_t lo = indptr[row], hi = indptr[row + 1];
      for (int j = 0; j < 4; ++j) {
        int32_t col = el[4 * e + j];
        int64_t a = lo, b = hi;
        while (a < b) {
          int64_t mid = (a + b) / 2;
          if (indices[mid] < col) a = mid + 1;
          else b = mid;
        }
        emap[16 * e + 4 * i + j] = (int32_t)a;
      }
    }
}

/* ---- pyfp3d/kernels/jacobian.py:105-129 assemble_matrix_data_colored -- */
void or_assemble_matrix_colored(int n_colors, const int64_t *off,
                                const int64_t *color_elems, const int32_t *emap,
                                const double *B, const double *V,
                                const double *rt, double *data) {
  for (int c = 0; c < n_colors; ++c) {
#pragma omp parallel for schedule(static)
    for (int64_t k = off[c]; k < off[c + 1]; ++k) {
      int64_t e = color_elems[k];
      double w = rt[e] * V[e];
      for (int i = 0; i < 4; ++i) {
        double bi0 = B[12 * e + 3 * i], bi1 = B[12 * e + 3 * i + 1],
               bi2 = B[12 * e + 3 * i + 2];
        for (int j = 0; j < 4; ++j)
          data[emap[16 * e + 4 * i + j]] +=
              w * (bi0 * B[12 * e + 3 * j] + bi1 * B[12 * e + 3 * j + 1] +
                   bi2 * B[12 * e + 3 * j + 2]);
      }
    }
  }
}

/* ---- pyfp3d/kernels/jacobian.py:132-180 assemble_jacobian_data_colored */
void or_assemble_jacobian_colored(int n_colors, const int64_t *off,
                                  const int64_t *color_elems,
                                  const int32_t *emap, const double *B,
                                  const double *V, const double *rt,
                                  const double *grad, const double *s_e,
                                  double inv_u2, double *data) {
  for (int c = 0; c < n_colors; ++c) {
#pragma omp parallel for schedule(static)
    for (int64_t k = off[c]; k < off[c + 1]; ++k) {
      int64_t e = color_elems[k];
      double w1 = rt[e] * V[e];
      double g0 = grad[3 * e], g1 = grad[3 * e + 1], g2 = grad[3 * e + 2];
      double w2 = 2.0 * inv_u2 * s_e[e] * V[e];
      for (int i = 0; i < 4; ++i) {
        double bi0 = B[12 * e + 3 * i], bi1 = B[12 * e + 3 * i + 1],
               bi2 = B[12 * e + 3 * i + 2];
        double gi = g0 * bi0 + g1 * bi1 + g2 * bi2;
        for (int j = 0; j < 4; ++j) {
          double bj0 = B[12 * e + 3 * j], bj1 = B[12 * e + 3 * j + 1],
                 bj2 = B[12 * e + 3 * j + 2];
          double val = w1 * (bi0 * bj0 + bi1 * bj1 + bi2 * bj2);
          if (w2 != 0.0) val += w2 * gi * (g0 * bj0 + g1 * bj1 + g2 * bj2);
          data[emap[16 * e + 4 * i + j]] += val;
        }
      }
    }
  }
}

/* ---- pyfp3d/kernels/jacobian.py:183-228 fill_term3_coo --------------- */
void or_fill_term3_coo(int64_t n_active, const int64_t *active,
                       const int32_t *el, const double *B, const double *V,
                       const double *grad, const int32_t *up, const double *s_u,
                       double inv_u2, int64_t *rows, int64_t *cols,
                       double *vals) {
#pragma omp parallel for schedule(static)
  for (int64_t a = 0; a < n_active; ++a) {
    int64_t e = active[a], u = up[e];
    double ge0 = grad[3 * e], ge1 = grad[3 * e + 1], ge2 = grad[3 * e + 2];
    double gu0 = grad[3 * u], gu1 = grad[3 * u + 1], gu2 = grad[3 * u + 2];
    double w = 2.0 * inv_u2 * s_u[e] * V[e];
    int64_t base = 16 * a;
    for (int i = 0; i < 4; ++i) {
      double gi = ge0 * B[12 * e + 3 * i] + ge1 * B[12 * e + 3 * i + 1] +
                  ge2 * B[12 * e + 3 * i + 2];
      int64_t ri = el[4 * e + i];
      double wi = w * gi;
      for (int k = 0; k < 4; ++k) {
        int64_t idx = base + 4 * i + k;
        rows[idx] = ri;
        cols[idx] = el[4 * u + k];
        vals[idx] = wi * (gu0 * B[12 * u + 3 * k] + gu1 * B[12 * u + 3 * k + 1] +
                          gu2 * B[12 * u + 3 * k + 2]);
      }
    }
  }
}

/* ---- pyfp3d/kernels/upwind.py:68-125 upstream_elements ---------------- */
void or_upstream_elements(int64_t n, const int32_t *fn, const double *cen,
                          const double *nodes, const int32_t *el,
                          const double *grad, int max_hops, int32_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    double gx = grad[3 * e], gy = grad[3 * e + 1], gz = grad[3 * e + 2];
    double pmin = 1e300, pmax = -1e300;
    for (int k = 0; k < 4; ++k) {
      int64_t nd = el[4 * e + k];
      double p = nodes[3 * nd] * gx + nodes[3 * nd + 1] * gy + nodes[3 * nd + 2] * gz;
      if (p < pmin) pmin = p;
      if (p > pmax) pmax = p;
    }
    double target = 0.8 * (pmax - pmin);
    int64_t cur = e;
    double reach = 0.0;
    int hops = 0;
    int walking = 1;
    while (walking && hops < max_hops && reach < target) {
      int64_t best = -1;
      double best_disp = 0.0;
      for (int f = 0; f < 4; ++f) {
        int64_t nb = fn[4 * cur + f];
        if (nb >= 0) {
          double d = (cen[3 * nb] - cen[3 * cur]) * gx +
                     (cen[3 * nb + 1] - cen[3 * cur + 1]) * gy +
                     (cen[3 * nb + 2] - cen[3 * cur + 2]) * gz;
          if (d < best_disp) {
            best_disp = d;
            best = nb;
          }
        }
      }
      if (best < 0) {
        walking = 0;
      } else {
        cur = best;
        reach += -best_disp;
        hops += 1;
      }
    }
    out[e] = (int32_t)cur;
  }
}

/* ---- pyfp3d/kernels/upwind.py:456-500 rho_tilde_sweep ----------------- */
void or_rho_tilde_sweep(int64_t n, const double *q2, const double *rho,
                        const int32_t *up, double M, double m_crit, double C,
                        double g, double rho_floor, double *nu_out,
                        double *rt_out) {
  double mc2 = m_crit * m_crit;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    int64_t u = up[e];
    double m2 = o_mach2(q2[e], M, g);
    double m2u = o_mach2(q2[u], M, g);
    double nu_e = o_nu(m2, mc2, C);
    double nu_u = o_nu(m2u, mc2, C);
    double nu = o_max(nu_e, nu_u);
    nu_out[e] = nu;
    double rt = rho[e] - nu * (rho[e] - rho[u]);
    rt_out[e] = rt > rho_floor ? rt : rho_floor;
  }
}

/* ---- pyfp3d/kernels/upwind.py:503-541 classify_upwind_branches -------- */
void or_classify_upwind_branches(int64_t n, const double *q2, const double *rho,
                                 const int32_t *up, double M, double m_crit,
                                 double C, double g, double rho_floor,
                                 int8_t *br) {
  double mc2 = m_crit * m_crit;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    int64_t u = up[e];
    if (u == e) {
      br[e] = rho[e] <= rho_floor ? 3 : 0;
      continue;
    }
    double nu_e = o_nu(o_mach2(q2[e], M, g), mc2, C);
    double nu_u = o_nu(o_mach2(q2[u], M, g), mc2, C);
    double nu = o_max(nu_e, nu_u);
    double rt = rho[e] - nu * (rho[e] - rho[u]);
    if (rt <= rho_floor) br[e] = 3;
    else if (nu == 0.0) br[e] = 0;
    else if (nu_e >= nu_u) br[e] = 1;
    else br[e] = 2;
  }
}

/* ---- pyfp3d/kernels/upwind.py:544-593 rho_tilde_frozen_sweep ---------- */
void or_rho_tilde_frozen_sweep(int64_t n, const double *q2, const double *rho,
                               const int32_t *up, const int8_t *br, double M,
                               double m_crit, double C, double g,
                               double rho_floor, double *nu_out,
                               double *rt_out) {
  double mc2 = m_crit * m_crit;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    int b = br[e];
    if (b == 3) {
      nu_out[e] = 0.0;
      rt_out[e] = rho_floor;
      continue;
    }
    if (b == 0) {
      nu_out[e] = 0.0;
      rt_out[e] = rho[e];
      continue;
    }
    int64_t u = up[e];
    double nu = (b == 1) ? o_nu(o_mach2(q2[e], M, g), mc2, C)
                         : o_nu(o_mach2(q2[u], M, g), mc2, C);
    nu_out[e] = nu;
    rt_out[e] = rho[e] - nu * (rho[e] - rho[u]);
  }
}

/* ---- pyfp3d/kernels/upwind.py:596-643 frozen sensitivities ------------ */
void or_rho_tilde_frozen_sensitivities(int64_t n, const double *q2,
                                       const double *rho, const int32_t *up,
                                       const int8_t *br, double M,
                                       double m_crit, double C, double g,
                                       double *se, double *su) {
  double mc2 = m_crit * m_crit;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    int b = br[e];
    if (b == 3) {
      se[e] = 0.0;
      su[e] = 0.0;
      continue;
    }
    double drho_e = o_drho(q2[e], M, g);
    if (b == 0) {
      se[e] = drho_e;
      su[e] = 0.0;
      continue;
    }
    int64_t u = up[e];
    double drho_u = o_drho(q2[u], M, g);
    double djump = rho[e] - rho[u];
    if (b == 1) {
      double m2 = o_mach2(q2[e], M, g);
      double nu = o_nu(m2, mc2, C);
      double dnu_e = m2 > mc2 ? C * (mc2 / (m2 * m2)) * o_dmach2(q2[e], M, g) : 0.0;
      se[e] = drho_e * (1.0 - nu) - djump * dnu_e;
      su[e] = nu * drho_u;
    } else {
      double m2u = o_mach2(q2[u], M, g);
      double nu = o_nu(m2u, mc2, C);
      double dnu_u = m2u > mc2 ? C * (mc2 / (m2u * m2u)) * o_dmach2(q2[u], M, g) : 0.0;
      se[e] = drho_e * (1.0 - nu);
      su[e] = nu * drho_u - djump * dnu_u;
    }
  }
}

/* ---- pyfp3d/kernels/upwind.py:646-730 rho_tilde_sensitivities_sweep -- */
void or_rho_tilde_sensitivities(int64_t n, const double *q2, const double *rho,
                                const int32_t *up, double M, double m_crit,
                                double C, double g, double rho_floor,
                                double *se, double *su) {
  double mc2 = m_crit * m_crit;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n; ++e) {
    int64_t u = up[e];
    double rho_e = rho[e];
    double drho_e = o_drho(q2[e], M, g);
    if (u == e) {
      if (rho_e <= rho_floor) {
        se[e] = 0.0;
        su[e] = 0.0;
      } else {
        se[e] = drho_e;
        su[e] = 0.0;
      }
      continue;
    }
    double m2 = o_mach2(q2[e], M, g);
    double m2u = o_mach2(q2[u], M, g);
    double nu_e = o_nu(m2, mc2, C);
    double nu_u = o_nu(m2u, mc2, C);
    double nu = o_max(nu_e, nu_u);
    double rho_u = rho[u];
    double rt = rho_e - nu * (rho_e - rho_u);
    if (rt <= rho_floor) {
      se[e] = 0.0;
      su[e] = 0.0;
      continue;
    }
    if (nu == 0.0) {
      se[e] = drho_e;
      su[e] = 0.0;
      continue;
    }
    double drho_u = o_drho(q2[u], M, g);
    double djump = rho_e - rho_u;
    if (nu_e >= nu_u) {
      double dnu_e = m2 > mc2 ? C * (mc2 / (m2 * m2)) * o_dmach2(q2[e], M, g) : 0.0;
      se[e] = drho_e * (1.0 - nu) - djump * dnu_e;
      su[e] = nu * drho_u;
    } else {
      double dnu_u = C * (mc2 / (m2u * m2u)) * o_dmach2(q2[u], M, g);
      se[e] = drho_e * (1.0 - nu);
      su[e] = nu * drho_u - djump * dnu_u;
    }
  }
}

/* ---- pyfp3d/mesh/coloring.py:65-95 _greedy_coloring_kernel ------------ */
int or_greedy_coloring(int64_t n, const int32_t *el, const int64_t *node_off,
                       const int64_t *node_elems, int32_t *colors) {
  const int max_colors = 256;
  int64_t mark[256];
  for (int c = 0; c < max_colors; ++c) mark[c] = -1;
  for (int64_t e = 0; e < n; ++e) colors[e] = -1;
  for (int64_t e = 0; e < n; ++e) {
    for (int i = 0; i < 4; ++i) {
      int64_t node = el[4 * e + i];
      for (int64_t k = node_off[node]; k < node_off[node + 1]; ++k) {
        int32_t c = colors[node_elems[k]];
        if (c >= 0) mark[c] = e;
      }
    }
    int c = 0;
    while (c < max_colors && mark[c] == e) ++c;
    if (c >= max_colors) return -1; /* the reference raises ValueError */
    colors[e] = c;
  }
  return 0;
}
