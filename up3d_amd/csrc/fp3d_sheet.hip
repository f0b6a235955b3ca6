// Level-set wake sheet terms on the device (gfx950, wave64).
//
// The multivalued operator's special elements (the two copies of each cut
// element, the lower copy of each below-TE fan element), the wake
// least-squares rows, the TE Kutta rows and the Newton Term 2 of the special
// copies all land in one extended CSR pattern that is fixed by the cut
// census, so the pattern and each value's list of contributions are built
// once on the host.  Per assembly one lane per CSR slot sums that slot's
// contributions in their fixed order (no atomics, bitwise repeatable):
//
//   kind 0  coef * wa[e]                 (upper-copy density)
//   kind 1  coef * wb[e]                 (lower-copy density)
//   kind 2  coef                         (static: wake least-squares rows)
//   kind 3  coef * kv[e]                 (per-call host values: Kutta rows)
//   kind 4  coef * sa[e] (g.B_a)(g.B_b)  (Term 2, upper copy: g = ga[e])
//   kind 5  coef * sb[e] (g.B_a)(g.B_b)  (Term 2, lower copy: g = gb[e])
//
// with e = src[c], (a, b) = (ab[c] >> 2, ab[c] & 3).  The Term 3 of special
// copies (and of plain elements whose upstream element is special) couples
// a row to the DOFs of the element the upwind walk selects, so its columns
// change with the selection: it is applied as an operator instead, per
// candidate element c (rows r_a, upstream DOFs col_b):
//
//   t_c = sum_b gu[c][b] x[col_b],   y[r] += sum over (c, a) with r_a = r of
//                                            W[c][a] t_c
//
// with W = 2 inv_u2 s_u V_e (g_e.B_a) and gu = g_u.B_u,b, the row gather in
// a fixed entry order.
#include "fp3d_common.h"

namespace fp3d {

__device__ __forceinline__ double gdot(const double *__restrict__ g,
                                       const double *__restrict__ B, int64_t e,
                                       int a) {
  const double *b = B + 12 * e + 3 * a;
  const double *h = g + 3 * e;
  return (h[0] * b[0] + h[1] * b[1]) + h[2] * b[2];
}

__global__ __launch_bounds__(kBlock) void k_sheet_fill(
    int64_t n_slots, const int32_t *__restrict__ seg,
    const uint8_t *__restrict__ kind, const int32_t *__restrict__ src,
    const uint8_t *__restrict__ ab, const double *__restrict__ coef,
    const double *__restrict__ wa, const double *__restrict__ wb,
    const double *__restrict__ kv, const double *__restrict__ sa,
    const double *__restrict__ sb, const double *__restrict__ ga,
    const double *__restrict__ gb, const double *__restrict__ B,
    const int32_t *__restrict__ srow, const int32_t *__restrict__ scol,
    const uint8_t *__restrict__ fmask, double *__restrict__ out) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  double acc = 0.0;
  for (int32_t c = seg[s]; c < seg[s + 1]; ++c) {
    const int64_t e = src[c];
    double t;
    switch (kind[c]) {
      case 0: t = coef[c] * wa[e]; break;
      case 1: t = coef[c] * wb[e]; break;
      case 2: t = coef[c]; break;
      case 3: t = coef[c] * kv[e]; break;
      case 4: {
        const int a = ab[c] >> 2, b = ab[c] & 3;
        t = ((coef[c] * sa[e]) * gdot(ga, B, e, a)) * gdot(ga, B, e, b);
        break;
      }
      default: {
        const int a = ab[c] >> 2, b = ab[c] & 3;
        t = ((coef[c] * sb[e]) * gdot(gb, B, e, a)) * gdot(gb, B, e, b);
        break;
      }
    }
    acc += t;
  }
  if (fmask && (!fmask[srow[s]] || !fmask[scol[s]])) acc = 0.0;
  out[s] = acc;
}

// per candidate: rows, W, gu, cols of the special Term 3
__global__ __launch_bounds__(kBlock) void k_sheet_t3_prepare(
    int64_t n_cand, const int32_t *__restrict__ cand,
    const uint8_t *__restrict__ cside, const double *__restrict__ g0,
    const double *__restrict__ g1, const int32_t *__restrict__ up0,
    const int32_t *__restrict__ up1, const double *__restrict__ su0,
    const double *__restrict__ su1, const uint8_t *__restrict__ lim0,
    const uint8_t *__restrict__ lim1, const int32_t *__restrict__ dof0,
    const int32_t *__restrict__ dof1, const int32_t *__restrict__ rd0,
    const int32_t *__restrict__ rd1, const int32_t *__restrict__ row_map,
    const uint8_t *__restrict__ fmask, const double *__restrict__ B,
    const double *__restrict__ V, double inv_u2, int32_t *__restrict__ rows,
    double *__restrict__ W, double *__restrict__ gu, int32_t *__restrict__ cols) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  const int64_t e = cand[c];
  const bool lo = cside[c] != 0;
  const double *g = lo ? g1 : g0;
  const int64_t u = (lo ? up1 : up0)[e];
  const uint8_t *lim = lo ? lim1 : lim0;
  const double su = (lo ? su1 : su0)[e] * (double)lim[u];
  const double w3 = ((2.0 * inv_u2) * su) * V[e];
  const int32_t *dof = lo ? dof1 : dof0;
  const int32_t *rd = lo ? rd1 : rd0;
  for (int a = 0; a < 4; ++a) {
    int32_t r = row_map[dof[4 * e + a]];
    if (r >= 0 && fmask && !fmask[r]) r = -1;
    rows[4 * c + a] = (w3 != 0.0) ? r : -1;
    W[4 * c + a] = w3 * gdot(g, B, e, a);
    const int32_t k = rd[4 * u + a];
    gu[4 * c + a] = (fmask && !fmask[k]) ? 0.0 : gdot(g, B, u, a);
    cols[4 * c + a] = k;
  }
}

__global__ __launch_bounds__(kBlock) void k_sheet_t3_tval(
    int64_t n_cand, const double *__restrict__ gu, const int32_t *__restrict__ cols,
    const double *__restrict__ x, double *__restrict__ t) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  const double *g = gu + 4 * c;
  const int32_t *k = cols + 4 * c;
  t[c] = ((g[0] * x[k[0]] + g[1] * x[k[1]]) + g[2] * x[k[2]]) + g[3] * x[k[3]];
}

// y[r] += sum_{q in [rptr[r], rptr[r+1])} W[rent[q]] t[rent[q] >> 2]
__global__ __launch_bounds__(kBlock) void k_sheet_t3_rows(
    int64_t n_rows, const int32_t *__restrict__ rptr,
    const int32_t *__restrict__ rent, const double *__restrict__ W,
    const double *__restrict__ t, double *__restrict__ y) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const int32_t q0 = rptr[r], q1 = rptr[r + 1];
  if (q0 == q1) return;
  double s = 0.0;
  for (int32_t q = q0; q < q1; ++q) {
    const int32_t ent = rent[q];
    s += W[ent] * t[ent >> 2];
  }
  y[r] += s;
}

}  // namespace fp3d

using namespace fp3d;

extern "C" {

int fp3d_sheet_fill(int64_t n_slots, const int32_t *seg, const uint8_t *kind,
                    const int32_t *src, const uint8_t *ab, const double *coef,
                    const double *wa, const double *wb, const double *kv,
                    const double *sa, const double *sb, const double *ga,
                    const double *gb, const double *B, const int32_t *srow,
                    const int32_t *scol, const uint8_t *fmask, double *out,
                    void *stream) {
  if (n_slots == 0) return FP3D_OK;
  FP3D_REQUIRE(n_slots > 0 && seg && kind && src && ab && coef && srow && scol && out,
               "sheet_fill: NULL pointer");
  hipLaunchKernelGGL(k_sheet_fill, dim3(grid_for(n_slots)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n_slots, seg, kind, src, ab, coef, wa, wb,
                     kv, sa, sb, ga, gb, B, srow, scol, fmask, out);
  return check_launch("sheet_fill");
}

int fp3d_sheet_t3_prepare(int64_t n_cand, const int32_t *cand, const uint8_t *cside,
                          const double *g0, const double *g1, const int32_t *up0,
                          const int32_t *up1, const double *su0, const double *su1,
                          const uint8_t *lim0, const uint8_t *lim1,
                          const int32_t *dof0, const int32_t *dof1,
                          const int32_t *rd0, const int32_t *rd1,
                          const int32_t *row_map, const uint8_t *fmask,
                          const double *B, const double *V, double inv_u2,
                          int32_t *rows, double *W, double *gu, int32_t *cols,
                          void *stream) {
  if (n_cand == 0) return FP3D_OK;
  FP3D_REQUIRE(n_cand > 0 && cand && cside && g0 && g1 && up0 && up1 && su0 && su1 &&
                   lim0 && lim1 && dof0 && dof1 && rd0 && rd1 && row_map && B && V &&
                   rows && W && gu && cols,
               "sheet_t3_prepare: NULL pointer");
  hipLaunchKernelGGL(k_sheet_t3_prepare, dim3(grid_for(n_cand)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n_cand, cand, cside, g0, g1, up0, up1, su0,
                     su1, lim0, lim1, dof0, dof1, rd0, rd1, row_map, fmask, B, V,
                     inv_u2, rows, W, gu, cols);
  return check_launch("sheet_t3_prepare");
}

int fp3d_sheet_t3_apply(int64_t n_cand, int64_t n_rows, const double *gu,
                        const int32_t *cols, const double *W, const int32_t *rptr,
                        const int32_t *rent, const double *x, double *t, double *y,
                        void *stream) {
  if (n_cand == 0 || n_rows == 0) return FP3D_OK;
  FP3D_REQUIRE(n_cand > 0 && n_rows > 0 && gu && cols && W && rptr && rent && x && t &&
                   y && x != y,
               "sheet_t3_apply: bad arguments");
  hipStream_t s = FP3D_STREAM(stream);
  hipLaunchKernelGGL(k_sheet_t3_tval, dim3(grid_for(n_cand)), dim3(kBlock), 0, s,
                     n_cand, gu, cols, x, t);
  hipLaunchKernelGGL(k_sheet_t3_rows, dim3(grid_for(n_rows)), dim3(kBlock), 0, s,
                     n_rows, rptr, rent, W, t, y);
  return check_launch("sheet_t3_apply");
}

}  // extern "C"
