// Sparse and dense linear-algebra kernels of the device Krylov/AMG solve
// (gfx950, wave64).  All HBM-bound:
//   * CSR SpMV with G lanes per row (G = 4..64 picked from nnz/row), float64
//     or float32 matrix values, float64 vectors and accumulation, with fused
//     epilogues (residual, prolongation-add, Chebyshev step, Term 3 gather);
//   * deterministic two-pass reductions (multi-dot, norms) for Arnoldi;
//   * the Newton Term-3 operator built once per Newton step as a row-gather
//     table (node -> active element incidences, in element order), so the
//     upstream coupling is applied inside the SpMV without atomics and
//     without rebuilding the CSR pattern.
#include <cmath>
#include <mutex>

#include "fp3d_common.h"

namespace fp3d {

// --------------------------------------------------------------- epilogues
// Every epilogue is templated on the vector type TV of the V-cycle (double
// or float; the arithmetic is double either way, so TV = double is exactly
// the float64 sequence).
template <typename TV>
struct EpStore {  // y = s
  TV *y;
  __device__ void operator()(int64_t r, double s) const { y[r] = (TV)s; }
};
template <typename TV>
struct EpResidual {  // r = b - s
  const TV *b;
  TV *r;
  __device__ void operator()(int64_t i, double s) const {
    r[i] = (TV)((double)b[i] - s);
  }
};
template <typename TV>
struct EpAdd {  // x += s
  TV *x;
  __device__ void operator()(int64_t i, double s) const {
    x[i] = (TV)((double)x[i] + s);
  }
};
// Chebyshev smoothing (three-term recurrence on D^-1 A).  x is never read
// by the SpMV of a kernel that writes it: each step adds the PREVIOUS
// direction to x, and the last direction is added by k_add.
// First step from a nonzero x: r = dinv (b - Ax); d = r/theta.
template <typename TV>
struct EpChebFirst {
  const TV *b, *dinv;
  TV *r, *d;
  double inv_theta;
  __device__ void operator()(int64_t i, double s) const {
    double ri = (double)dinv[i] * ((double)b[i] - s);
    r[i] = (TV)ri;
    d[i] = (TV)(ri * inv_theta);
  }
};
// Recurrence: x += d_old; r -= dinv A d_old; d_new = c1 d_old + c2 r
template <typename TV>
struct EpChebIter {
  const TV *dinv, *d_old;
  TV *x, *r, *d_new;
  double c1, c2;
  __device__ void operator()(int64_t i, double s) const {
    double dold = d_old[i];
    double ri = (double)r[i] - (double)dinv[i] * s;
    r[i] = (TV)ri;
    d_new[i] = (TV)(c1 * dold + c2 * ri);
    x[i] = (TV)((double)x[i] + dold);
  }
};

// Last recurrence step: as EpChebIter, then the new direction is added to
// x as well -- the same two additions, in the same order, that the step
// and a separate x += d_last pass make.
template <typename TV>
struct EpChebLast {
  const TV *dinv, *d_old;
  TV *x, *r, *d_new;
  double c1, c2;
  __device__ void operator()(int64_t i, double s) const {
    double dold = d_old[i];
    double ri = (double)r[i] - (double)dinv[i] * s;
    r[i] = (TV)ri;
    double dn = c1 * dold + c2 * ri;
    d_new[i] = (TV)dn;
    x[i] = (TV)(((double)x[i] + dold) + dn);
  }
};
// Restriction b_c = R r that also takes the coarse level's first smoothing
// step from x_c = 0 (as k_cheb_zero on b_c): r_c = dinv b_c,
// d_c = r_c / theta, and x_c = 0, or x_c = 0 + d_c when that step is the
// only one.
template <typename TV>
struct EpRestrictStart {
  TV *bc;
  const TV *dinv;
  TV *x, *r, *d;
  double inv_theta;
  int x_is_d;
  __device__ void operator()(int64_t i, double s) const {
    bc[i] = (TV)s;
    double ri = (double)dinv[i] * s;
    r[i] = (TV)ri;
    double di = ri * inv_theta;
    d[i] = (TV)di;
    x[i] = (TV)(x_is_d ? 0.0 + di : 0.0);
  }
};

// Column j of a CSR row: int32 absolute indices, or int16 offsets from the
// row (square matrices whose bandwidth fits: half the index bytes).
template <typename TI>
__device__ __forceinline__ int32_t csr_col(int64_t row, TI v);
template <>
__device__ __forceinline__ int32_t csr_col<int32_t>(int64_t, int32_t v) {
  return v;
}
template <>
__device__ __forceinline__ int32_t csr_col<int16_t>(int64_t row, int16_t v) {
  return (int32_t)row + (int32_t)v;
}

// Row `row` of A x on its G lanes: lane-strided products, then a fixed
// shuffle tree (lane 0 holds the sum).
template <int G, typename T, typename TI, typename TX>
__device__ __forceinline__ double spmv_row_sum(int64_t row, int lane,
                                               const int32_t *__restrict__ ptr,
                                               const TI *__restrict__ idx,
                                               const T *__restrict__ val,
                                               const TX *__restrict__ x) {
  int32_t j0 = ptr[row], j1 = ptr[row + 1];
  double s = 0.0;
  for (int32_t j = j0 + lane; j < j1; j += G)
    s += (double)val[j] * (double)x[csr_col<TI>(row, idx[j])];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_down(s, off, G);
  return s;
}

template <int G, typename T, typename TI, typename TX, typename Ep>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n_rows,
                                                 const int32_t *__restrict__ ptr,
                                                 const TI *__restrict__ idx,
                                                 const T *__restrict__ val,
                                                 const TX *__restrict__ x,
                                                 Ep ep) {
  constexpr int RPB = kBlock / G;
  int64_t row = (int64_t)blockIdx.x * RPB + threadIdx.x / G;
  int lane = threadIdx.x % G;
  if (row >= n_rows) return;
  const double s = spmv_row_sum<G, T, TI, TX>(row, lane, ptr, idx, val, x);
  if (lane == 0) ep(row, s);
}

// SpMV + Term-3 gather + Dirichlet identity rows (free_mask nullable)
// (the body takes its block index, so the batched kernel below runs the
// same code per instance)
template <int G, typename T, typename TI>
__device__ __forceinline__ void spmv_t3_body(
    int64_t bx, int64_t n_rows, const int32_t *__restrict__ ptr,
    const TI *__restrict__ idx, const T *__restrict__ val,
    const double *__restrict__ x, const int32_t *__restrict__ t3ptr,
    const int32_t *__restrict__ t3ent,
    const double *__restrict__ W,  // weight of entry q, beside t3ent[q]
    const double *__restrict__ tval, const uint8_t *__restrict__ fmask,
    double *__restrict__ y) {
  constexpr int RPB = kBlock / G;
  int64_t row = bx * RPB + threadIdx.x / G;
  int lane = threadIdx.x % G;
  if (row >= n_rows) return;
  int32_t j0 = ptr[row], j1 = ptr[row + 1];
  double s = 0.0;
  for (int32_t j = j0 + lane; j < j1; j += G)
    s += (double)val[j] * x[csr_col<TI>(row, idx[j])];
  // the row's Term-3 entries are strided over the same G lanes (fixed lane
  // assignment and reduction tree: deterministic)
  if (t3ptr) {
    const int32_t q1 = t3ptr[row + 1];
    for (int32_t q = t3ptr[row] + lane; q < q1; q += G)
      s += W[q] * tval[t3ent[q] >> 2];
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_down(s, off, G);
  if (lane == 0) {
    if (fmask && !fmask[row]) s = x[row];
    y[row] = s;
  }
}
template <int G, typename T, typename TI>
__global__ __launch_bounds__(kBlock) void k_spmv_t3(
    int64_t n_rows, const int32_t *__restrict__ ptr,
    const TI *__restrict__ idx, const T *__restrict__ val,
    const double *__restrict__ x, const int32_t *__restrict__ t3ptr,
    const int32_t *__restrict__ t3ent, const double *__restrict__ W,
    const double *__restrict__ tval, const uint8_t *__restrict__ fmask,
    double *__restrict__ y) {
  spmv_t3_body<G, T, TI>(blockIdx.x, n_rows, ptr, idx, val, x, t3ptr, t3ent, W, tval,
                         fmask, y);
}

inline int lanes_for(int64_t n_rows, int64_t nnz) {
  double avg = n_rows > 0 ? (double)nnz / (double)n_rows : 1.0;
  if (avg <= 6) return 4;
  if (avg <= 20) return 8;
  if (avg <= 48) return 16;
  if (avg <= 128) return 32;
  return 64;
}

template <typename T, typename TI, typename TX, typename Ep>
int launch_spmv(int64_t n_rows, int64_t nnz, const int32_t *ptr,
                const TI *idx, const T *val, const TX *x, Ep ep,
                hipStream_t s) {
  if (n_rows <= 0) return FP3D_OK;
  int G = lanes_for(n_rows, nnz);
  unsigned grid = grid_for(n_rows, kBlock / G);
  switch (G) {
    case 4: hipLaunchKernelGGL((k_spmv<4, T, TI, TX, Ep>), dim3(grid), dim3(kBlock), 0, s, n_rows, ptr, idx, val, x, ep); break;
    case 8: hipLaunchKernelGGL((k_spmv<8, T, TI, TX, Ep>), dim3(grid), dim3(kBlock), 0, s, n_rows, ptr, idx, val, x, ep); break;
    case 16: hipLaunchKernelGGL((k_spmv<16, T, TI, TX, Ep>), dim3(grid), dim3(kBlock), 0, s, n_rows, ptr, idx, val, x, ep); break;
    case 32: hipLaunchKernelGGL((k_spmv<32, T, TI, TX, Ep>), dim3(grid), dim3(kBlock), 0, s, n_rows, ptr, idx, val, x, ep); break;
    default: hipLaunchKernelGGL((k_spmv<64, T, TI, TX, Ep>), dim3(grid), dim3(kBlock), 0, s, n_rows, ptr, idx, val, x, ep); break;
  }
  return check_launch("spmv");
}

// Matrix format bits of the V-cycle and Jacobian entry points.
constexpr int kFmtF32Values = 1, kFmtF32Vectors = 2, kFmtIdx16 = 4;

// SpMV with the matrix format of `fmt` (values float32 or float64, int32 or
// int16 row-relative columns) and an x of the epilogue's vector type.
template <typename TX, typename Ep>
int spmv_fmt(int64_t n_rows, int64_t nnz, const int32_t *ptr, const void *idx,
             const void *val, int fmt, const TX *x, Ep ep, hipStream_t s) {
  const bool v32 = fmt & kFmtF32Values, i16 = fmt & kFmtIdx16;
  if (v32 && i16)
    return launch_spmv<float, int16_t, TX, Ep>(n_rows, nnz, ptr, (const int16_t *)idx,
                                               (const float *)val, x, ep, s);
  if (v32)
    return launch_spmv<float, int32_t, TX, Ep>(n_rows, nnz, ptr, (const int32_t *)idx,
                                               (const float *)val, x, ep, s);
  if (i16)
    return launch_spmv<double, int16_t, TX, Ep>(n_rows, nnz, ptr, (const int16_t *)idx,
                                                (const double *)val, x, ep, s);
  return launch_spmv<double, int32_t, TX, Ep>(n_rows, nnz, ptr, (const int32_t *)idx,
                                              (const double *)val, x, ep, s);
}

// ------------------------------------------------------------ BLAS-1 etc.
__global__ void k_axpby(int64_t n, double a, const double *x, double b,
                        double *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void k_mask_copy(int64_t n, const uint8_t *__restrict__ m,
                            const double *x, double *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = m[i] ? x[i] : 0.0;
}
__global__ void k_to_f32(int64_t n, const double *__restrict__ x,
                         float *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
__global__ void k_scale_cvt(int64_t n, double a, const double *x, float *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)(a * x[i]);
}
__global__ void k_widen(int64_t n, const float *x, double *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}
// y = x / nrm[0] with the divisor read on the device (so the next Arnoldi
// step can be queued before the host has seen the norm)
template <typename TY>
__global__ void k_scale_inv_dev(int64_t n, const double *__restrict__ nrm,
                                const double *__restrict__ x, TY *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (TY)((1.0 / nrm[0]) * x[i]);
}
__global__ void k_scale(int64_t n, double a, const double *x, double *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

// rows/cols of non-free dofs -> identity (principal submatrix embedding)
__global__ __launch_bounds__(kBlock) void k_mask_dirichlet(
    int64_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
    const uint8_t *__restrict__ fm, double *__restrict__ data,
    float *__restrict__ d32) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  bool rf = fm[r] != 0;
  for (int32_t j = ptr[r]; j < ptr[r + 1]; ++j) {
    int32_t c = idx[j];
    double v = data[j];
    if (!rf) v = (c == r) ? 1.0 : 0.0;
    else if (!fm[c]) v = 0.0;
    data[j] = v;
    if (d32) d32[j] = (float)v;
  }
}

// lumped nodal volume m_v = sum over incident (element, local) entries of
// V_e / 4, summed in the incidence order (ascending storage element)
__global__ __launch_bounds__(kBlock) void k_lumped_volume(
    int64_t nn, const int32_t *__restrict__ iptr,
    const int32_t *__restrict__ ient, const double *__restrict__ V,
    double *__restrict__ out) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  double m = 0.0;
  for (int32_t q = iptr[v]; q < iptr[v + 1]; ++q) m += 0.25 * V[ient[q] >> 2];
  out[v] = m;
}

// data[diag[r]] += s * w[r] on free rows (pseudo-transient diagonal)
__global__ __launch_bounds__(kBlock) void k_diag_shift(
    int64_t n, const int32_t *__restrict__ diag, const double *__restrict__ w,
    double s, const uint8_t *__restrict__ fm, double *__restrict__ data) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !fm[r]) return;
  data[diag[r]] += s * w[r];
}

__device__ __forceinline__ double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wsum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kBlock / 64; ++k) t += sh[k];
  return t;  // valid on thread 0
}

inline int rblocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b > kReduceBlocks) b = kReduceBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

// partial[k*nb + b] = sum over block b's row slice of V_k . w.  Block b
// owns the contiguous rows [b*seg, min(n, (b+1)*seg)); its slice of w is
// staged once in LDS (when seg <= kMdotLds) and each wave streams one basis
// vector's slice at a time (64 lanes x 8 B contiguous per load, four
// independent accumulators per lane), so the only block-wide barrier is the
// staging one.  Per (k, b) the sum is lane-strided then a fixed shuffle tree:
// deterministic.  cond != NULL makes the call a CGS2 re-orthogonalisation
// pass that runs only when the first pass cancelled (DGKS: ||w'||^2 =
// ||w||^2 - ||h||^2 < eta2 ||w||^2 with h = cond[0..nv), ||w||^2 =
// cond[nv]); otherwise the partials are written as exact zeros.
constexpr int kMdotLds = 2048;
constexpr int64_t kMdotSegMin = 1024;
constexpr int kMdotMinBlocks32 = 1024, kMdotMinBlocks64 = 768;
template <bool STAGED, typename TV>
__device__ __forceinline__ void mdot_p1_body(int64_t bx, int nb, int64_t n, int nv,
                                             const TV *__restrict__ V, int64_t ld,
                                             const double *__restrict__ w,
                                             double *__restrict__ part,
                                             const double *__restrict__ cond,
                                             double eta2, int64_t seg, int with_self,
                                             int group = 0, int n_groups = 1) {
  // STAGED reads the slice of w from LDS (a separate instantiation, so the
  // loads stay in the LDS address space rather than generic).  with_self:
  // row nv is w itself (w . w, summed exactly as a one-row call on w).
  // The rows k are dealt to (group, wave) round-robin: row k belongs to
  // wave k % nwave of group (k / nwave) % n_groups, so splitting the rows
  // over more blocks leaves every (k, bx) partial computed by the same code
  __shared__ double ws[STAGED ? kMdotLds : 1];
  if (cond) {
    double h2 = 0.0;
    for (int k = 0; k < nv; ++k) h2 += cond[k] * cond[k];
    double w2 = cond[nv];
    if (!(w2 - h2 < eta2 * w2)) {
      if (group == 0)
        for (int k = threadIdx.x; k < nv; k += blockDim.x)
          part[(int64_t)k * nb + bx] = 0.0;
      return;
    }
  }
  const int64_t r0 = (int64_t)bx * seg;
  const int64_t r1 = r0 + seg < n ? r0 + seg : n;
  const int len = (int)(r1 - r0);
  if (STAGED) {
    for (int i = threadIdx.x; i < len; i += blockDim.x) ws[i] = w[r0 + i];
    __syncthreads();
  }
  const double *__restrict__ wg = w + r0;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwave = blockDim.x >> 6;
  for (int k = group * nwave + wave; k < nv; k += nwave * n_groups) {
    const TV *__restrict__ v = V + (int64_t)k * ld + r0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = lane;
    for (; i + 192 < len; i += 256) {
      if (STAGED) {
        a0 += (double)v[i] * ws[i];
        a1 += (double)v[i + 64] * ws[i + 64];
        a2 += (double)v[i + 128] * ws[i + 128];
        a3 += (double)v[i + 192] * ws[i + 192];
      } else {
        a0 += (double)v[i] * wg[i];
        a1 += (double)v[i + 64] * wg[i + 64];
        a2 += (double)v[i + 128] * wg[i + 128];
        a3 += (double)v[i + 192] * wg[i + 192];
      }
    }
    for (; i < len; i += 64) a0 += (double)v[i] * (STAGED ? ws[i] : wg[i]);
    double t = wsum((a0 + a1) + (a2 + a3));
    if (lane == 0) part[(int64_t)k * nb + bx] = t;
  }
  if (with_self && wave == nv % nwave && group == (nv / nwave) % n_groups) {
    // the row after the last basis vector: w itself, read from global as
    // the V operand of a one-row call would be
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = lane;
    for (; i + 192 < len; i += 256) {
      if (STAGED) {
        a0 += wg[i] * ws[i];
        a1 += wg[i + 64] * ws[i + 64];
        a2 += wg[i + 128] * ws[i + 128];
        a3 += wg[i + 192] * ws[i + 192];
      } else {
        a0 += wg[i] * wg[i];
        a1 += wg[i + 64] * wg[i + 64];
        a2 += wg[i + 128] * wg[i + 128];
        a3 += wg[i + 192] * wg[i + 192];
      }
    }
    for (; i < len; i += 64) a0 += wg[i] * (STAGED ? ws[i] : wg[i]);
    double t = wsum((a0 + a1) + (a2 + a3));
    if (lane == 0) part[(int64_t)nv * nb + bx] = t;
  }
}
template <bool STAGED, typename TV>
__global__ __launch_bounds__(kBlock) void k_mdot_p1(int64_t n, int nv,
                                                   const TV *__restrict__ V,
                                                   int64_t ld,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ part,
                                                   const double *__restrict__ cond,
                                                   double eta2, int64_t seg,
                                                   int with_self) {
  mdot_p1_body<STAGED, TV>(blockIdx.x, gridDim.x, n, nv, V, ld, w, part, cond, eta2, seg,
                           with_self, blockIdx.y, gridDim.y);
}
inline int64_t mdot_seg(int64_t n) {
  int64_t seg = (n + kReduceBlocks - 1) / kReduceBlocks;
  return seg < kMdotSegMin ? kMdotSegMin : seg;
}
inline int mdot_blocks(int64_t n, int64_t seg) {
  int64_t b = (n + seg - 1) / seg;
  return (int)(b < 1 ? 1 : b);
}
// one block per k, sums the nb partials in index order (wave-parallel then
// fixed-order combine)
__device__ __forceinline__ void mdot_p2_body(int k, int nb,
                                             const double *__restrict__ part,
                                             double *__restrict__ out) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(int64_t)k * nb + b];
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) out[k] = t;
}
__global__ __launch_bounds__(kBlock) void k_mdot_p2(int nb,
                                                   const double *__restrict__ part,
                                                   double *__restrict__ out) {
  mdot_p2_body(blockIdx.x, nb, part, out);
}

__device__ __forceinline__ void norms_p1_body(int64_t bx, int nb, int64_t n,
                                              const double *__restrict__ x,
                                              double *__restrict__ part) {
  __shared__ double sh[kBlock / 64];
  __shared__ double shm[kBlock / 64];
  double s = 0.0, m = 0.0;
  for (int64_t i = bx * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)nb * blockDim.x) {
    double v = x[i];
    s += v * v;
    double a = fabs(v);
    m = a > m ? a : m;
  }
  m = wmax(m);
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) mm = shm[k] > mm ? shm[k] : mm;
    part[2 * bx] = t;
    part[2 * bx + 1] = mm;
  }
}
__global__ __launch_bounds__(kBlock) void k_norms_p1(int64_t n,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ part) {
  norms_p1_body(blockIdx.x, gridDim.x, n, x, part);
}
__device__ __forceinline__ void norms_p2_body(int nb, const double *__restrict__ part,
                                              double *__restrict__ out) {
  __shared__ double sh[kBlock / 64];
  __shared__ double shm[kBlock / 64];
  double s = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s += part[2 * b];
    m = part[2 * b + 1] > m ? part[2 * b + 1] : m;
  }
  m = wmax(m);
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) mm = shm[k] > mm ? shm[k] : mm;
    out[0] = sqrt(t);
    out[1] = mm;
  }
}
__global__ __launch_bounds__(kBlock) void k_norms_p2(
    int nb, const double *__restrict__ part, double *__restrict__ out) {
  norms_p2_body(nb, part, out);
}

// w -= sum_k h[k] V_k   (h in LDS; V float64 or float32, arithmetic float64)
template <typename TV>
__device__ __forceinline__ void maxpy_body(int64_t bx, int64_t n, int nv,
                                           const TV *__restrict__ V, int64_t ld,
                                           const double *__restrict__ h,
                                           double *__restrict__ w, double sign) {
  __shared__ double hs[256];
  __shared__ int any_nz;
  if (threadIdx.x == 0) any_nz = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < nv; k += blockDim.x) {
    hs[k] = h[k];
    if (hs[k] != 0.0) any_nz = 1;
  }
  __syncthreads();
  // all-zero coefficients (a skipped re-orthogonalisation): w unchanged
  if (!any_nz) return;
  int64_t i = bx * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = 0; k < nv; ++k) acc += hs[k] * (double)V[(int64_t)k * ld + i];
  w[i] += sign * acc;
}
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_maxpy(int64_t n, int nv,
                                                 const TV *__restrict__ V,
                                                 int64_t ld,
                                                 const double *__restrict__ h,
                                                 double *__restrict__ w,
                                                 double sign) {
  maxpy_body<TV>(blockIdx.x, n, nv, V, ld, h, w, sign);
}

// x = M b, M dense row-major: one wave per row, lanes stride the row
// (coalesced), fixed shuffle-tree sum
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_dense_mv(int64_t n,
                                                    const double *__restrict__ M,
                                                    const TV *__restrict__ b,
                                                    TV *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double *__restrict__ row = M + i * n;
  double s = 0.0;
  for (int64_t j = lane; j < n; j += 64) s += row[j] * (double)b[j];
  s = wsum(s);
  if (lane == 0) x[i] = (TV)s;
}

// Chebyshev first step from x = 0: x = 0; r = dinv b; d = r/theta (x = 0 + d
// when this step is the only one)
template <typename TV>
__global__ void k_cheb_zero(int64_t n, const TV *__restrict__ b,
                            const TV *__restrict__ dinv, double inv_theta,
                            TV *__restrict__ x, TV *__restrict__ r,
                            TV *__restrict__ d, int x_is_d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ri = (double)dinv[i] * (double)b[i];
  r[i] = (TV)ri;
  double di = ri * inv_theta;
  d[i] = (TV)di;
  x[i] = (TV)(x_is_d ? 0.0 + di : 0.0);
}
template <typename TV>
__global__ void k_add(int64_t n, const TV *__restrict__ d, TV *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (TV)((double)x[i] + (double)d[i]);
}

// ----------------------------------------------------------- Term 3 (device)
// Workspace layout (byte offsets computed by t3_layout).
struct T3Layout {
  size_t flag8, scan, bsum, act, W, Cc, tval, cnt, rowptr, ent, fq, Wq,
      total;
};
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
inline T3Layout t3_layout(int64_t nt, int64_t nn) {
  T3Layout L;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o = align256(o + bytes);
    return at;
  };
  int64_t nmax = nt > nn ? nt : nn;
  int64_t nbs = (nmax + 1 + kBlock - 1) / kBlock + 1;
  L.flag8 = take(nt + 1);
  L.scan = take(sizeof(int32_t) * (nmax + 2));
  L.bsum = take(sizeof(int32_t) * (nbs + 1));
  L.act = take(sizeof(int32_t) * (nt + 1));
  L.W = take(sizeof(double) * 4 * nt);
  L.Cc = take(sizeof(double) * 4 * nt);
  L.tval = take(sizeof(double) * nt);
  L.cnt = take(sizeof(int32_t) * (nn + 1));
  L.rowptr = take(sizeof(int32_t) * (nn + 2));
  L.ent = take(sizeof(int32_t) * 4 * nt);
  L.fq = take(4 * nt + 1);
  L.Wq = take(sizeof(double) * 4 * nt);
  L.total = o + 256;  // header: n_active (int64) lives in the last 256 B
  return L;
}

// per element: masked s_u, activity flag and the two coefficient sets (the
// weights W of the element's four entries in incidence order)
template <bool HAS_GB>
__global__ __launch_bounds__(kBlock) void k_t3_coef(
    int64_t nt, const int32_t *__restrict__ elements,
    const double *__restrict__ B, const double *__restrict__ V,
    const double *__restrict__ grad, const double *__restrict__ gdotb,
    const int32_t *__restrict__ up,
    const double *__restrict__ s_u, const uint8_t *__restrict__ lim,
    double inv_u2, const int32_t *__restrict__ pos,
    uint8_t *__restrict__ flag8,
    double *__restrict__ Wi, double *__restrict__ Cc) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  int64_t u = up[e];
  double s = s_u[e];
  if (lim) s = s * (lim[u] ? 1.0 : 0.0);
  bool act = s != 0.0;
  flag8[e] = act ? 1 : 0;
  if (!act) return;
  double w = 2.0 * inv_u2 * s * V[e];
  const int4 p = *reinterpret_cast<const int4 *>(pos + 4 * e);
  const int32_t pa[4] = {p.x, p.y, p.z, p.w};
  if (HAS_GB) {
    // grad . B_i of e and of u, from the state sweep (same bits)
    const double4 ge = *reinterpret_cast<const double4 *>(gdotb + 4 * e);
    const double4 gu = *reinterpret_cast<const double4 *>(gdotb + 4 * u);
    Wi[pa[0]] = w * ge.x;
    Wi[pa[1]] = w * ge.y;
    Wi[pa[2]] = w * ge.z;
    Wi[pa[3]] = w * ge.w;
    *reinterpret_cast<double4 *>(Cc + 4 * e) = gu;
    return;
  }
  double ge0 = grad[3 * e], ge1 = grad[3 * e + 1], ge2 = grad[3 * e + 2];
  double gu0 = grad[3 * u], gu1 = grad[3 * u + 1], gu2 = grad[3 * u + 2];
  for (int i = 0; i < 4; ++i) {
    double gi = ge0 * B[12 * e + 3 * i] + ge1 * B[12 * e + 3 * i + 1] +
                ge2 * B[12 * e + 3 * i + 2];
    // weight of entry (e, i), at its slot in node elements[e][i]'s
    // incidence list
    Wi[pa[i]] = w * gi;
    Cc[4 * e + i] = gu0 * B[12 * u + 3 * i] + gu1 * B[12 * u + 3 * i + 1] +
                    gu2 * B[12 * u + 3 * i + 2];
  }
}

// Block-wide exclusive scan of one int per lane: prefix inside each 64-lane
// wave by shuffles, then the wave totals through LDS (sh holds kBlock / 64
// ints; the caller syncs before sh is reused).  Integer sums: exact.
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *sh,
                                                   int32_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    const int32_t t = sh[k];
    pre += k < w ? t : 0;
    tot += t;
  }
  *total = tot;
  return pre + x - v;
}

// exclusive scan (reduce, scan the block totals, rescan), pass 1:
// kScanPer consecutive entries per lane, block totals out; the entries are
// read again in pass 3 instead of writing and re-reading a partial scan
constexpr int kScanPer = 4;
constexpr int kScanTile = kBlock * kScanPer;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_p1(int64_t n, const T *__restrict__ in,
                                                   int32_t *__restrict__ bsum) {
  __shared__ int32_t sh[kBlock / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  int32_t t = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) t += i0 + k < n ? (int32_t)in[i0 + k] : 0;
  int32_t tot;
  (void)block_excl_scan(t, sh, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
// pass 2: exclusive scan of the block totals in one workgroup: tiles of
// kScanP2Per * kBlock totals staged through LDS (coalesced), kScanP2Per
// consecutive totals per lane, one block scan per tile, running carry
constexpr int kScanP2Per = 16;
__global__ __launch_bounds__(kBlock) void k_scan_p2(int nb,
                                                   int32_t *__restrict__ bsum,
                                                   int32_t *__restrict__ total) {
  constexpr int kTile = kScanP2Per * kBlock;
  __shared__ int32_t tile[kTile];
  __shared__ int32_t sh[kBlock / 64];
  int32_t carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kTile) {
#pragma unroll
    for (int k = 0; k < kScanP2Per; ++k) {
      const int i = b0 + k * kBlock + threadIdx.x;
      tile[k * kBlock + threadIdx.x] = i < nb ? bsum[i] : 0;
    }
    __syncthreads();
    int32_t *mine = tile + threadIdx.x * kScanP2Per;
    int32_t t = 0;
#pragma unroll
    for (int k = 0; k < kScanP2Per; ++k) t += mine[k];
    int32_t tot;
    int32_t run = carry + block_excl_scan(t, sh, &tot);
#pragma unroll
    for (int k = 0; k < kScanP2Per; ++k) {
      const int32_t v = mine[k];
      mine[k] = run;
      run += v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanP2Per; ++k) {
      const int i = b0 + k * kBlock + threadIdx.x;
      if (i < nb) bsum[i] = tile[k * kBlock + threadIdx.x];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}
// pass 3: scan the block's entries again from its offset; with act != NULL
// also compact: the index of every nonzero entry goes to act[out[i]]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_p3(
    int64_t n, const T *__restrict__ in, int32_t *__restrict__ out,
    const int32_t *__restrict__ bsum, int32_t *__restrict__ act) {
  __shared__ int32_t sh[kBlock / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  int32_t v[kScanPer], t = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = i0 + k < n ? (int32_t)in[i0 + k] : 0;
    t += v[k];
  }
  int32_t tot;
  int32_t run = bsum[blockIdx.x] + block_excl_scan(t, sh, &tot);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t i = i0 + k;
    if (i < n) {
      out[i] = run;
      if (act && v[k]) act[run] = (int32_t)i;
    }
    run += v[k];
  }
}

// in: int32 counts or uint8 flags (the element selection's byte flags)
template <typename T>
int exclusive_scan(int64_t n, const T *in, int32_t *out, int32_t *bsum,
                   int32_t *total, int32_t *act, hipStream_t s) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_p1<T>, dim3(nb), dim3(kBlock), 0, s, n, in, bsum);
  hipLaunchKernelGGL(k_scan_p2, dim3(1), dim3(kBlock), 0, s, nb, bsum, total);
  hipLaunchKernelGGL(k_scan_p3<T>, dim3(nb), dim3(kBlock), 0, s, n, in, out, bsum,
                     act);
  return check_launch("exclusive_scan");
}

// Per block of kBlock nodes, whose incidence entries are the contiguous
// range [iptr[v0], iptr[v1]): the active flags of the entries (the
// element's flag, 0 on rows that fmask marks Dirichlet) are gathered in
// coalesced chunks through LDS, written in incidence order (fq), and
// counted per node (cnt).
constexpr int kT3Chunk = 2048;
__global__ __launch_bounds__(kBlock) void k_t3_count(
    int64_t nn, int64_t nb_pad, const int32_t *__restrict__ iptr,
    const int32_t *__restrict__ ient, const uint8_t *__restrict__ flag8,
    const uint8_t *__restrict__ fmask, uint8_t *__restrict__ fq,
    int32_t *__restrict__ cnt) {
  __shared__ uint8_t fl[kT3Chunk];
  constexpr int P = kT3Chunk / kBlock;
  const int64_t v0 = xcd_block(blockIdx.x, nb_pad) * kBlock;
  if (v0 >= nn) return;                        // uniform over the block
  const int64_t v1 = v0 + kBlock < nn ? v0 + kBlock : nn;
  const int64_t v = v0 + threadIdx.x;
  const int64_t qb = iptr[v0], qe = iptr[v1];
  int64_t q = 0, q1 = 0;
  bool on = false;
  if (v < v1) {
    q = iptr[v];
    q1 = iptr[v + 1];
    on = !fmask || fmask[v];
  }
  uint8_t nf[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t i = qb + threadIdx.x + k * kBlock;
    nf[k] = i < qe ? flag8[ient[i] >> 2] : (uint8_t)0;
  }
  int32_t c = 0;
  for (int64_t c0 = qb; c0 < qe; c0 += kT3Chunk) {
    __syncthreads();                           // previous chunk written out
#pragma unroll
    for (int k = 0; k < P; ++k) fl[threadIdx.x + k * kBlock] = nf[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int64_t i = c0 + kT3Chunk + threadIdx.x + k * kBlock;
      nf[k] = i < qe ? flag8[ient[i] >> 2] : (uint8_t)0;
    }
    const int64_t hi = q1 < c0 + kT3Chunk ? q1 : c0 + kT3Chunk;
    for (; q < hi; ++q) {
      if (on)
        c += fl[q - c0];
      else
        fl[q - c0] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int64_t i = c0 + threadIdx.x + k * kBlock;
      if (i < qe) fq[i] = fl[threadIdx.x + k * kBlock];
    }
  }
  if (v < v1) cnt[v] = c;
}
// Per block of kBlock nodes: the active entries of the block's incidence
// range, kept in incidence order (= by node, then ascending element), go
// to ent[rowptr[v0] + running count] with their weight beside them (Wq), so
// the matvec reads both contiguously.  Lane t of a chunk holds entries
// c0 + 256 k + t: each wave compacts its 64 entries by ballot, the 32
// (k, wave) counts are prefix-summed from LDS, and loads and stores stay
// coalesced.
constexpr int kT3FillP = 8;
constexpr int kT3FillChunk = kBlock * kT3FillP;
__global__ __launch_bounds__(kBlock) void k_t3_fill(
    int64_t nn, int64_t nb_pad, const int32_t *__restrict__ iptr,
    const int32_t *__restrict__ ient, const uint8_t *__restrict__ fq,
    const double *__restrict__ Wi, const int32_t *__restrict__ rowptr,
    int32_t *__restrict__ ent, double *__restrict__ Wq) {
  constexpr int NW = kBlock / 64;
  __shared__ int32_t wc[kT3FillP * NW];
  const int64_t v0 = xcd_block(blockIdx.x, nb_pad) * kBlock;
  if (v0 >= nn) return;                        // uniform over the block
  const int64_t v1 = v0 + kBlock < nn ? v0 + kBlock : nn;
  const int64_t qb = iptr[v0], qe = iptr[v1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int64_t run = rowptr[v0];
  for (int64_t c0 = qb; c0 < qe; c0 += kT3FillChunk) {
    uint64_t m[kT3FillP];
#pragma unroll
    for (int k = 0; k < kT3FillP; ++k) {
      const int64_t q = c0 + k * kBlock + threadIdx.x;
      const bool f = q < qe && fq[q] != 0;
      m[k] = __ballot(f);
      if (lane == 0) wc[k * NW + w] = __popcll(m[k]);
    }
    __syncthreads();
    // tot = active entries of the chunk before slot (k, j), slots taken in
    // q order (k-major, then wave)
    int32_t tot = 0;
#pragma unroll
    for (int k = 0; k < kT3FillP; ++k) {
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        if (j == w && ((m[k] >> lane) & 1ull)) {
          const int64_t q = c0 + k * kBlock + threadIdx.x;
          const int64_t o = run + tot + __popcll(m[k] & below);
          ent[o] = ient[q];
          Wq[o] = Wi[q];
        }
        tot += wc[k * NW + j];
      }
    }
    run += tot;
    __syncthreads();                           // wc reused next chunk
  }
}
// t_e = sum_k (grad_u . B_u,k) x[elements[u,k]] for the active elements
__device__ __forceinline__ void t3_tval_body(int64_t bx, int64_t n_active,
                                             const int32_t *__restrict__ act,
                                             const int32_t *__restrict__ elements,
                                             const int32_t *__restrict__ up,
                                             const double *__restrict__ Cc,
                                             const double *__restrict__ x,
                                             double *__restrict__ tval) {
  int64_t a = bx * blockDim.x + threadIdx.x;
  if (a >= n_active) return;
  int64_t e = act[a];
  int64_t u = up[e];
  const int4 nd = *reinterpret_cast<const int4 *>(elements + 4 * u);
  tval[e] = Cc[4 * e] * x[nd.x] + Cc[4 * e + 1] * x[nd.y] +
            Cc[4 * e + 2] * x[nd.z] + Cc[4 * e + 3] * x[nd.w];
}
__global__ void k_t3_tval(int64_t n_active, const int32_t *__restrict__ act,
                          const int32_t *__restrict__ elements,
                          const int32_t *__restrict__ up,
                          const double *__restrict__ Cc,
                          const double *__restrict__ x,
                          double *__restrict__ tval) {
  t3_tval_body(blockIdx.x, n_active, act, elements, up, Cc, x, tval);
}

// ------------------------------------------ batched residual evaluation
__global__ void k_be_mask(int64_t n, const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) j.sR[i] = j.fmask[i] ? j.R[i] : 0.0;
}
__global__ __launch_bounds__(kBlock) void k_be_norms_p1(int64_t n,
                                                       const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  norms_p1_body(blockIdx.x, gridDim.x, n, j.sR, j.work);
}
__global__ __launch_bounds__(kBlock) void k_be_norms_p2(int nb,
                                                       const EvalJob *__restrict__ jobs) {
  const EvalJob &j = jobs[blockIdx.y];
  norms_p2_body(nb, j.work, j.out + 4);
}

// ------------------------------------------------ batched Arnoldi steps
// One Krylov step of each of nb GMRES solves on the same mesh, one launch
// per kernel for all of them: blockIdx.y selects the solve, whose buffers
// come from its entry of a device-resident job table; every block then runs
// the same body, with the same grid, as the solve's own launch would, so
// each solve gets the same bits as when it runs alone.
struct BatchJob {          // 16 x 8 bytes; layout shared with the host
  const void *V;           // Krylov basis (float64 or float32 rows)
  int64_t ld;              // its row stride
  int64_t nv;              // basis vectors to orthogonalise against (col + 1)
  double *w;               // M^-1 A v_col, orthogonalised in place
  double *out;             // [0] ||w||, [1] max|w|, [2, 2+nv) h, [2+nv] ||w||^2,
                           // [3+nv, 3+2nv) second pass
  double *work;            // reduction partials
  void *vnext;             // basis row col + 1
  const float *src32;      // basis row col (float32 basis) or null
  double *x;               // Jacobian matvec input
  double *y;               // Jacobian matvec output
  const void *vals;        // Jacobian values (float64 or float32)
  const int32_t *up;       // upstream elements (Term 3)
  const void *t3;          // Term-3 workspace
  int64_t n_active;        // active Term-3 elements (0: no Term 3)
  const uint8_t *fmask;    // free mask
  int64_t reserved;
};
static_assert(sizeof(BatchJob) == 128, "BatchJob layout");
constexpr int kMaxBatch = 1024;

__global__ void k_b_widen(int64_t n, const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) j.x[i] = (double)j.src32[i];
}
__global__ void k_b_t3_tval(const BatchJob *__restrict__ jobs,
                            const int32_t *__restrict__ elements, size_t off_act,
                            size_t off_cc, size_t off_tval) {
  const BatchJob &j = jobs[blockIdx.y];
  if (j.n_active <= 0) return;
  const char *base = (const char *)j.t3;
  t3_tval_body(blockIdx.x, j.n_active, (const int32_t *)(base + off_act), elements, j.up,
               (const double *)(base + off_cc), j.x, (double *)(base + off_tval));
}
template <int G, typename T, typename TI>
__global__ __launch_bounds__(kBlock) void k_b_spmv_t3(
    int64_t n_rows, const int32_t *__restrict__ ptr, const TI *__restrict__ idx,
    const BatchJob *__restrict__ jobs, size_t off_rowptr, size_t off_ent, size_t off_wq,
    size_t off_tval) {
  const BatchJob &j = jobs[blockIdx.y];
  const char *base = (const char *)j.t3;
  const bool t3 = j.n_active > 0;
  spmv_t3_body<G, T, TI>(blockIdx.x, n_rows, ptr, idx, (const T *)j.vals, j.x,
                         t3 ? (const int32_t *)(base + off_rowptr) : nullptr,
                         t3 ? (const int32_t *)(base + off_ent) : nullptr,
                         t3 ? (const double *)(base + off_wq) : nullptr,
                         t3 ? (const double *)(base + off_tval) : nullptr, j.fmask, j.y);
}
// pass 0: h = V^T w and ||w||^2 into out[2..]; pass 1: the conditional
// second CGS pass into out[3 + nv..]
template <bool STAGED, typename TV>
__global__ __launch_bounds__(kBlock) void k_b_mdot_p1(int64_t n, int64_t seg, int pass,
                                                     double eta2,
                                                     const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  mdot_p1_body<STAGED, TV>(blockIdx.x, gridDim.x, n, (int)j.nv, (const TV *)j.V, j.ld,
                           j.w, j.work, pass ? j.out + 2 : nullptr, eta2, seg,
                           pass ? 0 : 1);
}
__global__ __launch_bounds__(kBlock) void k_b_mdot_p2(int nb, int pass,
                                                     const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  const int rows = (int)j.nv + (pass ? 0 : 1);
  if ((int)blockIdx.x >= rows) return;
  mdot_p2_body(blockIdx.x, nb, j.work, pass ? j.out + 3 + j.nv : j.out + 2);
}
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_b_maxpy(int64_t n, int pass,
                                                   const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  maxpy_body<TV>(blockIdx.x, n, (int)j.nv, (const TV *)j.V, j.ld,
                 pass ? j.out + 3 + j.nv : j.out + 2, j.w, -1.0);
}
__global__ __launch_bounds__(kBlock) void k_b_norms_p1(int64_t n,
                                                      const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  norms_p1_body(blockIdx.x, gridDim.x, n, j.w, j.work);
}
__global__ __launch_bounds__(kBlock) void k_b_norms_p2(int nb,
                                                      const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  norms_p2_body(nb, j.work, j.out);
}
template <typename TY>
__global__ void k_b_scale_inv(int64_t n, const BatchJob *__restrict__ jobs) {
  const BatchJob &j = jobs[blockIdx.y];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ((TY *)j.vnext)[i] = (TY)((1.0 / j.out[0]) * j.w[i]);
}

}  // namespace fp3d

using namespace fp3d;

// Chebyshev smoothing of one level on vectors of type TV
template <typename TV>
static int cheb_smooth_t(int64_t n, int64_t nnz, const int32_t *indptr,
                         const void *indices, const void *data, int fmt,
                         const TV *dinv, const TV *b, TV *x, TV *r, TV *d0, TV *d1,
                         int degree, double lmin, double lmax, int zero_init,
                         hipStream_t s) {
  double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  double sigma = theta / delta;
  double rho0 = 1.0 / sigma;
  int rc = FP3D_OK;
  // the last direction goes into x inside the last step (EpChebLast), or in
  // the first step itself from x = 0 when degree == 1; only the first step
  // from a nonzero x (which reads x in its SpMV) needs a separate x += d
  if (zero_init == 1) {
    hipLaunchKernelGGL(k_cheb_zero<TV>, dim3(grid_for(n)), dim3(kBlock), 0, s, n,
                       b, dinv, 1.0 / theta, x, r, d0, degree == 1 ? 1 : 0);
    rc = check_launch("cheb_zero");
  } else if (zero_init == 0) {
    rc = spmv_fmt(n, nnz, indptr, indices, data, fmt, (const TV *)x,
                  EpChebFirst<TV>{b, dinv, r, d0, 1.0 / theta}, s);
  }
  // zero_init == 2: the caller took the first step from x = 0 (x, r, d0 set,
  // as fp3d_amg_restrict_start does)
  if (rc) return rc;
  TV *dold = d0, *dnew = d1;
  for (int k = 1; k < degree; ++k) {
    double rho1 = 1.0 / (2.0 * sigma - rho0);
    double c1 = rho1 * rho0, c2 = 2.0 * rho1 / delta;
    if (k == degree - 1)
      rc = spmv_fmt(n, nnz, indptr, indices, data, fmt, (const TV *)dold,
                    EpChebLast<TV>{dinv, dold, x, r, dnew, c1, c2}, s);
    else
      rc = spmv_fmt(n, nnz, indptr, indices, data, fmt, (const TV *)dold,
                    EpChebIter<TV>{dinv, dold, x, r, dnew, c1, c2}, s);
    if (rc) return rc;
    rho0 = rho1;
    TV *t = dold;
    dold = dnew;
    dnew = t;
  }
  if (degree == 1 && zero_init == 0) {
    hipLaunchKernelGGL(k_add<TV>, dim3(grid_for(n)), dim3(kBlock), 0, s, n,
                       (const TV *)dold, x);
    return check_launch("cheb_smooth");
  }
  return FP3D_OK;
}

// ------------------------------------------------ grouped V-cycle steps
// One V-cycle operation for many solves at once (their own hierarchies,
// buffers and Chebyshev scalars), one launch: blockIdx.y picks the solve's
// 128-byte record, and every block runs the body, grid and epilogue of the
// solve's own launch (fp3d_amg_*, fp3d_dense_matvec_v, fp3d_scale_to_f32,
// fp3d_f32_to_f64), so each solve gets the bits of its own V-cycle.
struct GrpRec {              // 16 x 8 bytes; layout shared with the host
  int64_t n;                 // rows (SpMV) or entries (element-wise, dense)
  const int32_t *ptr;        // CSR row pointer
  const void *idx;           // CSR columns (int32, or int16 row offsets)
  const void *val;           // CSR values (float64 or float32)
  const void *x;             // SpMV operand
  int64_t kind;              // kGrp* below
  void *p[5];                // epilogue / operation pointers
  double d0, d1;             // epilogue scalars
  int64_t flag;
  int64_t reserved[2];
};
static_assert(sizeof(GrpRec) == 128, "GrpRec layout");

enum : int64_t {
  kGrpStore = 0,         // y[p0] = s
  kGrpResidual = 1,      // r[p1] = b[p0] - s
  kGrpAdd = 2,           // x[p0] += s
  kGrpChebFirst = 3,     // EpChebFirst{b p0, dinv p1, r p2, d p3, 1/theta d0}
  kGrpChebIter = 4,      // EpChebIter{dinv p0, d_old p1, x p2, r p3, d_new p4, c1 d0, c2 d1}
  kGrpChebLast = 5,      // EpChebLast, same fields
  kGrpRestrictStart = 6, // EpRestrictStart{bc p0, dinv p1, x p2, r p3, d p4, 1/theta d0, flag}
  kGrpChebZero = 10,     // k_cheb_zero{b p0, dinv p1, x p2, r p3, d p4, 1/theta d0, flag}
  kGrpAddVec = 11,       // k_add{d p0, x p1}
  kGrpToF32 = 12,        // k_scale_cvt{a d0, x p0, y p1}
  kGrpToF64 = 13,        // k_widen{x p0, y p1}
  kGrpDense = 14,        // k_dense_mv{M p0, b p1, x p2}
};

template <typename TV>
__device__ __forceinline__ void grp_epilogue(const GrpRec &r, int64_t row, double s) {
  switch (r.kind) {
    case kGrpStore:
      EpStore<TV>{(TV *)r.p[0]}(row, s);
      break;
    case kGrpResidual:
      EpResidual<TV>{(const TV *)r.p[0], (TV *)r.p[1]}(row, s);
      break;
    case kGrpAdd:
      EpAdd<TV>{(TV *)r.p[0]}(row, s);
      break;
    case kGrpChebFirst:
      EpChebFirst<TV>{(const TV *)r.p[0], (const TV *)r.p[1], (TV *)r.p[2], (TV *)r.p[3],
                      r.d0}(row, s);
      break;
    case kGrpChebIter:
      EpChebIter<TV>{(const TV *)r.p[0], (const TV *)r.p[1], (TV *)r.p[2], (TV *)r.p[3],
                     (TV *)r.p[4], r.d0, r.d1}(row, s);
      break;
    case kGrpChebLast:
      EpChebLast<TV>{(const TV *)r.p[0], (const TV *)r.p[1], (TV *)r.p[2], (TV *)r.p[3],
                     (TV *)r.p[4], r.d0, r.d1}(row, s);
      break;
    default:  // kGrpRestrictStart
      EpRestrictStart<TV>{(TV *)r.p[0], (const TV *)r.p[1], (TV *)r.p[2], (TV *)r.p[3],
                          (TV *)r.p[4], r.d0, (int)r.flag}(row, s);
      break;
  }
}

template <int G, typename T, typename TI, typename TV>
__global__ __launch_bounds__(kBlock) void k_grp_spmv(const GrpRec *__restrict__ recs) {
  const GrpRec &r = recs[blockIdx.y];
  constexpr int RPB = kBlock / G;
  int64_t row = (int64_t)blockIdx.x * RPB + threadIdx.x / G;
  int lane = threadIdx.x % G;
  if (row >= r.n) return;
  const double s = spmv_row_sum<G, T, TI, TV>(row, lane, r.ptr, (const TI *)r.idx,
                                              (const T *)r.val, (const TV *)r.x);
  if (lane == 0) grp_epilogue<TV>(r, row, s);
}

template <typename TV>
__global__ void k_grp_elem(const GrpRec *__restrict__ recs) {
  const GrpRec &r = recs[blockIdx.y];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r.n) return;
  switch (r.kind) {
    case kGrpChebZero: {
      const TV *b = (const TV *)r.p[0], *dinv = (const TV *)r.p[1];
      TV *x = (TV *)r.p[2], *rr = (TV *)r.p[3], *d = (TV *)r.p[4];
      double ri = (double)dinv[i] * (double)b[i];
      rr[i] = (TV)ri;
      double di = ri * r.d0;
      d[i] = (TV)di;
      x[i] = (TV)(r.flag ? 0.0 + di : 0.0);
      break;
    }
    case kGrpAddVec: {
      const TV *d = (const TV *)r.p[0];
      TV *x = (TV *)r.p[1];
      x[i] = (TV)((double)x[i] + (double)d[i]);
      break;
    }
    case kGrpToF32:
      ((float *)r.p[1])[i] = (float)(r.d0 * ((const double *)r.p[0])[i]);
      break;
    default:  // kGrpToF64
      ((double *)r.p[1])[i] = (double)((const float *)r.p[0])[i];
      break;
  }
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_grp_dense(const GrpRec *__restrict__ recs) {
  const GrpRec &r = recs[blockIdx.y];
  const int64_t n = r.n;
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double *__restrict__ row = (const double *)r.p[0] + i * n;
  const TV *__restrict__ b = (const TV *)r.p[1];
  double s = 0.0;
  for (int64_t j = lane; j < n; j += 64) s += row[j] * (double)b[j];
  s = wsum(s);
  if (lane == 0) ((TV *)r.p[2])[i] = (TV)s;
}

// ======================================================================
// C-ABI
// ======================================================================
template <typename TV>
static int multi_dot_any(int64_t n, int nv, const TV *V, int64_t ld, const double *w,
                         const double *cond, double eta2, double *work, double *out,
                         void *stream, const char *what, int with_self = 0) {
  FP3D_REQUIRE(nv >= 1 && nv <= 256 && V && w && work && out && ld >= n, what);
  hipStream_t s = FP3D_STREAM(stream);
  int64_t seg = mdot_seg(n);
  int nb = mdot_blocks(n, seg);
  // a short basis (few row segments) also splits its rows k over block
  // groups, so enough blocks keep loads in flight: 1,024 for float32 basis
  // vectors, 768 for float64 ones (whose wider loads need fewer; on
  // 815,425 rows the unsplit 797 blocks stream float64 faster than two
  // groups do)
  const int rows = nv + with_self, waves = kBlock / 64;
  const int min_blocks = sizeof(TV) == 8 ? kMdotMinBlocks64 : kMdotMinBlocks32;
  int ng = (min_blocks + nb - 1) / nb;
  const int ng_max = (rows + waves - 1) / waves;
  if (ng > ng_max) ng = ng_max;
  if (ng < 1) ng = 1;
  if (seg <= kMdotLds)
    hipLaunchKernelGGL((k_mdot_p1<true, TV>), dim3(nb, ng), dim3(kBlock), 0, s, n, nv,
                       V, ld, w, work, cond, eta2, seg, with_self);
  else
    hipLaunchKernelGGL((k_mdot_p1<false, TV>), dim3(nb, ng), dim3(kBlock), 0, s, n, nv,
                       V, ld, w, work, cond, eta2, seg, with_self);
  hipLaunchKernelGGL(k_mdot_p2, dim3(nv + with_self), dim3(kBlock), 0, s, nb, work,
                     out);
  return check_launch(what);
}

// ---------------------------------------------- allocation-free multi-dot
// Partial sums of fp3d_multi_dot_norm_noalloc live in this file-scope
// device array instead of a caller workspace: (kNoallocMaxNv + 1) rows of
// kReduceBlocks doubles = 2,105,344 bytes, held by every GPU the library
// runs on until the process ends.  Host code never passes `scratch` itself
// (in host code the name is a stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, once per device, and is cached in
// g_scratch_dev[device].
constexpr int kNoallocMaxNv = 256;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[(kNoallocMaxNv + 1) * kReduceBlocks];
static double *g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

// The current device's copy of `scratch` (looked up on the first call on
// that device, with that device current), or nullptr with the error set.
static double *scratch_on_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error(std::string("multi_dot_norm_noalloc: hipGetDevice: ") +
              hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("multi_dot_norm_noalloc: device ordinal beyond 63");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    double *p = nullptr;
    e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("multi_dot_norm_noalloc: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

template <typename TV>
static int multi_axpy_any(int64_t n, int nv, const TV *V, int64_t ld, const double *h,
                          double sign, double *w, void *stream, const char *what) {
  FP3D_REQUIRE(nv >= 1 && nv <= 256 && V && h && w && ld >= n, what);
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL((k_maxpy<TV>), dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, nv, V, ld, h, w, sign);
  return check_launch(what);
}

extern "C" {

int fp3d_csr_spmv(int64_t n_rows, int64_t nnz, const int32_t *indptr,
                  const int32_t *indices, const double *data, const double *x,
                  double *y, void *stream) {
  FP3D_REQUIRE(n_rows == 0 || (indptr && indices && data && x && y),
               "csr_spmv: NULL pointer");
  FP3D_REQUIRE(x != y, "csr_spmv: x and y must not alias");
  return launch_spmv<double, int32_t, double>(n_rows, nnz, indptr, indices, data, x,
                                              EpStore<double>{y},
                             FP3D_STREAM(stream));
}

int fp3d_csr_spmv_f32(int64_t n_rows, int64_t nnz, const int32_t *indptr,
                      const int32_t *indices, const float *data,
                      const double *x, double *y, void *stream) {
  FP3D_REQUIRE(n_rows == 0 || (indptr && indices && data && x && y),
               "csr_spmv_f32: NULL pointer");
  FP3D_REQUIRE(x != y, "csr_spmv_f32: x and y must not alias");
  return launch_spmv<float, int32_t, double>(n_rows, nnz, indptr, indices, data, x,
                                             EpStore<double>{y},
                            FP3D_STREAM(stream));
}

int64_t fp3d_term3_workspace_bytes(int64_t n_tets, int64_t n_nodes) {
  if (n_tets < 0 || n_nodes < 0) return -1;
  return (int64_t)t3_layout(n_tets, n_nodes).total;
}

int fp3d_term3_build(int64_t n_tets, int64_t n_nodes, const int32_t *elements,
                     const int32_t *inc_ptr, const int32_t *inc_ent,
                     const int32_t *inc_pos, const double *B,
                     const double *V, const double *grad, const double *gdotb,
                     const int32_t *upstream, const double *s_u,
                     const uint8_t *lim, const uint8_t *free_mask,
                     double inv_u2, void *t3, int64_t *n_active_out,
                     void *stream) {
  FP3D_REQUIRE(n_tets > 0 && n_nodes > 0, "term3_build: empty mesh");
  FP3D_REQUIRE(elements && inc_ptr && inc_ent && inc_pos && B && V && grad &&
                   upstream && s_u && t3,
               "term3_build: NULL pointer");
  FP3D_REQUIRE(FP3D_ALIGNED16(inc_pos), "inc_pos must be 16-byte aligned");
  FP3D_REQUIRE(((uintptr_t)gdotb & 31) == 0 && ((uintptr_t)t3 & 255) == 0,
               "term3_build: gdotb 32-byte / t3 256-byte alignment");
  hipStream_t s = FP3D_STREAM(stream);
  T3Layout L = t3_layout(n_tets, n_nodes);
  char *base = (char *)t3;
  uint8_t *flag8 = (uint8_t *)(base + L.flag8);
  int32_t *scan = (int32_t *)(base + L.scan);
  int32_t *bsum = (int32_t *)(base + L.bsum);
  int32_t *act = (int32_t *)(base + L.act);
  double *W = (double *)(base + L.W);
  double *Cc = (double *)(base + L.Cc);
  int32_t *cnt = (int32_t *)(base + L.cnt);
  int32_t *rowptr = (int32_t *)(base + L.rowptr);
  int32_t *ent = (int32_t *)(base + L.ent);
  uint8_t *fq = (uint8_t *)(base + L.fq);
  double *Wq = (double *)(base + L.Wq);
  int32_t *hdr = (int32_t *)(base + L.total - 256);
  if (gdotb)
    hipLaunchKernelGGL(k_t3_coef<true>, dim3(grid_for(n_tets)), dim3(kBlock), 0,
                       s, n_tets, elements, B, V, grad, gdotb, upstream, s_u,
                       lim, inv_u2, inc_pos, flag8, W, Cc);
  else
    hipLaunchKernelGGL(k_t3_coef<false>, dim3(grid_for(n_tets)), dim3(kBlock),
                       0, s, n_tets, elements, B, V, grad, gdotb, upstream,
                       s_u, lim, inv_u2, inc_pos, flag8, W, Cc);
  int rc = exclusive_scan(n_tets, flag8, scan, bsum, hdr, act, s);
  if (rc) return rc;
  const unsigned gn = xcd_grid((n_nodes + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_t3_count, dim3(gn), dim3(kBlock), 0, s, n_nodes,
                     (int64_t)gn, inc_ptr, inc_ent, flag8, free_mask, fq, cnt);
  rc = exclusive_scan(n_nodes, cnt, rowptr, bsum, rowptr + n_nodes, nullptr, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_t3_fill, dim3(gn), dim3(kBlock), 0, s, n_nodes,
                     (int64_t)gn, inc_ptr, inc_ent, fq, W, rowptr, ent, Wq);
  rc = check_launch("term3_build");
  if (rc) return rc;
  if (!n_active_out) return FP3D_OK;
  return fp3d_term3_active_count(t3, n_tets, n_nodes, n_active_out, stream);
}

int fp3d_term3_active_count(const void *t3, int64_t n_tets, int64_t n_nodes,
                            int64_t *n_active_out, void *stream) {
  FP3D_REQUIRE(t3 && n_active_out && n_tets > 0 && n_nodes > 0,
               "term3_active_count: bad argument");
  hipStream_t s = FP3D_STREAM(stream);
  const T3Layout L = t3_layout(n_tets, n_nodes);
  const int32_t *hdr = (const int32_t *)((const char *)t3 + L.total - 256);
  int32_t na = 0;
  FP3D_HIP(hipMemcpyAsync(&na, hdr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  FP3D_HIP(hipStreamSynchronize(s));
  *n_active_out = na;
  return FP3D_OK;
}

int fp3d_csr_spmv_term3(int64_t n_rows, int64_t nnz, const int32_t *indptr,
                        const void *indices, const void *data, int fmt,
                        const int32_t *elements, const int32_t *upstream,
                        const void *t3, int64_t n_active, int64_t n_tets,
                        const uint8_t *free_mask, const double *x, double *y,
                        void *stream) {
  FP3D_REQUIRE(n_rows > 0 && indptr && indices && data && x && y,
               "csr_spmv_term3: NULL pointer");
  FP3D_REQUIRE((fmt & ~(kFmtF32Values | kFmtIdx16)) == 0,
               "csr_spmv_term3: fmt takes bits 1 (float32 values) and 4 (int16 columns)");
  FP3D_REQUIRE(x != y, "csr_spmv_term3: x and y must not alias");
  hipStream_t s = FP3D_STREAM(stream);
  const int32_t *t3ptr = nullptr, *t3ent = nullptr;
  const double *W = nullptr, *tval = nullptr;
  if (t3 && n_active > 0) {
    FP3D_REQUIRE(elements && upstream, "csr_spmv_term3: Term 3 needs mesh");
    FP3D_REQUIRE(FP3D_ALIGNED16(elements), "elements must be 16-byte aligned");
    T3Layout L = t3_layout(n_tets, n_rows);
    const char *base = (const char *)t3;
    const int32_t *act = (const int32_t *)(base + L.act);
    const double *Cc = (const double *)(base + L.Cc);
    double *tv = (double *)(base + L.tval);
    hipLaunchKernelGGL(k_t3_tval, dim3(grid_for(n_active)), dim3(kBlock), 0, s,
                       n_active, act, elements, upstream, Cc, x, tv);
    t3ptr = (const int32_t *)(base + L.rowptr);
    t3ent = (const int32_t *)(base + L.ent);
    W = (const double *)(base + L.Wq);
    tval = tv;
  }
  int G = lanes_for(n_rows, nnz);
  unsigned grid = grid_for(n_rows, kBlock / G);
#define T3_LAUNCH(GG, T, TI)                                                 \
  hipLaunchKernelGGL((k_spmv_t3<GG, T, TI>), dim3(grid), dim3(kBlock), 0, s, \
                     n_rows, indptr, (const TI *)indices, (const T *)data, x, \
                     t3ptr, t3ent, W, tval, free_mask, y)
#define T3_SWITCH(T, TI)                       \
  switch (G) {                                 \
    case 4: T3_LAUNCH(4, T, TI); break;        \
    case 8: T3_LAUNCH(8, T, TI); break;        \
    case 16: T3_LAUNCH(16, T, TI); break;      \
    case 32: T3_LAUNCH(32, T, TI); break;      \
    default: T3_LAUNCH(64, T, TI); break;      \
  }
  const bool v32 = fmt & kFmtF32Values, i16 = fmt & kFmtIdx16;
  if (v32 && i16) {
    T3_SWITCH(float, int16_t)
  } else if (v32) {
    T3_SWITCH(float, int32_t)
  } else if (i16) {
    T3_SWITCH(double, int16_t)
  } else {
    T3_SWITCH(double, int32_t)
  }
#undef T3_SWITCH
#undef T3_LAUNCH
  return check_launch("csr_spmv_term3");
}

int fp3d_batch_job_bytes(void) { return (int)sizeof(BatchJob); }

int fp3d_batch_jacobian_matvec(int nb, const void *jobs, int64_t max_active,
                               int64_t n_rows, int64_t nnz, const int32_t *indptr,
                               const void *indices, int fmt, const int32_t *elements,
                               int64_t n_tets, int widen, void *stream) {
  FP3D_REQUIRE(nb >= 1 && nb <= kMaxBatch, "batch_jacobian_matvec: nb in [1, 1024]");
  FP3D_REQUIRE(jobs && n_rows > 0 && indptr && indices, "batch_jacobian_matvec: NULL");
  FP3D_REQUIRE((fmt & ~(kFmtF32Values | kFmtIdx16)) == 0,
               "batch_jacobian_matvec: fmt takes bits 1 and 4");
  FP3D_REQUIRE(max_active >= 0 && max_active <= n_tets,
               "batch_jacobian_matvec: max_active in [0, n_tets]");
  FP3D_REQUIRE(max_active == 0 || (elements && FP3D_ALIGNED16(elements)),
               "batch_jacobian_matvec: Term 3 needs 16-byte aligned elements");
  hipStream_t s = FP3D_STREAM(stream);
  const BatchJob *J = (const BatchJob *)jobs;
  if (widen)
    hipLaunchKernelGGL(k_b_widen, dim3(grid_for(n_rows), nb), dim3(kBlock), 0, s, n_rows,
                       J);
  const T3Layout L = t3_layout(n_tets, n_rows);
  if (max_active > 0)
    hipLaunchKernelGGL(k_b_t3_tval, dim3(grid_for(max_active), nb), dim3(kBlock), 0, s,
                       J, elements, L.act, L.Cc, L.tval);
  int G = lanes_for(n_rows, nnz);
  unsigned grid = grid_for(n_rows, kBlock / G);
#define B3_LAUNCH(GG, T, TI)                                                       \
  hipLaunchKernelGGL((k_b_spmv_t3<GG, T, TI>), dim3(grid, nb), dim3(kBlock), 0, s,  \
                     n_rows, indptr, (const TI *)indices, J, L.rowptr, L.ent, L.Wq, \
                     L.tval)
#define B3_SWITCH(T, TI)                  \
  switch (G) {                            \
    case 4: B3_LAUNCH(4, T, TI); break;   \
    case 8: B3_LAUNCH(8, T, TI); break;   \
    case 16: B3_LAUNCH(16, T, TI); break; \
    case 32: B3_LAUNCH(32, T, TI); break; \
    default: B3_LAUNCH(64, T, TI); break; \
  }
  const bool v32 = fmt & kFmtF32Values, i16 = fmt & kFmtIdx16;
  if (v32 && i16) {
    B3_SWITCH(float, int16_t)
  } else if (v32) {
    B3_SWITCH(float, int32_t)
  } else if (i16) {
    B3_SWITCH(double, int16_t)
  } else {
    B3_SWITCH(double, int32_t)
  }
#undef B3_SWITCH
#undef B3_LAUNCH
  return check_launch("batch_jacobian_matvec");
}

}  // extern "C"

int fp3d::batch_masked_norms(int nb, const EvalJob *jobs, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_be_mask, dim3(grid_for(n), nb), dim3(kBlock), 0, s, n, jobs);
  const int rb = rblocks(n);
  hipLaunchKernelGGL(k_be_norms_p1, dim3(rb, nb), dim3(kBlock), 0, s, n, jobs);
  hipLaunchKernelGGL(k_be_norms_p2, dim3(1, nb), dim3(kBlock), 0, s, rb, jobs);
  return check_launch("batch_masked_norms");
}

template <typename TV>
static void batch_ortho_t(int nb, const BatchJob *J, int max_nv, int64_t n, double eta2,
                          hipStream_t s) {
  const int64_t seg = mdot_seg(n);
  const int mb = mdot_blocks(n, seg);
  const unsigned gn = grid_for(n);
  for (int pass = 0; pass < 2; ++pass) {
    if (seg <= kMdotLds)
      hipLaunchKernelGGL((k_b_mdot_p1<true, TV>), dim3(mb, nb), dim3(kBlock), 0, s, n, seg,
                         pass, eta2, J);
    else
      hipLaunchKernelGGL((k_b_mdot_p1<false, TV>), dim3(mb, nb), dim3(kBlock), 0, s, n, seg,
                         pass, eta2, J);
    hipLaunchKernelGGL(k_b_mdot_p2, dim3(max_nv + 1, nb), dim3(kBlock), 0, s, mb, pass, J);
    hipLaunchKernelGGL((k_b_maxpy<TV>), dim3(gn, nb), dim3(kBlock), 0, s, n, pass, J);
  }
  const int rb = rblocks(n);
  hipLaunchKernelGGL(k_b_norms_p1, dim3(rb, nb), dim3(kBlock), 0, s, n, J);
  hipLaunchKernelGGL(k_b_norms_p2, dim3(1, nb), dim3(kBlock), 0, s, rb, J);
  hipLaunchKernelGGL((k_b_scale_inv<TV>), dim3(gn, nb), dim3(kBlock), 0, s, n, J);
}

extern "C" {

int fp3d_batch_orthonormalise(int nb, const void *jobs, int max_nv, int64_t n, int f32,
                              double eta2, void *stream) {
  FP3D_REQUIRE(nb >= 1 && nb <= kMaxBatch, "batch_orthonormalise: nb in [1, 1024]");
  FP3D_REQUIRE(jobs && n > 0, "batch_orthonormalise: NULL");
  FP3D_REQUIRE(max_nv >= 1 && max_nv <= 256, "batch_orthonormalise: max_nv in [1, 256]");
  FP3D_REQUIRE(eta2 > 0.0 && eta2 <= 1.0, "batch_orthonormalise: eta2 in (0, 1]");
  hipStream_t s = FP3D_STREAM(stream);
  if (f32)
    batch_ortho_t<float>(nb, (const BatchJob *)jobs, max_nv, n, eta2, s);
  else
    batch_ortho_t<double>(nb, (const BatchJob *)jobs, max_nv, n, eta2, s);
  return check_launch("batch_orthonormalise");
}

int fp3d_csr_mask_dirichlet(int64_t n_rows, const int32_t *indptr,
                            const int32_t *indices, const uint8_t *free_mask,
                            double *data, float *data32, void *stream) {
  FP3D_REQUIRE(n_rows == 0 || (indptr && indices && free_mask && data),
               "csr_mask_dirichlet: NULL pointer");
  if (n_rows == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_mask_dirichlet, dim3(grid_for(n_rows)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n_rows, indptr, indices, free_mask,
                     data, data32);
  return check_launch("csr_mask_dirichlet");
}

int fp3d_lumped_volume(int64_t n_nodes, const int32_t *inc_ptr,
                       const int32_t *inc_ent, const double *V, double *out,
                       void *stream) {
  FP3D_REQUIRE(n_nodes == 0 || (inc_ptr && inc_ent && V && out),
               "lumped_volume: NULL pointer");
  if (n_nodes == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_lumped_volume, dim3(grid_for(n_nodes)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n_nodes, inc_ptr, inc_ent, V, out);
  return check_launch("lumped_volume");
}

int fp3d_csr_diag_shift(int64_t n_rows, const int32_t *diag_pos,
                        const double *w, double scale,
                        const uint8_t *free_mask, double *data, void *stream) {
  FP3D_REQUIRE(n_rows == 0 || (diag_pos && w && free_mask && data),
               "csr_diag_shift: NULL pointer");
  if (n_rows == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_diag_shift, dim3(grid_for(n_rows)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n_rows, diag_pos, w, scale, free_mask,
                     data);
  return check_launch("csr_diag_shift");
}

int fp3d_csr_to_f32(int64_t nnz, const double *data, float *out, void *stream) {
  FP3D_REQUIRE(nnz == 0 || (data && out), "csr_to_f32: NULL pointer");
  if (nnz == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(nnz)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), nnz, data, out);
  return check_launch("csr_to_f32");
}

int fp3d_axpby(int64_t n, double a, const double *x, double b, double *y,
               void *stream) {
  FP3D_REQUIRE(n == 0 || (x && y), "axpby: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, a, x, b, y);
  return check_launch("axpby");
}

int fp3d_scale(int64_t n, double a, const double *x, double *y, void *stream) {
  FP3D_REQUIRE(n == 0 || (x && y), "scale: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, a, x, y);
  return check_launch("scale");
}

int fp3d_mask_copy(int64_t n, const uint8_t *mask, const double *x, double *y,
                   void *stream) {
  FP3D_REQUIRE(n == 0 || (mask && x && y), "mask_copy: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_mask_copy, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, mask, x, y);
  return check_launch("mask_copy");
}

int fp3d_multi_dot(int64_t n, int nv, const double *V, int64_t ld,
                   const double *w, double *work, double *out, void *stream) {
  return multi_dot_any(n, nv, V, ld, w, nullptr, 0.0, work, out, stream,
                       "multi_dot: bad arguments");
}

int fp3d_multi_dot_norm(int64_t n, int nv, const double *V, int64_t ld,
                        const double *w, double *work, double *out,
                        void *stream) {
  return multi_dot_any<double>(n, nv, V, ld, w, nullptr, 0.0, work, out, stream,
                               "multi_dot_norm: bad arguments", 1);
}

int fp3d_multi_dot_norm_f32v(int64_t n, int nv, const float *V, int64_t ld,
                             const double *w, double *work, double *out,
                             void *stream) {
  return multi_dot_any<float>(n, nv, V, ld, w, nullptr, 0.0, work, out, stream,
                              "multi_dot_norm_f32v: bad arguments", 1);
}

int fp3d_multi_dot_norm_noalloc(int64_t n, int nv, const double *V, int64_t ld,
                                const double *w, double *out, void *stream) {
  // every partial the second pass reads, (nv + 1) x nb with nb <=
  // kReduceBlocks, is written by the first pass of the same call, so no
  // clearing is needed; nv is checked here so the rows always fit
  FP3D_REQUIRE(nv >= 1 && nv <= kNoallocMaxNv,
               "multi_dot_norm_noalloc: nv must be in [1, 256] (scratch holds "
               "257 x 1024 partial sums)");
  double *work = scratch_on_current_device();
  if (!work) return FP3D_EHIP;
  return multi_dot_any<double>(n, nv, V, ld, w, nullptr, 0.0, work, out, stream,
                               "multi_dot_norm_noalloc: bad arguments", 1);
}

int fp3d_multi_dot_f32v(int64_t n, int nv, const float *V, int64_t ld,
                        const double *w, double *work, double *out, void *stream) {
  return multi_dot_any(n, nv, V, ld, w, nullptr, 0.0, work, out, stream,
                       "multi_dot_f32v: bad arguments");
}

int fp3d_multi_dot_reorth(int64_t n, int nv, const double *V, int64_t ld,
                          const double *w, const double *cond, double eta2,
                          double *work, double *out, void *stream) {
  FP3D_REQUIRE(cond && eta2 > 0.0 && eta2 <= 1.0,
               "multi_dot_reorth: cond and eta2 in (0, 1] required");
  return multi_dot_any(n, nv, V, ld, w, cond, eta2, work, out, stream,
                       "multi_dot_reorth: bad arguments");
}

int fp3d_multi_dot_reorth_f32v(int64_t n, int nv, const float *V, int64_t ld,
                               const double *w, const double *cond, double eta2,
                               double *work, double *out, void *stream) {
  FP3D_REQUIRE(cond && eta2 > 0.0 && eta2 <= 1.0,
               "multi_dot_reorth_f32v: cond and eta2 in (0, 1] required");
  return multi_dot_any(n, nv, V, ld, w, cond, eta2, work, out, stream,
                       "multi_dot_reorth_f32v: bad arguments");
}

int fp3d_multi_dot_work_doubles(int nv) { return nv * kReduceBlocks; }

int fp3d_multi_axpy(int64_t n, int nv, const double *V, int64_t ld,
                    const double *h, double sign, double *w, void *stream) {
  return multi_axpy_any(n, nv, V, ld, h, sign, w, stream, "multi_axpy: bad arguments");
}

int fp3d_multi_axpy_f32v(int64_t n, int nv, const float *V, int64_t ld,
                         const double *h, double sign, double *w, void *stream) {
  return multi_axpy_any(n, nv, V, ld, h, sign, w, stream,
                        "multi_axpy_f32v: bad arguments");
}

int fp3d_scale_to_f32(int64_t n, double a, const double *x, float *y, void *stream) {
  FP3D_REQUIRE(n == 0 || (x && y), "scale_to_f32: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_scale_cvt, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, a, x, y);
  return check_launch("scale_to_f32");
}

int fp3d_scale_inv_dev(int64_t n, const double *nrm, const double *x, double *y,
                       void *stream) {
  FP3D_REQUIRE(n == 0 || (nrm && x && y), "scale_inv_dev: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_scale_inv_dev<double>, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, nrm, x, y);
  return check_launch("scale_inv_dev");
}

int fp3d_scale_inv_dev_to_f32(int64_t n, const double *nrm, const double *x,
                              float *y, void *stream) {
  FP3D_REQUIRE(n == 0 || (nrm && x && y), "scale_inv_dev_to_f32: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_scale_inv_dev<float>, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, nrm, x, y);
  return check_launch("scale_inv_dev_to_f32");
}

int fp3d_f32_to_f64(int64_t n, const float *x, double *y, void *stream) {
  FP3D_REQUIRE(n == 0 || (x && y), "f32_to_f64: NULL pointer");
  if (n == 0) return FP3D_OK;
  hipLaunchKernelGGL(k_widen, dim3(grid_for(n)), dim3(kBlock), 0,
                     FP3D_STREAM(stream), n, x, y);
  return check_launch("f32_to_f64");
}

int fp3d_norms(int64_t n, const double *x, double *work, double *out2,
               void *stream) {
  FP3D_REQUIRE(x && work && out2, "norms: NULL pointer");
  hipStream_t s = FP3D_STREAM(stream);
  int nb = rblocks(n);
  hipLaunchKernelGGL(k_norms_p1, dim3(nb), dim3(kBlock), 0, s, n, x, work);
  hipLaunchKernelGGL(k_norms_p2, dim3(1), dim3(kBlock), 0, s, nb, work, out2);
  return check_launch("norms");
}

int fp3d_amg_cheb_smooth(int64_t n, int64_t nnz, const int32_t *indptr,
                         const void *indices, const void *data, int fmt,
                         const void *dinv, const void *b, void *x, void *r,
                         void *d0, void *d1, int degree, double lmin, double lmax,
                         int zero_init, void *stream) {
  FP3D_REQUIRE(n > 0 && indptr && indices && data && dinv && b && x && r &&
                   d0 && d1 && degree >= 1 && lmax > lmin && lmin > 0.0,
               "amg_cheb_smooth: bad arguments");
  FP3D_REQUIRE(zero_init >= 0 && zero_init <= 2,
               "amg_cheb_smooth: zero_init must be 0, 1 or 2");
  FP3D_REQUIRE((fmt & ~7) == 0, "amg_cheb_smooth: fmt takes bits 1, 2 and 4");
  hipStream_t s = FP3D_STREAM(stream);
  if (fmt & kFmtF32Vectors)
    return cheb_smooth_t<float>(n, nnz, indptr, indices, data, fmt, (const float *)dinv,
                                (const float *)b, (float *)x, (float *)r, (float *)d0,
                                (float *)d1, degree, lmin, lmax, zero_init, s);
  return cheb_smooth_t<double>(n, nnz, indptr, indices, data, fmt, (const double *)dinv,
                               (const double *)b, (double *)x, (double *)r,
                               (double *)d0, (double *)d1, degree, lmin, lmax,
                               zero_init, s);
}

int fp3d_amg_restrict_start(int64_t nc, int64_t nnz, const int32_t *indptr,
                            const void *indices, const void *data, int fmt,
                            const void *r, void *bc, const void *dinv_c,
                            int degree_c, double lmin_c, double lmax_c, void *x_c,
                            void *r_c, void *d_c, void *stream) {
  FP3D_REQUIRE(nc > 0 && indptr && indices && data && r && bc && dinv_c &&
                   x_c && r_c && d_c && degree_c >= 1 && lmax_c > lmin_c &&
                   lmin_c > 0.0,
               "amg_restrict_start: bad arguments");
  FP3D_REQUIRE((fmt & ~7) == 0, "amg_restrict_start: fmt takes bits 1, 2 and 4");
  const double theta = 0.5 * (lmax_c + lmin_c);
  const int xd = degree_c == 1 ? 1 : 0;
  if (fmt & kFmtF32Vectors)
    return spmv_fmt(nc, nnz, indptr, indices, data, fmt, (const float *)r,
                    EpRestrictStart<float>{(float *)bc, (const float *)dinv_c,
                                           (float *)x_c, (float *)r_c, (float *)d_c,
                                           1.0 / theta, xd},
                    FP3D_STREAM(stream));
  return spmv_fmt(nc, nnz, indptr, indices, data, fmt, (const double *)r,
                  EpRestrictStart<double>{(double *)bc, (const double *)dinv_c,
                                          (double *)x_c, (double *)r_c, (double *)d_c,
                                          1.0 / theta, xd},
                  FP3D_STREAM(stream));
}

int fp3d_amg_residual(int64_t n, int64_t nnz, const int32_t *indptr,
                      const void *indices, const void *data, int fmt,
                      const void *b, const void *x, void *r, void *stream) {
  FP3D_REQUIRE(n > 0 && indptr && indices && data && b && x && r,
               "amg_residual: bad arguments");
  FP3D_REQUIRE((fmt & ~7) == 0, "amg_residual: fmt takes bits 1, 2 and 4");
  if (fmt & kFmtF32Vectors)
    return spmv_fmt(n, nnz, indptr, indices, data, fmt, (const float *)x,
                    EpResidual<float>{(const float *)b, (float *)r}, FP3D_STREAM(stream));
  return spmv_fmt(n, nnz, indptr, indices, data, fmt, (const double *)x,
                  EpResidual<double>{(const double *)b, (double *)r}, FP3D_STREAM(stream));
}

int fp3d_amg_restrict(int64_t nc, int64_t nnz, const int32_t *indptr,
                      const void *indices, const void *data, int fmt,
                      const void *r, void *bc, void *stream) {
  FP3D_REQUIRE(nc > 0 && indptr && indices && data && r && bc,
               "amg_restrict: bad arguments");
  FP3D_REQUIRE((fmt & ~7) == 0, "amg_restrict: fmt takes bits 1, 2 and 4");
  if (fmt & kFmtF32Vectors)
    return spmv_fmt(nc, nnz, indptr, indices, data, fmt, (const float *)r,
                    EpStore<float>{(float *)bc}, FP3D_STREAM(stream));
  return spmv_fmt(nc, nnz, indptr, indices, data, fmt, (const double *)r,
                  EpStore<double>{(double *)bc}, FP3D_STREAM(stream));
}

int fp3d_amg_prolong_add(int64_t n, int64_t nnz, const int32_t *indptr,
                         const void *indices, const void *data, int fmt,
                         const void *xc, void *x, void *stream) {
  FP3D_REQUIRE(n > 0 && indptr && indices && data && xc && x,
               "amg_prolong_add: bad arguments");
  FP3D_REQUIRE((fmt & ~7) == 0, "amg_prolong_add: fmt takes bits 1, 2 and 4");
  if (fmt & kFmtF32Vectors)
    return spmv_fmt(n, nnz, indptr, indices, data, fmt, (const float *)xc,
                    EpAdd<float>{(float *)x}, FP3D_STREAM(stream));
  return spmv_fmt(n, nnz, indptr, indices, data, fmt, (const double *)xc,
                  EpAdd<double>{(double *)x}, FP3D_STREAM(stream));
}

int fp3d_dense_matvec_v(int fmt, int64_t n, const double *M, const void *b, void *x,
                        void *stream) {
  FP3D_REQUIRE(n > 0 && M && b && x, "dense_matvec_v: bad arguments");
  FP3D_REQUIRE((fmt & ~kFmtF32Vectors) == 0, "dense_matvec_v: fmt takes bit 2 only");
  const int64_t rows_per_block = kBlock / 64;
  const unsigned grid = (unsigned)((n + rows_per_block - 1) / rows_per_block);
  if (fmt & kFmtF32Vectors)
    hipLaunchKernelGGL(k_dense_mv<float>, dim3(grid), dim3(kBlock), 0,
                       FP3D_STREAM(stream), n, M, (const float *)b, (float *)x);
  else
    hipLaunchKernelGGL(k_dense_mv<double>, dim3(grid), dim3(kBlock), 0,
                       FP3D_STREAM(stream), n, M, (const double *)b, (double *)x);
  return check_launch("dense_matvec_v");
}

int fp3d_arnoldi_step(const int64_t *a, int col, void *stream) {
  FP3D_REQUIRE(a && col >= 0, "arnoldi_step: bad arguments");
  const int64_t n = a[0];
  const bool f32 = a[1] != 0;
  const int64_t ld = a[3];
  const int k = col + 1;
  FP3D_REQUIRE(n > 0 && k <= 256, "arnoldi_step: n > 0 and col <= 255");
  hipStream_t s = FP3D_STREAM(stream);
  double *v64 = (double *)a[4], *tmp = (double *)a[5];
  double *w = (double *)a[6 + (col & 1)];
  double *o = (double *)a[8 + (col & 1)];
  double *work = (double *)a[10];
  double *hb = (double *)a[11 + (col & 1)];
  const void *gexec = (const void *)a[24 + (col & 1)];
  int rc;
  // w = M^-1 A v_col
  const double *x;
  if (f32) {
    const float *V = (const float *)a[2];
    rc = fp3d_f32_to_f64(n, V + (int64_t)col * ld, v64, stream);
    if (rc) return rc;
    x = v64;
  } else {
    x = (const double *)a[2] + (int64_t)col * ld;
  }
  rc = fp3d_csr_spmv_term3(n, a[13], (const int32_t *)a[14], (const void *)a[15],
                           (const void *)a[16], (int)a[17], (const int32_t *)a[18],
                           (const int32_t *)a[19], (const void *)a[20], a[21], a[22],
                           (const uint8_t *)a[23], x, tmp, stream);
  if (rc) return rc;
  if (a[26]) {
    FP3D_REQUIRE(gexec, "arnoldi_step: no V-cycle graph for this column's parity");
    FP3D_HIP(hipGraphLaunch((hipGraphExec_t)gexec, s));
  } else {
    FP3D_HIP(hipMemcpyAsync(w, tmp, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  // h = V^T w, ||w||^2; w -= V h; the conditional second pass; ||w||; v_{col+1}
  if (f32) {
    const float *V = (const float *)a[2];
    if ((rc = fp3d_multi_dot_norm_f32v(n, k, V, ld, w, work, o + 2, stream))) return rc;
    if ((rc = fp3d_multi_axpy_f32v(n, k, V, ld, o + 2, -1.0, w, stream))) return rc;
    if ((rc = fp3d_multi_dot_reorth_f32v(n, k, V, ld, w, o + 2, 0.5, work, o + 3 + k,
                                         stream)))
      return rc;
    if ((rc = fp3d_multi_axpy_f32v(n, k, V, ld, o + 3 + k, -1.0, w, stream))) return rc;
    if ((rc = fp3d_norms(n, w, work, o, stream))) return rc;
    if ((rc = fp3d_scale_inv_dev_to_f32(n, o, w, (float *)a[2] + (int64_t)(col + 1) * ld,
                                        stream)))
      return rc;
  } else {
    const double *V = (const double *)a[2];
    if ((rc = fp3d_multi_dot_norm(n, k, V, ld, w, work, o + 2, stream))) return rc;
    if ((rc = fp3d_multi_axpy(n, k, V, ld, o + 2, -1.0, w, stream))) return rc;
    if ((rc = fp3d_multi_dot_reorth(n, k, V, ld, w, o + 2, 0.5, work, o + 3 + k, stream)))
      return rc;
    if ((rc = fp3d_multi_axpy(n, k, V, ld, o + 3 + k, -1.0, w, stream))) return rc;
    if ((rc = fp3d_norms(n, w, work, o, stream))) return rc;
    if ((rc = fp3d_scale_inv_dev(n, o, w, (double *)a[2] + (int64_t)(col + 1) * ld,
                                 stream)))
      return rc;
  }
  // the step's coefficients to pinned host memory
  FP3D_HIP(hipMemcpyAsync(hb, o, (size_t)(3 + 2 * k) * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  return FP3D_OK;
}

int fp3d_grp_vcycle_op(int nb, const void *recs, int op, int64_t max_n, int G, int fmt,
                       void *stream) {
  FP3D_REQUIRE(nb >= 1 && nb <= kMaxBatch, "grp_vcycle_op: nb in [1, 1024]");
  FP3D_REQUIRE(recs && max_n > 0, "grp_vcycle_op: bad arguments");
  FP3D_REQUIRE(op >= 0 && op <= 2, "grp_vcycle_op: op is 0 (SpMV), 1 (element-wise) or 2 (dense)");
  FP3D_REQUIRE((fmt & ~7) == 0, "grp_vcycle_op: fmt takes bits 1, 2 and 4");
  hipStream_t s = FP3D_STREAM(stream);
  const GrpRec *R = (const GrpRec *)recs;
  const bool v32 = fmt & kFmtF32Vectors;
  if (op == 1) {
    if (v32)
      hipLaunchKernelGGL(k_grp_elem<float>, dim3(grid_for(max_n), nb), dim3(kBlock), 0, s, R);
    else
      hipLaunchKernelGGL(k_grp_elem<double>, dim3(grid_for(max_n), nb), dim3(kBlock), 0, s, R);
    return check_launch("grp_vcycle_op");
  }
  if (op == 2) {
    const int64_t rpb = kBlock / 64;
    const unsigned grid = (unsigned)((max_n + rpb - 1) / rpb);
    if (v32)
      hipLaunchKernelGGL(k_grp_dense<float>, dim3(grid, nb), dim3(kBlock), 0, s, R);
    else
      hipLaunchKernelGGL(k_grp_dense<double>, dim3(grid, nb), dim3(kBlock), 0, s, R);
    return check_launch("grp_vcycle_op");
  }
  FP3D_REQUIRE(G == 4 || G == 8 || G == 16 || G == 32 || G == 64,
               "grp_vcycle_op: G must be 4, 8, 16, 32 or 64");
  const unsigned grid = grid_for(max_n, kBlock / G);
  const bool f32v = fmt & kFmtF32Values, i16 = fmt & kFmtIdx16;
#define GRP_L(GG, T, TI, TV) \
  hipLaunchKernelGGL((k_grp_spmv<GG, T, TI, TV>), dim3(grid, nb), dim3(kBlock), 0, s, R)
#define GRP_G(T, TI, TV)                     \
  switch (G) {                               \
    case 4: GRP_L(4, T, TI, TV); break;      \
    case 8: GRP_L(8, T, TI, TV); break;      \
    case 16: GRP_L(16, T, TI, TV); break;    \
    case 32: GRP_L(32, T, TI, TV); break;    \
    default: GRP_L(64, T, TI, TV); break;    \
  }
#define GRP_V(T, TI)             \
  if (v32) {                     \
    GRP_G(T, TI, float)          \
  } else {                       \
    GRP_G(T, TI, double)         \
  }
  if (f32v && i16) {
    GRP_V(float, int16_t)
  } else if (f32v) {
    GRP_V(float, int32_t)
  } else if (i16) {
    GRP_V(double, int16_t)
  } else {
    GRP_V(double, int32_t)
  }
#undef GRP_V
#undef GRP_G
#undef GRP_L
  return check_launch("grp_vcycle_op");
}

int fp3d_grp_vcycle_run(int nb, const void *recs, int n_ops, const int64_t *ops,
                        void *stream) {
  FP3D_REQUIRE(recs && n_ops >= 0 && (n_ops == 0 || ops), "grp_vcycle_run: bad arguments");
  FP3D_REQUIRE(nb >= 1 && nb <= kMaxBatch, "grp_vcycle_run: nb in [1, 1024]");
  const char *base = (const char *)recs;
  for (int i = 0; i < n_ops; ++i) {
    const int64_t *o = ops + 4 * (int64_t)i;
    const int rc = fp3d_grp_vcycle_op(nb, base + (size_t)i * nb * sizeof(GrpRec), (int)o[0],
                                      o[1], (int)o[2], (int)o[3], stream);
    if (rc) return rc;
  }
  return FP3D_OK;
}

int fp3d_dense_matvec(int64_t n, const double *M, const double *b, double *x,
                      void *stream) {
  FP3D_REQUIRE(n > 0 && M && b && x, "dense_matvec: bad arguments");
  const int64_t rows_per_block = kBlock / 64;
  hipLaunchKernelGGL(k_dense_mv<double>, dim3((n + rows_per_block - 1) / rows_per_block),
                     dim3(kBlock), 0, FP3D_STREAM(stream), n, M, b, x);
  return check_launch("dense_matvec");
}

}  // extern "C"
