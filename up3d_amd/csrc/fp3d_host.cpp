// Host-side, once-per-mesh set-up of the full-potential solve: element
// colouring, CSR pattern and scatter map, face adjacency, element geometry,
// and the aggregation step of the smoothed-aggregation AMG set-up.
// Plain C++ (no device code); exported through the same C-ABI as the
// kernels.
#include <algorithm>
#include <array>
#include <memory>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#include <climits>
#include <cstdlib>
#include <thread>

#include "../../include/fp3d.h"

namespace fp3d {
void set_error(const std::string &msg);
}

#define HREQ(cond, msg)      \
  do {                       \
    if (!(cond)) {           \
      fp3d::set_error(msg);  \
      return FP3D_EINVAL;    \
    }                        \
  } while (0)

namespace {

// node -> incident elements, each list in ascending element order
// (pyfp3d/mesh/coloring.py:45-62 node_to_element_csr: stable argsort)
void node_incidence(int64_t n_tets, const int32_t *el, int64_t n_nodes,
                    std::vector<int64_t> &off, std::vector<int32_t> &lst) {
  off.assign(n_nodes + 1, 0);
  for (int64_t e = 0; e < n_tets; ++e)
    for (int i = 0; i < 4; ++i) off[el[4 * e + i] + 1]++;
  for (int64_t v = 0; v < n_nodes; ++v) off[v + 1] += off[v];
  lst.resize(off[n_nodes]);
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < n_tets; ++e)
    for (int i = 0; i < 4; ++i) lst[pos[el[4 * e + i]]++] = (int32_t)e;
}

int check_elements(int64_t n_tets, const int32_t *el, int64_t n_nodes) {
  HREQ(n_tets >= 0 && (n_tets == 0 || el), "elements missing");
  for (int64_t k = 0; k < 4 * n_tets; ++k)
    HREQ(el[k] >= 0 && el[k] < n_nodes, "element node id out of range");
  return FP3D_OK;
}

// Host worker threads for the row-parallel set-up products: at most 16, or
// FP3D_HOST_THREADS.  Each output row is produced by one thread in a fixed
// order, so results do not depend on the thread count.
unsigned host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 1;
  if (t > 16) t = 16;
  if (const char *e = std::getenv("FP3D_HOST_THREADS")) {
    int v = std::atoi(e);
    if (v >= 1) t = (unsigned)v;
  }
  return t;
}

// Rows split over threads in proportion to the work (stored entries of
// the rows): a thread costs tens of microseconds to start, so one thread
// per 256K entries at most.
template <class F>
void parallel_rows(int64_t n, int64_t work, F f) {
  int64_t want = work / (1 << 18);
  unsigned T = host_threads();
  if (want < (int64_t)T) T = want < 1 ? 1u : (unsigned)want;
  if (T <= 1 || n < 2 * (int64_t)T) {
    f(0, n, 0u);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  for (unsigned t = 0; t < T; ++t)
    th.emplace_back(f, n * (int64_t)t / T, n * (int64_t)(t + 1) / T, t);
  for (auto &x : th) x.join();
}

// Multiply-adds of C = A B, estimated for the thread count: A's entries
// times the mean length of the B rows they select, sampled at up to 1,024
// evenly spaced entries of A (an O(1) estimate; the products themselves do
// not depend on it).
int64_t product_work(int64_t n_rows, const int32_t *Ap, const int32_t *Aj,
                     const int32_t *Bp) {
  const int64_t nnz = n_rows ? (int64_t)Ap[n_rows] : 0;
  if (nnz == 0) return 0;
  const int64_t samples = nnz < 1024 ? nnz : 1024;
  int64_t len = 0;
  for (int64_t t = 0; t < samples; ++t) {
    const int32_t j = Aj[t * nnz / samples];
    len += Bp[j + 1] - Bp[j];
  }
  return nnz * (len / samples > 1 ? len / samples : 1);
}

// One row of C = A B accumulated as scipy.sparse's csr_matmat does (A
// entries in stored order, then the B entries of that row, sums from 0);
// emit(k, value) is called for the nonzero sums in ascending column k.
struct RowProduct {
  std::vector<double> sums;
  std::vector<uint8_t> seen;
  std::vector<int32_t> touched;
  explicit RowProduct(int64_t n_cols) : sums(n_cols, 0.0), seen(n_cols, 0) {}
  template <class Emit>
  int32_t run(int64_t i, const int32_t *Ap, const int32_t *Aj, const double *Ax,
              const int32_t *Bp, const int32_t *Bj, const double *Bx, Emit emit) {
    touched.clear();
    for (int32_t jj = Ap[i]; jj < Ap[i + 1]; ++jj) {
      const int32_t j = Aj[jj];
      const double v = Ax[jj];
      for (int32_t kk = Bp[j]; kk < Bp[j + 1]; ++kk) {
        const int32_t k = Bj[kk];
        sums[k] += v * Bx[kk];
        if (!seen[k]) {
          seen[k] = 1;
          touched.push_back(k);
        }
      }
    }
    std::sort(touched.begin(), touched.end());
    int32_t cnt = 0;
    for (const int32_t k : touched) {
      if (sums[k] != 0.0) emit(cnt++, k, sums[k]);
      sums[k] = 0.0;
      seen[k] = 0;
    }
    return cnt;
  }
};

// Row i of the smoothed prolongator P = T - diag(s) (A T) of an aggregation
// (T: one entry tval[agg[j]] in column agg[j] of row j, none for agg[j] < 0),
// with the operations of the scipy recipe it replaces: (A T) row sums from 0
// in A's stored order, exact zeros dropped; X = s_i (A T); P = T - X merged
// by ascending column (t - x where both are stored, t or 0 - x where one
// is), exact zeros dropped.  emit(c, k, v) for the c-th entry, column k.
struct ProlongRow {
  std::vector<double> sums;
  std::vector<uint8_t> seen;
  std::vector<int32_t> touched;
  explicit ProlongRow(int64_t n_agg) : sums(n_agg, 0.0), seen(n_agg, 0) {}
  template <class Emit>
  int32_t run(int64_t i, const int32_t *Ap, const int32_t *Aj, const double *Ax,
              const int32_t *agg, const double *tval, const double *s, Emit emit) {
    touched.clear();
    for (int32_t jj = Ap[i]; jj < Ap[i + 1]; ++jj) {
      const int32_t a = agg[Aj[jj]];
      if (a < 0) continue;
      sums[a] += Ax[jj] * tval[a];
      if (!seen[a]) {
        seen[a] = 1;
        touched.push_back(a);
      }
    }
    std::sort(touched.begin(), touched.end());
    const int32_t ti = agg[i];
    const double t = ti >= 0 ? tval[ti] : 0.0;
    bool t_done = ti < 0;
    int32_t cnt = 0;
    auto put = [&](int32_t k, double v) {
      if (v != 0.0) emit(cnt++, k, v);
    };
    for (const int32_t k : touched) {
      const double at = sums[k];
      sums[k] = 0.0;
      seen[k] = 0;
      if (!t_done && ti < k) {
        put(ti, t);
        t_done = true;
      }
      const double x = at != 0.0 ? s[i] * at : 0.0;
      if (x == 0.0) {
        if (!t_done && ti == k) {
          put(ti, t);
          t_done = true;
        }
        continue;
      }
      if (!t_done && ti == k) {
        put(k, t - x);
        t_done = true;
      } else {
        put(k, 0.0 - x);
      }
    }
    if (!t_done) put(ti, t);
    return cnt;
  }
};

}  // namespace

extern "C" {

/* Row-parallel sparse product C = A B on the host, in two passes: the
 * first writes the row pointer of C (caller array, n_rows + 1) and its
 * entry count, the second fills the caller's indices and data.  Row i is
 * accumulated exactly as scipy.sparse's csr_matmat does (A entries in
 * stored order, then B entries, sums from 0) and exact zeros are dropped,
 * so C carries scipy's values; columns ascend. */
int fp3d_host_spgemm(int64_t n_rows, int64_t n_cols, const int32_t *Ap,
                     const int32_t *Aj, const double *Ax, const int32_t *Bp,
                     const int32_t *Bj, const double *Bx, int32_t *indptr,
                     int64_t *nnz_out) {
  HREQ(n_rows >= 0 && n_cols >= 0 && indptr && nnz_out, "spgemm: bad arguments");
  HREQ(n_rows == 0 || (Ap && Bp), "spgemm: NULL pointer");
  indptr[0] = 0;
  parallel_rows(n_rows, product_work(n_rows, Ap, Aj, Bp),
                [&](int64_t lo, int64_t hi, unsigned) {
    RowProduct rp(n_cols);
    for (int64_t i = lo; i < hi; ++i)
      indptr[i + 1] = rp.run(i, Ap, Aj, Ax, Bp, Bj, Bx,
                             [](int32_t, int32_t, double) {});
  });
  int64_t nnz = 0;
  for (int64_t i = 0; i < n_rows; ++i) {
    nnz += indptr[i + 1];
    HREQ(nnz <= INT32_MAX, "spgemm: product has more than 2^31 entries");
    indptr[i + 1] = (int32_t)nnz;
  }
  *nnz_out = nnz;
  return FP3D_OK;
}

int fp3d_host_spgemm_fill(int64_t n_rows, int64_t n_cols, const int32_t *Ap,
                          const int32_t *Aj, const double *Ax, const int32_t *Bp,
                          const int32_t *Bj, const double *Bx,
                          const int32_t *indptr, int32_t *indices, double *data) {
  HREQ(n_rows >= 0 && n_cols >= 0 && indptr, "spgemm_fill: bad arguments");
  HREQ(n_rows == 0 || indptr[n_rows] == 0 || (indices && data),
       "spgemm_fill: NULL output");
  parallel_rows(n_rows, product_work(n_rows, Ap, Aj, Bp),
                [&](int64_t lo, int64_t hi, unsigned) {
    RowProduct rp(n_cols);
    for (int64_t i = lo; i < hi; ++i) {
      int32_t *ci = indices + indptr[i];
      double *cx = data + indptr[i];
      rp.run(i, Ap, Aj, Ax, Bp, Bj, Bx, [ci, cx](int32_t c, int32_t k, double v) {
        ci[c] = k;
        cx[c] = v;
      });
    }
  });
  return FP3D_OK;
}

// Rotations 0..col-1 of the Hessenberg matrix, applied in order to one new
// column (a dependent chain: rotation j rewrites entry j + 1, which rotation
// j + 1 then reads).
int fp3d_host_apply_givens(int col, double *hrow, const double *givens) {
  HREQ(col >= 0 && (col == 0 || (hrow && givens)), "apply_givens: NULL");
  for (int j = 0; j < col; ++j) {
    const double c = givens[2 * j], s = givens[2 * j + 1];
    const double n0 = hrow[j], n1 = hrow[j + 1];
    hrow[j] = c * n0 + s * n1;
    hrow[j + 1] = -s * n0 + c * n1;
  }
  return FP3D_OK;
}

int fp3d_host_node_min_edge_length(int64_t n_tets, const int32_t *elements,
                                   int64_t n_nodes, const double *nodes, double *h) {
  int rc = check_elements(n_tets, elements, n_nodes);
  if (rc) return rc;
  HREQ(n_nodes == 0 || (nodes && h), "node_min_edge_length: NULL pointer");
  static const int ed[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int64_t v = 0; v < n_nodes; ++v) h[v] = INFINITY;
  for (int64_t e = 0; e < n_tets; ++e)
    for (int k = 0; k < 6; ++k) {
      const int32_t i = elements[4 * e + ed[k][0]], j = elements[4 * e + ed[k][1]];
      const double dx = nodes[3 * (int64_t)i] - nodes[3 * (int64_t)j];
      const double dy = nodes[3 * (int64_t)i + 1] - nodes[3 * (int64_t)j + 1];
      const double dz = nodes[3 * (int64_t)i + 2] - nodes[3 * (int64_t)j + 2];
      const double L = std::sqrt((dx * dx + dy * dy) + dz * dz);
      if (L < h[i]) h[i] = L;
      if (L < h[j]) h[j] = L;
    }
  return FP3D_OK;
}

/* Smoothed-aggregation prolongator P = T - diag(s) (A T) in two passes
 * (row pointer and count, then indices and data), row-parallel. */
int fp3d_host_sa_prolongator(int64_t n, const int32_t *Ap, const int32_t *Aj,
                             const double *Ax, const int32_t *agg, int64_t n_agg,
                             const double *tval, const double *s, int32_t *Pp,
                             int64_t *nnz_out) {
  HREQ(n >= 0 && n_agg >= 0 && Pp && nnz_out, "sa_prolongator: bad arguments");
  HREQ(n == 0 || (Ap && agg && s && (n_agg == 0 || tval)), "sa_prolongator: NULL pointer");
  for (int64_t i = 0; i < n; ++i)
    HREQ(agg[i] >= -1 && agg[i] < n_agg, "sa_prolongator: aggregate id out of range");
  const int64_t nnzA = n ? (int64_t)Ap[n] : 0;
  for (int64_t q = 0; q < nnzA; ++q)
    HREQ(Aj[q] >= 0 && Aj[q] < n, "sa_prolongator: column out of range");
  Pp[0] = 0;
  parallel_rows(n, nnzA, [&](int64_t lo, int64_t hi, unsigned) {
    ProlongRow pr(n_agg);
    for (int64_t i = lo; i < hi; ++i)
      Pp[i + 1] = pr.run(i, Ap, Aj, Ax, agg, tval, s, [](int32_t, int32_t, double) {});
  });
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    nnz += Pp[i + 1];
    HREQ(nnz <= INT32_MAX, "sa_prolongator: more than 2^31 entries");
    Pp[i + 1] = (int32_t)nnz;
  }
  *nnz_out = nnz;
  return FP3D_OK;
}

int fp3d_host_sa_prolongator_fill(int64_t n, const int32_t *Ap, const int32_t *Aj,
                                  const double *Ax, const int32_t *agg, int64_t n_agg,
                                  const double *tval, const double *s,
                                  const int32_t *Pp, int32_t *Pj, double *Px) {
  HREQ(n >= 0 && n_agg >= 0 && Pp, "sa_prolongator_fill: bad arguments");
  HREQ(n == 0 || Pp[n] == 0 || (Pj && Px), "sa_prolongator_fill: NULL output");
  parallel_rows(n, n ? (int64_t)Ap[n] : 0, [&](int64_t lo, int64_t hi, unsigned) {
    ProlongRow pr(n_agg);
    for (int64_t i = lo; i < hi; ++i) {
      int32_t *pj = Pj + Pp[i];
      double *px = Px + Pp[i];
      const int32_t c = pr.run(i, Ap, Aj, Ax, agg, tval, s,
                               [pj, px](int32_t c, int32_t k, double v) {
                                 pj[c] = k;
                                 px[c] = v;
                               });
      (void)c;
    }
  });
  return FP3D_OK;
}

/* Transpose of an (n_rows x n_cols) CSR: counting sort by column, rows
 * visited in ascending order, so each row of the result has ascending
 * columns (scipy's csr -> csc). */
int fp3d_host_csr_transpose(int64_t n_rows, int64_t n_cols, const int32_t *Ap,
                            const int32_t *Aj, const double *Ax, int32_t *Tp,
                            int32_t *Tj, double *Tx) {
  HREQ(n_rows >= 0 && n_cols >= 0 && Tp, "csr_transpose: bad arguments");
  const int64_t nnz = n_rows ? (int64_t)Ap[n_rows] : 0;
  HREQ(nnz == 0 || (Aj && Ax && Tj && Tx), "csr_transpose: NULL pointer");
  std::vector<int64_t> cnt(n_cols + 1, 0);
  for (int64_t q = 0; q < nnz; ++q) {
    HREQ(Aj[q] >= 0 && Aj[q] < n_cols, "csr_transpose: column out of range");
    ++cnt[Aj[q] + 1];
  }
  for (int64_t k = 0; k < n_cols; ++k) cnt[k + 1] += cnt[k];
  for (int64_t k = 0; k <= n_cols; ++k) Tp[k] = (int32_t)cnt[k];
  for (int64_t i = 0; i < n_rows; ++i)
    for (int32_t q = Ap[i]; q < Ap[i + 1]; ++q) {
      const int64_t dst = cnt[Aj[q]]++;
      Tj[dst] = (int32_t)i;
      Tx[dst] = Ax[q];
    }
  return FP3D_OK;
}

/* y = A x on the host, row-parallel; row i sums from 0 in stored order, as
 * scipy.sparse's csr_matvec. */
int fp3d_host_csr_matvec(int64_t n_rows, const int32_t *Ap, const int32_t *Aj,
                         const double *Ax, const double *x, double *y) {
  HREQ(n_rows >= 0 && (n_rows == 0 || (Ap && x && y)), "csr_matvec: bad arguments");
  parallel_rows(n_rows, n_rows ? (int64_t)Ap[n_rows] : 0,
                [&](int64_t lo, int64_t hi, unsigned) {
    for (int64_t i = lo; i < hi; ++i) {
      double s = 0.0;
      for (int32_t jj = Ap[i]; jj < Ap[i + 1]; ++jj) s += Ax[jj] * x[Aj[jj]];
      y[i] = s;
    }
  });
  return FP3D_OK;
}

int fp3d_host_greedy_coloring(int64_t n_tets, const int32_t *elements,
                              int64_t n_nodes, int32_t *colors,
                              int32_t *n_colors_out) {
  int rc = check_elements(n_tets, elements, n_nodes);
  if (rc) return rc;
  HREQ(colors && n_colors_out, "colors output missing");
  std::vector<int64_t> off;
  std::vector<int32_t> lst;
  node_incidence(n_tets, elements, n_nodes, off, lst);
  const int kMax = 256;
  int64_t mark[kMax];
  for (int c = 0; c < kMax; ++c) mark[c] = -1;
  for (int64_t e = 0; e < n_tets; ++e) colors[e] = -1;
  int32_t ncol = 0;
  for (int64_t e = 0; e < n_tets; ++e) {
    for (int i = 0; i < 4; ++i) {
      int32_t v = elements[4 * e + i];
      for (int64_t k = off[v]; k < off[v + 1]; ++k) {
        int32_t c = colors[lst[k]];
        if (c >= 0) mark[c] = e;
      }
    }
    int c = 0;
    while (c < kMax && mark[c] == e) ++c;
    if (c >= kMax) {
      fp3d::set_error(
          "greedy coloring exceeded 256 colors -- pathological mesh "
          "connectivity (a node shared by >255 elements)");
      return FP3D_ELIMIT;
    }
    colors[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  *n_colors_out = ncol;
  return FP3D_OK;
}

int fp3d_host_csr_pattern(int64_t n_tets, const int32_t *elements,
                          int64_t n_nodes, int32_t *indptr, int32_t *indices,
                          int64_t *nnz_out) {
  int rc = check_elements(n_tets, elements, n_nodes);
  if (rc) return rc;
  HREQ(indptr && nnz_out, "indptr/nnz output missing");
  std::vector<int64_t> off;
  std::vector<int32_t> lst;
  node_incidence(n_tets, elements, n_nodes, off, lst);
  // row v: the sorted, deduplicated nodes of v's incident elements.  Rows
  // are independent: counted, then (with indices) written, row-parallel
  auto build_row = [&](int64_t v, std::vector<int32_t> &row) {
    row.clear();
    for (int64_t k = off[v]; k < off[v + 1]; ++k)
      for (int j = 0; j < 4; ++j) row.push_back(elements[4 * (int64_t)lst[k] + j]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
  };
  const int64_t work = 16 * n_tets;
  std::vector<int64_t> len(n_nodes);
  parallel_rows(n_nodes, work, [&](int64_t lo, int64_t hi, unsigned) {
    std::vector<int32_t> row;
    for (int64_t v = lo; v < hi; ++v) {
      build_row(v, row);
      len[v] = (int64_t)row.size();
    }
  });
  int64_t nnz = 0;
  indptr[0] = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    nnz += len[v];
    HREQ(nnz < INT32_MAX, "CSR pattern exceeds int32 nnz");
    indptr[v + 1] = (int32_t)nnz;
  }
  if (indices)
    parallel_rows(n_nodes, work, [&](int64_t lo, int64_t hi, unsigned) {
      std::vector<int32_t> row;
      for (int64_t v = lo; v < hi; ++v) {
        build_row(v, row);
        std::memcpy(indices + indptr[v], row.data(), row.size() * sizeof(int32_t));
      }
    });
  *nnz_out = nnz;
  return FP3D_OK;
}

int fp3d_host_elem_to_csr(int64_t n_tets, const int32_t *elements,
                          const int32_t *indptr, const int32_t *indices,
                          int32_t *out) {
  HREQ(n_tets == 0 || (elements && indptr && indices && out),
       "elem_to_csr: NULL pointer");
  // elements are independent: element-parallel, each thread noting an
  // element with an entry absent from the pattern
  std::vector<int64_t> bad(host_threads(), -1);
  parallel_rows(n_tets, 16 * n_tets, [&](int64_t lo, int64_t hi, unsigned t) {
    for (int64_t e = lo; e < hi; ++e)
      for (int i = 0; i < 4; ++i) {
        int32_t r = elements[4 * e + i];
        const int32_t *lo_p = indices + indptr[r], *hi_p = indices + indptr[r + 1];
        for (int j = 0; j < 4; ++j) {
          int32_t c = elements[4 * e + j];
          const int32_t *p = std::lower_bound(lo_p, hi_p, c);
          if (p == hi_p || *p != c) {
            if (bad[t] < 0) bad[t] = e;
            out[16 * e + 4 * i + j] = -1;
          } else {
            out[16 * e + 4 * i + j] = (int32_t)(p - indices);
          }
        }
      }
  });
  for (int64_t b : bad) HREQ(b < 0, "elem_to_csr: entry missing from pattern");
  return FP3D_OK;
}

int fp3d_host_face_adjacency(int64_t n_tets, const int32_t *elements,
                             int32_t *face_neighbors) {
  HREQ(n_tets == 0 || (elements && face_neighbors), "face_adjacency: NULL");
  static const int fd[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  struct Face {
    int32_t b, c;    // the face's two larger node ids (the smallest is its bucket)
    int64_t packed;  // e*4 + f, the order the reference visits faces in
  };
  if (n_tets == 0) return FP3D_OK;
  int64_t n_nodes = 0;
  for (int64_t q = 0; q < 4 * n_tets; ++q) {
    HREQ(elements[q] >= 0, "face_adjacency: negative node id");
    if (elements[q] + 1 > n_nodes) n_nodes = elements[q] + 1;
  }
  // faces bucketed by their smallest node (a counting sort), then each
  // bucket sorted by (b, c, packed): the order of one global sort by
  // (a, b, c, packed), built in linear time and sorted bucket-parallel
  std::vector<int64_t> start(n_nodes + 1, 0);
  auto sorted3 = [&](int64_t e, int f, int32_t t[3]) {
    t[0] = elements[4 * e + fd[f][0]];
    t[1] = elements[4 * e + fd[f][1]];
    t[2] = elements[4 * e + fd[f][2]];
    std::sort(t, t + 3);
  };
  for (int64_t e = 0; e < n_tets; ++e)
    for (int f = 0; f < 4; ++f) {
      int32_t t[3];
      sorted3(e, f, t);
      ++start[t[0] + 1];
      face_neighbors[4 * e + f] = -1;
    }
  for (int64_t v = 0; v < n_nodes; ++v) start[v + 1] += start[v];
  // every slot is written below: no value-initialisation pass
  std::unique_ptr<Face[]> faces(new Face[(size_t)(4 * n_tets)]);
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t e = 0; e < n_tets; ++e)
      for (int f = 0; f < 4; ++f) {
        int32_t t[3];
        sorted3(e, f, t);
        faces[fill[t[0]]++] = Face{t[1], t[2], 4 * e + f};
      }
  }
  // pyfp3d/mesh/metrics.py:248-276: the first owner waits in the map, the
  // next occurrence links to it and clears the entry -> consecutive pairs
  parallel_rows(n_nodes, 4 * n_tets, [&](int64_t lo, int64_t hi, unsigned) {
    for (int64_t v = lo; v < hi; ++v) {
      Face *f0 = faces.get() + start[v], *f1 = faces.get() + start[v + 1];
      std::sort(f0, f1, [](const Face &x, const Face &y) {
        if (x.b != y.b) return x.b < y.b;
        if (x.c != y.c) return x.c < y.c;
        return x.packed < y.packed;
      });
      for (Face *i = f0; i < f1;) {
        Face *j = i;
        while (j < f1 && j->b == i->b && j->c == i->c) ++j;
        for (Face *k = i; k + 1 < j; k += 2) {
          const int64_t p = k->packed, q = (k + 1)->packed;
          face_neighbors[p] = (int32_t)(q / 4);
          face_neighbors[q] = (int32_t)(p / 4);
        }
        i = j;
      }
    }
  });
  return FP3D_OK;
}

int fp3d_host_element_geometry(int64_t n_tets, const int32_t *elements,
                               const double *nodes, double *B, double *V,
                               int64_t *bad_elem) {
  HREQ(n_tets == 0 || (elements && nodes && B && V), "geometry: NULL");
  if (bad_elem) *bad_elem = -1;
  for (int64_t e = 0; e < n_tets; ++e) {
    const double *x0 = nodes + 3 * (int64_t)elements[4 * e];
    double J[3][3];  // columns = edge vectors x_k - x_0
    for (int k = 0; k < 3; ++k) {
      const double *xk = nodes + 3 * (int64_t)elements[4 * e + k + 1];
      for (int d = 0; d < 3; ++d) J[d][k] = xk[d] - x0[d];
    }
    // pyfp3d/mesh/metrics.py:58-66: det = v1 . (v2 x v3), V = |det| / 6
    double c0 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
    double c1 = J[2][1] * J[0][2] - J[0][1] * J[2][2];
    double c2 = J[0][1] * J[1][2] - J[1][1] * J[0][2];
    double det = J[0][0] * c0 + J[1][0] * c1 + J[2][0] * c2;
    double s1 = std::sqrt(J[0][0] * J[0][0] + J[1][0] * J[1][0] + J[2][0] * J[2][0]);
    double s2 = std::sqrt(J[0][1] * J[0][1] + J[1][1] * J[1][1] + J[2][1] * J[2][1]);
    double s3 = std::sqrt(J[0][2] * J[0][2] + J[1][2] * J[1][2] + J[2][2] * J[2][2]);
    double scale = s1 * s2 * s3;
    // pyfp3d/mesh/metrics.py:155-164 degenerate guard
    if (scale == 0.0 || std::fabs(det) < 1e-12 * scale) {
      if (bad_elem) *bad_elem = e;
      fp3d::set_error(
          "element_gradients: degenerate tetrahedron (|det J| < 1e-12 x "
          "edge-length^3 scale) -- fix the mesh");
      return FP3D_EINVAL;
    }
    V[e] = std::fabs(det) / 6.0;
    // grad N_i = J^{-T} grad_ref N_i; J^{-1} = adj(J) / det.  Row k of
    // J^{-1} (k = 0..2) is grad N_{k+1}; grad N_0 = -(sum of the others).
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    double *b = B + 12 * e;
    for (int d = 0; d < 3; ++d) {
      b[3 + d] = inv[0][d];
      b[6 + d] = inv[1][d];
      b[9 + d] = inv[2][d];
      b[d] = -inv[0][d] - inv[1][d] - inv[2][d];
    }
  }
  return FP3D_OK;
}

int fp3d_host_standard_aggregation(int64_t n, const int32_t *indptr,
                                   const int32_t *indices, const double *data,
                                   int32_t *agg, int64_t *n_agg) {
  HREQ(n >= 0 && indptr && indices && agg && n_agg, "aggregation: NULL");
  // strong connection = nonzero off-diagonal entry (symmetric strength at
  // theta = 0); explicit zeros (e.g. Dirichlet-masked couplings) are weak.
  auto strong = [&](int64_t q, int64_t i) {
    return indices[q] != i && (!data || data[q] != 0.0);
  };
  std::vector<int64_t> x((size_t)n, 0);  // 0 = free, >0 agg+1, <0 pass-2
  int64_t next = 1;
  const int64_t kIsolated = INT64_MIN;
  // pass 1: seed aggregates at nodes whose strong neighbours are all free
  for (int64_t i = 0; i < n; ++i) {
    if (x[i] != 0) continue;
    bool has_nb = false, nb_aggregated = false;
    for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q) {
      if (!strong(q, i)) continue;
      has_nb = true;
      if (x[indices[q]] != 0) {
        nb_aggregated = true;
        break;
      }
    }
    if (!has_nb) {
      x[i] = kIsolated;
    } else if (!nb_aggregated) {
      x[i] = next;
      for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q)
        if (strong(q, i)) x[indices[q]] = next;
      ++next;
    }
  }
  // pass 2: attach leftovers to a neighbouring pass-1 aggregate
  for (int64_t i = 0; i < n; ++i) {
    if (x[i] != 0) continue;
    for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q) {
      if (!strong(q, i)) continue;
      int64_t xj = x[indices[q]];
      if (xj > 0) {
        x[i] = -xj;
        break;
      }
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (x[i] < 0 && x[i] != kIsolated) x[i] = -x[i];
  // pass 3: remaining free nodes seed new aggregates with free neighbours
  for (int64_t i = 0; i < n; ++i) {
    if (x[i] != 0) continue;
    x[i] = next;
    for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q)
      if (strong(q, i) && x[indices[q]] == 0) x[indices[q]] = next;
    ++next;
  }
  for (int64_t i = 0; i < n; ++i)
    agg[i] = x[i] == kIsolated ? -1 : (int32_t)(x[i] - 1);
  *n_agg = next - 1;
  return FP3D_OK;
}


// Point location for P1 transfer between meshes (no reference counterpart):
// a visibility walk through the face graph, starting from the previous
// point's element (queries in mesh order are spatially coherent), moving
// across the face opposite the most negative barycentric coordinate.  A walk
// that reaches the boundary while still outside keeps the last element (a
// point just outside the polyhedral hull extrapolates linearly from it).
int fp3d_host_locate_points(int64_t n_tets, const int32_t *elements,
                            const double *nodes, const int32_t *face_neighbors,
                            int64_t n_pts, const double *pts,
                            int32_t *out_elem, double *out_bary) {
  HREQ(n_tets > 0 && elements && nodes && face_neighbors,
       "locate_points: mesh missing");
  HREQ(n_pts >= 0 && (n_pts == 0 || (pts && out_elem && out_bary)),
       "locate_points: points missing");
  auto bary = [&](int64_t e, const double *p, double *l) {
    const double *a = nodes + 3 * (int64_t)elements[4 * e];
    double m[3][3];
    for (int j = 0; j < 3; ++j) {
      const double *b = nodes + 3 * (int64_t)elements[4 * e + 1 + j];
      for (int d = 0; d < 3; ++d) m[d][j] = b[d] - a[d];
    }
    double r[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                 m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double x[3];
    for (int j = 0; j < 3; ++j) {
      double c[3][3];
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 3; ++k) c[d][k] = (k == j) ? r[d] : m[d][k];
      x[j] = (c[0][0] * (c[1][1] * c[2][2] - c[1][2] * c[2][1]) -
              c[0][1] * (c[1][0] * c[2][2] - c[1][2] * c[2][0]) +
              c[0][2] * (c[1][0] * c[2][1] - c[1][1] * c[2][0])) / det;
    }
    l[1] = x[0];
    l[2] = x[1];
    l[3] = x[2];
    l[0] = 1.0 - x[0] - x[1] - x[2];
  };
  int64_t cur = 0;
  const int64_t max_steps = 4096;
  for (int64_t q = 0; q < n_pts; ++q) {
    const double *p = pts + 3 * q;
    double l[4];
    bool lost = false;
    for (int64_t step = 0;; ++step) {
      bary(cur, p, l);
      int worst = 0;
      for (int i = 1; i < 4; ++i)
        if (l[i] < l[worst]) worst = i;
      if (l[worst] >= -1e-12) break;
      if (step >= max_steps) {
        lost = true;   // a walk that cycles: fall back to a full scan
        break;
      }
      // face f of the reference's adjacency is opposite local node f
      int32_t nb = face_neighbors[4 * cur + worst];
      if (nb < 0) {
        // at the boundary: try the other negative faces before stopping
        int alt = -1;
        for (int i = 0; i < 4; ++i)
          if (i != worst && l[i] < -1e-12 && face_neighbors[4 * cur + i] >= 0 &&
              (alt < 0 || l[i] < l[alt]))
            alt = i;
        if (alt < 0) break;
        nb = face_neighbors[4 * cur + alt];
      }
      cur = nb;
    }
    if (lost) {
      double best = -1e300, lb[4];
      for (int64_t e = 0; e < n_tets; ++e) {
        bary(e, p, lb);
        double mn = std::min(std::min(lb[0], lb[1]), std::min(lb[2], lb[3]));
        if (mn > best) {
          best = mn;
          cur = e;
          for (int i = 0; i < 4; ++i) l[i] = lb[i];
        }
      }
    }
    out_elem[q] = (int32_t)cur;
    for (int i = 0; i < 4; ++i) out_bary[4 * q + i] = l[i];
  }
  return FP3D_OK;
}

}  // extern "C"
