/*
 * fp3d.h -- C-ABI of libfp3d_hip.so: the MI355X (gfx950) kernels of the
 * full-potential nonlinear solve on unstructured P1 tetrahedra.
 *
 * Conventions (every entry point):
 *   - returns int status: 0 = ok, FP3D_EINVAL = bad argument, FP3D_EHIP = a
 *     HIP runtime error, FP3D_ELIMIT = a documented limit exceeded (e.g.
 *     more than 256 colours).  fp3d_last_error() returns the message of the
 *     last failure on the calling thread.
 *   - array arguments of the device entry points are DEVICE pointers (HBM),
 *     owned by the caller; outputs are overwritten in place (the reference's
 *     "preallocated output" convention); nothing is allocated on the device
 *     except where a function says so.
 *   - `stream` is a hipStream_t passed as void* (0 = the null stream).
 *   - connectivity is int32 (elements (n_tets,4), face neighbours (n_tets,4),
 *     upstream ids, CSR column indices, elem_to_csr (n_tets,16)); the CSR
 *     row pointer is int32 as well (nnz < 2^31 for every supported mesh).
 *   - floating point state is float64; the *_f32 variants take float32
 *     matrix values (vectors stay float64).
 *   - `elem_index` (nullable) maps a launch position k to the element id
 *     e = elem_index[k]; NULL means e = k (colour-major storage).
 *
 * Reference interface replaced by each function is cited as
 * pyfp3d/<file>:<line> of swayli94/UP3D.
 */
#ifndef FP3D_H
#define FP3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP3D_OK 0
#define FP3D_EINVAL (-1)
#define FP3D_EHIP (-2)
#define FP3D_ELIMIT (-3)

/* ---------------------------------------------------------------- misc */
const char *fp3d_last_error(void);
int fp3d_version(void);
/* number of HIP devices visible (0 on a machine without a GPU) */
int fp3d_device_count(int *count);
int fp3d_stream_sync(void *stream);
/* Launch an instantiated HIP graph (hipGraphExec_t, e.g. a captured AMG
 * V-cycle) on a stream: the one launch, without a framework's replay
 * bookkeeping.  No reference counterpart. */
int fp3d_graph_launch(void *graph_exec, void *stream);

/* ------------------------------------------------- physics (elementwise)
 * pyfp3d/physics/isentropic.py:235-258 density_field,
 * :292-314 limit_q2_field, :317-334 mach_squared_field,
 * :366-380 density_derivative_wrt_q_sq_field. */
int fp3d_density_field(int64_t n, const double *q2, double m_inf, double gamma,
                       double *out, void *stream);
int fp3d_mach_squared_field(int64_t n, const double *q2, double m_inf,
                            double gamma, double *out, void *stream);
/* q2cap = q2_at_mach(m_cap, m_inf, gamma) computed by the host (+inf at
 * m_inf == 0, pyfp3d/physics/isentropic.py:261-289). */
int fp3d_limit_q2_field(int64_t n, const double *q2, double q2cap, double *out,
                        void *stream);
int fp3d_density_derivative_field(int64_t n, const double *q2, double m_inf,
                                  double gamma, double *out, void *stream);

/* ------------------------------------------------------ element sweeps */
/* grad(phi)_e = sum_i phi[elem[e,i]] B[e,i,:], q2_e = |grad|^2.
 * pyfp3d/kernels/gradient.py:32-64 element_velocity_q2. */
int fp3d_element_velocity_q2(int64_t n_tets, const int32_t *elements,
                             const double *B, const double *phi,
                             double *grad_out, double *q2_out, void *stream);

/* Fused state sweep of NewtonWorkspace.eval_residual
 * (pyfp3d/solve/newton.py:345-351): grad, q2n = q2/u2, q2l = min(q2n, q2cap),
 * rho = density(q2l), lim = (q2l == q2n).  lim_out may be NULL.
 * gdotb_out (nullable, float64[4 n_tets], 32-byte aligned) receives
 * grad_e . B[e,a,:] for a = 0..3, which fp3d_element_residual_terms and
 * fp3d_term3_build take in place of grad and B (bitwise the same values
 * they would compute). */
int fp3d_element_state(int64_t n_tets, const int32_t *elements, const double *B,
                       const double *phi, double u2, double m_inf, double gamma,
                       double q2cap, double *grad_out, double *q2n_out,
                       double *q2l_out, double *rho_out, uint8_t *lim_out,
                       double *gdotb_out, void *stream);

/* ------------------------------------------------------------ upwinding */
/* Multi-hop upstream walk u(e); pyfp3d/kernels/upwind.py:68-125
 * upstream_elements (face_neighbors int32 (n,4), -1 = boundary). */
int fp3d_upstream_elements(int64_t n_tets, const int32_t *face_neighbors,
                           const double *centroids, const double *nodes,
                           const int32_t *elements, const double *grad,
                           int max_hops, int32_t *out, void *stream);
/* The same walk reading a per-face centroid displacement table
 * face_disp[e][f][:] = centroid(face_neighbors[e][f]) - centroid(e)
 * (float64 (n,4,3), any value where the neighbour is -1): two contiguous
 * rows per hop instead of five scattered centroid gathers; identical u(e).
 * elements, face_neighbors and face_disp must be 16-byte aligned. */
int fp3d_upstream_elements_d(int64_t n_tets, const int32_t *face_neighbors,
                             const double *face_disp, const double *nodes,
                             const int32_t *elements, const double *grad,
                             int max_hops, int32_t *out, void *stream);
int fp3d_upwind_rho_tilde_d(int64_t n_tets, const int32_t *face_neighbors,
                            const double *face_disp, const double *nodes,
                            const int32_t *elements, const double *grad,
                            const double *q2, const double *rho, int max_hops,
                            double m_inf, double m_crit, double C, double gamma,
                            double rho_floor, int32_t *upstream_out,
                            double *nu_out, double *rho_tilde_out,
                            void *stream);
/* The same sweep plus the exact sensitivities of the selection it makes
 * (s_e, s_u: pyfp3d/kernels/upwind.py:733-920, as
 * fp3d_rho_tilde_sensitivities on upstream_out), from one read of q2 / rho
 * of e and of u. */
int fp3d_upwind_rho_tilde_sens_d(
    int64_t n_tets, const int32_t *face_neighbors, const double *face_disp,
    const double *nodes, const int32_t *elements, const double *grad,
    const double *q2, const double *rho, int max_hops, double m_inf,
    double m_crit, double C, double gamma, double rho_floor,
    int32_t *upstream_out, double *nu_out, double *rho_tilde_out,
    double *s_e_out, double *s_u_out, void *stream);
/* pyfp3d/kernels/upwind.py:456-500 rho_tilde_sweep. */
int fp3d_rho_tilde_sweep(int64_t n, const double *q2, const double *rho,
                         const int32_t *upstream, double m_inf, double m_crit,
                         double C, double gamma, double rho_floor,
                         double *nu_out, double *rho_tilde_out, void *stream);
/* Walk + sweep fused in one launch (UpwindOperator.rho_tilde walk mode,
 * pyfp3d/kernels/upwind.py:811-815); upstream_out may be NULL. */
int fp3d_upwind_rho_tilde(int64_t n_tets, const int32_t *face_neighbors,
                          const double *centroids, const double *nodes,
                          const int32_t *elements, const double *grad,
                          const double *q2, const double *rho, int max_hops,
                          double m_inf, double m_crit, double C, double gamma,
                          double rho_floor, int32_t *upstream_out,
                          double *nu_out, double *rho_tilde_out, void *stream);
/* pyfp3d/kernels/upwind.py:503-541 classify_upwind_branches. */
int fp3d_classify_upwind_branches(int64_t n, const double *q2, const double *rho,
                                  const int32_t *upstream, double m_inf,
                                  double m_crit, double C, double gamma,
                                  double rho_floor, int8_t *branch_out,
                                  void *stream);
/* pyfp3d/kernels/upwind.py:544-593 rho_tilde_frozen_sweep. */
int fp3d_rho_tilde_frozen_sweep(int64_t n, const double *q2, const double *rho,
                                const int32_t *upstream, const int8_t *branch,
                                double m_inf, double m_crit, double C,
                                double gamma, double rho_floor, double *nu_out,
                                double *rho_tilde_out, void *stream);
/* pyfp3d/kernels/upwind.py:596-643 rho_tilde_frozen_sensitivities_sweep. */
int fp3d_rho_tilde_frozen_sensitivities(int64_t n, const double *q2,
                                        const double *rho,
                                        const int32_t *upstream,
                                        const int8_t *branch, double m_inf,
                                        double m_crit, double C, double gamma,
                                        double rho_floor, double *se_out,
                                        double *su_out, void *stream);
/* pyfp3d/kernels/upwind.py:646-730 rho_tilde_sensitivities_sweep. */
int fp3d_rho_tilde_sensitivities(int64_t n, const double *q2, const double *rho,
                                 const int32_t *upstream, double m_inf,
                                 double m_crit, double C, double gamma,
                                 double rho_floor, double *se_out,
                                 double *su_out, void *stream);
/* Monitors of UpwindOperator.rho_tilde (pyfp3d/kernels/upwind.py:816-818):
 * out3 = {nu_max, n_nu_positive, n_rho_tilde_eq_floor} (doubles), computed
 * by a deterministic two-pass reduction; `work` is device scratch of
 * fp3d_reduce_work_doubles() doubles. */
int fp3d_upwind_monitors(int64_t n, const double *nu, const double *rho_tilde,
                         double rho_floor, double *work, double *out3,
                         void *stream);
/* count of lim == 0 (speed-limited elements) into out1 (double) */
int fp3d_count_limited(int64_t n, const uint8_t *lim, double *work,
                       double *out1, void *stream);
int fp3d_reduce_work_doubles(void);

/* ------------------------------------------------- coloured assembly
 * Colour classes are launched in order; within a class no two elements
 * share a node, so every scatter is a plain read-modify-write (no atomics)
 * and the per-node / per-nnz accumulation order is the colour sequence:
 * bit-deterministic.  color_offsets is a HOST array (n_colors+1). */
/* pyfp3d/kernels/residual.py:43-76 assemble_residual_colored (R zeroed). */
int fp3d_assemble_residual_colored(int n_colors, const int64_t *color_offsets,
                                   const int32_t *elem_index,
                                   const int32_t *elements, const double *B,
                                   const double *V, const double *rho_tilde,
                                   const double *grad, int64_t n_nodes,
                                   double *R, void *stream);
/* pyfp3d/kernels/jacobian.py:105-129 assemble_matrix_data_colored
 * (data must be zeroed by the caller, as in the reference). */
int fp3d_assemble_matrix_colored(int n_colors, const int64_t *color_offsets,
                                 const int32_t *elem_index,
                                 const int32_t *elem_to_csr, const double *B,
                                 const double *V, const double *rho_tilde,
                                 double *data, void *stream);
/* pyfp3d/kernels/jacobian.py:132-180 assemble_jacobian_data_colored; s_e is
 * multiplied by lim[e] first when lim != NULL
 * (pyfp3d/kernels/jacobian.py:323-324). */
int fp3d_assemble_jacobian_colored(int n_colors, const int64_t *color_offsets,
                                   const int32_t *elem_index,
                                   const int32_t *elem_to_csr, const double *B,
                                   const double *V, const double *rho_tilde,
                                   const double *grad, const double *s_e,
                                   const uint8_t *lim, double inv_u2,
                                   double *data, void *stream);
/* SPD preconditioner matrix (no reference counterpart): Terms 1+2 with the
 * Term-2 weight clamped from below at -clamp * rho_tilde V / q2, so every
 * element matrix stays positive semi-definite (0 < clamp < 1).  Used to
 * build the AMG hierarchy that preconditions the Newton GMRES. */
int fp3d_assemble_jacobian_spd_colored(int n_colors,
                                       const int64_t *color_offsets,
                                       const int32_t *elem_index,
                                       const int32_t *elem_to_csr,
                                       const double *B, const double *V,
                                       const double *rho_tilde,
                                       const double *grad, const double *s_e,
                                       const uint8_t *lim, double inv_u2,
                                       double clamp, double *data,
                                       void *stream);
/* ------------------------------------------------- node-gather assembly
 * The coloured sums above computed row by row (no colour passes, no
 * atomics): node v adds its incident (element, local) entries in ascending
 * element order, the accumulation order of the reference's serial assembly
 * (pyfp3d/kernels/residual.py:144-189 assemble_residual; the Picard/Newton
 * matrices of pyfp3d/kernels/jacobian.py:105-180 summed as the serial
 * triplets of pyfp3d/kernels/residual.py:106-129 + COO->CSR).  inc_ptr /
 * inc_ent: node -> (4 e + i) incidence sorted by e.  Used by the operator
 * classes (pyfp3d/kernels/jacobian.py:231-375 PicardOperator).
 *
 * Two stages: the element pass writes terms[4e + a] = rho_tilde V
 * (grad . B_a) (grad . B_a read from gdotb when that is given, B and grad
 * may then be NULL), or the 16 matrix values K[16e + 4a + b]; the node pass sums
 * them per node / per CSR slot.  With inc_pos (int32[4 n_tets], the inverse
 * of inc_ent, 16-byte aligned) the element pass writes in incidence order
 * instead: terms[inc_pos[4e + a]], matrix row a at K[4 inc_pos[4e + a]];
 * the node pass then takes inc_ent == NULL and reads terms / K in order
 * (same sums, same order: bitwise equal results).  Matrix terms: s_e == NULL gives the Picard
 * matrix A = sum rho_tilde V B_i.B_j (grad unused); else Terms 1+2 with s_e
 * masked by lim (lim may be NULL); clamp in (0, 1) gives the SPD
 * preconditioner variant (as fp3d_assemble_jacobian_spd_colored), 0 the
 * exact Terms 1+2.  gather_csr_terms writes every CSR value (no zero fill);
 * max_seg >= max over row blocks of indptr[r0 + rows] - indptr[r0] (rows =
 * fp3d_gather_rows_per_block()), at most 8192.  terms and K 32-byte
 * aligned. */
int fp3d_gather_rows_per_block(void);
int fp3d_element_residual_terms(int64_t n_tets, const double *B,
                                const double *V, const double *rho_tilde,
                                const double *grad, const double *gdotb,
                                const int32_t *inc_pos, double *terms,
                                void *stream);
int fp3d_gather_node_terms(int64_t n_nodes, const int32_t *inc_ptr,
                           const int32_t *inc_ent, const double *terms,
                           double *R, void *stream);
int fp3d_element_matrix_terms(int64_t n_tets, const double *B, const double *V,
                              const double *rho_tilde, const double *grad,
                              const double *s_e, const uint8_t *lim,
                              double inv_u2, double clamp,
                              const int32_t *inc_pos, double *K, void *stream);
/* rel8[4q + b] = elem_to_csr[e][a][b] - indptr[elements[e][a]] where
 * inc_ent[q] = 4e + a (incidence order; uint8: every row is at most 256
 * long), 4-byte aligned.  inc_ent == NULL: K holds row q at K[4q]. */
int fp3d_gather_csr_terms(int64_t n_nodes, int max_seg, const int32_t *indptr,
                          const int32_t *inc_ptr, const int32_t *inc_ent,
                          const uint8_t *rel8, const double *K, double *data,
                          void *stream);
/* pyfp3d/kernels/jacobian.py:183-228 fill_term3_coo (rows/cols int32). */
int fp3d_fill_term3_coo(int64_t n_active, const int32_t *active,
                        const int32_t *elements, const double *B,
                        const double *V, const double *grad,
                        const int32_t *upstream, const double *s_u,
                        double inv_u2, int32_t *rows, int32_t *cols,
                        double *vals, void *stream);

/* ----------------------------- Term 3 as a device operator (no COO merge)
 * The upstream coupling of the Newton Jacobian (pyfp3d/kernels/jacobian.py
 * :183-228, 340-357; there a COO added to J with scipy every Newton step)
 * applied matrix-free inside the Krylov matvec.  fp3d_term3_build selects
 * the active elements (s_u != 0 after the limiter mask lim[upstream[e]]),
 * stores their coefficients and a row-gather table built from the static
 * node -> (4*element + local) incidence CSR (inc_ptr int32[n_nodes+1],
 * inc_ent int32[4*n_tets], entries of a node in ascending element order;
 * inc_pos int32[4*n_tets] its inverse, 16-byte aligned; gdotb nullable:
 * the state sweep's grad . B_i per element, used in place of grad and B),
 * so fp3d_csr_spmv_term3 adds Term 3 row by row in a fixed order: no
 * atomics, bitwise repeatable.  Rows with free_mask[row] == 0 (nullable)
 * get no Term-3 entries.  `t3` is caller-owned device memory of
 * fp3d_term3_workspace_bytes bytes.  With n_active_out != NULL the call
 * synchronises the stream once and stores the active count there; with
 * NULL it returns without synchronising, and fp3d_term3_active_count
 * reads the count of the last build later (one synchronisation). */
int64_t fp3d_term3_workspace_bytes(int64_t n_tets, int64_t n_nodes);
int fp3d_term3_build(int64_t n_tets, int64_t n_nodes, const int32_t *elements,
                     const int32_t *inc_ptr, const int32_t *inc_ent,
                     const int32_t *inc_pos, const double *B,
                     const double *V, const double *grad, const double *gdotb,
                     const int32_t *upstream, const double *s_u,
                     const uint8_t *lim, const uint8_t *free_mask,
                     double inv_u2, void *t3, int64_t *n_active_out,
                     void *stream);
int fp3d_term3_active_count(const void *t3, int64_t n_tets, int64_t n_nodes,
                            int64_t *n_active_out, void *stream);

/* ------------------------------------------------------ sparse / BLAS-1
 * CSR SpMV, G lanes of a wave64 per row (G chosen from nnz / n_rows). */
int fp3d_csr_spmv(int64_t n_rows, int64_t nnz, const int32_t *indptr,
                  const int32_t *indices, const double *data, const double *x,
                  double *y, void *stream);
int fp3d_csr_spmv_f32(int64_t n_rows, int64_t nnz, const int32_t *indptr,
                      const int32_t *indices, const float *data,
                      const double *x, double *y, void *stream);
/* y = A x + Term3 x (t3 from fp3d_term3_build, skipped when t3 == NULL or
 * n_active == 0); rows with free_mask[row] == 0 give y[row] = x[row].
 * fmt (no reference counterpart; the reference's matvec is scipy's): bit 1
 * = data float32 (else float64), bit 4 = indices are int16 offsets from
 * the row (col = row + indices[j], a square matrix whose bandwidth fits),
 * else int32 columns.  Vectors are float64. */
int fp3d_csr_spmv_term3(int64_t n_rows, int64_t nnz, const int32_t *indptr,
                        const void *indices, const void *data, int fmt,
                        const int32_t *elements, const int32_t *upstream,
                        const void *t3, int64_t n_active, int64_t n_tets,
                        const uint8_t *free_mask, const double *x, double *y,
                        void *stream);
/* Dirichlet principal-submatrix embedding (pyfp3d/solve/linear.py:20-58,
 * apply_dirichlet) in place on CSR data: rows and columns of non-free dofs
 * become identity, so Krylov vectors with zero Dirichlet entries see
 * exactly A[free][:, free]; optional float32 copy into data32 (nullable). */
int fp3d_csr_mask_dirichlet(int64_t n_rows, const int32_t *indptr,
                            const int32_t *indices, const uint8_t *free_mask,
                            double *data, float *data32, void *stream);
/* Lumped nodal volume m_v = sum_e V_e / 4 over the node -> (4k+i)
 * incidence (fixed order).  No reference counterpart: the pseudo-transient
 * Newton globalisation is an addition of this library. */
int fp3d_lumped_volume(int64_t n_nodes, const int32_t *inc_ptr,
                       const int32_t *inc_ent, const double *V, double *out,
                       void *stream);
/* data[diag_pos[r]] += scale * w[r] on free rows (pseudo-transient term
 * m_v / dt on the Jacobian or preconditioner diagonal).  No reference
 * counterpart. */
int fp3d_csr_diag_shift(int64_t n_rows, const int32_t *diag_pos,
                        const double *w, double scale,
                        const uint8_t *free_mask, double *data, void *stream);
int fp3d_csr_to_f32(int64_t nnz, const double *data, float *out, void *stream);
int fp3d_fill(int64_t n, double value, double *x, void *stream);
/* y = a x + b y */
int fp3d_axpby(int64_t n, double a, const double *x, double b, double *y,
               void *stream);
/* y = a x */
int fp3d_scale(int64_t n, double a, const double *x, double *y, void *stream);
/* y = mask ? x : 0 */
int fp3d_mask_copy(int64_t n, const uint8_t *mask, const double *x, double *y,
                   void *stream);
/* out[k] = v_k . w, k < nv (v_k = V + k*ld), deterministic two-pass
 * reduction; work = fp3d_multi_dot_work_doubles(nv) doubles, out on device. */
int fp3d_multi_dot(int64_t n, int nv, const double *V, int64_t ld,
                   const double *w, double *work, double *out, void *stream);
int fp3d_multi_dot_work_doubles(int nv);
/* As fp3d_multi_dot, and out[nv] = w . w in the same pass (the value a
 * one-row call on w gives; no reference counterpart: the Arnoldi step's
 * ||w||^2 for the DGKS test).  work = fp3d_multi_dot_work_doubles(nv + 1);
 * _f32v reads float32 basis vectors (products and sums in float64). */
int fp3d_multi_dot_norm(int64_t n, int nv, const double *V, int64_t ld,
                        const double *w, double *work, double *out,
                        void *stream);
int fp3d_multi_dot_norm_f32v(int64_t n, int nv, const float *V, int64_t ld,
                             const double *w, double *work, double *out,
                             void *stream);
/* Allocation-free fp3d_multi_dot_norm (no reference counterpart): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments without
 * `work` and the same bits; the partial sums live in a library-owned device
 * array of 257 x 1024 doubles (2,105,344 bytes per GPU, held until the
 * process ends), so nv > 256 is rejected (FP3D_EINVAL).  The array's
 * address is looked up on the first call on each device (make that call
 * outside graph capture).  Nothing needs clearing between calls, but all
 * calls on one GPU share the array: calls that overlap in time on one GPU
 * (two streams, two host threads) race and must be serialised by the
 * caller. */
int fp3d_multi_dot_norm_noalloc(int64_t n, int nv, const double *V, int64_t ld,
                                const double *w, double *out, void *stream);
/* Batched Arnoldi steps (no reference counterpart: the reference solves one
 * flow condition per call, pyfp3d/solve/newton.py:1144-1307).  One Krylov
 * step of each of nb GMRES solves on the same mesh, one launch per kernel for
 * all of them (blockIdx.y = solve).  `jobs` is a device array of nb job
 * records of fp3d_batch_job_bytes() = 128 bytes, sixteen 8-byte words each:
 *   0 V (basis rows, float64 or float32)   1 ld (row stride)
 *   2 nv (= col + 1)                        3 w (float64, orthogonalised in place)
 *   4 out (float64[3 + 2 nv]: [0] ||w|| after, [1] max|w|, [2, 2+nv) h,
 *     [2+nv] ||w||^2 before, [3+nv, 3+2nv) second pass)
 *   5 work (fp3d_multi_dot_work_doubles(nv + 1) doubles)
 *   6 vnext (basis row col + 1)             7 src32 (basis row col, float32 basis)
 *   8 x (float64 matvec input)              9 y (float64 matvec output)
 *  10 Jacobian values (fmt)                11 upstream (int32[n_tets])
 *  12 Term-3 workspace of fp3d_term3_build 13 its active count (0: none)
 *  14 free mask (uint8[n_rows])            15 reserved (0).
 * batch_jacobian_matvec: widen ? x = (double) src32 : -, then y = J x as
 * fp3d_csr_spmv_term3 (the mesh arguments are shared; max_active = the
 * largest word 13).  batch_orthonormalise (after the caller has applied its
 * preconditioner, w = M^-1 y): the CGS pass into out[2..], w -= V h, the
 * DGKS-conditional second pass into out[3+nv..] (eta2), w -= V h2,
 * out[0..1] = norms of w, vnext = w / out[0] (float32 when f32); max_nv = the
 * largest word 2, at most 256.  Per solve the same kernel bodies on the same
 * grids as the single-solve calls: the same bits. */
/* Batched live residual evaluation (no reference counterpart, as the
 * batched Arnoldi steps): for nb solves on one mesh, the chain of the
 * Newton driver's residual evaluation -- fp3d_element_state (with grad.B),
 * fp3d_count_limited, fp3d_upwind_rho_tilde_sens_d, fp3d_upwind_monitors,
 * fp3d_element_residual_terms (incidence order) + fp3d_gather_node_terms,
 * fp3d_mask_copy and fp3d_norms -- one launch per kernel for all solves.
 * `jobs`: device array of nb records of fp3d_batch_eval_job_bytes() = 192
 * bytes, 24 words: 0 phi, 1 grad, 2 q2n, 3 q2l, 4 rho, 5 lim (uint8),
 * 6 grad.B, 7 upstream (int32), 8 nu, 9 rho_tilde, 10 s_e, 11 s_u,
 * 12 element terms (4 n_tets), 13 R, 14 masked R, 15 free mask (uint8),
 * 16 work (3 * 1024 doubles), 17 out (double[6]: limited count, nu max,
 * count nu > 0, floored count, ||masked R||_2, ||masked R||_inf),
 * 18 m_inf and 19 q2 cap (float64 bit patterns), 20-23 reserved (0).  The
 * mesh arrays and u2, gamma, m_crit, C, rho_floor, max_hops are shared.  Per
 * solve the same kernel bodies on the same grids as the single calls: the
 * same bits. */
int fp3d_batch_eval_job_bytes(void);
int fp3d_batch_residual_eval(int nb, const void *jobs, int64_t n_tets, int64_t n_nodes,
                             const int32_t *elements, const double *B, const double *V,
                             const int32_t *face_neighbors, const double *face_disp,
                             const double *nodes, const int32_t *inc_ptr,
                             const int32_t *inc_pos, double u2, double gamma,
                             double m_crit, double C, double rho_floor, int max_hops,
                             void *stream);
int fp3d_batch_job_bytes(void);
int fp3d_batch_jacobian_matvec(int nb, const void *jobs, int64_t max_active,
                               int64_t n_rows, int64_t nnz, const int32_t *indptr,
                               const void *indices, int fmt, const int32_t *elements,
                               int64_t n_tets, int widen, void *stream);
int fp3d_batch_orthonormalise(int nb, const void *jobs, int max_nv, int64_t n, int f32,
                              double eta2, void *stream);
/* w += sign * sum_k h[k] v_k  (h on device, nv <= 256) */
/* Re-orthogonalisation pass of CGS2 (no reference counterpart; the
 * reference's GMRES is scipy's): out[k] = V_k . w only when the first pass
 * cancelled, judged on the device by the DGKS test
 * ||w||^2 - sum h_k^2 < eta2 ||w||^2 with h = cond[0..nv), ||w||^2 =
 * cond[nv] (the first pass's outputs); otherwise out[0..nv) = 0 exactly and
 * fp3d_multi_axpy with those coefficients returns without touching w. */
int fp3d_multi_dot_reorth(int64_t n, int nv, const double *V, int64_t ld,
                          const double *w, const double *cond, double eta2,
                          double *work, double *out, void *stream);
int fp3d_multi_axpy(int64_t n, int nv, const double *V, int64_t ld,
                    const double *h, double sign, double *w, void *stream);
/* FP32 Krylov basis (no reference counterpart): the same three operations
 * with the basis V stored as float32 and every product and sum in float64
 * (w, h and out stay float64), for GMRES in the FP32-preconditioned solve.
 * fp3d_scale_to_f32: y = (float)(a x); fp3d_f32_to_f64: y = (double)x. */
int fp3d_multi_dot_f32v(int64_t n, int nv, const float *V, int64_t ld,
                        const double *w, double *work, double *out, void *stream);
int fp3d_multi_dot_reorth_f32v(int64_t n, int nv, const float *V, int64_t ld,
                               const double *w, const double *cond, double eta2,
                               double *work, double *out, void *stream);
int fp3d_multi_axpy_f32v(int64_t n, int nv, const float *V, int64_t ld,
                         const double *h, double sign, double *w, void *stream);
int fp3d_scale_to_f32(int64_t n, double a, const double *x, float *y, void *stream);
int fp3d_f32_to_f64(int64_t n, const float *x, double *y, void *stream);
/* One Arnoldi step of the device GMRES in one call (no reference
 * counterpart; pyfp3d/solve/linear.py:182-193 runs scipy's GMRES, whose
 * step this is): w = M^-1 J v_col (fp3d_csr_spmv_term3, then the V-cycle as
 * an instantiated HIP graph, or w = J v_col without one), w orthogonalised
 * against v_0..v_col (fp3d_multi_dot_norm, fp3d_multi_axpy, the
 * conditional second pass, the f32v variants for a float32 basis),
 * normalised into v_{col+1} (fp3d_norms, fp3d_scale_inv_dev[_to_f32]), and
 * the step's coefficients out[0 .. 3 + 2(col+1)) copied to pinned host
 * memory.  The same launches, in the same order, as issued one by one, so
 * the same bits.  a: 27 int64 words -- n, f32 basis flag, V, ld, v64 scratch,
 * tmp, w[0], w[1], out[0], out[1], reduction work, pinned hbuf[0], hbuf[1];
 * the Jacobian matvec's nnz, indptr, columns, values, fmt, elements,
 * upstream, Term-3 workspace, n_active, n_tets, free_mask; V-cycle graph
 * execs for w[0] and w[1]; 1 when a preconditioner is applied.  Buffers of
 * parity col & 1 are used.  col <= 255. */
int fp3d_arnoldi_step(const int64_t *a, int col, void *stream);
/* y = x / nrm[0], the divisor read on the device (no reference
 * counterpart): GMRES normalises the new Arnoldi vector without a host
 * round trip, so the next step is queued while the host applies Givens. */
int fp3d_scale_inv_dev(int64_t n, const double *nrm, const double *x, double *y,
                       void *stream);
int fp3d_scale_inv_dev_to_f32(int64_t n, const double *nrm, const double *x,
                              float *y, void *stream);
/* out2 = {||x||_2, ||x||_inf} on device */
int fp3d_norms(int64_t n, const double *x, double *work, double *out2,
               void *stream);

/* ----------------------------------------------------- AMG V-cycle kernels
 * (the reference applies pyamg's V-cycle, pyfp3d/solve/linear.py:61-81)
 * fmt: bit 1 = matrix values float32 (else float64); bit 2 = every vector
 * argument (dinv, b, x, r, d*, xc, bc) float32 (else float64), arithmetic
 * float64 either way; bit 4 = indices are int16 offsets from the row (a
 * square matrix whose bandwidth fits in int16), else int32 columns.
 *
 * Chebyshev smoothing (degree >= 1) of A x = b on D^-1 A with eigenvalue
 * interval [lmin, lmax]; zero_init = 1 starts from x = 0 (x is written),
 * 0 from the given x, 2 continues after a first step from x = 0 that the
 * caller already took (x, r, d0 as fp3d_amg_restrict_start leaves them).
 * r, d0, d1 are n-vectors of scratch. */
int fp3d_amg_cheb_smooth(int64_t n, int64_t nnz, const int32_t *indptr,
                         const void *indices, const void *data, int fmt,
                         const void *dinv, const void *b, void *x, void *r,
                         void *d0, void *d1, int degree, double lmin, double lmax,
                         int zero_init, void *stream);
/* bc = R r, and the coarse level's first Chebyshev step from x_c = 0 on
 * bc (dinv_c, degree_c, [lmin_c, lmax_c]): r_c, d_c and x_c as
 * fp3d_amg_cheb_smooth(zero_init = 1) leaves them before its second step.
 * Then fp3d_amg_cheb_smooth(zero_init = 2) on the coarse level. */
int fp3d_amg_restrict_start(int64_t nc, int64_t nnz, const int32_t *indptr,
                            const void *indices, const void *data, int fmt,
                            const void *r, void *bc, const void *dinv_c,
                            int degree_c, double lmin_c, double lmax_c, void *x_c,
                            void *r_c, void *d_c, void *stream);
/* r = b - A x */
int fp3d_amg_residual(int64_t n, int64_t nnz, const int32_t *indptr,
                      const void *indices, const void *data, int fmt,
                      const void *b, const void *x, void *r, void *stream);
/* bc = R r (R = P^T stored as CSR, nc rows) */
int fp3d_amg_restrict(int64_t nc, int64_t nnz, const int32_t *indptr,
                      const void *indices, const void *data, int fmt,
                      const void *r, void *bc, void *stream);
/* x += P xc */
int fp3d_amg_prolong_add(int64_t n, int64_t nnz, const int32_t *indptr,
                         const void *indices, const void *data, int fmt,
                         const void *xc, void *x, void *stream);
/* x = M b, M dense row-major n x n (the coarsest level's dense inverse) */
int fp3d_dense_matvec(int64_t n, const double *M, const double *b, double *x,
                      void *stream);
/* the same with float32 b and x when fmt bit 2 is set (M and the sums stay
 * float64: the coarsest solve is float64 in both precisions) */
int fp3d_dense_matvec_v(int fmt, int64_t n, const double *M, const void *b, void *x,
                        void *stream);
/* One V-cycle operation for nb solves at once (no reference counterpart:
 * the reference applies one pyamg hierarchy per call,
 * pyfp3d/solve/linear.py:61-81): recs points to nb 128-byte device records,
 * one per solve (row count, CSR arrays, operand, operation kind, five
 * pointers, two scalars, a flag; the layout up3d_amd/solve/amg.py writes).
 * op 0: SpMV with G lanes per row and an epilogue (store, residual, add,
 * the three Chebyshev steps, restriction + coarse first step), fmt bits
 * 1 float32 values, 2 float32 vectors, 4 int16 row offsets; op 1:
 * element-wise (Chebyshev first step from 0, x += d, float32 / float64
 * conversions); op 2: dense coarsest matvec.  Each solve runs the body,
 * grid and arithmetic of its own fp3d_amg_* / fp3d_dense_matvec_v /
 * conversion launch (max_n = the largest row or entry count), so the
 * results are bitwise those of its own V-cycle. */
int fp3d_grp_vcycle_op(int nb, const void *recs, int op, int64_t max_n, int G, int fmt,
                       void *stream);
/* A whole grouped V-cycle: n_ops operations, operation i being
 * fp3d_grp_vcycle_op(nb, recs + i * nb records, ops[4i], ops[4i + 1],
 * ops[4i + 2], ops[4i + 3]) (ops: host array of (op, max_n, G, fmt)
 * words), launched in order on the stream; one call instead of one per
 * operation. */
int fp3d_grp_vcycle_run(int nb, const void *recs, int n_ops, const int64_t *ops,
                        void *stream);

/* ---------------------------------------------------- discrete adjoint
 * (new functionality: the reference has no adjoint.)  At a frozen upstream
 * selection and branch (pyfp3d/kernels/upwind.py:544-643):
 * out = d rho_tilde / d M_inf per element (branch-wise, at fixed q2). */
int fp3d_drho_tilde_dmach(int64_t n, const double *q2, const double *rho,
                          const int32_t *upstream, const int8_t *branch,
                          double m_inf, double m_crit, double C, double gamma,
                          double *out, void *stream);
/* out (n_nodes, 3) = d(lambda^T R)/d(node coordinates) for the residual
 * R_i = sum_e rho_tilde_e V_e (g_e . B_e,i) with the rho_tilde chain
 * through (s_e, s_u); down_ptr/down_list: for each element the elements
 * whose upstream it is (ascending); inc_ptr/inc_ent: node -> (4e+i).
 * work: fp3d_adjoint_dx_work_doubles(n_tets) doubles. */
int fp3d_adjoint_dx(int64_t n_tets, int64_t n_nodes, const int32_t *elements,
                    const double *B, const double *V, const double *grad,
                    const double *rho_tilde, const double *s_e,
                    const double *s_u, const int32_t *down_ptr,
                    const int32_t *down_list, const int32_t *inc_ptr,
                    const int32_t *inc_ent, const double *lam, double *work,
                    double *out, void *stream);
int64_t fp3d_adjoint_dx_work_doubles(int64_t n_tets);

/* ------------------------------------------------ level-set wake sheet
 * Replaces the per-assembly host COO -> CSR of the sheet terms: the
 * special-copy mass conservation (pyfp3d/kernels/cut_assembly.py:119-227
 * mass_conservation_coo, the two copies of each cut element and the fan's
 * lower copy, pyfp3d/wake/multivalued.py:276-379), the wake least-squares
 * rows (cut_assembly.py:230-318 wake_ls_coo), the TE pressure Kutta rows
 * and their Jacobian (cut_assembly.py:321-398, 478-515) and the Newton
 * Term 2 of the special copies (cut_assembly.py:401-475).  Their extended
 * CSR pattern is fixed by the cut census; slot s sums its contributions
 * c in [seg[s], seg[s+1]) in that order, contribution c being
 *   kind 0: coef*wa[e]   1: coef*wb[e]   2: coef   3: coef*kv[e]
 *   4: coef*sa[e]*(ga[e].B_a)*(ga[e].B_b)   5: the same with sb, gb
 * with e = src[c], a = ab[c] >> 2, b = ab[c] & 3, B (n_tets,4,3), ga / gb
 * (n_tets,3); the slot value is 0 where fmask (nullable) is 0 at its row
 * srow[s] or column scol[s].  Pointers a kind does not use may be NULL. */
int fp3d_sheet_fill(int64_t n_slots, const int32_t *seg, const uint8_t *kind,
                    const int32_t *src, const uint8_t *ab, const double *coef,
                    const double *wa, const double *wb, const double *kv,
                    const double *sa, const double *sb, const double *ga,
                    const double *gb, const double *B, const int32_t *srow,
                    const int32_t *scol, const uint8_t *fmask, double *out,
                    void *stream);
/* Term 3 of the special copies and of plain elements whose upstream element
 * is special (cut_assembly.py:401-475, the upstream part): per candidate
 * element cand[c] on side cside[c] (0 upper, 1 lower; per-side grad g,
 * upstream up, s_u su, limiter lim, DOF table dof and read table rd, all
 * (n_tets[,4|3])), rows[4c+a] = row_map[dof[e][a]] (-1 when s_u lim[u] = 0,
 * the row is dropped or fmask (nullable) is 0 there), W[4c+a] =
 * 2 inv_u2 s_u lim[u] V_e (g_e.B_a), gu[4c+b] = g_u.B_u,b (0 where fmask is 0
 * at the column), cols[4c+b] = rd[u][b]. */
int fp3d_sheet_t3_prepare(int64_t n_cand, const int32_t *cand, const uint8_t *cside,
                          const double *g0, const double *g1, const int32_t *up0,
                          const int32_t *up1, const double *su0, const double *su1,
                          const uint8_t *lim0, const uint8_t *lim1,
                          const int32_t *dof0, const int32_t *dof1,
                          const int32_t *rd0, const int32_t *rd1,
                          const int32_t *row_map, const uint8_t *fmask,
                          const double *B, const double *V, double inv_u2,
                          int32_t *rows, double *W, double *gu, int32_t *cols,
                          void *stream);
/* y += T x for that Term 3: t[c] = sum_b gu[4c+b] x[cols[4c+b]], then per
 * row r the entries rent[rptr[r]..rptr[r+1]) (entry q = 4c+a, in a fixed
 * order) add W[q] t[q >> 2].  t is n_cand doubles of scratch; x != y. */
int fp3d_sheet_t3_apply(int64_t n_cand, int64_t n_rows, const double *gu,
                        const int32_t *cols, const double *W, const int32_t *rptr,
                        const int32_t *rent, const double *x, double *t, double *y,
                        void *stream);

/* --------------------------------------------------- host-side set-up
 * (plain host pointers; run once per mesh, as in the reference) */
/* pyfp3d/mesh/coloring.py:65-95 _greedy_coloring_kernel + :129-145
 * color_partition_csr.  colors int32[n_tets]; *n_colors_out; returns
 * FP3D_ELIMIT above 256 colours (the reference's ValueError). */
int fp3d_host_greedy_coloring(int64_t n_tets, const int32_t *elements,
                              int64_t n_nodes, int32_t *colors,
                              int32_t *n_colors_out);
/* pyfp3d/kernels/jacobian.py:60-74 build_csr_pattern: sorted, deduplicated
 * node-node pattern.  Call with indices == NULL to get *nnz_out and
 * indptr, then again with indices allocated. */
int fp3d_host_csr_pattern(int64_t n_tets, const int32_t *elements,
                          int64_t n_nodes, int32_t *indptr, int32_t *indices,
                          int64_t *nnz_out);
/* pyfp3d/kernels/jacobian.py:77-102 build_elem_to_csr (int32 (n,16)). */
int fp3d_host_elem_to_csr(int64_t n_tets, const int32_t *elements,
                          const int32_t *indptr, const int32_t *indices,
                          int32_t *out);
/* Row-parallel host products of the AMG set-up (pyfp3d/solve/linear.py:
 * 61-81 builds the hierarchy with pyamg on scipy.sparse).  C = A B (CSR,
 * int32 indices) in two passes: fp3d_host_spgemm writes C's row pointer
 * (n_rows + 1, caller array) and *nnz_out, fp3d_host_spgemm_fill the
 * caller's indices / data.  Each row is accumulated exactly as
 * scipy.sparse's csr_matmat (A entries in stored order, then B entries,
 * sums from 0; exact zeros dropped), so C carries scipy's values; columns
 * ascend.  fp3d_host_csr_matvec: y = A x, each row summed from 0 in stored
 * order (scipy's csr_matvec).  Threads: at most 16, or FP3D_HOST_THREADS;
 * results do not depend on the count. */
int fp3d_host_spgemm(int64_t n_rows, int64_t n_cols, const int32_t *Ap,
                     const int32_t *Aj, const double *Ax, const int32_t *Bp,
                     const int32_t *Bj, const double *Bx, int32_t *indptr,
                     int64_t *nnz_out);
int fp3d_host_spgemm_fill(int64_t n_rows, int64_t n_cols, const int32_t *Ap,
                          const int32_t *Aj, const double *Ax, const int32_t *Bp,
                          const int32_t *Bj, const double *Bx,
                          const int32_t *indptr, int32_t *indices, double *data);
int fp3d_host_csr_matvec(int64_t n_rows, const int32_t *Ap, const int32_t *Aj,
                         const double *Ax, const double *x, double *y);
/* The smoothed-aggregation prolongator of the same set-up (pyamg's
 * smoothed_aggregation_solver, called at pyfp3d/solve/linear.py:61-81):
 * P = T - diag(s) (A T) for the tentative prolongator T of an aggregation
 * (row j of T: tval[agg[j]] in column agg[j]; no entry when agg[j] = -1)
 * and the per-row damping s_i = (omega / rho) / a_ii, with the values of
 * the scipy recipe (A T accumulated as fp3d_host_spgemm, X = s_i (A T),
 * T - X merged by column, exact zeros dropped).  Two passes as
 * fp3d_host_spgemm (Pp: n + 1).  fp3d_host_csr_transpose: R = P^T (Tp:
 * n_cols + 1), rows of R with ascending columns, values copied. */
int fp3d_host_sa_prolongator(int64_t n, const int32_t *Ap, const int32_t *Aj,
                             const double *Ax, const int32_t *agg, int64_t n_agg,
                             const double *tval, const double *s, int32_t *Pp,
                             int64_t *nnz_out);
int fp3d_host_sa_prolongator_fill(int64_t n, const int32_t *Ap, const int32_t *Aj,
                                  const double *Ax, const int32_t *agg, int64_t n_agg,
                                  const double *tval, const double *s,
                                  const int32_t *Pp, int32_t *Pj, double *Px);
int fp3d_host_csr_transpose(int64_t n_rows, int64_t n_cols, const int32_t *Ap,
                            const int32_t *Aj, const double *Ax, int32_t *Tp,
                            int32_t *Tj, double *Tx);
/* GMRES Hessenberg update (pyfp3d/solve/linear.py:182-193 calls
 * scipy.sparse.linalg.gmres, whose Arnoldi loop rotates each new column by
 * the previous Givens rotations): rotations j = 0 .. col-1 applied in order
 * to hrow, the new column's entries h[0 .. col+1]:
 * (hrow[j], hrow[j+1]) <- (c_j hrow[j] + s_j hrow[j+1],
 *                          -s_j hrow[j] + c_j hrow[j+1]),
 * with (c_j, s_j) = givens[2j], givens[2j + 1].  The same IEEE operations,
 * in the same order, as the host loop it replaces. */
int fp3d_host_apply_givens(int col, double *hrow, const double *givens);
/* pyfp3d/mesh/metrics.py:207-278 build_face_adjacency (int32 (n,4)). */
int fp3d_host_face_adjacency(int64_t n_tets, const int32_t *elements,
                             int32_t *face_neighbors);
/* pyfp3d/mesh/metrics.py:176-204 precompute_element_geometry; returns
 * FP3D_EINVAL with the failing element in *bad_elem on a degenerate tet. */
int fp3d_host_element_geometry(int64_t n_tets, const int32_t *elements,
                               const double *nodes, double *B, double *V,
                               int64_t *bad_elem);
/* Smoothed-aggregation set-up helpers (pyfp3d/solve/linear.py:61-81 uses
 * pyamg.smoothed_aggregation_solver): standard aggregation on the pattern
 * of A (symmetric strength, theta = 0).  agg int32[n] (-1 = isolated);
 * returns the aggregate count in *n_agg. */
int fp3d_host_standard_aggregation(int64_t n, const int32_t *indptr,
                                   const int32_t *indices, const double *data,
                                   int32_t *agg, int64_t *n_agg);

/* Point location for mesh-to-mesh P1 transfer (no reference counterpart;
 * grid sequencing): for each point, the containing tet (visibility walk over
 * face_neighbors, face f opposite local node f) and its 4 barycentric
 * coordinates; a point outside the mesh gets the boundary tet where the
 * walk stopped (coordinates then extrapolate). */
int fp3d_host_locate_points(int64_t n_tets, const int32_t *elements,
                            const double *nodes, const int32_t *face_neighbors,
                            int64_t n_pts, const double *pts,
                            int32_t *out_elem, double *out_bary);

/* Shortest incident edge per node (pyfp3d/wake/cut_elements.py's node
 * length scale for the level-set tolerances): h[v] = min over the six
 * edges (i, j) of every tet containing v of |x_i - x_j|, the length summed
 * as (dx^2 + dy^2) + dz^2 then square-rooted; +inf for a node in no tet. */
int fp3d_host_node_min_edge_length(int64_t n_tets, const int32_t *elements,
                                   int64_t n_nodes, const double *nodes, double *h);

/* Threads per block of the element-wise kernels (state, walk, upwind,
 * sensitivity and element-term sweeps): 64, 128, 192 or 256, process-wide
 * (no reference counterpart; set by up3d_amd.tune). */
int fp3d_set_element_block(int threads);
int fp3d_get_element_block(void);
/* Rows per staged chunk of the incidence-ordered CSR gather
 * (fp3d_gather_csr_terms with inc_ent == NULL): 256, 512, 1024 (default)
 * or 2048.  Changes only the LDS staging; the sums and their order are the
 * same.  No reference counterpart (autotuner). */
int fp3d_set_gather_chunk(int rows);
int fp3d_get_gather_chunk(void);
/* Upstream walks per lane in fp3d_upwind_rho_tilde_sens_d: 1, 2 (default),
 * 3 or 4.  Each lane advances that many elements' walks together so their
 * random reads overlap; the selections and values are the same for every
 * width.  No reference counterpart (autotuner). */
int fp3d_set_walk_width(int walks);
int fp3d_get_walk_width(void);

#ifdef __cplusplus
}
#endif
#endif /* FP3D_H */
