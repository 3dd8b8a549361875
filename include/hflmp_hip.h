/*
 * hflmp_hip.h -- C ABI of the gfx950 HIP platform for HiFlow3's local
 * multi-platform linear algebra ("lmp").
 *
 * Everything here is extern "C" with plain pointers and sizes.  Device
 * pointers are HIP device allocations (hflmp_malloc or any hipMalloc'd
 * memory on the context's device).  All routines enqueue on the context's
 * stream; the ones returning a host scalar synchronise that stream.
 *
 * Return value: HFLMP_OK (0) or a negative HFLMP_E* code; the message of
 * the last error is hflmp_ctx_last_error(ctx).  The C++ platform classes
 * turn a failure into LOG_ERROR + exit, the reference's convention
 * (linear_algebra/lmp/cuda/cuda_utils.h:25-55, common/macros.h:79-84).
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to HiFlow3's src/).
 */
#ifndef HFLMP_HIP_H
#define HFLMP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFLMP_OK 0
#define HFLMP_EINVAL -1   /* bad argument (size, null pointer, shape)      */
#define HFLMP_EHIP -2     /* HIP runtime error                             */
#define HFLMP_ENOMEM -3   /* allocation failed / admission budget exceeded */
#define HFLMP_ESTATE -4   /* object not built / wrong state                */
#define HFLMP_ENUMERIC -5 /* zero diagonal, NaN, breakdown                 */
#define HFLMP_EFORMAT -6  /* matrix does not suit the requested layout     */

typedef struct hflmp_ctx hflmp_ctx;

/* ---- platform / context -------------------------------------------------
 * Replaces init_platform / stop_platform (linear_algebra/lmp/
 * platform_management.h:32-35, cuda/platform_management_gpu.cu:78-143). */
int hflmp_device_count(int *count);
int hflmp_ctx_create(int device, hflmp_ctx **ctx);
int hflmp_ctx_destroy(hflmp_ctx *ctx);
/* Synchronise the context's stream (lVector::Sync, lvector.h:266). */
int hflmp_ctx_sync(hflmp_ctx *ctx);
const char *hflmp_ctx_last_error(const hflmp_ctx *ctx);
/* The stream the context enqueues on (a hipStream_t).  set_stream(NULL)
 * returns to the context's own stream. */
int hflmp_ctx_get_stream(hflmp_ctx *ctx, void **stream);
int hflmp_ctx_set_stream(hflmp_ctx *ctx, void *stream);
int hflmp_ctx_device(const hflmp_ctx *ctx, int *device);

/* HIP events on the context's stream, for timing kernels where they run
 * (the reference times operations on the host only: common/timer.h:30-100,
 * tools/ScopedTimer.h:33-135).  elapsed_ms waits for `stop`. */
int hflmp_event_create(hflmp_ctx *ctx, void **event);
int hflmp_event_record(hflmp_ctx *ctx, void *event);
int hflmp_event_elapsed_ms(hflmp_ctx *ctx, void *start, void *stop, float *ms);
int hflmp_event_destroy(hflmp_ctx *ctx, void *event);

/* ---- memory ----------------------------------------------------------------
 * Replaces lmp_mem / mem_cuda (cuda/mem_cuda.cu:42-135: memcpy2dev,
 * memcpy2host, memcpydev, memsetdev). */
int hflmp_malloc(hflmp_ctx *ctx, size_t bytes, void **dptr);
int hflmp_free(hflmp_ctx *ctx, void *dptr);
int hflmp_memcpy_h2d(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes);
int hflmp_memcpy_d2h(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes);
int hflmp_memcpy_d2d(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes);
/* zero fill by a kernel (no hipMemset node, safe under graph capture) */
int hflmp_fill_zero(hflmp_ctx *ctx, void *dptr, size_t bytes);

/* ---- indexed access (cuda/mem_cuda.h:21-92: kernel_cudasetvalues,
 *      kernel_cudagetvalues, kernel_cudaaddblockvalues) -------------------- */
/* dst[k] = src[idx[k]], k < n  (GetIndexedValues, lvector_gpu.cc:231-255) */
int hflmp_gather_f64(hflmp_ctx *ctx, int n, const int *idx, const double *src,
                     double *dst);
int hflmp_gather_f32(hflmp_ctx *ctx, int n, const int *idx, const float *src,
                     float *dst);
/* dst[idx[k]] = src[k]  (SetValues / SetIndexedValues, lvector_gpu.cc:125-186) */
int hflmp_scatter_f64(hflmp_ctx *ctx, int n, const int *idx, const double *src,
                      double *dst);
int hflmp_scatter_f32(hflmp_ctx *ctx, int n, const int *idx, const float *src,
                      float *dst);
/* dst[idx[k]] += src[k]  (add_values, lvector.h:76-83).  The adds are
 * device atomics: with distinct indices the result equals the sequential
 * loop's; where an index repeats, the order of its additions is not fixed,
 * so the sum may differ from the sequential one in the last bits (and from
 * run to run).  Integer-valued data and single hits are exact. */
int hflmp_scatter_add_f64(hflmp_ctx *ctx, int n, const int *idx,
                          const double *src, double *dst);
int hflmp_scatter_add_f32(hflmp_ctx *ctx, int n, const int *idx,
                          const float *src, float *dst);

/* ---- BLAS-1 (cuda/GPUblas_lvector.cu:75-433, cuBLAS calls) --------------
 * Scalar results: the plain form writes a HOST scalar and synchronises the
 * stream (lVector::Dot returns by value); the _dev form writes a DEVICE
 * scalar and does not synchronise. */
int hflmp_dot_f64(hflmp_ctx *ctx, int n, const double *x, const double *y,
                  double *result);
int hflmp_dot_f32(hflmp_ctx *ctx, int n, const float *x, const float *y,
                  float *result);
int hflmp_dot_dev_f64(hflmp_ctx *ctx, int n, const double *x, const double *y,
                      double *d_result);
int hflmp_dot_dev_f32(hflmp_ctx *ctx, int n, const float *x, const float *y,
                      float *d_result);
/* sqrt(Dot(x,x)) (CPUsimple_lvector.cc:144-153; CoupledVector::Norm2,
 * coupled_vector.cc:392-395) */
int hflmp_nrm2_f64(hflmp_ctx *ctx, int n, const double *x, double *result);
int hflmp_nrm2_f32(hflmp_ctx *ctx, int n, const float *x, float *result);
int hflmp_norm1_f64(hflmp_ctx *ctx, int n, const double *x, double *result);
int hflmp_norm1_f32(hflmp_ctx *ctx, int n, const float *x, float *result);
/* max |x_i| and the first index attaining it (NormMax / ArgMax);
 * argmin gives the first index of min |x_i| (ArgMin) */
int hflmp_normmax_f64(hflmp_ctx *ctx, int n, const double *x, double *result);
int hflmp_normmax_f32(hflmp_ctx *ctx, int n, const float *x, float *result);
int hflmp_argmax_abs_f64(hflmp_ctx *ctx, int n, const double *x, int *index);
int hflmp_argmax_abs_f32(hflmp_ctx *ctx, int n, const float *x, int *index);
int hflmp_argmin_abs_f64(hflmp_ctx *ctx, int n, const double *x, int *index);
int hflmp_argmin_abs_f32(hflmp_ctx *ctx, int n, const float *x, int *index);
/* number of NaN entries (ContainsNaN, lvector_cpu.h:142-145) */
int hflmp_count_nan_f64(hflmp_ctx *ctx, int n, const double *x, int *count);
int hflmp_count_nan_f32(hflmp_ctx *ctx, int n, const float *x, int *count);
/* y += a x  (Axpy) */
int hflmp_axpy_f64(hflmp_ctx *ctx, int n, double a, const double *x, double *y);
int hflmp_axpy_f32(hflmp_ctx *ctx, int n, float a, const float *x, float *y);
/* y = a y + x  (ScaleAdd, lvector.h:282-286) */
int hflmp_scaleadd_f64(hflmp_ctx *ctx, int n, double a, const double *x,
                       double *y);
int hflmp_scaleadd_f32(hflmp_ctx *ctx, int n, float a, const float *x, float *y);
/* y *= a  (Scale) */
int hflmp_scal_f64(hflmp_ctx *ctx, int n, double a, double *y);
int hflmp_scal_f32(hflmp_ctx *ctx, int n, float a, float *y);
/* y[i] *= x[i]  (ElementWiseMult, kernel_cudamultvalues mem_cuda.h:83) */
int hflmp_ewmul_f64(hflmp_ctx *ctx, int n, const double *x, double *y);
int hflmp_ewmul_f32(hflmp_ctx *ctx, int n, const float *x, float *y);
/* y = x */
int hflmp_copy_f64(hflmp_ctx *ctx, int n, const double *x, double *y);
int hflmp_copy_f32(hflmp_ctx *ctx, int n, const float *x, float *y);
/* (Rot, lvector.h:306) x' = c x + s y, y' = c y - s x */
int hflmp_rot_f64(hflmp_ctx *ctx, int n, double *x, double *y, double c,
                  double s);
int hflmp_rot_f32(hflmp_ctx *ctx, int n, float *x, float *y, float c, float s);
/* precision casts (CastFrom/CastTo, kernel_cudacast* mem_cuda.h:94-132) */
int hflmp_cast_f64_to_f32(hflmp_ctx *ctx, int n, const double *x, float *y);
int hflmp_cast_f32_to_f64(hflmp_ctx *ctx, int n, const float *x, double *y);

/* ---- CSR matrix (zero-based, int32 row_ptr[nrow+1], col[nnz]) -----------
 * Replaces GPUscalar_CSR_lMatrix::VectorMult / VectorMultAdd /
 * VectorMultNoDiag (cuda/GPUscalar_csr_lmatrix.cu:67-210) and
 * GPUcublas2_CSR_lMatrix::VectorMult (cuSPARSE csrmv).  Each row is summed
 * in ascending storage order from 0 with rounded products, the order of
 * CPUsimple_CSR_lMatrix::VectorMult (CPUsimple_csr_lmatrix.cc:76-111). */
int hflmp_csr_spmv_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                       const int *row_ptr, const int *col, const double *val,
                       const double *x, double *y);
int hflmp_csr_spmv_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                       const int *row_ptr, const int *col, const float *val,
                       const float *x, float *y);
/* y += A x (VectorMultAdd, lmatrix_csr_cpu.cc:1424-1450) */
int hflmp_csr_spmv_add_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                           const int *row_ptr, const int *col,
                           const double *val, const double *x, double *y);
int hflmp_csr_spmv_add_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                           const int *row_ptr, const int *col, const float *val,
                           const float *x, float *y);
/* y = (A - D) x (kernel_spmvnd_csr_scalar, cuda_spmv_kernel_csr_scalar.h:48) */
int hflmp_csr_spmv_nodiag_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                              const int *row_ptr, const int *col,
                              const double *val, const double *x, double *y);
int hflmp_csr_spmv_nodiag_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                              const int *row_ptr, const int *col,
                              const float *val, const float *x, float *y);
/* invdiag[i] = 1/a_ii (extract_invdiagelements, lmatrix_csr_cpu.cc:
 * 2420-2442); *n_bad (host) = rows whose diagonal is missing or zero. */
int hflmp_csr_extract_invdiag_f64(hflmp_ctx *ctx, int nrow, const int *row_ptr,
                                  const int *col, const double *val,
                                  double *invdiag, int *n_bad);
int hflmp_csr_extract_invdiag_f32(hflmp_ctx *ctx, int nrow, const int *row_ptr,
                                  const int *col, const float *val,
                                  float *invdiag, int *n_bad);
/* diag_ptr[i] = position of a_ii, -1 if absent (get_diag_ptr,
 * lmatrix_formats.h:489-532) */
int hflmp_csr_diag_ptr(hflmp_ctx *ctx, int nrow, const int *row_ptr,
                       const int *col, int *diag_ptr);
/* Dirichlet rows: val = 0 in listed rows, a_ii = alpha (ZeroRows,
 * lmatrix_csr_cpu.cc:908-927) */
int hflmp_csr_zero_rows_f64(hflmp_ctx *ctx, int nrows_zero, const int *rows,
                            const int *row_ptr, const int *col, double *val,
                            double alpha);
int hflmp_csr_zero_rows_f32(hflmp_ctx *ctx, int nrows_zero, const int *rows,
                            const int *row_ptr, const int *col, float *val,
                            float alpha);
/* val_f32 = (float) val_f64 (CastFrom on matrices, lmatrix.h) */
int hflmp_cast_csr_f64_to_f32(hflmp_ctx *ctx, int nnz, const double *val,
                              float *val32);

/* ---- SELL-64 execution layout ---------------------------------------------
 * A device copy of a CSR matrix in sliced ELL: slices of 64 rows (one
 * wave64, lane = row), each padded to its longest row and interleaved in
 * chunks of P entries per lane (P = 2 for FP64, 4 for FP32), so every
 * lane-load instruction reads one contiguous block.  It fills the
 * reference's ELL slot (la_global.h:56, MATRIX_FORMAT::ELL; parsed in
 * init_vec_mat.cc:135-137) for matrices the caller assembles in CSR.
 * Products are bit-identical to hflmp_csr_spmv_* (each row summed in
 * ascending storage order from zero; VectorMult, CPUsimple_csr_lmatrix.cc:
 * 76-111, and VectorMultAdd, lmatrix_csr_cpu.cc:1424-1450).
 * create() builds on the device from the DEVICE CSR arrays (which it does
 * not keep).  flags HFLMP_SELL_COL16 stores columns as int16 offsets from
 * the slice's first row when every offset fits (else int32).  Returns
 * HFLMP_EFORMAT if a row has more than 255 entries or the stored entries
 * (nonzeros plus padding) would exceed 3 nnz + 64 P (unless flags has
 * HFLMP_SELL_ANY_PADDING); use the CSR kernels then.  Each lane stops at
 * its row's length: padding past a row's last chunk of P is never read. */
#define HFLMP_SELL_COL16 1
#define HFLMP_SELL_ANY_PADDING 2
typedef struct hflmp_sell hflmp_sell;
typedef struct {
  int nrow, ncol, nnz;
  int chunk;          /* P: entries per lane load                         */
  int col_bytes;      /* 4 (int32) or 2 (int16 slice offsets)             */
  int value_bytes;    /* 8 or 4                                           */
  long long entries;  /* stored entries including padding                 */
  size_t device_bytes;
} hflmp_sell_info;
int hflmp_sell_create_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                          const int *row_ptr, const int *col,
                          const double *val, int flags, hflmp_sell **out);
int hflmp_sell_create_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                          const int *row_ptr, const int *col, const float *val,
                          int flags, hflmp_sell **out);
/* New values for the same pattern (row_ptr, val: device CSR arrays).
 * Synchronises the stream; HFLMP_EFORMAT if some row's length differs from
 * the pattern the layout was built for (those rows' values are not
 * copied: build a new layout for a new pattern). */
int hflmp_sell_set_values_f64(hflmp_ctx *ctx, hflmp_sell *m,
                              const int *row_ptr, const double *val);
int hflmp_sell_set_values_f32(hflmp_ctx *ctx, hflmp_sell *m,
                              const int *row_ptr, const float *val);
int hflmp_sell_get_info(const hflmp_sell *m, hflmp_sell_info *info);
void hflmp_sell_destroy(hflmp_sell *m);
/* y = A x, y += A x */
int hflmp_sell_spmv_f64(hflmp_ctx *ctx, const hflmp_sell *m, const double *x,
                        double *y);
int hflmp_sell_spmv_f32(hflmp_ctx *ctx, const hflmp_sell *m, const float *x,
                        float *y);
int hflmp_sell_spmv_add_f64(hflmp_ctx *ctx, const hflmp_sell *m,
                            const double *x, double *y);
int hflmp_sell_spmv_add_f32(hflmp_ctx *ctx, const hflmp_sell *m,
                            const float *x, float *y);

/* ---- matrix-free stencil operator ------------------------------------------
 * StencilOperator (linear_algebra/stencil_operator.h:43-330) on a
 * structured mesh of ncell[0] x ncell[1] (x ncell[2]) cells of Q1 (2D, 3D)
 * or Q2 (2D) elements: nodes numbered lexicographically (x fastest), local
 * nodes of a cell likewise.  Cell c contributes f_c E_m (E_m: one of nmat
 * nloc x nloc local matrices, m = cell_lm[c]; the reference's
 * CellMatrixHandler with one matrix per cell); fixed dofs get y_d = x_d
 * (stencil_operator.h:319-325).  Products follow the reference's cell loop
 * (per cell tmp = E_m x_c summed from 0 over ascending local columns, lhs =
 * 0 + (alpha f_c) tmp, added into y in ascending cell order) and are
 * bit-identical to it; no matrix is stored.
 * create() copies the local matrices (HOST array, nmat * nloc * nloc
 * values, row-major), the cell factors and cell_lm (DEVICE arrays, one
 * entry per cell, lexicographic; cell_lm may be NULL when nmat == 1) and
 * the fixed dof list (DEVICE int array).  Vectors and the cell array stay
 * below 2 GiB each (nnode and ncell times value_bytes < 2^31).  Interface
 * ("cross") matrices of discontinuous elements are not supported. */
typedef struct hflmp_stencil hflmp_stencil;
typedef struct {
  int order;     /* 1 or 2                                                 */
  int dim;       /* 2 or 3 (order 2: 2)                                    */
  int ncell[3];  /* cells per direction (ncell[2] unused in 2D)            */
  int nmat;      /* local matrices, 1..256                                 */
} hflmp_stencil_desc;
typedef struct {
  int order, dim, nloc, nmat;
  int ncell[3];
  long long nnode, ncell_total;
  int nfixed;
  int value_bytes;
  size_t device_bytes;
} hflmp_stencil_info;
int hflmp_stencil_create_f64(hflmp_ctx *ctx, const hflmp_stencil_desc *desc,
                             const double *local_matrices,
                             const unsigned char *cell_lm,
                             const double *cell_factor, int nfixed,
                             const int *fixed, hflmp_stencil **out);
int hflmp_stencil_create_f32(hflmp_ctx *ctx, const hflmp_stencil_desc *desc,
                             const float *local_matrices,
                             const unsigned char *cell_lm,
                             const float *cell_factor, int nfixed,
                             const int *fixed, hflmp_stencil **out);
/* New cell factors (DEVICE array, one per cell) */
int hflmp_stencil_set_factors_f64(hflmp_ctx *ctx, hflmp_stencil *m,
                                  const double *cell_factor);
int hflmp_stencil_set_factors_f32(hflmp_ctx *ctx, hflmp_stencil *m,
                                  const float *cell_factor);
int hflmp_stencil_get_info(const hflmp_stencil *m, hflmp_stencil_info *info);
void hflmp_stencil_destroy(hflmp_stencil *m);
/* y = A x (VectorMult, stencil_operator.h:87-91; y is not read) */
int hflmp_stencil_apply_f64(hflmp_ctx *ctx, const hflmp_stencil *m,
                            const double *x, double *y);
int hflmp_stencil_apply_f32(hflmp_ctx *ctx, const hflmp_stencil *m,
                            const float *x, float *y);
/* y = beta y + alpha A x, fixed dofs y_d = x_d (VectorMultAdd,
 * stencil_operator.h:153-330); x and y must differ */
int hflmp_stencil_apply_add_f64(hflmp_ctx *ctx, const hflmp_stencil *m,
                                double alpha, const double *x, double beta,
                                double *y);
int hflmp_stencil_apply_add_f32(hflmp_ctx *ctx, const hflmp_stencil *m,
                                float alpha, const float *x, float beta,
                                float *y);
/* invdiag[i] = 1 / sum_c f_c E_m[l][l] over the node's cells in ascending
 * order (the diagonal cell-by-cell assembly produces), 1 on fixed dofs;
 * the Jacobi preconditioner of the operator (lpreconditioner.cc:124-181 on
 * the assembled matrix).  *n_bad (host) = zero diagonals. */
int hflmp_stencil_extract_invdiag_f64(hflmp_ctx *ctx, const hflmp_stencil *m,
                                      double *invdiag, int *n_bad);
int hflmp_stencil_extract_invdiag_f32(hflmp_ctx *ctx, const hflmp_stencil *m,
                                      float *invdiag, int *n_bad);

/* ---- finite-element assembly into CSR ---------------------------------------
 * The reference assembles on the host: the pattern of every dof pair that
 * shares a cell, columns ascending (compute_sparsity_structure,
 * assembly/assembly_utils.cc:144; coo_2_csr, lmp/lmatrix_formats.h:
 * 320-456), then the values cell by cell in ascending cell order into a
 * zeroed matrix (add_values, lmp/lmatrix_csr_cpu.cc:1102-1153).  These do
 * both on the device for any mesh given as a cell -> dof table (DEVICE
 * int array, ncell x nloc, nloc <= 32), with values bit-identical to that
 * cell loop.
 *   mesh_create: node -> cell incidence and the pattern's size;
 *   mesh_pattern: row_ptr[nnode + 1] and col[nnz] (DEVICE);
 *   assemble: val[nnz] (DEVICE) = sum over cells c of the entries
 *     factor[c] * E[a][b] (E: one nloc x nloc DEVICE matrix, factor NULL
 *     for 1) or cell_mats[c][a][b] (ncell x nloc x nloc, when not NULL),
 *     added in ascending cell order from 0.
 * Dirichlet rows afterwards: hflmp_csr_zero_rows_*. */
typedef struct hflmp_fe_mesh hflmp_fe_mesh;
int hflmp_fe_mesh_create(hflmp_ctx *ctx, int nnode, int ncell, int nloc,
                         const int *cell_dofs, hflmp_fe_mesh **out);
int hflmp_fe_mesh_nnz(const hflmp_fe_mesh *m, long long *nnz);
int hflmp_fe_mesh_pattern(hflmp_ctx *ctx, const hflmp_fe_mesh *m, int *row_ptr,
                          int *col);
int hflmp_fe_assemble_f64(hflmp_ctx *ctx, const hflmp_fe_mesh *m,
                          const double *cell_mats, const double *E,
                          const double *factor, const int *row_ptr,
                          const int *col, double *val);
int hflmp_fe_assemble_f32(hflmp_ctx *ctx, const hflmp_fe_mesh *m,
                          const float *cell_mats, const float *E,
                          const float *factor, const int *row_ptr,
                          const int *col, float *val);
void hflmp_fe_mesh_destroy(hflmp_fe_mesh *m);

/* ---- preconditioners ------------------------------------------------------
 * Jacobi apply z = invdiag .* r fused into one pass
 * (lPreconditioner_Jacobi::ApplylPreconditioner, lpreconditioner.cc:184-200:
 * CopyFrom + ElementWiseMult). */
int hflmp_jacobi_apply_f64(hflmp_ctx *ctx, int n, const double *invdiag,
                           const double *r, double *z);
int hflmp_jacobi_apply_f32(hflmp_ctx *ctx, int n, const float *invdiag,
                           const float *r, float *z);

/* Multicoloured symmetric Gauss-Seidel
 * (lPreconditioner_MultiColoring_SymmetricGaussSeidel,
 *  lpreconditioner_mc.cc:147-723, lpreconditioner_mc_sgs.cc:44-90;
 *  colouring CPU_CSR_lMatrix::Multicoloring lmatrix_csr_cpu.cc:2195-2300).
 * create() takes the HOST colouring (h_perm: rows grouped by colour,
 * h_color_sizes[ncolors]) and the DEVICE matrix, checks that no two rows of
 * one colour couple, and builds the per-colour split; apply() computes
 * z = M^-1 r with M = w(2-w) (D + wL) D^-1 (D + wU) in the colour order.
 * The object keeps pointers to col/val, which must outlive it; new values
 * in val are picked up without a rebuild. */
typedef struct hflmp_mcsgs hflmp_mcsgs;
int hflmp_mcsgs_create_f64(hflmp_ctx *ctx, int nrow, int nnz,
                           const int *row_ptr, const int *col,
                           const double *val, int ncolors,
                           const int *h_color_sizes, const int *h_perm,
                           double omega, hflmp_mcsgs **out);
int hflmp_mcsgs_apply_f64(hflmp_ctx *ctx, const hflmp_mcsgs *m,
                          const double *r, double *z);
int hflmp_mcsgs_destroy(hflmp_ctx *ctx, hflmp_mcsgs *m);

/* Symmetric Gauss-Seidel in the sequential row order
 * (CPU_CSR_lMatrix::Psgauss_seidel, lmatrix_csr_cpu.cc:495-565, applied by
 * lPreconditioner_SymmetricGaussSeidel, lpreconditioner.cc:248-257) and its
 * block form (BlocksPsgauss_seidel, lmatrix_csr_cpu.cc:568-641, and the
 * CUDA platform's GPUscalar_CSR_lMatrix::BlocksPsgauss_seidel,
 * cuda/GPUscalar_csr_lmatrix.cu:314-363; applied by
 * lPreconditioner_BlocksSymmetricGaussSeidel, lpreconditioner.cc:285-293):
 * z = M^-1 r with M = (D + L) D^-1 (D + U).  num_blocks row ranges of
 * nrow / num_blocks rows (the last takes the rest; num_blocks = 1 is
 * Psgauss_seidel) are swept independently over the entries whose column
 * lies in the range.  Per row the reference's operations in its order:
 * forward (rows ascending) z_i = r_i, z_i -= z_c a_ij over the lower
 * entries in storage order up to the first diagonal entry, z_i /= a_ii;
 * then z_i *= a_ii; backward (rows descending) z_i -= z_c a_ij over every
 * upper entry, z_i /= a_ii (last diagonal entry).  Rows run by dependency
 * level, so z is bit-identical to the sequential sweeps.
 * create() analyses the pattern on the host (h_row_ptr/h_col: HOST copies
 * of it, or both NULL to download the DEVICE arrays) and keeps the DEVICE
 * row_ptr, col and val pointers, which must outlive the object; new values
 * are used without a rebuild.  HFLMP_EFORMAT if a row has no diagonal
 * entry.  r and z may be the same array. */
typedef struct hflmp_sgs hflmp_sgs;
typedef struct {
  int nrow, num_blocks;
  int levels_forward, levels_backward;  /* dependency levels of the sweeps */
  int launches;                         /* kernel launches per apply       */
} hflmp_sgs_info;
int hflmp_sgs_create_f64(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                         const int *col, const double *val,
                         const int *h_row_ptr, const int *h_col, int num_blocks,
                         hflmp_sgs **out);
int hflmp_sgs_create_f32(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                         const int *col, const float *val, const int *h_row_ptr,
                         const int *h_col, int num_blocks, hflmp_sgs **out);
int hflmp_sgs_apply_f64(hflmp_ctx *ctx, const hflmp_sgs *m, const double *r,
                        double *z);
int hflmp_sgs_apply_f32(hflmp_ctx *ctx, const hflmp_sgs *m, const float *r,
                        float *z);
int hflmp_sgs_get_info(const hflmp_sgs *m, hflmp_sgs_info *info);
void hflmp_sgs_destroy(hflmp_sgs *m);

/* ---- device-resident Krylov solves ----------------------------------------
 * Jacobi-preconditioned CG with the arithmetic of CG::SolvePrecond
 * (linear_solver/cg.h:165-246) and the stopping rule of IterateControl::
 * Check (common/iterate_control.cc:80-107) on the recursive residual, with
 * every scalar kept on the device: one iteration is three kernels (SpMV
 * fused with p.Ap; x/r/z update fused with r.r and r.z; p update) and the
 * host reads the state only every `poll` iterations.  x holds x0 on entry.
 * Workspace: hflmp_pcg_workspace_bytes() bytes of device memory. */
typedef struct {
  int iters;          /* iterations taken                                  */
  int status;         /* 0 converged (abs or rel), 1 max iterations,
                         2 divergence, 3 breakdown (p.Ap <= 0 or NaN)     */
  double res_init;    /* ||b - A x0||                                      */
  double res_final;   /* recursive residual at exit                        */
  double seconds;     /* wall time of the iteration loop                   */
} hflmp_solve_info;

size_t hflmp_pcg_workspace_bytes(int nrow, int nnz, int value_bytes);
int hflmp_pcg_jacobi_f64(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                         const int *col, const double *val,
                         const double *invdiag, const double *b, double *x,
                         int maxit, double atol, double rtol, double dtol,
                         void *work, size_t work_bytes, hflmp_solve_info *info);
int hflmp_pcg_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                         const int *col, const float *val,
                         const float *invdiag, const float *b, float *x,
                         int maxit, double atol, double rtol, double dtol,
                         void *work, size_t work_bytes, hflmp_solve_info *info);
/* The same solve with the operator in SELL-64 layout.  The products are
 * identical; the dot products group their per-workgroup partials over 256
 * rows instead of one CSR tile, so iterates agree to rounding.  Workspace
 * as hflmp_pcg_workspace_bytes(nrow, nnz, value_bytes). */
int hflmp_pcg_jacobi_sell_f64(hflmp_ctx *ctx, const hflmp_sell *A,
                              const double *invdiag, const double *b, double *x,
                              int maxit, double atol, double rtol, double dtol,
                              void *work, size_t work_bytes,
                              hflmp_solve_info *info);
int hflmp_pcg_jacobi_sell_f32(hflmp_ctx *ctx, const hflmp_sell *A,
                              const float *invdiag, const float *b, float *x,
                              int maxit, double atol, double rtol, double dtol,
                              void *work, size_t work_bytes,
                              hflmp_solve_info *info);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hflmp_pcg_jacobi_sell_f64 with its workspace in a
 * static device array of the library instead of `work` -- the same kernels
 * and partial-sum grids, so x, the iteration count and the residuals are
 * bit-identical to it.  It calls no hipMalloc/hipFree and creates no
 * events (it polls the state by a copy into the context's pinned staging
 * and a stream wait every 16 iterations).  At most 262144 rows (the array
 * is 8.4 MB, held on each GPU the library runs on until the process ends);
 * larger systems are refused with HFLMP_EINVAL.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int hflmp_pcg_jacobi_sell_noalloc_f64(hflmp_ctx *ctx, const hflmp_sell *A,
                                      const double *invdiag, const double *b,
                                      double *x, int maxit, double atol,
                                      double rtol, double dtol,
                                      hflmp_solve_info *info);
/* The same solve with the matrix-free stencil operator (CG<LAD> over a
 * StencilOperator, stencil_operator.h:43-330): Ap from the cells, every
 * other step as above.  Workspace: hflmp_pcg_stencil_workspace_bytes(). */
size_t hflmp_pcg_stencil_workspace_bytes(long long nnode, int value_bytes);
int hflmp_pcg_jacobi_stencil_f64(hflmp_ctx *ctx, const hflmp_stencil *A,
                                 const double *invdiag, const double *b,
                                 double *x, int maxit, double atol,
                                 double rtol, double dtol, void *work,
                                 size_t work_bytes, hflmp_solve_info *info);
int hflmp_pcg_jacobi_stencil_f32(hflmp_ctx *ctx, const hflmp_stencil *A,
                                 const float *invdiag, const float *b, float *x,
                                 int maxit, double atol, double rtol,
                                 double dtol, void *work, size_t work_bytes,
                                 hflmp_solve_info *info);

/* ---- block-asynchronous Jacobi ---------------------------------------------
 * AsynchronousIterationGPU (linear_solver/gpu-based/async_iter_gpu.h:32-157,
 * async_iter_gpu.cc:493-820; kernels cuda/cuda_async_iter_kernels.h:17-283)
 * made deterministic on one device.  Rows are cut into blocks of block_rows
 * consecutive rows (block_rows >= nrow: one block), each in the part of
 * the reference's device-local row set.  One launch (block iteration): for
 * every row i of block B, c_i = b_i - sum of a_ij x_j over its entries
 * outside B (x at the start of the launch, storage order; the reference's
 * c = b - A_offdiag x_ghost), then n_inner synchronous sweeps over B,
 *   tmp = c_i - sum of a_ij y_j over the entries in B with j != i,
 *   y_i = D^-1_i tmp (kernel_async_inner_iter) or, damped,
 *   y_i = (w D^-1_i tmp) + ((1 - w) y_i) (kernel_async_inner_damped_iter),
 * from y = x on B.  The reference's in-place updates within a device read
 * whatever the neighbouring rows hold at the time; here every sweep of a
 * block reads the previous sweep, so results do not depend on scheduling.
 * Blocks of at most 2048 rows run in one workgroup each (values in LDS);
 * larger ones sweep through global memory with the same arithmetic.
 * mode (SetSolveMode): 0 solve, 1 solve damped -- repeat n_global * n_block
 * launches, then the residual r_i = b_i - sum_{j != i} a_ij x_j - a_ii x_i
 * (kernel_compute_block_squared_residual_no_offdiag), ||r|| and
 * IterateControl::Check on the number of inner sweeps, until it stops;
 * 2 smooth, 3 smooth damped -- the launches once, no residual.  D^-1 = 1 /
 * the first diagonal entry; HFLMP_ENUMERIC if one is missing or zero.
 * x holds x0 on entry.  Workspace: hflmp_async_workspace_bytes(). */
typedef struct {
  int block_rows;                 /* rows per block, >= 1                   */
  int n_global, n_block, n_inner; /* SetNumIter                             */
  int mode;                       /* 0..3 as above                          */
  double w;                       /* damping of modes 1 and 3 (> 0)         */
} hflmp_async_params;
size_t hflmp_async_workspace_bytes(int nrow, int value_bytes);
int hflmp_async_jacobi_f64(hflmp_ctx *ctx, int nrow, int nnz,
                           const int *row_ptr, const int *col,
                           const double *val, const double *b, double *x,
                           const hflmp_async_params *prm, int maxit,
                           double atol, double rtol, double dtol, void *work,
                           size_t work_bytes, hflmp_solve_info *info);
int hflmp_async_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz,
                           const int *row_ptr, const int *col, const float *val,
                           const float *b, float *x,
                           const hflmp_async_params *prm, int maxit,
                           double atol, double rtol, double dtol, void *work,
                           size_t work_bytes, hflmp_solve_info *info);

/* ---- geometric multigrid ----------------------------------------------------
 * GMG (linear_solver/gmg.h:979-1230, GMGBase::SolveImpl / SolveImplLevel)
 * over nlevels CSR operators A[0] (finest) .. A[nlevels-1] with
 * restrictions R[l] (n_{l+1} x n_l) and prolongations P[l] (n_l x n_{l+1});
 * all arrays are DEVICE CSR arrays, copied at create() into SELL-64
 * layouts (any padding accepted).  One V-cycle z = M r from level l with
 * right-hand side b_l (b_0 = r):
 *   x_l = w D^-1 b_l, then pre_smooth-1 damped Jacobi sweeps
 *     x = (w D^-1 (b - (A - D) x)) + ((1 - w) x)
 *     (kernel_async_inner_damped_iter_no_offdiag, cuda_async_iter_kernels.h:
 *     116-144; the diagonal is skipped, D^-1 from hflmp_csr_extract_invdiag);
 *   r_l = b_l - A x_l, subtracted term by term in row order
 *     (kernel_gmg_compute_residual_no_offdiag, cuda_gmg_kernel.h:46-64);
 *   b_{l+1} = R r_l (kernel_gmg_compute_restriction_no_offdiag, :89-108);
 *   the cycle from level l+1;
 *   x_l = x_l + P x_{l+1}, added term by term onto x_l
 *     (kernel_gmg_compute_updated_solution_no_offdiag, :137-157);
 *   post_smooth sweeps.
 * The coarsest level takes coarse_sweeps sweeps with coarse_omega.  With
 * pre_smooth == post_smooth and R = P^T the cycle is a symmetric operator
 * and preconditions CG.  Dirichlet (unit) rows with zero right-hand side
 * stay zero on every level when R has zero rows for coarse Dirichlet nodes
 * and P zero rows for fine ones. */
typedef struct {
  int pre_smooth;       /* sweeps before the coarse correction, >= 1       */
  int post_smooth;      /* sweeps after it, >= 0                           */
  int coarse_sweeps;    /* sweeps on the coarsest level, >= 1              */
  double omega;         /* damping on the other levels, > 0                */
  double coarse_omega;  /* damping on the coarsest level, > 0              */
} hflmp_mg_params;
typedef struct {
  int nrow, ncol, nnz;
  const int *row_ptr, *col;
  const double *val;
} hflmp_csr_f64;
typedef struct hflmp_mg hflmp_mg;
int hflmp_mg_create_f64(hflmp_ctx *ctx, int nlevels, const hflmp_csr_f64 *A,
                        const hflmp_csr_f64 *R, const hflmp_csr_f64 *P,
                        const hflmp_mg_params *prm, hflmp_mg **out);
/* The same with the finest operator matrix-free: S (FP64 stencil, which
 * must outlive the hierarchy) replaces A[0] (not read) on level 0, and
 * D^-1 comes from S.  Level 0 then smooths with the reference's Richardson
 * iteration with the Jacobi preconditioner (linear_solver/richardson.h:
 * 136-251, SolveImplFullResidual): r = b, S.VectorMultAdd(-1, x, 1, r),
 * x = x + w (r .* D^-1); its residual is that VectorMultAdd, the first step
 * x = w (b .* D^-1). */
int hflmp_mg_create_stencil_f64(hflmp_ctx *ctx, const hflmp_stencil *S,
                                int nlevels, const hflmp_csr_f64 *A,
                                const hflmp_csr_f64 *R, const hflmp_csr_f64 *P,
                                const hflmp_mg_params *prm, hflmp_mg **out);
void hflmp_mg_destroy(hflmp_mg *m);
/* Coarsest level solved exactly: x = C b with C (DEVICE, n x n row-major,
 * n <= 4096 rows of the coarsest level) its operator's inverse, each row
 * summed from 0 in column order, instead of coarse_sweeps sweeps (the
 * reference's GMG takes any LinearSolver as its coarse solver cS_,
 * gmg.h:1159).  C is copied; NULL returns to the sweeps. */
int hflmp_mg_set_coarse_inverse_f64(hflmp_ctx *ctx, hflmp_mg *m,
                                    const double *inv);
/* No reference counterpart (the reference's GMG always applies its R and P
 * matrices, cuda_gmg_kernel.h:89-157): the levels are the halved
 * structured Q1 meshes of ncells, ncells/2, ... cells per direction (nodes
 * lexicographic, x fastest) and R, P the Q1 interpolation with Dirichlet
 * rows and columns removed (as built by ba_amd/mg.py); restriction and
 * prolongation then run without reading R and P, bit-identical to the
 * products over them.  dim 0 switches back.  A hierarchy runs one cycle at
 * a time: its level vectors are shared by every call. */
int hflmp_mg_set_structured_q1(hflmp_ctx *ctx, hflmp_mg *m, int dim,
                               int ncells);
/* z = M r (one V-cycle from z = 0) */
int hflmp_mg_apply_f64(hflmp_ctx *ctx, const hflmp_mg *m, const double *r,
                       double *z);
/* CG::SolvePrecond (linear_solver/cg.h:165-246) with z = M r, every scalar
 * on the device as in hflmp_pcg_jacobi_f64; A[0] is the operator. */
size_t hflmp_mg_workspace_bytes(const hflmp_mg *m);
int hflmp_pcg_mg_f64(hflmp_ctx *ctx, const hflmp_mg *m, const double *b,
                     double *x, int maxit, double atol, double rtol,
                     double dtol, void *work, size_t work_bytes,
                     hflmp_solve_info *info);
/* Multigrid as the solver (GMGBase::SolveImpl, gmg.h:979-1066): r = b - A x,
 * then x = x + M r and a new residual per iteration until IterateControl
 * stops on ||r||. */
int hflmp_mg_solve_f64(hflmp_ctx *ctx, const hflmp_mg *m, const double *b,
                       double *x, int maxit, double atol, double rtol,
                       double dtol, void *work, size_t work_bytes,
                       hflmp_solve_info *info);

/* ---- several right-hand sides (no reference counterpart: the reference
 * solves one right-hand side per CG::Solve, linear_solver/cg.h:165-246) ---
 * Blocks of nrhs <= 64 vectors are row-major: X[i*nrhs + c] is entry i of
 * column c.  Y = A X sums every column in ascending storage order, so
 * column c is bit-identical to hflmp_csr_spmv of column c. */
int hflmp_csr_spmm_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz, int nrhs,
                       const int *row_ptr, const int *col, const double *val,
                       const double *X, double *Y);
int hflmp_csr_spmm_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz, int nrhs,
                       const int *row_ptr, const int *col, const float *val,
                       const float *X, float *Y);
/* Row-block layout for the same product (no reference counterpart): rows
 * in blocks, each block's distinct columns listed once in ascending order
 * and every entry's column given as its position in that list.  A
 * workgroup loads a block's X rows into LDS once and all the block's
 * entries read them there, the next block's rows loading while one is
 * computed.  A block is a run of a consecutive rows (a = 64..4), or an
 * a x nl patch: the rows at positions c a .. c a + a - 1 of nl consecutive
 * lines of L rows, where L is read off the pattern (the offset of the
 * band next to the diagonal band, as a structured mesh's numbering gives).
 * Of the geometries whose blocks keep to 80 (FP64) or 128 (FP32) distinct
 * columns within a window of 65,536 columns, 1024 entries, and lds_budget
 * bytes of LDS (0 = 64 KiB, at most 160 KiB) for products of nrhs vectors
 * of value_bytes (4 or 8), the plan takes the one with the fewest staged X
 * rows; HFLMP_EFORMAT if none fits, HFLMP_EINVAL for a column outside
 * [0, ncol).  The plan keeps no reference to row_ptr/col; the products
 * take the matrix's row_ptr and values, which must be those the plan was
 * built from, and any nrhs <= 64 whose LDS need stays within 160 KiB.
 * Products are bit-identical to hflmp_csr_spmm_*. */
typedef struct hflmp_spmm_plan hflmp_spmm_plan;
int hflmp_spmm_plan_create(hflmp_ctx *ctx, int nrow, int ncol, int nnz,
                           const int *row_ptr, const int *col, int nrhs,
                           int value_bytes, int lds_budget,
                           hflmp_spmm_plan **plan);
void hflmp_spmm_plan_destroy(hflmp_spmm_plan *plan);
int hflmp_spmm_plan_info(const hflmp_spmm_plan *plan, int *rows_per_block,
                         int *max_cols_per_block, long long *list_entries,
                         long long *device_bytes);
/* the chosen geometry: line length L (nrow for runs), rows per run a, runs
 * (lines) per block nl */
int hflmp_spmm_plan_geometry(const hflmp_spmm_plan *plan, int *line_rows,
                             int *run_rows, int *runs);
int hflmp_plan_spmm_f64(hflmp_ctx *ctx, const hflmp_spmm_plan *plan,
                        const int *row_ptr, const double *val,
                        const double *X, double *Y, int nrhs);
int hflmp_plan_spmm_f32(hflmp_ctx *ctx, const hflmp_spmm_plan *plan,
                        const int *row_ptr, const float *val, const float *X,
                        float *Y, int nrhs);
/* Jacobi-PCG on nrhs systems with the same matrix, advanced in lockstep:
 * every column follows the single-vector iteration with its own alpha,
 * beta and stopping test and stops updating when it converges.  Per-column
 * outputs (host arrays of nrhs, any may be NULL): iterations, status (as
 * hflmp_solve_info), initial and final recursive residuals; *seconds is the
 * wall time.  Workspace: hflmp_batched_pcg_workspace_bytes(). */
size_t hflmp_batched_pcg_workspace_bytes(hflmp_ctx *ctx, int nrow, int nrhs,
                                         int value_bytes);
int hflmp_batched_pcg_jacobi_f64(hflmp_ctx *ctx, int nrow, int nnz, int nrhs,
                                 const int *row_ptr, const int *col,
                                 const double *val, const double *invdiag,
                                 const double *B, double *X, int maxit,
                                 double atol, double rtol, double dtol,
                                 void *work, size_t work_bytes, int *iters,
                                 int *status, double *res_init,
                                 double *res_final, double *seconds);
int hflmp_batched_pcg_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz, int nrhs,
                                 const int *row_ptr, const int *col,
                                 const float *val, const float *invdiag,
                                 const float *B, float *X, int maxit,
                                 double atol, double rtol, double dtol,
                                 void *work, size_t work_bytes, int *iters,
                                 int *status, double *res_init,
                                 double *res_final, double *seconds);
/* The same iteration with its products Y = A P through a block layout of
 * the matrix (hflmp_spmm_plan_create on row_ptr/col with this nrhs and
 * value size): the products are bit-identical to the CSR walk's, the
 * per-column p.Ap partial sums are added in the layout's block order (the
 * iterates match the CSR form's to rounding).  Same workspace. */
int hflmp_batched_pcg_jacobi_plan_f64(hflmp_ctx *ctx, const hflmp_spmm_plan *plan,
                                      int nrow, int nnz, int nrhs,
                                      const int *row_ptr, const int *col,
                                      const double *val, const double *invdiag,
                                      const double *B, double *X, int maxit,
                                      double atol, double rtol, double dtol,
                                      void *work, size_t work_bytes, int *iters,
                                      int *status, double *res_init,
                                      double *res_final, double *seconds);
int hflmp_batched_pcg_jacobi_plan_f32(hflmp_ctx *ctx, const hflmp_spmm_plan *plan,
                                      int nrow, int nnz, int nrhs,
                                      const int *row_ptr, const int *col,
                                      const float *val, const float *invdiag,
                                      const float *B, float *X, int maxit,
                                      double atol, double rtol, double dtol,
                                      void *work, size_t work_bytes, int *iters,
                                      int *status, double *res_init,
                                      double *res_final, double *seconds);

/* The same refinement for a block of nrhs <= 64 right-hand sides (B, X
 * row-major FP64; no reference counterpart for the block): MPIR<
 * LADescriptorCoupledD, LADescriptorCoupledS> (linear_solver/mpir.cc:
 * 52-296) per column.  Each outer step forms R = B - A X in FP64 (one SpMM
 * over the block) and every column's ||r||; a column stops when ||r|| <=
 * atol or < rtol ||r_0|| (its initial residual) or after max_outer
 * corrections; the others get R32 = (float) R and the FP32 lockstep
 * Jacobi-PCG of hflmp_batched_pcg_jacobi_f32 (val32, invdiag32) solves
 * A32 C = R32 from C = 0 to inner_rtol (at most inner_maxit iterations),
 * then X += (double) C.  Per-column outputs (host arrays of nrhs, any may
 * be NULL): corrections, inner iterations over all corrections, status,
 * initial and final FP64 residual norms.  Workspace:
 * hflmp_batched_mpcg_workspace_bytes(). */
size_t hflmp_batched_mpcg_workspace_bytes(hflmp_ctx *ctx, int nrow, int nrhs);
int hflmp_batched_mpcg_jacobi(hflmp_ctx *ctx, int nrow, int nnz, int nrhs,
                              const int *row_ptr, const int *col,
                              const double *val, const float *val32,
                              const float *invdiag32, const double *B,
                              double *X, int max_outer, double atol,
                              double rtol, int inner_maxit, double inner_rtol,
                              void *work, size_t work_bytes, int *outer_iters,
                              int *inner_iters, int *status, double *res_init,
                              double *res_final, double *seconds);
/* The same with the inner FP32 solves' products through a block layout
 * of the FP32 matrix (hflmp_spmm_plan_create with value_bytes 4 and this
 * nrhs), as hflmp_batched_pcg_jacobi_plan_f32. */
int hflmp_batched_mpcg_jacobi_plan(hflmp_ctx *ctx, const hflmp_spmm_plan *plan32,
                                   int nrow, int nnz, int nrhs,
                                   const int *row_ptr, const int *col,
                                   const double *val, const float *val32,
                                   const float *invdiag32, const double *B,
                                   double *X, int max_outer, double atol,
                                   double rtol, int inner_maxit,
                                   double inner_rtol, void *work,
                                   size_t work_bytes, int *outer_iters,
                                   int *inner_iters, int *status,
                                   double *res_init, double *res_final,
                                   double *seconds);

/* FP32-inner / FP64-refined Jacobi-CG: the refinement of
 * MPIR<LADescriptorCoupledD, LADescriptorCoupledS> (linear_solver/mpir.cc:
 * 52-296) with every step on the device.  Per outer step: r = b - A x and
 * ||r|| in FP64 (one fused SpMV kernel that also writes r32 = (float) r and
 * starts the inner solve), then Jacobi-PCG in FP32 on val32 / invdiag32
 * (FP32 values, vectors and reductions) from c = 0 to ||r32 - A32 c|| <
 * inner_rtol ||r32|| or inner_maxit, then x += (double) c.  Stops when
 * ||r|| <= atol or ||r|| < rtol ||r_0|| (r_0 the initial residual, kept for
 * the whole run; the reference resets it every step, mpir.cc:240), or after
 * max_outer corrections.  val32 shares row_ptr/col with val.
 * Each correction's inner tolerance is max(inner_rtol, rtol ||r_0|| /
 * (2 ||r||)): no solve goes further than the outer target needs.
 * The FP32 operator reads its columns as int16 offsets from each row
 * tile's first row (built in the workspace at the start of the call) when
 * every offset fits, else as int32; flags HFLMP_MP_INT32_COLS forces int32.
 * The FP64 residual always uses the int32 CSR arrays. */
#define HFLMP_MP_INT32_COLS 1
/* Reliable updates: after the first outer step the inner CG is not
 * restarted; its recursive residual is replaced by the FP64 true residual
 * and its search direction kept (beta from the replaced residual), which
 * keeps the Krylov space built so far.  Without the flag each correction
 * restarts the inner solve, as MPIR's inner solver does. */
#define HFLMP_MP_RELIABLE 2
typedef struct {
  int outer_iters;  /* corrections applied                                */
  int inner_iters;  /* FP32 CG iterations over all corrections           */
  int status;       /* 0 converged, 1 max_outer reached, 3 breakdown      */
  double res_init;  /* ||b - A x0|| (FP64)                                 */
  double res_final; /* ||b - A x|| (FP64, true residual)                   */
  double seconds;
} hflmp_mp_info;

size_t hflmp_mpcg_workspace_bytes(int nrow, int nnz);
int hflmp_mpcg_jacobi(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                      const int *col, const double *val, const float *val32,
                      const float *invdiag32, const double *b, double *x,
                      int max_outer, double atol, double rtol, int inner_maxit,
                      double inner_rtol, int flags, void *work, size_t work_bytes,
                      hflmp_mp_info *info);
/* The same refinement with both operators in SELL-64 layout: A64 (FP64
 * values) for the residual and A32 (FP32 values, int16 columns where they
 * fit) for the inner solve; flags: HFLMP_MP_RELIABLE.  Workspace:
 * hflmp_mpcg_sell_workspace_bytes. */
size_t hflmp_mpcg_sell_workspace_bytes(int nrow);
int hflmp_mpcg_jacobi_sell(hflmp_ctx *ctx, const hflmp_sell *A64,
                           const hflmp_sell *A32, const float *invdiag32,
                           const double *b, double *x, int max_outer,
                           double atol, double rtol, int inner_maxit,
                           double inner_rtol, int flags, void *work,
                           size_t work_bytes, hflmp_mp_info *info);

/* ---- solve server (no reference counterpart: the reference runs one
 * CG::Solve at a time, linear_solver/cg.h:165-246) -------------------------
 * Many Jacobi-PCG solves of different sizes advanced together on one
 * device, each in FP64 or in the FP32-inner / FP64-refined mode of
 * hflmp_mpcg_jacobi (MPIR, linear_solver/mpir.cc:52-296: FP64 residual,
 * correction and outer test; FP32 inner CG; at most 50 corrections).
 * One arena of hbm_budget bytes holds the admitted systems; a submitted
 * system waits in a first-come-first-served queue until a slot
 * is free and its footprint (hflmp_server_footprint: CSR, SELL-64 copy,
 * inverse diagonal(s), b, x and the Krylov vectors) fits.  hflmp_server_step
 * queues the next chunk of chunk_iters iterations of every active system
 * in shared launches (one launch per kernel for all systems, captured as a
 * hipGraph per active set when use_graphs) behind the chunk still running;
 * while they run it admits what fits (the server's admission thread does
 * the uploads and queues the layout kernels on a second stream); then it
 * waits for the older chunk and retires the systems that stopped in it
 * (their x is copied to the caller's buffer), and the admitted systems
 * whose admission work has finished join, starting with the next chunk
 * queued (the others at a later step; with nothing running, step waits
 * for them).  So a step returns with one chunk in flight, and a system's
 * result arrives in the step after the chunk in which it stopped.  A
 * server is driven from one thread at a time.  Each system keeps its own scalars and partial sums
 * in a fixed order, so its result is bit-identical alone and in any mix.
 * Solves start from x0 = 0.  FP64 solves stop as IterateControl::Check
 * (common/iterate_control.cc:80-107) on the recursive residual; mixed
 * solves stop on the FP64 true residual ||b - A x|| <= atol or
 * < rtol ||b||, with maxit bounding each inner solve.  precision 0 at
 * submission picks the mixed mode for mixed_min_nnz <= nnz < mixed_max_nnz
 * (mixed_min_nnz 0: never; mixed_max_nnz 0: no upper bound).
 * Both SELL-64 operators (FP64, and the FP32 one of a mixed system) hold
 * int16 column offsets from each slice's first row when every column fits
 * (decided on the device at admission; the footprint assumes int32), else
 * int32 columns.
 * FP64 systems with at least solo_min_nnz nonzeros run alone instead: the
 * single-system iteration of hflmp_pcg_jacobi_sell_f64 (three launches per
 * iteration, its grids) on a third stream of the server, concurrently
 * with the grouped launches, chunk_iters iterations per step; their
 * iterates are that solver's bit for bit (with dtol unbounded), and their
 * footprint holds a partial-sum area of max(2 x units, 4096) values.
 * The caller's host arrays (row_ptr, col, val, b, x) must stay valid until
 * the ticket is done.  Rows may hold 0..255 entries. */
typedef struct hflmp_server hflmp_server;
typedef struct {
  size_t hbm_budget; /* device bytes for admitted systems (>= 1 MiB)       */
  int max_active;    /* systems advanced together, 1..4096                */
  int chunk_iters;   /* CG iterations per step, 1..256                    */
  int use_graphs;    /* 1: replay each step's launches as a hipGraph      */
  int mixed_min_nnz;    /* precision 0: mixed from this many nonzeros (0: off) */
  double inner_rtol;    /* mixed: residual drop between replacements (0: 1e-2) */
  int mixed_max_nnz;    /* precision 0: mixed below this many nonzeros (0: no bound) */
  int solo_min_nnz;     /* FP64 systems from this many nonzeros run alone
                           on the second stream (0: none)                 */
} hflmp_server_config;
typedef struct {
  int state;              /* 0 queued, 1 active, 2 done                   */
  int iters;
  int status;             /* as hflmp_solve_info.status once done, else -1;
                             4: refused, a column index lies outside
                             0..nrow-1 (checked on the device at admission
                             before any product; x is set to 0)           */
  double res_init, res_final;
  int zero_diagonals;     /* rows whose diagonal is missing or zero        */
  size_t footprint;       /* arena bytes that admission takes; the CSR
                             arrays' share (12 nnz + 4 (nrow+1) bytes,
                             rounded) returns to the arena as soon as the
                             SELL-64 layout is built                       */
  double queued_seconds;  /* submit -> admission (-1 while queued)         */
  double latency_seconds; /* submit -> x in the caller's buffer (-1 before) */
  int precision;          /* 64 or 32 (FP32-inner / FP64-refined)          */
  int corrections;        /* mixed: outer corrections applied              */
  int col_bytes;          /* done: column width of the FP64 operator, 2 (int16
                             offsets, when every column is within int16 of its
                             slice's first row) or 4                         */
} hflmp_server_result;
typedef struct {
  int queued, active, active_units;
  size_t arena_free, arena_largest_free;
  long long steps, graph_captures;
  int peak_active;          /* most systems active at once so far          */
  size_t peak_arena_in_use; /* most arena bytes held at once so far        */
  /* host seconds spent in step() so far: admitting (slots, blocks and
   * descriptors on the calling thread; the uploads and layout kernels are
   * the admission thread's), rebuilding tables, capturing graphs and
   * launching, and waiting for the chunks' kernels */
  double admit_seconds, capture_seconds, wait_seconds;
  long long solo_runs; /* systems started alone so far (solo_min_nnz)     */
  double join_seconds; /* joining admitted systems (waiting for their
                        * admission work only when nothing runs)          */
  double retire_seconds; /* copying retired systems' x to the callers      */
  /* device seconds of the grouped chunks so far (HIP events around each
   * chunk's work on the context stream: table uploads and its launches);
   * against the host's clock it gives the time the device spent idle
   * between chunks */
  double chunk_device_seconds;
} hflmp_server_stats_t;
int hflmp_server_create(hflmp_ctx *ctx, const hflmp_server_config *cfg,
                        hflmp_server **out);
void hflmp_server_destroy(hflmp_server *sv);
/* footprint for precision 64 or 32, for a system in the grouped launches
 * (int32 columns assumed; col is not read and may be NULL) */
size_t hflmp_server_footprint(int nrow, int nnz, const int *row_ptr,
                              const int *col, int precision);
int hflmp_server_submit(hflmp_server *sv, int nrow, int nnz, const int *row_ptr,
                        const int *col, const double *val, const double *b,
                        double *x, double atol, double rtol, int maxit,
                        int precision, long long *ticket);
int hflmp_server_step(hflmp_server *sv, int *n_done);
int hflmp_server_poll(hflmp_server *sv, long long ticket,
                      hflmp_server_result *res);
/* forget a done ticket */
int hflmp_server_release(hflmp_server *sv, long long ticket);
int hflmp_server_stats(const hflmp_server *sv, hflmp_server_stats_t *out);

#ifdef __cplusplus
}
#endif
#endif
