/*
 * hflmp_oracle.h -- CPU restatement of the HiFlow3 "lmp" CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only the test suite, the smoke check and the CPU
 * baseline leg of the benchmark may load it, and only as the checker.  The
 * product libraries never link or call it.
 *
 * Every routine follows a reference file (paths relative to the HiFlow3
 * source tree, src/):
 *   SpMV          linear_algebra/lmp/CPUsimple_csr_lmatrix.cc:76-111
 *   SpMV-add      linear_algebra/lmp/lmatrix_csr_cpu.cc:1424-1450
 *   SpMV no-diag  linear_algebra/lmp/cuda/cuda_spmv_kernel_csr_scalar.h:48
 *   OpenMP SpMV   linear_algebra/lmp/openmp/CPUopenmp_csr_lmatrix.cc:145-165
 *   COO->CSR      linear_algebra/lmp/lmatrix_formats.h:320-456
 *   diag ptr      linear_algebra/lmp/lmatrix_formats.h:489-532
 *   ZeroRows      linear_algebra/lmp/lmatrix_csr_cpu.cc:908-927
 *   inv diag      linear_algebra/lmp/lmatrix_csr_cpu.cc:2420-2442
 *   SGS           linear_algebra/lmp/lmatrix_csr_cpu.cc:496-565
 *   block SGS     linear_algebra/lmp/lmatrix_csr_cpu.cc:568-641
 *   colouring     linear_algebra/lmp/lmatrix_csr_cpu.cc:2195-2300
 *   MC-SGS        linear_algebra/lmp/lpreconditioner_mc.cc:449-723,
 *                 linear_algebra/lmp/lpreconditioner_mc_sgs.cc:44-90,
 *                 linear_algebra/lmp/lpreconditioner_mc_gs.cc:83-102
 *   BLAS-1        linear_algebra/lmp/CPUsimple_lvector.cc:144-153,224-227,
 *                 304-306,390-391,463-465,638-700; lvector_cpu.cc:485-493
 *   Jacobi        linear_algebra/lmp/lpreconditioner.cc:125-200
 *   CG            linear_solver/cg.h:87-246
 *   convergence   common/iterate_control.cc:80-107
 *   GMRES         linear_solver/gmres.h:97-116,464-718
 *   FGMRES        linear_solver/fgmres.h:119-320
 *   MPIR          linear_solver/mpir.cc:52-296 (res0 kept at the initial
 *                 residual; the reference resets it after every correction,
 *                 mpir.cc:240)
 *   stencil       linear_algebra/stencil_operator.h:153-330 (matrix-free
 *                 VectorMultAdd, cell by cell)
 *   BiCGSTAB      linear_solver/bicgstab.h:127-483
 *   assembly      assembly/assembly_utils.cc:144 (pattern),
 *                 lmatrix_csr_cpu.cc:1102-1153 (add_values, cell by cell)
 *   GMG residual  linear_algebra/lmp/cuda/cuda_gmg_kernel.h:46-64
 *   damped Jacobi linear_algebra/lmp/cuda/cuda_async_iter_kernels.h:116-144
 *                 (one synchronous sweep); the V-cycle itself is in
 *                 oracle.py (linear_solver/gmg.h:979-1187)
 *   async Jacobi  linear_solver/gpu-based/async_iter_gpu.cc:493-874,
 *                 cuda_async_iter_kernels.h:17-193 (blocks of rows, each
 *                 launch: outside couplings from the launch's input, then
 *                 synchronous inner sweeps; deterministic)
 *
 * Problem generator: structured Q1 (2D/3D) and Q2 (2D) Poisson/diffusion on
 * the unit square/cube, assembled element by element in lexicographic cell
 * order into a sorted CSR pattern, then Dirichlet rows via ZeroRows.  The
 * random streams are SplitMix64 (Steele, Lea, Flood, OOPSLA 2014).
 */
#ifndef HFLMP_ORACLE_H
#define HFLMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random streams ---------------------------------------------------- */
uint64_t or_splitmix64_at(uint64_t seed, uint64_t i);
double or_uniform_pm1(uint64_t seed, uint64_t i);
double or_gauss(uint64_t seed, uint64_t c);

/* ---- structured FE systems --------------------------------------------- */
/* order: 1 = Q1, 2 = Q2;  dim: 2 or 3;  n: cells per side. */
int or_fe_sizes(int order, int dim, int n, int64_t *nrow, int64_t *nnz);
/* kappa: per cell (n^dim entries, lexicographic) or NULL for kappa = 1. */
int or_fe_assemble(int order, int dim, int n, const double *kappa,
                   int *rowptr, int *col, double *val);
void or_fe_kappa(int dim, int n, double sigma, uint64_t seed, double *kappa);
void or_fe_rhs(int order, int dim, int n, uint64_t seed, double *b);
void or_fe_boundary_mask(int order, int dim, int n, uint8_t *mask);
/* element stiffness of the generator (nloc x nloc, local nodes
 * lexicographic, h included; the cell factor is kappa) */
void or_fe_element(int order, int dim, int n, double *E);

/* ---- CSR utilities ----------------------------------------------------- */
void or_coo_to_csr_d(int nrow, int nnz, const int *ri, const int *ci,
                     const double *v, int *rowptr, int *col, double *val);
int or_csr_diag_ptr(int nrow, const int *rowptr, const int *col, int *diag);
void or_csr_zero_rows_d(const int *rows, int nrows_zero, const int *rowptr,
                        const int *col, double *val, double alpha);
int or_multicoloring(int nrow, const int *rowptr, const int *col,
                     int *ncolors, int *color_sizes, int *perm, int *color);

/* ---- vector / matrix kernels (d = double, s = float) ------------------- */
#define OR_DECL(T, S)                                                         \
  void or_spmv_##S(int nrow, const int *rp, const int *ci, const T *v,        \
                   const T *x, T *y);                                         \
  void or_spmv_add_##S(int nrow, const int *rp, const int *ci, const T *v,    \
                       const T *x, T *y);                                     \
  void or_spmv_nodiag_##S(int nrow, const int *rp, const int *ci,             \
                          const T *v, const T *x, T *y);                      \
  void or_spmv_omp_##S(int nrow, const int *rp, const int *ci, const T *v,    \
                       const T *x, T *y, int nthreads);                       \
  void or_residual_sub_##S(int nrow, const int *rp, const int *ci,            \
                           const T *v, const T *b, const T *x, T *r);         \
  void or_jacobi_damped_##S(int nrow, const int *rp, const int *ci,           \
                            const T *v, const T *invd, T w, const T *b,       \
                            const T *xi, T *xo);                              \
  T or_dot_##S(int n, const T *x, const T *y);                                \
  T or_nrm2_##S(int n, const T *x);                                           \
  T or_norm1_##S(int n, const T *x);                                          \
  T or_normmax_##S(int n, const T *x);                                        \
  void or_axpy_##S(int n, T a, const T *x, T *y);                             \
  void or_scaleadd_##S(int n, T a, const T *x, T *y);                         \
  void or_scale_##S(int n, T a, T *y);                                        \
  void or_ewmult_##S(int n, const T *x, T *y);                                \
  void or_invdiag_##S(int nrow, const int *rp, const int *ci, const T *v,     \
                      T *invd);                                               \
  void or_jacobi_apply_##S(int n, const T *invd, const T *r, T *z);           \
  void or_sgs_apply_##S(int nrow, const int *rp, const int *ci, const T *v,   \
                        const T *r, T *z);                                    \
  void or_block_sgs_apply_##S(int nrow, const int *rp, const int *ci,         \
                              const T *v, int num_blocks, const T *r, T *z);  \
  void or_mcsgs_apply_##S(int nrow, const int *rp, const int *ci,             \
                          const T *v, int ncolors, const int *color_sizes,    \
                          const int *perm, const int *color, T omega,         \
                          const T *r, T *z);                                  \
  int or_cg_##S(int nrow, const int *rp, const int *ci, const T *v,           \
                int prec, const T *invd, int ncolors, const int *csz,         \
                const int *perm, const int *color, const T *b, T *x,          \
                int maxit, double atol, double rtol, double dtol,             \
                int *iters, double *res_init, double *res_final,              \
                double *res_hist);                                            \
  int or_gmres_##S(int nrow, const int *rp, const int *ci, const T *v,        \
                   int prec, const T *invd, int ncolors, const int *csz,      \
                   const int *perm, const int *color, int basis,              \
                   int flexible, const T *b, T *x, int maxit, double atol,    \
                   double rtol, double dtol, int *iters, double *res_init,    \
                   double *res_final);                                        \
  void or_stencil_vmadd_##S(int order, int dim, const int *nc, const T *E,    \
                            const uint8_t *cell_lm, const T *cell_factor,     \
                            int nfixed, const int *fixed, T alpha,            \
                            const T *in, T beta, T *out);                  \
  int or_fe_assemble_cells_##S(int nnode, int ncell, int nloc,                \
                               const int *dofs, const T *cell_mats,           \
                               const T *E, const T *factor, int *rp,          \
                               int *col, T *val);                             \
  int or_async_jacobi_##S(int nrow, const int *rp, const int *ci,             \
                          const T *v, const T *b, T *x, int block_rows,       \
                          int n_global, int n_block, int n_inner, int mode,   \
                          T w, int maxit, double atol, double rtol,           \
                          double dtol, int *iters, double *res_init,          \
                          double *res_final);                                 \
  int or_bicgstab_##S(int nrow, const int *rp, const int *ci, const T *v,     \
                      int prec, const T *invd, int ncolors, const int *csz,   \
                      const int *perm, const int *color, const T *b, T *x,    \
                      int maxit, double atol, double rtol, double dtol,       \
                      int *iters, double *res_init, double *res_final);

OR_DECL(double, d)
OR_DECL(float, s)

/* preconditioner selector for or_cg_* / or_gmres_* */
#define OR_PREC_NONE 0
#define OR_PREC_JACOBI 1
#define OR_PREC_SGS 2
#define OR_PREC_MCSGS 3
/* block SGS (BlocksPsgauss_seidel); the ncolors argument carries num_blocks */
#define OR_PREC_BSGS 4

/* casts (CPUsimple_lvector.cc:638-700) */
void or_cast_d2s(int n, const double *x, float *y);
void or_cast_s2d(int n, const float *x, double *y);

/* Mixed-precision refinement: FP64 residual/update/norm, FP32 Jacobi-PCG
 * correction solve.  Returns 0 on success. */
/* threads of the solvers' SpMV row loop (1 = sequential; same results) */
void or_set_solver_threads(int n);
int or_get_solver_threads(void);
/* threads of or_cg_*'s vector operations (1 = sequential; dot then sums
 * per-thread blocks: deterministic, not the sequential bits) */
void or_set_blas_threads(int n);
int or_get_blas_threads(void);

int or_mpir_ds(int nrow, const int *rp, const int *ci, const double *vd,
               const float *vs, const double *b, double *x, int max_outer,
               double atol, double rtol, int inner_maxit, double inner_rtol,
               int *outer_iters, int *inner_iters_total, double *res_init,
               double *res_final);

#ifdef __cplusplus
}
#endif
#endif
