/*
 * hflmp_lmp.h -- C entry points into the C++ lmp layer (libhflmp_lmp.so).
 *
 * These drive the unchanged solver templates (CG, GMRES, FGMRES, MPIR over
 * CoupledMatrix / CoupledVector) on the HIP platform for a system given as
 * host CSR arrays.  They exist for embedding and testing: the reference's
 * own solver code talks to the lmp classes directly.
 * Errors inside the C++ layer are returned as negative codes with a
 * message (hflmp_lmp_last_error) instead of ending the process.
 */
#ifndef HFLMP_LMP_H
#define HFLMP_LMP_H

#include "hflmp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HFLMP_LMP_CG 0
#define HFLMP_LMP_GMRES 1
#define HFLMP_LMP_FGMRES 2
#define HFLMP_LMP_MPIR 3 /* FP64 refinement around an FP32 inner solver */
#define HFLMP_LMP_BICGSTAB 4 /* linear_solver/bicgstab.h:127-483 */

#define HFLMP_LMP_PREC_NONE 0
#define HFLMP_LMP_PREC_JACOBI 1
#define HFLMP_LMP_PREC_MCSGS 2
#define HFLMP_LMP_PREC_GMG 3 /* hflmp_lmp_solve_gmg only */
/* Init_SymmetricGaussSeidel: Psgauss_seidel, the sequential row order */
#define HFLMP_LMP_PREC_SGS 4
/* lPreconditioner_BlocksSymmetricGaussSeidel: BlocksPsgauss_seidel on
 * sgs_blocks row ranges */
#define HFLMP_LMP_PREC_BSGS 5

typedef struct {
  int solver;         /* HFLMP_LMP_* (CG, GMRES, FGMRES, MPIR, BICGSTAB)   */
  int precond;        /* HFLMP_LMP_PREC_*                                  */
  int precision;      /* 64 or 32: value type of CG/GMRES/FGMRES           */
  int maxit;
  double atol, rtol, dtol;
  int basis;          /* GMRES/FGMRES basis size                           */
  double omega;       /* MC-SGS relaxation                                 */
  int mpir_max_outer; /* MPIR: outer corrections; inner solver = `solver_inner` */
  double mpir_atol, mpir_rtol;
  int solver_inner;   /* MPIR inner solver: CG/GMRES/FGMRES/BICGSTAB       */
  int sgs_blocks;     /* HFLMP_LMP_PREC_BSGS: row ranges (>= 1)            */
} hflmp_lmp_options;

typedef struct {
  int iters;       /* iterations (MPIR: accumulated inner iterations)       */
  int status;      /* 0 success, 1 not converged, 2 solver error            */
  int outer;       /* MPIR corrections                                      */
  double res_init; /* initial residual norm                                 */
  double res_final;
  double seconds;  /* Solve() wall time                                     */
} hflmp_lmp_result;

void hflmp_lmp_default_options(hflmp_lmp_options *opt);
/* Solve A x = b; x holds x0 on entry.  Host arrays, FP64. */
int hflmp_lmp_solve(int device, int nrow, int nnz, const int *row_ptr, const int *col,
                    const double *val, const double *b, double *x,
                    const hflmp_lmp_options *opt, hflmp_lmp_result *res);
/* Solve A x = b with opt->solver (CG, GMRES, FGMRES or BiCGSTAB; opt->precond
 * is ignored) preconditioned by one geometric-multigrid V-cycle per
 * application: PreconditionerGMG, the reference's GMG used as a
 * preconditioner (linear_solver/gmg.h:1070-1187), over hflmp_mg_*.  A, b, x
 * as in hflmp_lmp_solve (the finest level); coarse[l], R[l], P[l]
 * (l < nlevels - 1) are HOST CSR arrays of level l + 1's operator and of the
 * restriction / prolongation between levels l and l + 1; coarse_inverse
 * (optional, host, n x n row-major for the coarsest level's n rows; needs
 * nlevels >= 2) replaces the coarsest level's sweeps.  The host levels are
 * checked (row_ptr from 0 to nnz and non-decreasing, columns in [0, ncol))
 * before anything reaches the device. */
int hflmp_lmp_solve_gmg(int device, int nrow, int nnz, const int *row_ptr, const int *col,
                        const double *val, const double *b, double *x, int nlevels,
                        const hflmp_csr_f64 *coarse, const hflmp_csr_f64 *R,
                        const hflmp_csr_f64 *P, const hflmp_mg_params *prm,
                        const double *coarse_inverse, const hflmp_lmp_options *opt,
                        hflmp_lmp_result *res);
/* y = A x through CoupledMatrix::VectorMult on the HIP platform. */
int hflmp_lmp_vector_mult(int device, int nrow, int nnz, const int *row_ptr, const int *col,
                          const double *val, const double *x, double *y);
/* MatrixMarket coordinate files (lMatrix::ReadFile / WriteFile,
 * lmatrix_csr_cpu.cc:1781-1810 through lmatrix_coo_cpu.cc:123-279), host
 * only.  read_coo: the entries as the reader produces them (0-based, file
 * order, a symmetric file's mirrors after their entry); with null arrays it
 * only reports the sizes.  write_csr: "coordinate real general", entries in
 * CSR order, 17 significant digits. */
int hflmp_lmp_mm_read_coo(const char *path, int *nrow, int *ncol, int *nentries, int *rows,
                          int *cols, double *vals, int capacity);
int hflmp_lmp_mm_write_csr(const char *path, int nrow, int ncol, const int *row_ptr,
                           const int *col, const double *val);
/* y = A x with A read by HIP_CSR_lMatrix::ReadFile and x by
 * lVector::ReadFile (one value per line), y written by lVector::WriteFile;
 * matrix_out (optional) receives HIP_CSR_lMatrix::WriteFile. */
int hflmp_lmp_file_spmv(int device, const char *matrix_file, const char *x_file,
                        const char *y_file, const char *matrix_out, int *nrow, int *nnz);
const char *hflmp_lmp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
