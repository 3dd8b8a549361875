/*
 * oracle_la.c -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
 * Instantiates the type-generic kernels and solvers for double and float,
 * and the FP64/FP32 mixed-precision refinement.
 */
#include "hflmp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* IterateControl states (common/iterate_control.h:28-34) */
enum {
  OR_ITERATE = 0,
  OR_SUCCESS_ABS = 1,
  OR_SUCCESS_REL = 2,
  OR_FAIL_DIV = 3,
  OR_FAIL_MAXIT = 4
};

/* Threads of the row loop of the solvers' SpMV (or_cg_*, or_mpir_ds).  1:
 * CPUsimple_CSR_lMatrix::VectorMult; more: CPUopenmp_CSR_lMatrix::VectorMult
 * (openmp/CPUopenmp_csr_lmatrix.cc:145-165), the same per-row sums, so the
 * same bits.  The vector operations stay on one thread, as the CPUopenmp
 * platform's (openmp/CPUopenmp_lvector.cc:314-370 and :397-540 are omp simd
 * loops, not parallel ones). */
static int or_solver_threads = 1;
/* Threads of the CG's vector operations (or_cg_*; 1 = sequential). */
#define OR_MAX_BLAS_THREADS 256
static int or_blas_threads = 1;

void or_set_solver_threads(int n) { or_solver_threads = n > 1 ? n : 1; }
int or_get_solver_threads(void) { return or_solver_threads; }
void or_set_blas_threads(int n) {
  or_blas_threads = n < 1 ? 1 : n > OR_MAX_BLAS_THREADS ? OR_MAX_BLAS_THREADS : n;
}
int or_get_blas_threads(void) { return or_blas_threads; }

static int cmp_int_or(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

#define T double
#define S d
#include "oracle_real.inc"
#undef T
#undef S

#define T float
#define S s
#include "oracle_real.inc"
#undef T
#undef S

/* MPIR<LADescriptorCoupledD, LADescriptorCoupledS>::SolveImpl
 * (linear_solver/mpir.cc:52-296) with the low-precision error correction
 * only: resh = b - A x (FP64); while the outer test holds: resl = (float)
 * resh, corl = 0, FP32 Jacobi-PCG on A_s corl = resl, x += (double) corl,
 * resh = b - A x, res = ||resh||.  The relative outer test uses the initial
 * residual throughout (the reference overwrites res0_ after every step,
 * mpir.cc:240, which makes the tolerance relative to the previous step). */
int or_mpir_ds(int nrow, const int *rp, const int *ci, const double *vd,
               const float *vs, const double *b, double *x, int max_outer,
               double atol, double rtol, int inner_maxit, double inner_rtol,
               int *outer_iters, int *inner_iters_total, double *res_init,
               double *res_final) {
  const int n = nrow;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double *resh = (double *)malloc(sizeof(double) * nn);
  double *corh = (double *)malloc(sizeof(double) * nn);
  float *resl = (float *)malloc(sizeof(float) * nn);
  float *corl = (float *)malloc(sizeof(float) * nn);
  float *invd = (float *)malloc(sizeof(float) * nn);
  or_invdiag_s(n, rp, ci, vs, invd);

  or_spmv_solver_d(n, rp, ci, vd, x, resh);
  or_scaleadd_d(n, -1.0, b, resh);
  double res = or_nrm2_d(n, resh);
  const double res0 = res;
  int outer = 0, inner_total = 0;
  while (res > atol && res > rtol * res0 && outer < max_outer) {
    ++outer;
    or_cast_d2s(n, resh, resl);
    memset(corl, 0, sizeof(float) * nn);
    int it = 0;
    double ri, rf;
    or_cg_s(n, rp, ci, vs, OR_PREC_JACOBI, invd, 0, NULL, NULL, NULL, resl,
            corl, inner_maxit, 0.0, inner_rtol, 1e6, &it, &ri, &rf, NULL);
    inner_total += it;
    or_cast_s2d(n, corl, corh);
    or_axpy_d(n, 1.0, corh, x);
    or_spmv_solver_d(n, rp, ci, vd, x, resh);
    or_scaleadd_d(n, -1.0, b, resh);
    res = or_nrm2_d(n, resh);
  }
  *outer_iters = outer;
  *inner_iters_total = inner_total;
  *res_init = res0;
  *res_final = res;
  free(resh);
  free(corh);
  free(resl);
  free(corl);
  free(invd);
  return (res <= atol || res < rtol * res0) ? 0 : 1;
}
