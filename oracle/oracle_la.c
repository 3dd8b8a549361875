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

  or_spmv_d(n, rp, ci, vd, x, resh);
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
    or_spmv_d(n, rp, ci, vd, x, resh);
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
