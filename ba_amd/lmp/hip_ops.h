// Value-type dispatch onto the _f64 / _f32 entry points of libhflmp_hip.
#pragma once

#include "hflmp_hip.h"

namespace hiflow {
namespace la {

template <typename T>
struct HipOps;

#define HFLMP_OPS(T, S)                                                                      \
  template <>                                                                                \
  struct HipOps<T> {                                                                         \
    static int dot(hflmp_ctx *c, int n, const T *x, const T *y, T *r) {                      \
      return hflmp_dot_##S(c, n, x, y, r);                                                   \
    }                                                                                        \
    static int nrm2(hflmp_ctx *c, int n, const T *x, T *r) { return hflmp_nrm2_##S(c, n, x, r); } \
    static int norm1(hflmp_ctx *c, int n, const T *x, T *r) {                                \
      return hflmp_norm1_##S(c, n, x, r);                                                    \
    }                                                                                        \
    static int normmax(hflmp_ctx *c, int n, const T *x, T *r) {                              \
      return hflmp_normmax_##S(c, n, x, r);                                                  \
    }                                                                                        \
    static int argmax(hflmp_ctx *c, int n, const T *x, int *i) {                             \
      return hflmp_argmax_abs_##S(c, n, x, i);                                               \
    }                                                                                        \
    static int argmin(hflmp_ctx *c, int n, const T *x, int *i) {                             \
      return hflmp_argmin_abs_##S(c, n, x, i);                                               \
    }                                                                                        \
    static int count_nan(hflmp_ctx *c, int n, const T *x, int *k) {                          \
      return hflmp_count_nan_##S(c, n, x, k);                                                \
    }                                                                                        \
    static int axpy(hflmp_ctx *c, int n, T a, const T *x, T *y) {                            \
      return hflmp_axpy_##S(c, n, a, x, y);                                                  \
    }                                                                                        \
    static int scaleadd(hflmp_ctx *c, int n, T a, const T *x, T *y) {                        \
      return hflmp_scaleadd_##S(c, n, a, x, y);                                              \
    }                                                                                        \
    static int scal(hflmp_ctx *c, int n, T a, T *y) { return hflmp_scal_##S(c, n, a, y); }   \
    static int ewmul(hflmp_ctx *c, int n, const T *x, T *y) {                                \
      return hflmp_ewmul_##S(c, n, x, y);                                                    \
    }                                                                                        \
    static int copy(hflmp_ctx *c, int n, const T *x, T *y) {                                 \
      return hflmp_copy_##S(c, n, x, y);                                                     \
    }                                                                                        \
    static int rot(hflmp_ctx *c, int n, T *x, T *y, T cs, T sn) {                            \
      return hflmp_rot_##S(c, n, x, y, cs, sn);                                              \
    }                                                                                        \
    static int gather(hflmp_ctx *c, int n, const int *i, const T *s, T *d) {                 \
      return hflmp_gather_##S(c, n, i, s, d);                                                \
    }                                                                                        \
    static int scatter(hflmp_ctx *c, int n, const int *i, const T *s, T *d) {                \
      return hflmp_scatter_##S(c, n, i, s, d);                                               \
    }                                                                                        \
    static int scatter_add(hflmp_ctx *c, int n, const int *i, const T *s, T *d) {            \
      return hflmp_scatter_add_##S(c, n, i, s, d);                                           \
    }                                                                                        \
    static int jacobi(hflmp_ctx *c, int n, const T *d, const T *r, T *z) {                   \
      return hflmp_jacobi_apply_##S(c, n, d, r, z);                                          \
    }                                                                                        \
    static int spmv(hflmp_ctx *c, int nr, int nc, int nnz, const int *rp, const int *ci,     \
                    const T *v, const T *x, T *y) {                                          \
      return hflmp_csr_spmv_##S(c, nr, nc, nnz, rp, ci, v, x, y);                            \
    }                                                                                        \
    static int spmv_add(hflmp_ctx *c, int nr, int nc, int nnz, const int *rp, const int *ci, \
                        const T *v, const T *x, T *y) {                                      \
      return hflmp_csr_spmv_add_##S(c, nr, nc, nnz, rp, ci, v, x, y);                        \
    }                                                                                        \
    static int spmv_nodiag(hflmp_ctx *c, int nr, int nc, int nnz, const int *rp,             \
                           const int *ci, const T *v, const T *x, T *y) {                    \
      return hflmp_csr_spmv_nodiag_##S(c, nr, nc, nnz, rp, ci, v, x, y);                     \
    }                                                                                        \
    static int invdiag(hflmp_ctx *c, int nr, const int *rp, const int *ci, const T *v, T *d, \
                       int *bad) {                                                           \
      return hflmp_csr_extract_invdiag_##S(c, nr, rp, ci, v, d, bad);                        \
    }                                                                                        \
    static int zero_rows(hflmp_ctx *c, int m, const int *rows, const int *rp, const int *ci, \
                         T *v, T alpha) {                                                    \
      return hflmp_csr_zero_rows_##S(c, m, rows, rp, ci, v, alpha);                          \
    }                                                                                        \
    static int sell_create(hflmp_ctx *c, int nr, int nc, int nnz, const int *rp,             \
                           const int *ci, const T *v, int flags, hflmp_sell **m) {           \
      return hflmp_sell_create_##S(c, nr, nc, nnz, rp, ci, v, flags, m);                     \
    }                                                                                        \
    static int sell_set_values(hflmp_ctx *c, hflmp_sell *m, const int *rp, const T *v) {     \
      return hflmp_sell_set_values_##S(c, m, rp, v);                                         \
    }                                                                                        \
    static int sell_spmv(hflmp_ctx *c, const hflmp_sell *m, const T *x, T *y) {              \
      return hflmp_sell_spmv_##S(c, m, x, y);                                                \
    }                                                                                        \
    static int sell_spmv_add(hflmp_ctx *c, const hflmp_sell *m, const T *x, T *y) {          \
      return hflmp_sell_spmv_add_##S(c, m, x, y);                                            \
    }                                                                                        \
    static int sgs_create(hflmp_ctx *c, int nr, int nnz, const int *rp, const int *ci,       \
                          const T *v, const int *hrp, const int *hci, int nb,                \
                          hflmp_sgs **m) {                                                   \
      return hflmp_sgs_create_##S(c, nr, nnz, rp, ci, v, hrp, hci, nb, m);                   \
    }                                                                                        \
    static int sgs_apply(hflmp_ctx *c, const hflmp_sgs *m, const T *r, T *z) {               \
      return hflmp_sgs_apply_##S(c, m, r, z);                                                \
    }                                                                                        \
  };

HFLMP_OPS(double, f64)
HFLMP_OPS(float, f32)
#undef HFLMP_OPS

}  // namespace la
}  // namespace hiflow
