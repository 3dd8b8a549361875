// CSR SpMV family and CSR utilities.  Replaces GPUscalar_CSR_lMatrix
// (linear_algebra/lmp/cuda/GPUscalar_csr_lmatrix.cu:67-210, kernels in
// cuda/cuda_spmv_kernel_csr_scalar.h:21-60) and the cuSPARSE csrmv path
// (cuda/GPUcublas2_CSR_lMatrix.cc:85-151).
#include "hflmp_spmv_tile.hpp"

namespace {

enum SpmvMode { kAssign = 0, kAdd = 1, kNoDiag = 2 };

template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_kernel(int nrow, int rows_per_tile, const int *__restrict__ rp,
                    const int *__restrict__ ci, const T *__restrict__ val,
                    const T *__restrict__ x, T *__restrict__ y) {
  __shared__ T prod[TileCap<T>::value];
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  // y += A x keeps the reference's accumulation into y[i]
  // (CPU_CSR_lMatrix::VectorMultAdd, lmatrix_csr_cpu.cc:1424-1450).
  T s = 0;
  if (MODE == kAdd) {
    const long long r = tile * (long long)rows_per_tile + threadIdx.x;
    if (threadIdx.x < rows_per_tile && r < nrow) s = y[r];
  }
  TileRow tr;
  s = tile_row_sum<T, MODE == kNoDiag>(tile, rows_per_tile, nrow, rp, ci, val, x, s, tr, prod);
  if (tr.has_row) __builtin_nontemporal_store(s, y + tr.row);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    invdiag_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                   const T *__restrict__ val, T *__restrict__ invd, unsigned int *bad) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    T d = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] == i) {
        d = val[j];
        break;
      }
    // 1./a_ii formed in double, as extract_invdiagelements does.
    if (d == T(0)) {
      invd[i] = T(0);
      atomicAdd(bad, 1u);
    } else {
      invd[i] = static_cast<T>(1.0 / static_cast<double>(d));
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    diag_ptr_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                    int *__restrict__ dp) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    int p = -1;
    for (int j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] == i) {
        p = j;
        break;
      }
    dp[i] = p;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    zero_rows_kernel(int m, const int *__restrict__ rows, const int *__restrict__ rp,
                     const int *__restrict__ ci, T *__restrict__ val, T alpha) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < m; k += gridDim.x * kBlock) {
    const int r = rows[k];
    for (int j = rp[r]; j < rp[r + 1]; ++j) val[j] = (ci[j] == r) ? alpha : T(0);
  }
}

template <typename T, int MODE>
int spmv(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *rp, const int *ci,
         const T *val, const T *x, T *y) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow >= 0 && ncol >= 0 && nnz >= 0, "spmv: negative size");
  if (nrow == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, rp && y, "spmv: null row_ptr or y");
  HFLMP_REQUIRE(ctx, nnz == 0 || (ci && val && x), "spmv: null col/val/x");
  const int R = tile_rows(nrow, nnz, TileCap<T>::value);
  const long long ntiles = ((long long)nrow + R - 1) / R;
  const int grid = tile_grid(ntiles);
  csr_spmv_kernel<T, MODE><<<grid, kBlock, 0, ctx->stream>>>(nrow, R, rp, ci, val, x, y);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

template <typename T>
int invdiag(hflmp_ctx *ctx, int nrow, const int *rp, const int *ci, const T *val, T *invd,
            int *n_bad) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow >= 0, "invdiag: negative size");
  if (nrow == 0) {
    if (n_bad) *n_bad = 0;
    return HFLMP_OK;
  }
  HFLMP_REQUIRE(ctx, rp && ci && val && invd, "invdiag: null pointer");
  unsigned int *bad = ctx->red_counter + 32;
  int rc = hflmp_fill_zero(ctx, bad, sizeof(unsigned int));
  if (rc) return rc;
  invdiag_kernel<T><<<grid_for(nrow, 1), kBlock, 0, ctx->stream>>>(nrow, rp, ci, val, invd, bad);
  HFLMP_LAUNCH_CHECK(ctx);
  unsigned int h = 0;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(ctx->host_pinned, bad, sizeof(unsigned int),
                                      hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(&h, ctx->host_pinned, sizeof(h));
  if (n_bad) *n_bad = (int)h;
  if (h) return set_error(ctx, HFLMP_ENUMERIC, "extract_invdiag: %u rows with zero or missing diagonal", h);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

#define DEFINE_SPMV(T, S)                                                                     \
  int hflmp_csr_spmv_##S(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *rp,         \
                         const int *ci, const T *val, const T *x, T *y) {                    \
    return spmv<T, kAssign>(ctx, nrow, ncol, nnz, rp, ci, val, x, y);                        \
  }                                                                                           \
  int hflmp_csr_spmv_add_##S(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *rp,     \
                             const int *ci, const T *val, const T *x, T *y) {                \
    return spmv<T, kAdd>(ctx, nrow, ncol, nnz, rp, ci, val, x, y);                           \
  }                                                                                           \
  int hflmp_csr_spmv_nodiag_##S(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *rp,  \
                                const int *ci, const T *val, const T *x, T *y) {             \
    return spmv<T, kNoDiag>(ctx, nrow, ncol, nnz, rp, ci, val, x, y);                        \
  }                                                                                           \
  int hflmp_csr_extract_invdiag_##S(hflmp_ctx *ctx, int nrow, const int *rp, const int *ci,  \
                                    const T *val, T *invd, int *n_bad) {                     \
    return invdiag<T>(ctx, nrow, rp, ci, val, invd, n_bad);                                  \
  }

DEFINE_SPMV(double, f64)
DEFINE_SPMV(float, f32)

int hflmp_csr_diag_ptr(hflmp_ctx *ctx, int nrow, const int *rp, const int *ci, int *dp) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow >= 0, "diag_ptr: negative size");
  if (nrow == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, rp && ci && dp, "diag_ptr: null pointer");
  diag_ptr_kernel<<<grid_for(nrow, 1), kBlock, 0, ctx->stream>>>(nrow, rp, ci, dp);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

#define DEFINE_ZERO_ROWS(T, S)                                                              \
  int hflmp_csr_zero_rows_##S(hflmp_ctx *ctx, int m, const int *rows, const int *rp,        \
                              const int *ci, T *val, T alpha) {                             \
    if (!ctx) return HFLMP_EINVAL;                                                          \
    HFLMP_REQUIRE(ctx, m >= 0, "zero_rows: negative size");                                \
    if (m == 0) return HFLMP_OK;                                                            \
    HFLMP_REQUIRE(ctx, rows && rp && ci && val, "zero_rows: null pointer");                \
    zero_rows_kernel<T><<<grid_for(m, 1), kBlock, 0, ctx->stream>>>(m, rows, rp, ci, val,   \
                                                                    alpha);                 \
    HFLMP_LAUNCH_CHECK(ctx);                                                                \
    return HFLMP_OK;                                                                        \
  }

DEFINE_ZERO_ROWS(double, f64)
DEFINE_ZERO_ROWS(float, f32)

}  // extern "C"
