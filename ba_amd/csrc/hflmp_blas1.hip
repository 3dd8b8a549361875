// BLAS-1, casts and indexed access on device vectors.  Replaces the cuBLAS
// calls of GPUblas_lVector (linear_algebra/lmp/cuda/GPUblas_lvector.cu:
// 75-433) and the mem_cuda kernels (cuda/mem_cuda.h:21-132).
//
// Reductions are single-launch and deterministic: every workgroup writes
// one partial, the workgroup that finishes last sums the partials in index
// order and writes the result to a device scalar (no atomics on values).
#include "hflmp_internal.hpp"

namespace {

// ---- elementwise ------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) axpy_kernel(int n, T a, const T *__restrict__ x,
                                                      T *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] += a * x[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) scaleadd_kernel(int n, T a, const T *__restrict__ x,
                                                          T *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] = a * y[i] + x[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) scal_kernel(int n, T a, T *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] *= a;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) ewmul_kernel(int n, const T *__restrict__ x,
                                                       T *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] *= x[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) mul_kernel(int n, const T *__restrict__ a,
                                                     const T *__restrict__ r,
                                                     T *__restrict__ z) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    z[i] = r[i] * a[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) copy_kernel(int n, const T *__restrict__ x,
                                                      T *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = x[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) rot_kernel(int n, T *__restrict__ x, T *__restrict__ y,
                                                     T c, T s) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const T xi = x[i], yi = y[i];
    x[i] = c * xi + s * yi;
    y[i] = c * yi - s * xi;
  }
}

template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) cast_kernel(int n, const S *__restrict__ x,
                                                      D *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] = static_cast<D>(x[i]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) gather_kernel(int n, const int *__restrict__ idx,
                                                        const T *__restrict__ src,
                                                        T *__restrict__ dst) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    dst[i] = src[idx[i]];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) scatter_kernel(int n, const int *__restrict__ idx,
                                                         const T *__restrict__ src,
                                                         T *__restrict__ dst) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    dst[idx[i]] = src[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) scatter_add_kernel(int n, const int *__restrict__ idx,
                                                             const T *__restrict__ src,
                                                             T *__restrict__ dst) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    atomicAdd(&dst[idx[i]], src[i]);
}

// ---- reductions ---------------------------------------------------------------
enum RedOp { kDot = 0, kSumSq = 1, kSumAbs = 2, kCountNan = 3 };

template <typename T, int OP>
__device__ __forceinline__ T red_term(const T *__restrict__ x, const T *__restrict__ y, int i) {
  if (OP == kDot) return x[i] * y[i];
  if (OP == kSumSq) return x[i] * x[i];
  if (OP == kSumAbs) return fabs(x[i]);
  return (x[i] != x[i]) ? T(1) : T(0);
}

// partials[b] = sum over this workgroup's grid-stride share of term(i)
template <typename T, int OP>
__global__ void __launch_bounds__(kBlock) reduce_kernel(int n, const T *__restrict__ x,
                                                        const T *__restrict__ y,
                                                        T *__restrict__ partials) {
  __shared__ T lds[kBlock / kWave];
  T acc = 0;
  const int stride = gridDim.x * kBlock;
  int i = blockIdx.x * kBlock + threadIdx.x;
  // four independent accumulators keep four loads in flight per lane
  T a1 = 0, a2 = 0, a3 = 0;
  for (; i + 3 * stride < n; i += 4 * stride) {
    acc += red_term<T, OP>(x, y, i);
    a1 += red_term<T, OP>(x, y, i + stride);
    a2 += red_term<T, OP>(x, y, i + 2 * stride);
    a3 += red_term<T, OP>(x, y, i + 3 * stride);
  }
  for (; i < n; i += stride) acc += red_term<T, OP>(x, y, i);
  acc = (acc + a1) + (a2 + a3);
  T s = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// out = sum of the partials; FINISH = 1 takes the square root.
template <typename T, int FINISH>
__global__ void __launch_bounds__(kFinalBlock)
    reduce_final_kernel(const T *__restrict__ partials, int count, T *out) {
  __shared__ T lds[kFinalBlock / kWave];
  const T tot = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) out[0] = FINISH ? sqrt(tot) : tot;
}

// max |x_i| (MODE 0) or min |x_i| (MODE 1) with the first index attaining it.
template <typename T, int MODE>
__device__ __forceinline__ bool ext_better(T ov, int oi, T mv, int mi) {
  if (oi < 0) return false;
  if (mi < 0) return true;
  if (MODE == 0 ? ov > mv : ov < mv) return true;
  return ov == mv && oi < mi;
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) argext_final_kernel(const T *__restrict__ pval,
                                                              const int *__restrict__ pidx,
                                                              int count, T *out_val,
                                                              int *out_idx) {
  if (threadIdx.x == 0) {
    T mv = 0;
    int mi = -1;
    for (int k = 0; k < count; ++k)
      if (ext_better<T, MODE>(pval[k], pidx[k], mv, mi)) {
        mv = pval[k];
        mi = pidx[k];
      }
    out_val[0] = mv;
    out_idx[0] = mi;
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) argext_kernel(int n, const T *__restrict__ x,
                                                        T *__restrict__ pval,
                                                        int *__restrict__ pidx) {
  __shared__ T sv[kBlock];
  __shared__ int si[kBlock];
  T best = MODE == 0 ? T(-1) : T(0);
  int bi = -1;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const T a = fabs(x[i]);
    const bool take = (bi < 0) || (MODE == 0 ? a > best : a < best);
    if (take) {
      best = a;
      bi = i;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const T ov = sv[threadIdx.x + off];
      const int oi = si[threadIdx.x + off];
      if (ext_better<T, MODE>(ov, oi, sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pval[blockIdx.x] = sv[0];
    pidx[blockIdx.x] = si[0];
  }
}

// ---- host helpers ----------------------------------------------------------------
template <typename T, int OP, int FINISH>
int launch_reduce(hflmp_ctx *ctx, int n, const T *x, const T *y, T *d_out) {
  const int g = red_blocks(ctx, n);
  T *partials = static_cast<T *>(ctx->red_partials);
  reduce_kernel<T, OP><<<g, kBlock, 0, ctx->stream>>>(n, x, y, partials);
  reduce_final_kernel<T, FINISH><<<1, kFinalBlock, 0, ctx->stream>>>(partials, g, d_out);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

template <typename T>
int read_scalar(hflmp_ctx *ctx, const T *d_val, T *h_out) {
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(ctx->host_pinned, d_val, sizeof(T), hipMemcpyDeviceToHost,
                                      ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(h_out, ctx->host_pinned, sizeof(T));
  return HFLMP_OK;
}

template <typename T, int OP, int FINISH>
int reduce_to_host(hflmp_ctx *ctx, int n, const T *x, const T *y, T *result) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, result, "reduction: null result");
  HFLMP_REQUIRE(ctx, n >= 0, "reduction: n < 0");
  if (n == 0) {
    *result = 0;
    return HFLMP_OK;
  }
  HFLMP_REQUIRE(ctx, x && y, "reduction: null vector");
  T *d_out = reinterpret_cast<T *>(ctx->red_result);
  int rc = launch_reduce<T, OP, FINISH>(ctx, n, x, y, d_out);
  if (rc) return rc;
  return read_scalar(ctx, d_out, result);
}

template <typename T, int MODE>
int argext(hflmp_ctx *ctx, int n, const T *x, T *val, int *idx) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, n > 0 && x, "argmax/argmin: empty or null vector");
  const int g = red_blocks(ctx, n);
  T *pval = static_cast<T *>(ctx->red_partials);
  int *pidx = reinterpret_cast<int *>(static_cast<char *>(ctx->red_partials) +
                                      8 * (size_t)kMaxRedBlocks);
  T *d_val = reinterpret_cast<T *>(ctx->red_result);
  int *d_idx = reinterpret_cast<int *>(ctx->red_result + 1);
  argext_kernel<T, MODE><<<g, kBlock, 0, ctx->stream>>>(n, x, pval, pidx);
  argext_final_kernel<T, MODE><<<1, kBlock, 0, ctx->stream>>>(pval, pidx, g, d_val, d_idx);
  HFLMP_LAUNCH_CHECK(ctx);
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(ctx->host_pinned, ctx->red_result, 16,
                                      hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const char *h = static_cast<const char *>(ctx->host_pinned);
  if (val) std::memcpy(val, h, sizeof(T));
  if (idx) std::memcpy(idx, h + 8, sizeof(int));
  return HFLMP_OK;
}

#define ELEMWISE_GUARD(ctx, n)                                                \
  if (!(ctx)) return HFLMP_EINVAL;                                            \
  HFLMP_REQUIRE(ctx, (n) >= 0, "negative size");                             \
  if ((n) == 0) return HFLMP_OK;

}  // namespace

extern "C" {

#define DEFINE_BLAS1(T, S)                                                                   \
  int hflmp_dot_##S(hflmp_ctx *ctx, int n, const T *x, const T *y, T *result) {             \
    return reduce_to_host<T, kDot, 0>(ctx, n, x, y, result);                                \
  }                                                                                          \
  int hflmp_dot_dev_##S(hflmp_ctx *ctx, int n, const T *x, const T *y, T *d_result) {       \
    if (!ctx) return HFLMP_EINVAL;                                                           \
    HFLMP_REQUIRE(ctx, n > 0 && x && y && d_result, "dot_dev: bad arguments");              \
    return launch_reduce<T, kDot, 0>(ctx, n, x, y, d_result);                               \
  }                                                                                          \
  int hflmp_nrm2_##S(hflmp_ctx *ctx, int n, const T *x, T *result) {                        \
    return reduce_to_host<T, kSumSq, 1>(ctx, n, x, x, result);                              \
  }                                                                                          \
  int hflmp_norm1_##S(hflmp_ctx *ctx, int n, const T *x, T *result) {                       \
    return reduce_to_host<T, kSumAbs, 0>(ctx, n, x, x, result);                             \
  }                                                                                          \
  int hflmp_count_nan_##S(hflmp_ctx *ctx, int n, const T *x, int *count) {                  \
    T c = 0;                                                                                 \
    int rc = reduce_to_host<T, kCountNan, 0>(ctx, n, x, x, &c);                             \
    if (rc == HFLMP_OK && count) *count = (int)c;                                            \
    return rc;                                                                               \
  }                                                                                          \
  int hflmp_normmax_##S(hflmp_ctx *ctx, int n, const T *x, T *result) {                     \
    if (ctx && n == 0 && result) {                                                           \
      *result = 0;                                                                           \
      return HFLMP_OK;                                                                       \
    }                                                                                        \
    return argext<T, 0>(ctx, n, x, result, nullptr);                                         \
  }                                                                                          \
  int hflmp_argmax_abs_##S(hflmp_ctx *ctx, int n, const T *x, int *index) {                 \
    return argext<T, 0>(ctx, n, x, nullptr, index);                                          \
  }                                                                                          \
  int hflmp_argmin_abs_##S(hflmp_ctx *ctx, int n, const T *x, int *index) {                 \
    return argext<T, 1>(ctx, n, x, nullptr, index);                                          \
  }                                                                                          \
  int hflmp_axpy_##S(hflmp_ctx *ctx, int n, T a, const T *x, T *y) {                        \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    axpy_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, a, x, y);                 \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_scaleadd_##S(hflmp_ctx *ctx, int n, T a, const T *x, T *y) {                    \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    scaleadd_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, a, x, y);             \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_scal_##S(hflmp_ctx *ctx, int n, T a, T *y) {                                    \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    scal_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, a, y);                    \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_ewmul_##S(hflmp_ctx *ctx, int n, const T *x, T *y) {                            \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    ewmul_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, x, y);                   \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_jacobi_apply_##S(hflmp_ctx *ctx, int n, const T *invdiag, const T *r, T *z) {   \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    mul_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, invdiag, r, z);            \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_copy_##S(hflmp_ctx *ctx, int n, const T *x, T *y) {                             \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    copy_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, x, y);                    \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_rot_##S(hflmp_ctx *ctx, int n, T *x, T *y, T c, T s) {                          \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    rot_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, x, y, c, s);               \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_gather_##S(hflmp_ctx *ctx, int n, const int *idx, const T *src, T *dst) {       \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    gather_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, idx, src, dst);         \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_scatter_##S(hflmp_ctx *ctx, int n, const int *idx, const T *src, T *dst) {      \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    scatter_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, idx, src, dst);        \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }                                                                                          \
  int hflmp_scatter_add_##S(hflmp_ctx *ctx, int n, const int *idx, const T *src, T *dst) {  \
    ELEMWISE_GUARD(ctx, n)                                                                   \
    scatter_add_kernel<T><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, idx, src, dst);    \
    HFLMP_LAUNCH_CHECK(ctx);                                                                 \
    return HFLMP_OK;                                                                         \
  }

DEFINE_BLAS1(double, f64)
DEFINE_BLAS1(float, f32)

int hflmp_cast_f64_to_f32(hflmp_ctx *ctx, int n, const double *x, float *y) {
  ELEMWISE_GUARD(ctx, n)
  cast_kernel<double, float><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, x, y);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

int hflmp_cast_f32_to_f64(hflmp_ctx *ctx, int n, const float *x, double *y) {
  ELEMWISE_GUARD(ctx, n)
  cast_kernel<float, double><<<grid_for(n, 4), kBlock, 0, ctx->stream>>>(n, x, y);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

int hflmp_cast_csr_f64_to_f32(hflmp_ctx *ctx, int nnz, const double *val, float *val32) {
  return hflmp_cast_f64_to_f32(ctx, nnz, val, val32);
}

}  // extern "C"
