// Block-asynchronous Jacobi iteration: AsynchronousIterationGPU
// (linear_solver/gpu-based/async_iter_gpu.h:32-157, async_iter_gpu.cc:
// 493-820) with its kernels (cuda/cuda_async_iter_kernels.h:17-283), made
// deterministic on one device.
//
// The reference splits the rows over devices and processes.  Within a
// device each thread runs n_inner Jacobi updates of its row in place,
// reading whatever its neighbours hold at that moment (so results depend
// on timing); couplings to other processes' rows enter through
// c = b - A_offdiag x_ghost, computed once per launch from the ghost values
// of the last exchange.  Here the rows are cut into blocks of block_rows
// consecutive rows, and a block plays the part of the reference's
// device-local row set:
//   per launch (one "block iteration"), for every block B and row i in B:
//     c_i = b_i - sum a_ij x_j over the entries with j outside B
//           (x: the iterate at the start of the launch; storage order)
//     n_inner sweeps, each over all rows of B at once (synchronous):
//       tmp = c_i - sum a_ij y_j over the entries with j in B, j != i
//       undamped: y_i = D^-1_i tmp    (kernel_async_inner_iter)
//       damped:   y_i = (w D^-1_i tmp) + ((1 - w) y_i)
//                                      (kernel_async_inner_damped_iter)
//     with y = x on B at the start; the launch's result is y.
// A block of at most kLdsRows rows is one workgroup holding y in LDS; larger
// blocks (down to one block of all rows) run each sweep as a launch over
// global ping-pong buffers with the same arithmetic.  Results do not depend
// on scheduling, and blocks of one device step together between launches,
// which is what the reference's BlockUpdate / GlobalUpdate exchanges do
// between devices and processes (nothing to exchange on one device).
//
// Solve modes (SetSolveMode, async_iter_gpu.cc:661-676): 0 / 1 iterate
// n_global * n_block launches, then the residual and IterateControl::Check
// on the inner-iteration count (SolveNormal / SolveDamped, :722-820), until
// it stops; 2 / 3 run the launches once with no residual (SmoothNormal /
// SmoothDamped, :822-874).  The residual is the reference's device form
// (kernel_compute_block_squared_residual_no_offdiag, :173-193): r_i = b_i -
// sum over the off-diagonal entries of a_ij x_j, then r_i -= a_ii x_i,
// squared, and the squares summed in a fixed order.
#include <chrono>
#include <cmath>

#include "hflmp_internal.hpp"

namespace {

constexpr int kLdsRows = 2048;                 // rows of a block held in LDS
constexpr int kRowsPerThread = kLdsRows / kBlock;

template <typename T>
struct AsyncMat {
  int nrow;
  const int *rp, *ci;
  const T *val;
  const T *b, *invd;
};

__device__ __forceinline__ void block_range(int i, int nrow, int brows, int &lo, int &hi) {
  lo = (i / brows) * brows;
  hi = lo + brows < nrow ? lo + brows : nrow;
}

// c_i: b_i minus the couplings to rows outside the block, in storage order
template <typename T>
__device__ __forceinline__ T outside_part(const AsyncMat<T> &m, int i, int lo, int hi,
                                          const T *__restrict__ x) {
#pragma clang fp contract(off)
  T c = m.b[i];
  for (int j = m.rp[i]; j < m.rp[i + 1]; ++j) {
    const int k = m.ci[j];
    if (k != i && (k < lo || k >= hi)) c -= m.val[j] * x[k];
  }
  return c;
}

// one Jacobi update of row i from c_i and the block's current values y
// (y indexed from the block's first row lo)
template <typename T, bool DAMPED>
__device__ __forceinline__ T row_update(const AsyncMat<T> &m, int i, int lo, int hi, T c,
                                        const T *__restrict__ y, T w) {
#pragma clang fp contract(off)
  T tmp = c;
  for (int j = m.rp[i]; j < m.rp[i + 1]; ++j) {
    const int k = m.ci[j];
    if (k != i && k >= lo && k < hi) tmp -= m.val[j] * y[k - lo];
  }
  if (DAMPED) return (w * m.invd[i] * tmp) + ((T(1) - w) * y[i - lo]);
  return m.invd[i] * tmp;
}

// One launch for blocks of at most kLdsRows rows: workgroup = block.
template <typename T, bool DAMPED>
__global__ void __launch_bounds__(kBlock)
    async_lds_kernel(AsyncMat<T> m, int brows, int n_inner, T w, const T *__restrict__ xin,
                     T *__restrict__ xout) {
  extern __shared__ unsigned char lds_raw[];
  T *cur = reinterpret_cast<T *>(lds_raw);
  T *nxt = cur + brows;
  const int blk = (int)xcd_tile(blockIdx.x, gridDim.x);
  const int lo = blk * brows;
  if (lo >= m.nrow) return;  // whole workgroup: no barrier is skipped by part of it
  const int hi = lo + brows < m.nrow ? lo + brows : m.nrow;
  T c[kRowsPerThread];
#pragma unroll
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int i = lo + threadIdx.x + k * kBlock;
    if (i < hi) {
      cur[i - lo] = xin[i];
      c[k] = outside_part(m, i, lo, hi, xin);
    }
  }
  __syncthreads();
  for (int s = 0; s < n_inner; ++s) {
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
      const int i = lo + threadIdx.x + k * kBlock;
      if (i < hi) nxt[i - lo] = row_update<T, DAMPED>(m, i, lo, hi, c[k], cur, w);
    }
    __syncthreads();
    T *t = cur;
    cur = nxt;
    nxt = t;
  }
#pragma unroll
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int i = lo + threadIdx.x + k * kBlock;
    if (i < hi) xout[i] = cur[i - lo];
  }
}

// Larger blocks: c for every row, then one launch per sweep.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    async_outside_kernel(AsyncMat<T> m, int brows, const T *__restrict__ x, T *__restrict__ c) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m.nrow; i += gridDim.x * kBlock) {
    int lo, hi;
    block_range(i, m.nrow, brows, lo, hi);
    c[i] = outside_part(m, i, lo, hi, x);
  }
}

template <typename T, bool DAMPED>
__global__ void __launch_bounds__(kBlock)
    async_sweep_kernel(AsyncMat<T> m, int brows, T w, const T *__restrict__ c,
                       const T *__restrict__ y, T *__restrict__ ynew) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m.nrow; i += gridDim.x * kBlock) {
    int lo, hi;
    block_range(i, m.nrow, brows, lo, hi);
    ynew[i] = row_update<T, DAMPED>(m, i, lo, hi, c[i], y + lo, w);
  }
}

// D^-1 and D from the first diagonal entry of each row; *bad counts rows
// without a nonzero one (the reference asserts a normal 1 / a_ii)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    async_diag_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                      const T *__restrict__ val, T *__restrict__ invd, T *__restrict__ d,
                      unsigned int *bad) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    T dg = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] == i) {
        dg = val[j];
        break;
      }
    d[i] = dg;
    invd[i] = dg == T(0) ? T(0) : T(1) / dg;
    if (dg == T(0)) atomicAdd(bad, 1u);
  }
}

// per-workgroup sums of the squared residuals
template <typename T>
__global__ void __launch_bounds__(kBlock)
    async_residual_kernel(AsyncMat<T> m, const T *__restrict__ d, const T *__restrict__ x,
                          T *__restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ T lds[kBlock / kWave];
  T acc = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m.nrow; i += gridDim.x * kBlock) {
    T r = m.b[i];
    for (int j = m.rp[i]; j < m.rp[i + 1]; ++j) {
      const int k = m.ci[j];
      if (k != i) r -= m.val[j] * x[k];
    }
    r -= d[i] * x[i];
    r *= r;
    acc += r;
  }
  const T s = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ void __launch_bounds__(kFinalBlock)
    async_residual_final(const T *__restrict__ partials, int count, double *out) {
  __shared__ T lds[kFinalBlock / kWave];
  const T s = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) *out = (double)s;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct AsyncLayout {
  size_t xa, xb, c, invd, d, part, scal, total;
  int grid;
};

AsyncLayout async_layout(int nrow, int vb) {
  AsyncLayout L;
  const size_t v = align256((size_t)nrow * vb);
  size_t o = 0;
  L.xa = o;
  o += v;
  L.xb = o;
  o += v;
  L.c = o;
  o += v;
  L.invd = o;
  o += v;
  L.d = o;
  o += v;
  long long g = ((long long)nrow + 4LL * kBlock - 1) / (4LL * kBlock);
  if (g > kMaxRedBlocks) g = kMaxRedBlocks;
  if (g < 1) g = 1;
  L.grid = (int)g;
  L.part = o;
  o += align256((size_t)L.grid * vb);
  L.scal = o;
  o += 256;
  L.total = o;
  return L;
}

template <typename T>
int async_jacobi(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                 const T *val, const T *b, T *x, const hflmp_async_params *prm, int maxit,
                 double atol, double rtol, double dtol, void *work, size_t work_bytes,
                 hflmp_solve_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz >= 0, "async: empty system");
  HFLMP_REQUIRE(ctx, rp && (nnz == 0 || (ci && val)) && b && x && prm && work && info,
                "async: null pointer");
  HFLMP_REQUIRE(ctx, prm->block_rows >= 1 && prm->n_global >= 1 && prm->n_block >= 1 &&
                         prm->n_inner >= 1 && prm->mode >= 0 && prm->mode <= 3,
                "async: need block_rows, n_global, n_block, n_inner >= 1 and mode 0..3");
  const bool damped = prm->mode == 1 || prm->mode == 3;
  const bool smooth = prm->mode >= 2;
  HFLMP_REQUIRE(ctx, !damped || prm->w > 0, "async: damping must be positive");
  HFLMP_REQUIRE(ctx, smooth || (maxit > 0 && atol >= 0 && rtol >= 0 && dtol > 1),
                "async: bad control parameters (maxit>0, atol>=0, rtol>=0, dtol>1)");
  const AsyncLayout L = async_layout(nrow, sizeof(T));
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "async: workspace %zu < %zu bytes", work_bytes,
                L.total);
  char *wb = static_cast<char *>(work);
  T *xa = reinterpret_cast<T *>(wb + L.xa), *xb = reinterpret_cast<T *>(wb + L.xb);
  T *c = reinterpret_cast<T *>(wb + L.c);
  T *invd = reinterpret_cast<T *>(wb + L.invd), *d = reinterpret_cast<T *>(wb + L.d);
  T *part = reinterpret_cast<T *>(wb + L.part);
  double *scal = reinterpret_cast<double *>(wb + L.scal);
  unsigned int *bad = reinterpret_cast<unsigned int *>(scal + 1);
  hipStream_t s = ctx->stream;
  const int brows = prm->block_rows < nrow ? prm->block_rows : nrow;
  const T w = (T)prm->w;
  const AsyncMat<T> m{nrow, rp, ci, val, b, invd};
  const unsigned gr = grid_for(nrow, 1);

  const auto t0 = std::chrono::steady_clock::now();
  int rc = hflmp_fill_zero(ctx, bad, sizeof(unsigned int));
  if (rc) return rc;
  async_diag_kernel<T><<<gr, kBlock, 0, s>>>(nrow, rp, ci, val, invd, d, bad);
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(xa, x, sizeof(T) * (size_t)nrow,
                                      hipMemcpyDeviceToDevice, s));
  HFLMP_LAUNCH_CHECK(ctx);
  unsigned int hbad = 0;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (hbad)
    return set_error(ctx, HFLMP_ENUMERIC, "async: %u rows have a missing or zero diagonal",
                     hbad);

  T *cur = xa, *oth = xb;
  auto residual = [&](double *res) -> int {
    async_residual_kernel<T><<<L.grid, kBlock, 0, s>>>(m, d, cur, part);
    async_residual_final<T><<<1, kFinalBlock, 0, s>>>(part, L.grid, scal);
    HFLMP_LAUNCH_CHECK(ctx);
    double sq = 0;
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&sq, scal, sizeof(sq), hipMemcpyDeviceToHost, s));
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    *res = std::sqrt(sq);
    return HFLMP_OK;
  };
  auto launch = [&]() {  // one block iteration: cur -> result in cur
    if (brows <= kLdsRows) {
      // a multiple of the XCD count, so that xcd_tile reaches every block;
      // the workgroups past the last block return at once
      const int nblk = (nrow + brows - 1) / brows;
      const int grid = (nblk + kXcds - 1) / kXcds * kXcds;
      const size_t lds = 2 * (size_t)brows * sizeof(T);
      if (damped)
        async_lds_kernel<T, true><<<grid, kBlock, lds, s>>>(m, brows, prm->n_inner, w, cur, oth);
      else
        async_lds_kernel<T, false><<<grid, kBlock, lds, s>>>(m, brows, prm->n_inner, w, cur, oth);
      T *t = cur;
      cur = oth;
      oth = t;
      return;
    }
    async_outside_kernel<T><<<gr, kBlock, 0, s>>>(m, brows, cur, c);
    for (int it = 0; it < prm->n_inner; ++it) {
      if (damped)
        async_sweep_kernel<T, true><<<gr, kBlock, 0, s>>>(m, brows, w, c, cur, oth);
      else
        async_sweep_kernel<T, false><<<gr, kBlock, 0, s>>>(m, brows, w, c, cur, oth);
      T *t = cur;
      cur = oth;
      oth = t;
    }
  };
  auto run = [&]() -> int {
    for (int g = 0; g < prm->n_global; ++g)
      for (int k = 0; k < prm->n_block; ++k) launch();
    HFLMP_LAUNCH_CHECK(ctx);
    return HFLMP_OK;
  };

  info->iters = 0;
  info->status = 0;
  info->res_init = info->res_final = 0;
  if (smooth) {
    if ((rc = run())) return rc;
    info->iters = prm->n_global * prm->n_block * prm->n_inner;
  } else {
    double res = 0;
    if ((rc = residual(&res))) return rc;
    const double res0 = res;
    info->res_init = res0;
    // IterateControl::Check (common/iterate_control.cc:80-107)
    auto check = [&](int it, double r) {
      if (r <= atol) return 0;
      if (r < rtol * res0) return 0;
      if (r > dtol * res0) return 2;
      if (it >= maxit) return 1;
      if (!(r == r)) return 3;
      return -1;
    };
    int iter = 0, st = check(0, res);
    while (st == -1) {
      if ((rc = run())) return rc;
      iter += prm->n_global * prm->n_block * prm->n_inner;
      if ((rc = residual(&res))) return rc;
      st = check(iter, res);
    }
    info->iters = iter;
    info->status = st;
    info->res_final = res;
  }
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(x, cur, sizeof(T) * (size_t)nrow,
                                      hipMemcpyDeviceToDevice, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  info->seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return HFLMP_OK;
}

}  // namespace

extern "C" {

size_t hflmp_async_workspace_bytes(int nrow, int value_bytes) {
  if (nrow <= 0 || (value_bytes != 8 && value_bytes != 4)) return 0;
  return async_layout(nrow, value_bytes).total;
}

int hflmp_async_jacobi_f64(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                           const int *col, const double *val, const double *b, double *x,
                           const hflmp_async_params *prm, int maxit, double atol, double rtol,
                           double dtol, void *work, size_t work_bytes,
                           hflmp_solve_info *info) {
  return async_jacobi<double>(ctx, nrow, nnz, row_ptr, col, val, b, x, prm, maxit, atol, rtol,
                              dtol, work, work_bytes, info);
}

int hflmp_async_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr,
                           const int *col, const float *val, const float *b, float *x,
                           const hflmp_async_params *prm, int maxit, double atol, double rtol,
                           double dtol, void *work, size_t work_bytes,
                           hflmp_solve_info *info) {
  return async_jacobi<float>(ctx, nrow, nnz, row_ptr, col, val, b, x, prm, maxit, atol, rtol,
                             dtol, work, work_bytes, info);
}

}  // extern "C"
