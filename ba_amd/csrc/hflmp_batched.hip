// Several right-hand sides at once: CSR times a row-major block of vectors
// (SpMM) and Jacobi-PCG advancing up to 64 independent systems with the
// same matrix in lockstep.  No reference counterpart: the reference solves
// one right-hand side per CG::Solve call (linear_solver/cg.h:165-246); each
// column here follows exactly that arithmetic with its own alpha, beta and
// IterateControl test (common/iterate_control.cc:80-107).
//
// Layout: a block of k <= 64 vectors is row-major, X[i*k + c] is entry i of
// column c, so a row of the block is k contiguous values.  One wave handles
// one matrix row at a time with lane c on column c: the row's col/val are
// loaded by the lanes and broadcast with readlane, and every nonzero gathers
// one contiguous k-value row of X (512 B for k = 64 in FP64).  Each column
// is summed in ascending storage order from zero, as the single-vector SpMV
// (CPUsimple_csr_lmatrix.cc:76-111), so column c of Y = A X is bit-identical
// to A times column c.  Rows are walked so that the waves of one XCD cover a
// contiguous band of rows, whose X rows stay in that XCD's L2.
#include <chrono>

#include "hflmp_solver_common.hpp"

namespace {

using namespace hflmp_solver;

constexpr int kMaxRhs = 64;
constexpr int kWavesPerBlock = kBlock / kWave;

// wave w of the grid walks rows band_start + w, + nwaves_in_band, ... so
// that at any time the waves on one XCD are on neighbouring rows
struct RowWalk {
  long long first, end, step;
};

__device__ __forceinline__ RowWalk row_walk(int nrow) {
  const long long nb = gridDim.x;
  const long long per_xcd_blocks = nb / kXcds;  // grid is a multiple of kXcds
  const long long xcd = blockIdx.x % kXcds, bx = blockIdx.x / kXcds;
  const long long band = (nrow + kXcds - 1) / kXcds;
  const long long start = xcd * band;
  const long long end = start + band < nrow ? start + band : nrow;
  // the wave index is uniform across the wave: say so, so that row-indexed
  // loads go through the scalar unit
  const long long wave =
      bx * kWavesPerBlock + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  return {start + wave, end, per_xcd_blocks * kWavesPerBlock};
}

// broadcast lane e's value to the wave (v_readlane)
__device__ __forceinline__ int bcast(int v, int e) { return __builtin_amdgcn_readlane(v, e); }
__device__ __forceinline__ float bcast(float v, int e) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), e));
}
__device__ __forceinline__ double bcast(double v, int e) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), e);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), e);
  return __hiloint2double(hi, lo);
}

// y_c = sum_j a_ij x[col_j][c] for lane c < k, in ascending j
template <typename T>
__device__ __forceinline__ T spmm_row(long long i, int k, const int *__restrict__ rp,
                                      const int *__restrict__ ci, const T *__restrict__ val,
                                      const T *__restrict__ X) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const int a = rp[i], b = rp[i + 1];
  T s = 0;
  for (int base = a; base < b; base += kWave) {
    const int cnt = min(kWave, b - base);
    int my_c = 0;
    T my_v = 0;
    if (lane < cnt) {
      my_c = ld_stream(ci + base + lane);
      my_v = ld_stream(val + base + lane);
    }
    // eight X-row gathers in flight, then their products added in order
    int e = 0;
    for (; e + 8 <= cnt; e += 8) {
      T xv[8], vv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = bcast(my_c, e + u);
        vv[u] = bcast(my_v, e + u);
        xv[u] = lane < k ? X[(size_t)c * k + lane] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) s += vv[u] * xv[u];
    }
    for (; e < cnt; ++e) {
      const int c = bcast(my_c, e);
      const T v = bcast(my_v, e);
      if (lane < k) s += v * X[(size_t)c * k + lane];
    }
  }
  return s;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    spmm_kernel(int nrow, int k, const int *__restrict__ rp, const int *__restrict__ ci,
                const T *__restrict__ val, const T *__restrict__ X, T *__restrict__ Y) {
  const int lane = threadIdx.x & (kWave - 1);
  const RowWalk w = row_walk(nrow);
  for (long long i = w.first; i < w.end; i += w.step) {
    const T s = spmm_row(i, k, rp, ci, val, X);
    // streamed past L2, which keeps the X rows the band reuses
    if (lane < k) __builtin_nontemporal_store(s, Y + (size_t)i * k + lane);
  }
}

inline int spmm_grid(hflmp_ctx *ctx) {
  const int cus = ctx && ctx->num_cus > 0 ? ctx->num_cus : 256;
  return ((8 * cus + kXcds - 1) / kXcds) * kXcds;  // 8 workgroups (32 waves) per CU
}

// ---- batched Jacobi-PCG -----------------------------------------------------------
template <typename T>
struct BState {
  T rz[kMaxRhs], alpha[kMaxRhs], beta[kMaxRhs], res[kMaxRhs], res0[kMaxRhs];
  int iter[kMaxRhs], status[kMaxRhs];
  int any_iterate;  // polled by the host: 1 while some column iterates, else 0
  int n_live;       // columns still iterating; the one that stops the last clears any_iterate
  int maxit;
  double atol, rtol, dtol;
};

// per-column partial sums of one or two quantities, stored [q][column][block]
// so that each column's partials are contiguous for its finalize workgroup
template <typename T, int NQ>
__device__ __forceinline__ void block_col_partials(const T (&a)[NQ], int k, T *__restrict__ part) {
  __shared__ T lds[NQ][kWavesPerBlock][kWave];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NQ; ++q) lds[q][wid][lane] = a[q];
  __syncthreads();
  if (wid == 0 && lane < k) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      T s = 0;
#pragma unroll
      for (int w = 0; w < kWavesPerBlock; ++w) s += lds[q][w][lane];
      part[((size_t)q * k + lane) * gridDim.x + blockIdx.x] = s;
    }
  }
}

// The finalize kernels run one workgroup of kBlock threads per column
// (blockIdx.x = column).  col_total: the column's total of quantity q over
// `count` workgroup partials, eight loads in flight per thread, fixed order;
// valid in thread 0.
template <typename T>
__device__ __forceinline__ T col_total(const T *__restrict__ part, int count, int k, int q) {
  __shared__ T lds[kBlock / kWave];
  const T *__restrict__ src = part + ((size_t)q * k + blockIdx.x) * count;
  T a[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) a[f] = 0;
  for (int b0 = threadIdx.x; b0 < count; b0 += 8 * kBlock) {
    T v[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const int b = b0 + f * kBlock;
      v[f] = b < count ? src[b] : T(0);
    }
#pragma unroll
    for (int f = 0; f < 8; ++f) a[f] += v[f];
  }
  const T t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  return block_sum(t, lds);
}

// A column leaves the iteration: the last one to leave clears any_iterate.
template <typename T>
__device__ __forceinline__ void column_stops(BState<T> *st) {
  if (atomicSub(&st->n_live, 1) == 1) st->any_iterate = 0;
}

template <typename T>
__global__ void binit_state_kernel(BState<T> *st, int k, int maxit, double atol, double rtol,
                                   double dtol) {
  const int c = threadIdx.x;
  if (c < kMaxRhs) {
    st->rz[c] = st->alpha[c] = st->beta[c] = st->res[c] = st->res0[c] = T(0);
    st->iter[c] = 0;
    st->status[c] = c < k ? kIterate : kConverged;
  }
  if (c == 0) {
    st->any_iterate = 1;
    st->n_live = k;
    st->maxit = maxit;
    st->atol = atol;
    st->rtol = rtol;
    st->dtol = dtol;
  }
}

// R = B - A X, Z = D^-1 R, P = Z; partials r.r, r.z per column
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_init_kernel(int nrow, int k, const int *__restrict__ rp, const int *__restrict__ ci,
                    const T *__restrict__ val, const T *__restrict__ invd,
                    const T *__restrict__ B, const T *__restrict__ X, T *__restrict__ R,
                    T *__restrict__ Z, T *__restrict__ P, T *__restrict__ part) {
  const int lane = threadIdx.x & (kWave - 1);
  T acc[2] = {T(0), T(0)};
  const RowWalk w = row_walk(nrow);
  for (long long i = w.first; i < w.end; i += w.step) {
    const T s = spmm_row(i, k, rp, ci, val, X);
    if (lane < k) {
      const size_t e = (size_t)i * k + lane;
      const T ri = T(-1) * s + B[e];
      const T zi = ri * invd[i];
      R[e] = ri;
      Z[e] = zi;
      P[e] = zi;
      acc[0] += ri * ri;
      acc[1] += ri * zi;
    }
  }
  block_col_partials<T, 2>(acc, k, part);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_init_final(BState<T> *st, const T *__restrict__ part, int count, int k) {
  const T rr = col_total(part, count, k, 0);
  const T rz = col_total(part, count, k, 1);
  if (threadIdx.x == 0) {
    const int c = blockIdx.x;
    const T r0 = sqrt(rr);
    st->rz[c] = rz;
    st->res0[c] = r0;
    st->res[c] = r0;
    st->iter[c] = 0;
    const int s = check_state(0, (double)r0, (double)r0, st->maxit, st->atol, st->rtol,
                              st->dtol);
    st->status[c] = s;
    if (s != kIterate) column_stops(st);
  }
}

// AP = A P with per-column partials of p.Ap
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_spmm_kernel(int nrow, int k, const int *__restrict__ rp, const int *__restrict__ ci,
                    const T *__restrict__ val, const T *__restrict__ P, T *__restrict__ AP,
                    const BState<T> *st, T *__restrict__ part) {
  if (!st->any_iterate) return;
  const int lane = threadIdx.x & (kWave - 1);
  T acc[1] = {T(0)};
  const RowWalk w = row_walk(nrow);
  for (long long i = w.first; i < w.end; i += w.step) {
    const T s = spmm_row(i, k, rp, ci, val, P);
    if (lane < k) {
      const size_t e = (size_t)i * k + lane;
      __builtin_nontemporal_store(s, AP + e);
      acc[0] += s * P[e];
    }
  }
  block_col_partials<T, 1>(acc, k, part);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_alpha_final(BState<T> *st, const T *__restrict__ part, int count, int k) {
  const int c = blockIdx.x;
  if (!st->any_iterate || st->status[c] != kIterate) return;  // uniform per workgroup
  const T pap = col_total(part, count, k, 0);
  if (threadIdx.x == 0) {
    st->alpha[c] = st->rz[c] / pap;
    if (pap == T(0) || !(pap == pap)) {
      st->status[c] = kBreakdown;
      st->alpha[c] = T(0);
      column_stops(st);
    }
  }
}

// x += alpha p, r -= alpha Ap, z = D^-1 r for iterating columns; partials
// of r.r and r.z
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_update_kernel(int nrow, int k, const T *__restrict__ invd, const T *__restrict__ P,
                      const T *__restrict__ AP, T *__restrict__ X, T *__restrict__ R,
                      T *__restrict__ Z, const BState<T> *st, T *__restrict__ part) {
  __shared__ T s_alpha[kMaxRhs];
  __shared__ int s_live[kMaxRhs];
  if (!st->any_iterate) return;
  if (threadIdx.x < kMaxRhs) {
    s_alpha[threadIdx.x] = st->alpha[threadIdx.x];
    s_live[threadIdx.x] = threadIdx.x < k && st->status[threadIdx.x] == kIterate;
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  T acc[2] = {T(0), T(0)};
  if (lane < k && s_live[lane]) {
    const T a = s_alpha[lane], ma = -a;
    const RowWalk w = row_walk(nrow);
    for (long long i = w.first; i < w.end; i += w.step) {
      const size_t e = (size_t)i * k + lane;
      X[e] += a * P[e];
      const T ri = R[e] + ma * AP[e];
      const T zi = ri * invd[i];
      R[e] = ri;
      Z[e] = zi;
      acc[0] += ri * ri;
      acc[1] += ri * zi;
    }
  }
  block_col_partials<T, 2>(acc, k, part);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_beta_final(BState<T> *st, const T *__restrict__ part, int count, int k) {
  const int c = blockIdx.x;
  if (!st->any_iterate || st->status[c] != kIterate) return;  // uniform per workgroup
  const T rr = col_total(part, count, k, 0);
  const T rz = col_total(part, count, k, 1);
  if (threadIdx.x == 0) {
    const int it = st->iter[c] + 1;
    const T res = sqrt(rr);
    st->iter[c] = it;
    st->res[c] = res;
    const int s = check_state(it, (double)res, (double)st->res0[c], st->maxit, st->atol,
                              st->rtol, st->dtol);
    if (s == kIterate) {
      st->beta[c] = rz / st->rz[c];
      st->rz[c] = rz;
    }
    st->status[c] = s;
    if (s != kIterate) column_stops(st);
  }
}

// p = beta p + z for iterating columns
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bcg_p_kernel(int nrow, int k, const T *__restrict__ Z, T *__restrict__ P,
                 const BState<T> *st) {
  __shared__ T s_beta[kMaxRhs];
  __shared__ int s_live[kMaxRhs];
  if (!st->any_iterate) return;
  if (threadIdx.x < kMaxRhs) {
    s_beta[threadIdx.x] = st->beta[threadIdx.x];
    s_live[threadIdx.x] = threadIdx.x < k && st->status[threadIdx.x] == kIterate;
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  if (lane >= k || !s_live[lane]) return;
  const T bt = s_beta[lane];
  const RowWalk w = row_walk(nrow);
  for (long long i = w.first; i < w.end; i += w.step) {
    const size_t e = (size_t)i * k + lane;
    P[e] = bt * P[e] + Z[e];
  }
}

struct BLayout {
  size_t state, part, vec, vstride, total;
  int grid;
};

inline BLayout b_layout(hflmp_ctx *ctx, int nrow, int k, int vb) {
  BLayout L;
  L.grid = spmm_grid(ctx);
  L.state = 0;
  L.part = 4096;  // BState < 4 KiB
  L.vec = align256(L.part + 2 * (size_t)L.grid * k * vb);
  L.vstride = align256((size_t)nrow * k * vb);
  L.total = L.vec + 4 * L.vstride;  // R, P, Z, AP
  return L;
}

template <typename T>
int batched_pcg(hflmp_ctx *ctx, int nrow, int nnz, int k, const int *rp, const int *ci,
                const T *val, const T *invd, const T *B, T *X, int maxit, double atol,
                double rtol, double dtol, void *work, size_t work_bytes, int *iters,
                int *status, double *res_init, double *res_final, double *seconds) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0 && k >= 1 && k <= kMaxRhs,
                "batched_pcg: need nrow, nnz > 0 and 1 <= nrhs <= 64");
  HFLMP_REQUIRE(ctx, rp && ci && val && invd && B && X && work, "batched_pcg: null pointer");
  HFLMP_REQUIRE(ctx, maxit > 0 && atol >= 0 && rtol >= 0 && dtol > 1,
                "batched_pcg: bad control parameters");
  const BLayout L = b_layout(ctx, nrow, k, sizeof(T));
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "batched_pcg: workspace %zu < %zu bytes",
                work_bytes, L.total);
  char *w = static_cast<char *>(work);
  auto *st = reinterpret_cast<BState<T> *>(w + L.state);
  T *part = reinterpret_cast<T *>(w + L.part);
  T *R = reinterpret_cast<T *>(w + L.vec);
  T *P = reinterpret_cast<T *>(w + L.vec + L.vstride);
  T *Z = reinterpret_cast<T *>(w + L.vec + 2 * L.vstride);
  T *AP = reinterpret_cast<T *>(w + L.vec + 3 * L.vstride);
  hipStream_t s = ctx->stream;
  const int g = L.grid;
  const auto t0 = std::chrono::steady_clock::now();
  binit_state_kernel<T><<<1, kMaxRhs, 0, s>>>(st, k, maxit, atol, rtol, dtol);
  bcg_init_kernel<T><<<g, kBlock, 0, s>>>(nrow, k, rp, ci, val, invd, B, X, R, Z, P, part);
  bcg_init_final<T><<<k, kBlock, 0, s>>>(st, part, g, k);
  HFLMP_LAUNCH_CHECK(ctx);
  auto iteration = [&]() {
    bcg_spmm_kernel<T><<<g, kBlock, 0, s>>>(nrow, k, rp, ci, val, P, AP, st, part);
    bcg_alpha_final<T><<<k, kBlock, 0, s>>>(st, part, g, k);
    bcg_update_kernel<T><<<g, kBlock, 0, s>>>(nrow, k, invd, P, AP, X, R, Z, st, part);
    bcg_beta_final<T><<<k, kBlock, 0, s>>>(st, part, g, k);
    bcg_p_kernel<T><<<g, kBlock, 0, s>>>(nrow, k, Z, P, st);
  };
  // the poll reads any_iterate: stop when it drops to 0 (kConverged)
  int rc = run_polled(ctx, maxit, &st->any_iterate, iteration, 16, 1);
  if (rc) return rc;
  BState<T> h;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  for (int c = 0; c < k; ++c) {
    if (iters) iters[c] = h.iter[c];
    if (status) status[c] = h.status[c] == kIterate ? kMaxIts : h.status[c];
    if (res_init) res_init[c] = (double)h.res0[c];
    if (res_final) res_final[c] = (double)h.res[c];
  }
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return HFLMP_OK;
}

template <typename T>
int spmm(hflmp_ctx *ctx, int nrow, int ncol, int nnz, int k, const int *rp, const int *ci,
         const T *val, const T *X, T *Y) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow >= 0 && ncol >= 0 && nnz >= 0 && k >= 1 && k <= kMaxRhs,
                "spmm: bad sizes (1 <= nrhs <= 64)");
  if (nrow == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, rp && Y && (nnz == 0 || (ci && val && X)), "spmm: null pointer");
  spmm_kernel<T><<<spmm_grid(ctx), kBlock, 0, ctx->stream>>>(nrow, k, rp, ci, val, X, Y);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_csr_spmm_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz, int nrhs, const int *rp,
                       const int *ci, const double *val, const double *X, double *Y) {
  return spmm<double>(ctx, nrow, ncol, nnz, nrhs, rp, ci, val, X, Y);
}

int hflmp_csr_spmm_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz, int nrhs, const int *rp,
                       const int *ci, const float *val, const float *X, float *Y) {
  return spmm<float>(ctx, nrow, ncol, nnz, nrhs, rp, ci, val, X, Y);
}

size_t hflmp_batched_pcg_workspace_bytes(hflmp_ctx *ctx, int nrow, int nrhs, int value_bytes) {
  if (nrow <= 0 || nrhs < 1 || nrhs > kMaxRhs || (value_bytes != 4 && value_bytes != 8))
    return 0;
  return b_layout(ctx, nrow, nrhs, value_bytes).total;
}

int hflmp_batched_pcg_jacobi_f64(hflmp_ctx *ctx, int nrow, int nnz, int nrhs, const int *rp,
                                 const int *ci, const double *val, const double *invdiag,
                                 const double *B, double *X, int maxit, double atol,
                                 double rtol, double dtol, void *work, size_t work_bytes,
                                 int *iters, int *status, double *res_init, double *res_final,
                                 double *seconds) {
  return batched_pcg<double>(ctx, nrow, nnz, nrhs, rp, ci, val, invdiag, B, X, maxit, atol,
                             rtol, dtol, work, work_bytes, iters, status, res_init, res_final,
                             seconds);
}

int hflmp_batched_pcg_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz, int nrhs, const int *rp,
                                 const int *ci, const float *val, const float *invdiag,
                                 const float *B, float *X, int maxit, double atol,
                                 double rtol, double dtol, void *work, size_t work_bytes,
                                 int *iters, int *status, double *res_init, double *res_final,
                                 double *seconds) {
  return batched_pcg<float>(ctx, nrow, nnz, nrhs, rp, ci, val, invdiag, B, X, maxit, atol,
                            rtol, dtol, work, work_bytes, iters, status, res_init, res_final,
                            seconds);
}

}  // extern "C"
