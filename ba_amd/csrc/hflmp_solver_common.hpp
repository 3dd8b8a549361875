// Pieces shared by the device-resident Krylov solvers: the device state,
// the stopping test, per-workgroup partial sums, the workspace layout and
// the host loop that queues iterations ahead and polls the state.
#pragma once

#include "hflmp_spmv_tile.hpp"

namespace hflmp_solver {

constexpr int kIterate = -1;
constexpr int kConverged = 0;
constexpr int kMaxIts = 1;
constexpr int kDiverged = 2;
constexpr int kBreakdown = 3;

template <typename T>
struct CgState {
  T rz;     // r.z of the current iterate ("beta" in the reference)
  T pAp;
  T alpha;
  T beta;   // gamma / beta for the p update
  T rr;
  T res;
  T res0;
  int iter;
  int status;  // read by the host poll together with `maxit`
  int maxit;
  double atol, rtol, dtol;
  // the three-launch iteration: r.z and the iteration count as the update
  // kernel found them, read by the p kernel while its workgroup 0 writes
  // rz and iter
  T rz_prev;
  int iter_prev;
};

// IterateControl::Check (common/iterate_control.cc:80-107); a NaN residual
// that would iterate forever is reported as a breakdown.
__device__ __forceinline__ int check_state(int iter, double res, double res0, int maxit,
                                           double atol, double rtol, double dtol) {
  if (res <= atol) return kConverged;
  if (res < rtol * res0) return kConverged;
  if (res > dtol * res0) return kDiverged;
  if (iter >= maxit) return kMaxIts;
  if (!(res == res)) return kBreakdown;
  return kIterate;
}

// Workgroup sums of one / two values stored at partials[b] / [2b], [2b+1].
template <typename T>
__device__ __forceinline__ void block_partial1(T a, T *partials) {
  __shared__ T lds[kBlock / kWave];
  const T s = block_sum(a, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__device__ __forceinline__ void block_partial2(T a, T b, T *partials) {
  __shared__ T lds[2 * (kBlock / kWave)];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) {
    lds[wid] = a;
    lds[kBlock / kWave + wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T sa = 0, sb = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      sa += lds[w];
      sb += lds[kBlock / kWave + w];
    }
    partials[2 * blockIdx.x] = sa;
    partials[2 * blockIdx.x + 1] = sb;
  }
}

// Sum of `count` strided partials src[k*stride + off] over one workgroup of
// kBlock threads, returned to every thread: the same fixed order in every
// workgroup, so all workgroups of a kernel get the same value.
template <typename T>
__device__ __forceinline__ T block_total(const T *__restrict__ src, int count, int stride,
                                         int off, T *lds) {
  constexpr int F = 8;
  T a[F];
#pragma unroll
  for (int f = 0; f < F; ++f) a[f] = 0;
  for (int k0 = threadIdx.x; k0 < count; k0 += F * kBlock) {
    T v[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const int kk = k0 + f * kBlock;
      v[f] = kk < count ? src[(size_t)kk * stride + off] : T(0);
    }
#pragma unroll
    for (int f = 0; f < F; ++f) a[f] += v[f];
  }
  const T v = wave_sum(((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7])));
  __syncthreads();  // a previous call may still be reading lds
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) s += lds[w];
  return s;
}

template <typename T>
__global__ void init_state_kernel(CgState<T> *st, int maxit, double atol, double rtol,
                                  double dtol) {
  if (threadIdx.x == 0) {
    CgState<T> s;
    s.rz = s.pAp = s.alpha = s.beta = s.rr = s.res = s.res0 = s.rz_prev = T(0);
    s.iter_prev = 0;
    s.iter = 0;
    s.status = kIterate;
    s.maxit = maxit;
    s.atol = atol;
    s.rtol = rtol;
    s.dtol = dtol;
    *st = s;
  }
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Workspace: [state | partials | nvec vectors of nrow values].
struct SolverLayout {
  size_t state, partials, vec, vstride, total;
  int R, grid1, grid2;
};

inline SolverLayout solver_layout(int nrow, int nnz, int vb, int nvec) {
  SolverLayout L;
  L.R = tile_rows(nrow, nnz, vb == 8 ? TileCap<double>::value : TileCap<float>::value);
  const long long ntiles = ((long long)nrow + L.R - 1) / L.R;
  L.grid1 = tile_grid(ntiles);
  // elementwise update kernels: enough workgroups to fill 256 CUs, capped
  long long g2 = ((long long)nrow + 4LL * kBlock - 1) / (4LL * kBlock);
  if (g2 > kMaxRedBlocks) g2 = kMaxRedBlocks;
  if (g2 < 1) g2 = 1;
  L.grid2 = (int)g2;
  const int gmax = L.grid1 > L.grid2 ? L.grid1 : L.grid2;
  L.state = 0;
  L.partials = 256;
  L.vec = align256(L.partials + 2 * (size_t)gmax * vb);
  L.vstride = align256((size_t)nrow * vb);
  L.total = L.vec + (size_t)nvec * L.vstride;
  return L;
}

// Queue `poll` iterations per chunk and keep one chunk queued ahead of the
// one whose state the host reads.  `status` points at the device state's
// status word, which holds `iterate_value` while the solve continues; every
// kernel of an iteration returns at once once it changes, so chunks queued
// past convergence cost only launches.
template <typename F>
int run_polled(hflmp_ctx *ctx, int maxit, const int *status, F &&iteration, int poll = 16,
               int iterate_value = kIterate) {
  hipStream_t s = ctx->stream;
  hipEvent_t ev[2];
  HFLMP_HIP_CHECK(ctx, hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  HFLMP_HIP_CHECK(ctx, hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int *slot[2] = {reinterpret_cast<int *>(static_cast<char *>(ctx->host_pinned) + 1024),
                  reinterpret_cast<int *>(static_cast<char *>(ctx->host_pinned) + 1088)};
  int launched = 0, chunk = 0, rc = HFLMP_OK;
  auto enqueue_chunk = [&](int k) -> int {
    const int m = (maxit - launched) < poll ? (maxit - launched) : poll;
    for (int it = 0; it < m; ++it) iteration();
    launched += m;
    if (hipGetLastError() != hipSuccess) return HFLMP_EHIP;
    if (hipMemcpyAsync(slot[k], status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
      return HFLMP_EHIP;
    if (hipEventRecord(ev[k], s) != hipSuccess) return HFLMP_EHIP;
    return HFLMP_OK;
  };
  rc = enqueue_chunk(0);
  while (rc == HFLMP_OK) {
    const int cur = chunk & 1;
    if (launched < maxit) rc = enqueue_chunk(cur ^ 1);
    if (rc) break;
    if (hipEventSynchronize(ev[cur]) != hipSuccess) {
      rc = HFLMP_EHIP;
      break;
    }
    // Once every iteration is queued the caller's final synchronisation
    // waits for the rest (the state cannot stay "iterate" past maxit).
    if (slot[cur][0] != iterate_value || launched >= maxit) break;
    ++chunk;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return set_error(ctx, rc, "solver: HIP error while iterating");
  return HFLMP_OK;
}

// The same loop without events (nothing is created): each chunk of `poll`
// iterations is followed by a copy of the status word into the context's
// pinned staging and a wait for the stream, so the device idles while the
// host reads the status and queues the next chunk.
template <typename F>
int run_polled_sync(hflmp_ctx *ctx, int maxit, const int *status, F &&iteration, int poll = 16,
                    int iterate_value = kIterate) {
  hipStream_t s = ctx->stream;
  int *slot = reinterpret_cast<int *>(static_cast<char *>(ctx->host_pinned) + 1152);
  int launched = 0;
  while (launched < maxit) {
    const int m = (maxit - launched) < poll ? (maxit - launched) : poll;
    for (int it = 0; it < m; ++it) iteration();
    launched += m;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(slot, status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return set_error(ctx, HFLMP_EHIP, "solver: HIP error while iterating");
    if (slot[0] != iterate_value) break;
  }
  return HFLMP_OK;
}

}  // namespace hflmp_solver
