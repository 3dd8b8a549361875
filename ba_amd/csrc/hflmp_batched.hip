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
#include <cmath>

#include "hflmp_solver_common.hpp"
#include "hflmp_spmm_plan.hpp"

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

// A wave's walk over its rows with the next row's entries loaded while the
// current row's X rows are gathered.  The row pointers come one row further
// ahead, so every load of a step is issued without waiting for another:
// the step's entries (lane e holds entry e of the row, e < 64) and the
// following row's pointer.
template <typename T>
struct RowPipe {
  long long i;     // current row
  int a, cnt;      // its first entry and length
  int c;           // lane e: column of entry e (e < 64)
  T v;             //          value of entry e
  int a1, cnt1;    // the next row's first entry and length (0 past the end)
};

template <typename T>
__device__ __forceinline__ void pipe_entries(int a, int cnt, const int *__restrict__ ci,
                                             const T *__restrict__ val, int &c, T &v) {
  const int lane = threadIdx.x & (kWave - 1);
  c = 0;
  v = T(0);
  if (lane < cnt && lane < kWave) {
    c = ld_stream(ci + a + lane);
    v = ld_stream(val + a + lane);
  }
}

__device__ __forceinline__ void pipe_ptrs(long long i, long long end, const int *__restrict__ rp,
                                          int &a, int &cnt) {
  a = 0;
  cnt = 0;
  if (i < end) {
    a = rp[i];
    cnt = rp[i + 1] - a;
  }
}

template <typename T>
__device__ __forceinline__ void pipe_start(RowPipe<T> &p, const RowWalk &w,
                                           const int *__restrict__ rp, const int *__restrict__ ci,
                                           const T *__restrict__ val) {
  p.i = w.first;
  pipe_ptrs(w.first, w.end, rp, p.a, p.cnt);
  pipe_entries(p.a, p.cnt, ci, val, p.c, p.v);
  pipe_ptrs(w.first + w.step, w.end, rp, p.a1, p.cnt1);
}

// y_c = sum_j a_ij x[col_j][c] for lane c < k over the current row, in
// ascending j, and the pipe advanced to the next row.  Rows of at most 32
// entries gather all their X rows at once; longer ones in batches of 32
// (entries past 64 through spmm_row).
template <typename T>
__device__ __forceinline__ T pipe_row(RowPipe<T> &p, const RowWalk &w, int k,
                                      const int *__restrict__ rp, const int *__restrict__ ci,
                                      const T *__restrict__ val, const T *__restrict__ X) {
#pragma clang fp contract(off)
  constexpr int U = 32;
  const int lane = threadIdx.x & (kWave - 1);
  // Load counters count in issue order, so a batch's gathers go first:
  // waiting for them then leaves the next row's entries (issued right
  // after them, in flight during the adds) and the pointers of the row
  // after it pending.  Gathers come in groups of four with no lane masking
  // (lanes >= k read column k - 1; entries past the row read the column of
  // lane list entry 0, which pipe_entries set to 0): every load is in bounds
  // and nothing is zeroed before its load, so nothing waits between loads.
  int c_n = 0;
  T v_n = 0;
  int a2 = 0, cnt2 = 0;
  bool fetched = false;
  T s = 0;
  const int lk = lane < k ? lane : k - 1;
  if (p.cnt <= kWave) {
    for (int base = 0; base < p.cnt; base += U) {
      const int m = p.cnt - base;
      T xv[U];
#pragma unroll
      for (int g = 0; g < U; g += 4)
        if (g < m) {
#pragma unroll
          for (int u = g; u < g + 4; ++u) {
            const int c = bcast(p.c, (base + u) & (kWave - 1));
            xv[u] = X[(size_t)c * k + lk];
          }
        }
      if (!fetched) {
        pipe_entries(p.a1, p.cnt1, ci, val, c_n, v_n);
        pipe_ptrs(p.i + 2 * w.step, w.end, rp, a2, cnt2);
        fetched = true;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (u < m) s += bcast(p.v, base + u) * xv[u];
    }
  }
  if (!fetched) {
    pipe_entries(p.a1, p.cnt1, ci, val, c_n, v_n);
    pipe_ptrs(p.i + 2 * w.step, w.end, rp, a2, cnt2);
  }
  if (p.cnt > kWave) s = spmm_row(p.i, k, rp, ci, val, X);
  p.i += w.step;
  p.a = p.a1;
  p.cnt = p.cnt1;
  p.c = c_n;
  p.v = v_n;
  p.a1 = a2;
  p.cnt1 = cnt2;
  return s;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    spmm_kernel(int nrow, int k, const int *__restrict__ rp, const int *__restrict__ ci,
                const T *__restrict__ val, const T *__restrict__ X, T *__restrict__ Y) {
  const int lane = threadIdx.x & (kWave - 1);
  const RowWalk w = row_walk(nrow);
  RowPipe<T> p;
  pipe_start(p, w, rp, ci, val);
  while (p.i < w.end) {
    const long long i = p.i;
    const T s = pipe_row(p, w, k, rp, ci, val, X);
    // streamed past L2, which keeps the X rows the band reuses
    if (lane < k) __builtin_nontemporal_store(s, Y + (size_t)i * k + lane);
  }
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
  RowPipe<T> pp;
  pipe_start(pp, w, rp, ci, val);
  while (pp.i < w.end) {
    const long long i = pp.i;
    const T s = pipe_row(pp, w, k, rp, ci, val, X);
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
  RowPipe<T> pp;
  pipe_start(pp, w, rp, ci, val);
  while (pp.i < w.end) {
    const long long i = pp.i;
    const T s = pipe_row(pp, w, k, rp, ci, val, P);
    if (lane < k) {
      const size_t e = (size_t)i * k + lane;
      __builtin_nontemporal_store(s, AP + e);
      acc[0] += s * P[e];
    }
  }
  block_col_partials<T, 1>(acc, k, part);
}

// The same through a block layout (hflmp_spmm_plan.hpp): AP = A P with
// per-column partials of p.Ap, one set per workgroup of the walk.
template <typename T, int KC>
__global__ void __launch_bounds__(kBlock)
    bcg_plan_spmm_kernel(PlanView pv, int k, const int *__restrict__ rp,
                         const T *__restrict__ val, const T *__restrict__ P,
                         T *__restrict__ AP, const BState<T> *st, T *__restrict__ part) {
  extern __shared__ __align__(16) unsigned char plan_lds[];
  if (!st->any_iterate) return;
  T acc[1] = {T(0)};
  const int lane = threadIdx.x & (kWave - 1);
  const int lk = lane < k ? lane : k - 1;
  plan_walk<T, KC>(
      pv, k, rp, val, P, plan_lds, [&](int i) { return P[(size_t)i * k + lk]; },
      [&](int i, int c, T s, T p) {
        if (c < k) {
          __builtin_nontemporal_store(s, AP + (size_t)i * k + c);
          acc[0] += s * p;
        }
      });
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

// Workgroups per CU of the row-walk kernels.  Measured on the 2D Q2 512^2
// SpMM with 64 columns (all bit-identical): FP64 -- 8 per CU 832 us, 6 per
// CU 956 us, 5 per CU (all resident at 92 VGPRs) 986 us, 4 per CU 875 us;
// FP32 -- 6 per CU 490 us, 5 per CU 540 us, 8 per CU 574 us.
inline int row_walk_blocks_per_cu(int vb) { return vb == 8 ? 8 : 6; }

inline int spmm_grid(hflmp_ctx *ctx, int vb) {
  const int cus = ctx && ctx->num_cus > 0 ? ctx->num_cus : 256;
  const int per = row_walk_blocks_per_cu(vb);
  return ((per * cus + kXcds - 1) / kXcds) * kXcds;
}

struct BLayout {
  size_t state, part, vec, vstride, total;
  int grid;
};

inline BLayout b_layout(hflmp_ctx *ctx, int nrow, int k, int vb) {
  BLayout L;
  L.grid = spmm_grid(ctx, vb);
  L.state = 0;
  L.part = 4096;  // BState < 4 KiB
  L.vec = align256(L.part + 2 * (size_t)L.grid * k * vb);
  L.vstride = align256((size_t)nrow * k * vb);
  L.total = L.vec + 4 * L.vstride;  // R, P, Z, AP
  return L;
}

// plan: the products of the iteration go through this block layout of the
// matrix (null: the CSR row walk).
template <typename T>
int batched_pcg(hflmp_ctx *ctx, int nrow, int nnz, int k, const int *rp, const int *ci,
                const T *val, const T *invd, const T *B, T *X, int maxit, double atol,
                double rtol, double dtol, void *work, size_t work_bytes, int *iters,
                int *status, double *res_init, double *res_final, double *seconds,
                const hflmp_spmm_plan *plan = nullptr) {
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
  // the block layout's walk: its grid (at most the partial area's) and LDS
  int gp = g;
  size_t lds = 0;
  auto plan_kern = k == 64 ? &bcg_plan_spmm_kernel<T, 64> : &bcg_plan_spmm_kernel<T, 0>;
  if (plan) {
    HFLMP_REQUIRE(ctx, plan->nrow == nrow && plan->nnz == nnz && plan->device == ctx->device,
                  "batched_pcg: the plan was built for another matrix or device");
    HFLMP_REQUIRE(ctx, plan->umax <= plan_max_cols((int)sizeof(T)),
                  "batched_pcg: the plan's blocks have up to %d columns, above %d for %zu-byte "
                  "values", plan->umax, plan_max_cols((int)sizeof(T)), sizeof(T));
    lds = plan_lds_bytes(plan->R, plan->umax, plan->emax, k, (int)sizeof(T));
    HFLMP_REQUIRE(ctx, lds <= 128 * 1024,
                  "batched_pcg: %zu bytes of LDS for %d x %zu-byte rows; build the plan for "
                  "these rows", lds, k, sizeof(T));
    if (lds > 64 * 1024)
      HFLMP_HIP_CHECK(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(plan_kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds));
    gp = plan_grid(ctx, reinterpret_cast<const void *>(plan_kern), plan->nblk, lds);
    if (gp > g) gp = g / kXcds * kXcds;
  }
  const auto t0 = std::chrono::steady_clock::now();
  binit_state_kernel<T><<<1, kMaxRhs, 0, s>>>(st, k, maxit, atol, rtol, dtol);
  bcg_init_kernel<T><<<g, kBlock, 0, s>>>(nrow, k, rp, ci, val, invd, B, X, R, Z, P, part);
  bcg_init_final<T><<<k, kBlock, 0, s>>>(st, part, g, k);
  HFLMP_LAUNCH_CHECK(ctx);
  auto iteration = [&]() {
    if (plan) {
      plan_kern<<<gp, kBlock, lds, s>>>(plan_view(plan), k, rp, val, P, AP, st, part);
      bcg_alpha_final<T><<<k, kBlock, 0, s>>>(st, part, gp, k);
    } else {
      bcg_spmm_kernel<T><<<g, kBlock, 0, s>>>(nrow, k, rp, ci, val, P, AP, st, part);
      bcg_alpha_final<T><<<k, kBlock, 0, s>>>(st, part, g, k);
    }
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

// ---- FP32-inner / FP64-refined blocked CG -------------------------------------
// MPIR<LADescriptorCoupledD, LADescriptorCoupledS> (linear_solver/mpir.cc:
// 52-296) applied column by column to a block of right-hand sides: per outer
// step R = B - A X in FP64 (one SpMM over the block) with each column's
// ||r||, the outer test per column (initial residual kept, as
// hflmp_mpcg_jacobi), R32 = (float) R for the columns still refining (0 for
// the others, whose inner solve then stops at once), the FP32 lockstep
// Jacobi-PCG above on A32 C = R32 from C = 0 to inner_rtol, and
// X += (double) C.

// R = B - A X (FP64); R32 = (float) R for live columns, else 0; per-column
// partials of r.r
__global__ void __launch_bounds__(kBlock)
    bmp_residual_kernel(int nrow, int k, const int *__restrict__ rp, const int *__restrict__ ci,
                        const double *__restrict__ val, const double *__restrict__ B,
                        const double *__restrict__ X, const int *__restrict__ live,
                        float *__restrict__ R32, double *__restrict__ part) {
  const int lane = threadIdx.x & (kWave - 1);
  double acc[1] = {0.0};
  const bool on = lane < k && live[lane];
  const RowWalk w = row_walk(nrow);
  RowPipe<double> pp;
  pipe_start(pp, w, rp, ci, val);
  while (pp.i < w.end) {
    const long long i = pp.i;
    const double s = pipe_row(pp, w, k, rp, ci, val, X);
    if (lane < k) {
      const size_t e = (size_t)i * k + lane;
      const double ri = -1.0 * s + B[e];
      R32[e] = on ? (float)ri : 0.0f;
      acc[0] += ri * ri;
    }
  }
  block_col_partials<double, 1>(acc, k, part);
}

__global__ void __launch_bounds__(kBlock)
    bmp_norm_final(const double *__restrict__ part, int count, int k, double *__restrict__ rr) {
  const double t = col_total(part, count, k, 0);
  if (threadIdx.x == 0) rr[blockIdx.x] = t;
}

// R32 = 0 in the columns that are not live
__global__ void __launch_bounds__(kBlock)
    bmp_mask_kernel(long long n, int k, const int *__restrict__ live, float *__restrict__ R32) {
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < n;
       e += (long long)gridDim.x * kBlock)
    if (!live[e % k]) R32[e] = 0.0f;
}

// X += (double) C for live columns
__global__ void __launch_bounds__(kBlock)
    bmp_correct_kernel(long long n, int k, const float *__restrict__ C,
                       const int *__restrict__ live, double *__restrict__ X) {
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < n;
       e += (long long)gridDim.x * kBlock)
    if (live[e % k]) X[e] += (double)C[e];
}

struct BmpLayout {
  size_t inner, r32, c32, part, rr, live, total;
};

inline BmpLayout bmp_layout(hflmp_ctx *ctx, int nrow, int k) {
  BmpLayout L;
  const BLayout I = b_layout(ctx, nrow, k, 4);
  const int g = spmm_grid(ctx, 8);
  size_t o = 0;
  L.inner = o;
  o = align256(o + I.total);
  L.r32 = o;
  o = align256(o + (size_t)nrow * k * 4);
  L.c32 = o;
  o = align256(o + (size_t)nrow * k * 4);
  L.part = o;
  o = align256(o + (size_t)g * k * 8);
  L.rr = o;
  o = align256(o + (size_t)kMaxRhs * 8);
  L.live = o;
  o = align256(o + (size_t)kMaxRhs * 4);
  L.total = o;
  return L;
}

// plan32: the inner FP32 solves' products through this block layout (null:
// the CSR walk).
int batched_mpcg(hflmp_ctx *ctx, int nrow, int nnz, int k, const int *rp, const int *ci,
                 const double *val, const float *val32, const float *invd32, const double *B,
                 double *X, int max_outer, double atol, double rtol, int inner_maxit,
                 double inner_rtol, void *work, size_t work_bytes, int *outer, int *inner,
                 int *status, double *res_init, double *res_final, double *seconds,
                 const hflmp_spmm_plan *plan32 = nullptr) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0 && k >= 1 && k <= kMaxRhs,
                "batched_mpcg: need nrow, nnz > 0 and 1 <= nrhs <= 64");
  HFLMP_REQUIRE(ctx, rp && ci && val && val32 && invd32 && B && X && work,
                "batched_mpcg: null pointer");
  HFLMP_REQUIRE(ctx, max_outer >= 1 && atol >= 0 && rtol >= 0 && inner_maxit > 0 &&
                         inner_rtol > 0 && inner_rtol < 1,
                "batched_mpcg: need max_outer >= 1, inner_maxit > 0, 0 < inner_rtol < 1");
  const BmpLayout L = bmp_layout(ctx, nrow, k);
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "batched_mpcg: workspace %zu < %zu bytes",
                work_bytes, L.total);
  char *w = static_cast<char *>(work);
  float *R32 = reinterpret_cast<float *>(w + L.r32);
  float *C32 = reinterpret_cast<float *>(w + L.c32);
  double *part = reinterpret_cast<double *>(w + L.part);
  double *rr = reinterpret_cast<double *>(w + L.rr);
  int *live = reinterpret_cast<int *>(w + L.live);
  hipStream_t s = ctx->stream;
  const int g = spmm_grid(ctx, 8);
  const long long nk = (long long)nrow * k;
  const auto t0 = std::chrono::steady_clock::now();
  int h_live[kMaxRhs], h_out[kMaxRhs], h_in[kMaxRhs];
  double h_rr[kMaxRhs], r0[kMaxRhs], rcur[kMaxRhs];
  for (int c = 0; c < kMaxRhs; ++c) {
    h_live[c] = c < k;
    h_out[c] = h_in[c] = 0;
  }
  for (int step = 0;; ++step) {
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(live, h_live, sizeof(int) * kMaxRhs,
                                        hipMemcpyHostToDevice, s));
    bmp_residual_kernel<<<g, kBlock, 0, s>>>(nrow, k, rp, ci, val, B, X, live, R32, part);
    bmp_norm_final<<<k, kBlock, 0, s>>>(part, g, k, rr);
    HFLMP_LAUNCH_CHECK(ctx);
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(h_rr, rr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    int n_live = 0;
    for (int c = 0; c < k; ++c) {
      if (!h_live[c]) continue;
      const double res = std::sqrt(h_rr[c]);
      if (step == 0) r0[c] = res;
      rcur[c] = res;
      // the outer test of MPIR with the initial residual kept
      if (res <= atol || res < rtol * r0[c] || !(res == res) || h_out[c] >= max_outer) {
        h_live[c] = 0;
        continue;
      }
      ++n_live;
    }
    if (!n_live) break;
    // the columns that stopped in this test: R32 = 0, so their inner solve
    // stops at once (and X is no longer corrected)
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(live, h_live, sizeof(int) * kMaxRhs,
                                        hipMemcpyHostToDevice, s));
    bmp_mask_kernel<<<grid_for(nk, 4), kBlock, 0, s>>>(nk, k, live, R32);
    int rc = hflmp_fill_zero(ctx, C32, sizeof(float) * (size_t)nk);
    if (rc) return rc;
    int it[kMaxRhs];
    rc = batched_pcg<float>(ctx, nrow, nnz, k, rp, ci, val32, invd32, R32, C32, inner_maxit,
                            0.0, inner_rtol, 1e6, w + L.inner, L.r32 - L.inner, it, nullptr,
                            nullptr, nullptr, nullptr, plan32);
    if (rc) return rc;
    bmp_correct_kernel<<<grid_for(nk, 4), kBlock, 0, s>>>(nk, k, C32, live, X);
    HFLMP_LAUNCH_CHECK(ctx);
    for (int c = 0; c < k; ++c)
      if (h_live[c]) {
        ++h_out[c];
        h_in[c] += it[c];
      }
  }
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  for (int c = 0; c < k; ++c) {
    const bool conv = rcur[c] <= atol || rcur[c] < rtol * r0[c];
    if (outer) outer[c] = h_out[c];
    if (inner) inner[c] = h_in[c];
    if (status) status[c] = conv ? kConverged : (rcur[c] == rcur[c] ? kMaxIts : kBreakdown);
    if (res_init) res_init[c] = r0[c];
    if (res_final) res_final[c] = rcur[c];
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
  spmm_kernel<T><<<spmm_grid(ctx, (int)sizeof(T)), kBlock, 0, ctx->stream>>>(nrow, k, rp, ci, val, X, Y);
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

size_t hflmp_batched_mpcg_workspace_bytes(hflmp_ctx *ctx, int nrow, int nrhs) {
  if (nrow <= 0 || nrhs < 1 || nrhs > kMaxRhs) return 0;
  return bmp_layout(ctx, nrow, nrhs).total;
}

int hflmp_batched_mpcg_jacobi(hflmp_ctx *ctx, int nrow, int nnz, int nrhs, const int *rp,
                              const int *ci, const double *val, const float *val32,
                              const float *invdiag32, const double *B, double *X, int max_outer,
                              double atol, double rtol, int inner_maxit, double inner_rtol,
                              void *work, size_t work_bytes, int *outer_iters,
                              int *inner_iters, int *status, double *res_init,
                              double *res_final, double *seconds) {
  return batched_mpcg(ctx, nrow, nnz, nrhs, rp, ci, val, val32, invdiag32, B, X, max_outer,
                      atol, rtol, inner_maxit, inner_rtol, work, work_bytes, outer_iters,
                      inner_iters, status, res_init, res_final, seconds);
}

int hflmp_batched_pcg_jacobi_plan_f64(hflmp_ctx *ctx, const hflmp_spmm_plan *plan, int nrow,
                                      int nnz, int nrhs, const int *rp, const int *ci,
                                      const double *val, const double *invdiag,
                                      const double *B, double *X, int maxit, double atol,
                                      double rtol, double dtol, void *work, size_t work_bytes,
                                      int *iters, int *status, double *res_init,
                                      double *res_final, double *seconds) {
  if (!plan) return ctx ? set_error(ctx, HFLMP_EINVAL, "batched_pcg: null plan") : HFLMP_EINVAL;
  return batched_pcg<double>(ctx, nrow, nnz, nrhs, rp, ci, val, invdiag, B, X, maxit, atol,
                             rtol, dtol, work, work_bytes, iters, status, res_init, res_final,
                             seconds, plan);
}

int hflmp_batched_pcg_jacobi_plan_f32(hflmp_ctx *ctx, const hflmp_spmm_plan *plan, int nrow,
                                      int nnz, int nrhs, const int *rp, const int *ci,
                                      const float *val, const float *invdiag, const float *B,
                                      float *X, int maxit, double atol, double rtol,
                                      double dtol, void *work, size_t work_bytes, int *iters,
                                      int *status, double *res_init, double *res_final,
                                      double *seconds) {
  if (!plan) return ctx ? set_error(ctx, HFLMP_EINVAL, "batched_pcg: null plan") : HFLMP_EINVAL;
  return batched_pcg<float>(ctx, nrow, nnz, nrhs, rp, ci, val, invdiag, B, X, maxit, atol,
                            rtol, dtol, work, work_bytes, iters, status, res_init, res_final,
                            seconds, plan);
}

int hflmp_batched_mpcg_jacobi_plan(hflmp_ctx *ctx, const hflmp_spmm_plan *plan32, int nrow,
                                   int nnz, int nrhs, const int *rp, const int *ci,
                                   const double *val, const float *val32,
                                   const float *invdiag32, const double *B, double *X,
                                   int max_outer, double atol, double rtol, int inner_maxit,
                                   double inner_rtol, void *work, size_t work_bytes,
                                   int *outer_iters, int *inner_iters, int *status,
                                   double *res_init, double *res_final, double *seconds) {
  if (!plan32)
    return ctx ? set_error(ctx, HFLMP_EINVAL, "batched_mpcg: null plan") : HFLMP_EINVAL;
  return batched_mpcg(ctx, nrow, nnz, nrhs, rp, ci, val, val32, invdiag32, B, X, max_outer,
                      atol, rtol, inner_maxit, inner_rtol, work, work_bytes, outer_iters,
                      inner_iters, status, res_init, res_final, seconds, plan32);
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
