// Geometric multigrid on the device: V-cycles over a hierarchy of CSR
// operators with CSR restriction and prolongation, as a preconditioner of
// CG and as a solver.  The reference's GMG (linear_solver/gmg.h:979-1230,
// GMGBase::SolveImpl / SolveImplLevel) runs pre-smoothing, the residual,
// restriction, the cycle on the coarser level, prolongation and
// post-smoothing; its GPU kernels (lmp/cuda/cuda_gmg_kernel.h) compute the
// residual as b minus the row's terms one by one, the restriction as a
// product from zero and the prolongation as the terms added one by one
// onto the solution.  The smoother is damped Jacobi with the arithmetic of
// kernel_async_inner_damped_iter_no_offdiag (cuda_async_iter_kernels.h:
// 116-144), run synchronously (every sweep reads the previous iterate).
//
// Every operator is a SELL-64 copy; each row is processed in ascending
// storage order, so every kernel is bit-identical to a sequential loop over
// the CSR rows.  One cycle from level l (right-hand side b_l):
//   x = (w invd) b, then pre-1 sweeps x' = (w invd (b - (A-D) x)) + ((1-w) x)
//   r = b - A x;  b_{l+1} = R r;  cycle on level l+1;  x = x + P x_{l+1};
//   post sweeps.  Level L-1: the first step and coarse-1 sweeps.
// The first step is a sweep from x = 0 without the product with zeros.
// With pre = post and R = P^T the cycle is a symmetric operator, so it can
// precondition CG.  Dirichlet rows are unit rows whose right-hand side is
// zero on every level (R has zero rows for coarse boundary nodes, P zero
// rows for fine ones), so they stay zero without fixed-dof handling.
#include <chrono>
#include <utility>

#include "hflmp_sell.hpp"
#include "hflmp_solver_common.hpp"
#include "hflmp_stencil.hpp"

namespace {

using namespace hflmp_solver;
using Op = SellOp<double, int>;
constexpr int kMaxLevels = 16;

__device__ __forceinline__ bool stopped(const int *status) {
  return status && *status != kIterate;
}

// x = (w invd) b: the sweep from x = 0
__global__ void __launch_bounds__(kBlock)
    mg_first_kernel(int n, double w, const double *__restrict__ invd,
                    const double *__restrict__ b, double *__restrict__ x, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    x[i] = w * invd[i] * b[i];
}

// xo = (w invd (b - (A - D) xi)) + ((1 - w) xi)
__global__ void __launch_bounds__(kBlock)
    mg_sweep_kernel(Op A, double w, double one_minus_w, const double *__restrict__ invd,
                    const double *__restrict__ b, const double *__restrict__ xi,
                    double *__restrict__ xo, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  const long long r = A.thread_row(unit);
  const double b0 = r < A.nrow ? b[r] : 0.0;
  TileRow tr;
  const double t = A.row_sum<true, true>(unit, xi, b0, tr, nullptr);
  if (tr.has_row) {
    const int i = tr.row;
    xo[i] = (w * invd[i] * t) + (one_minus_w * xi[i]);
  }
}

// r = b - A x, term by term
__global__ void __launch_bounds__(kBlock)
    mg_residual_kernel(Op A, const double *__restrict__ b, const double *__restrict__ x,
                       double *__restrict__ r, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  const long long row = A.thread_row(unit);
  const double b0 = row < A.nrow ? b[row] : 0.0;
  TileRow tr;
  const double s = A.row_sum<true>(unit, x, b0, tr, nullptr);
  if (tr.has_row) r[tr.row] = s;
}

// y = M x from zero (ADD: y = y + M x, term by term onto y)
template <bool ADD>
__global__ void __launch_bounds__(kBlock)
    mg_apply_kernel(Op M, const double *__restrict__ x, double *__restrict__ y, const int *status) {
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  double s = 0;
  if (ADD) {
    const long long r = M.thread_row(unit);
    if (r < M.nrow) s = y[r];
  }
  TileRow tr;
  s = M.row_sum(unit, x, s, tr, nullptr);
  if (tr.has_row) y[tr.row] = s;
}

// ---- the finest level given matrix-free (a StencilOperator) --------------
// An operator without stored entries has no off-diagonal product, so its
// smoother is the reference's Richardson iteration with the Jacobi
// preconditioner (linear_solver/richardson.h:136-251, SolveImplFullResidual):
//   r = b; A.VectorMultAdd(-1, x, 1, r);  z = r .* invd;  x = x + w z
// and its residual is that VectorMultAdd (gmg.h:1110-1119).  The products
// have the stencil's VectorMultAdd arithmetic: each row starts from
// b * 1 + 0 and adds (-1 f_c) (E x)_c cell by cell.

// x = w (b .* invd): the sweep from x = 0
__global__ void __launch_bounds__(kBlock)
    mg_st_first_kernel(int n, double w, const double *__restrict__ invd,
                       const double *__restrict__ b, double *__restrict__ x, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    x[i] = w * (b[i] * invd[i]);
}

// (every thread of the workgroup calls it: x is staged through lds)
template <class SOp>
__device__ __forceinline__ double st_residual(const SOp &A, long long unit,
                                              const double *__restrict__ b,
                                              const double *__restrict__ x, TileRow &tr,
                                              double *lds) {
#pragma clang fp contract(off)
  const long long row = A.thread_row(unit);
  const double s = row < A.nrow ? b[row] * 1.0 + 0.0 : 0.0;
  return A.template row_sum_scaled<true>(unit, x, s, -1.0, tr, lds);
}

// xo = xi + w ((b - A xi) .* invd)
template <class SOp>
__global__ void __launch_bounds__(kBlock)
    mg_st_sweep_kernel(SOp A, double w, const double *__restrict__ invd,
                       const double *__restrict__ b, const double *__restrict__ xi,
                       double *__restrict__ xo, const int *status) {
#pragma clang fp contract(off)
  __shared__ double lds[SOp::kLds];
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const double r = st_residual(A, unit, b, xi, tr, lds);
  if (tr.has_row) xo[tr.row] = xi[tr.row] + w * (r * invd[tr.row]);
}

// r = b - A x (PART: and the partials of r.r)
template <class SOp, bool PART>
__global__ void __launch_bounds__(kBlock)
    mg_st_residual_kernel(SOp A, const double *__restrict__ b, const double *__restrict__ x,
                          double *__restrict__ r, double *__restrict__ partials,
                          const int *status) {
#pragma clang fp contract(off)
  __shared__ double lds[SOp::kLds];
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const double ri = st_residual(A, unit, b, x, tr, lds);
  double rr = 0;
  if (tr.has_row) {
    r[tr.row] = ri;
    rr = ri * ri;
  }
  if (PART) block_partial1(rr, partials);
}

// ---- CG with the cycle as preconditioner (CG::SolvePrecond, cg.h:165-246) --
// init: r = b - (A x) (VectorMult, then ScaleAdd(b, -1)), z = cycle(r), then
// the r.r and r.z partials
template <class AOp>
__global__ void __launch_bounds__(kBlock)
    mgcg_residual_kernel(AOp A, const double *__restrict__ b, const double *__restrict__ x,
                         double *__restrict__ r) {
#pragma clang fp contract(off)
  __shared__ double lds[AOp::kLds];
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const double s = A.row_sum(unit, x, 0.0, tr, lds);
  if (tr.has_row) r[tr.row] = b[tr.row] - s;
}

__global__ void __launch_bounds__(kBlock)
    mgcg_dots_kernel(int n, const double *__restrict__ r, const double *__restrict__ z,
                     double *__restrict__ partials, const int *status) {
  if (stopped(status)) return;
  double rr = 0, rz = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    rr += r[i] * r[i];
    rz += r[i] * z[i];
  }
  block_partial2(rr, rz, partials);
}

__global__ void __launch_bounds__(kFinalBlock)
    mgcg_init_final(CgState<double> *st, const double *__restrict__ partials, int count) {
  __shared__ double lds[kFinalBlock / kWave];
  const double rr = sum_partials(partials, count, 2, 0, lds);
  const double rz = sum_partials(partials, count, 2, 1, lds);
  if (threadIdx.x == 0) {
    const double res0 = sqrt(rr);
    st->rr = rr;
    st->rz = rz;
    st->res0 = res0;
    st->res = res0;
    st->iter = 0;
    st->status = check_state(0, res0, res0, st->maxit, st->atol, st->rtol, st->dtol);
  }
}

// Ap = A p, partials of p.Ap
template <class AOp>
__global__ void __launch_bounds__(kBlock)
    mgcg_spmv_kernel(AOp A, const double *__restrict__ p, double *__restrict__ Ap,
                     double *__restrict__ partials, const int *status) {
  __shared__ double lds[AOp::kLds];
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const double s = A.row_sum(unit, p, 0.0, tr, lds);
  double part = 0;
  if (tr.has_row) {
    Ap[tr.row] = s;
    part = s * p[tr.row];
  }
  block_partial1(part, partials);
}

__global__ void __launch_bounds__(kFinalBlock)
    mgcg_alpha_final(CgState<double> *st, const double *__restrict__ partials, int count) {
  __shared__ double lds[kFinalBlock / kWave];
  if (st->status != kIterate) return;
  const double pAp = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) {
    st->pAp = pAp;
    st->alpha = st->rz / pAp;
    if (pAp == 0.0 || !(pAp == pAp)) st->status = kBreakdown;
  }
}

// x += alpha p; r -= alpha Ap; partials of r.r
__global__ void __launch_bounds__(kBlock)
    mgcg_update_kernel(int n, const double *__restrict__ p, const double *__restrict__ Ap,
                       double *__restrict__ x, double *__restrict__ r, const CgState<double> *st,
                       double *__restrict__ partials) {
  if (st->status != kIterate) return;
  const double alpha = st->alpha, malpha = -alpha;
  double rr = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] + malpha * Ap[i];
    r[i] = ri;
    rr += ri * ri;
  }
  block_partial1(rr, partials);
}

// res = sqrt(r.r), IterateControl::Check
__global__ void __launch_bounds__(kFinalBlock)
    mgcg_check_final(CgState<double> *st, const double *__restrict__ partials, int count) {
  __shared__ double lds[kFinalBlock / kWave];
  if (st->status != kIterate) return;
  const double rr = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) {
    const int iter = st->iter + 1;
    const double res = sqrt(rr);
    st->iter = iter;
    st->rr = rr;
    st->res = res;
    st->status = check_state(iter, res, st->res0, st->maxit, st->atol, st->rtol, st->dtol);
  }
}

// partials of r.z
__global__ void __launch_bounds__(kBlock)
    mgcg_rz_kernel(int n, const double *__restrict__ r, const double *__restrict__ z,
                   const CgState<double> *st, double *__restrict__ partials) {
  if (st->status != kIterate) return;
  double rz = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    rz += r[i] * z[i];
  block_partial1(rz, partials);
}

// beta = gamma / beta
__global__ void __launch_bounds__(kFinalBlock)
    mgcg_beta_final(CgState<double> *st, const double *__restrict__ partials, int count) {
  __shared__ double lds[kFinalBlock / kWave];
  if (st->status != kIterate) return;
  const double rz = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) {
    st->beta = rz / st->rz;
    st->rz = rz;
  }
}

// p = beta p + z (init: p = z)
__global__ void __launch_bounds__(kBlock)
    mgcg_p_kernel(int n, const double *__restrict__ z, double *__restrict__ p,
                  const CgState<double> *st, int first) {
  if (st->status != kIterate) return;
  const double beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = first ? z[i] : beta * p[i] + z[i];
}

// ---- the multigrid solver: r = b - A x, check ||r||, x += cycle(r) ---------
__global__ void __launch_bounds__(kBlock)
    mgs_residual_kernel(Op A, const double *__restrict__ b, const double *__restrict__ x,
                        double *__restrict__ r, double *__restrict__ partials, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  const long long row = A.thread_row(unit);
  const double b0 = row < A.nrow ? b[row] : 0.0;
  TileRow tr;
  const double ri = A.row_sum<true>(unit, x, b0, tr, nullptr);
  double rr = 0;
  if (tr.has_row) {
    r[tr.row] = ri;
    rr = ri * ri;
  }
  block_partial1(rr, partials);
}

__global__ void __launch_bounds__(kFinalBlock)
    mgs_check_final(CgState<double> *st, const double *__restrict__ partials, int count,
                    int first) {
  __shared__ double lds[kFinalBlock / kWave];
  if (st->status != kIterate) return;
  const double rr = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) {
    const double res = sqrt(rr);
    if (first) {
      st->res0 = res;
      st->iter = 0;
    } else {
      st->iter = st->iter + 1;
    }
    st->res = res;
    st->status = check_state(st->iter, res, st->res0, st->maxit, st->atol, st->rtol, st->dtol);
  }
}

__global__ void __launch_bounds__(kBlock)
    mgs_add_kernel(int n, const double *__restrict__ z, double *__restrict__ x, const int *status) {
  if (stopped(status)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] += z[i];
}

}  // namespace

// ---- transfers of the structured Q1 hierarchy without stored entries -------
// For the hierarchy of halved structured Q1 meshes (nodes lexicographic, x
// fastest), R and P are the Q1 interpolation with Dirichlet rows and
// columns removed.  Their entries are products of 1 and 1/2, exact in any
// order, so summing the same terms in the CSR rows' order (ascending
// column index: z, then y, then x) gives results bit-identical to the
// SELL-64 products over R and P, while reading only the vectors.

// bc = R r: coarse node (I,J,K) interior sums its 3^dim fine neighbours.
// Node counts stay below 2^31 (checked when the hierarchy is set), so the
// index arithmetic is 32-bit with multiply-shift division by pf and pc.
template <int DIM>
__global__ void __launch_bounds__(kBlock)
    mg_restrict_q1_kernel(int pf, int pc, DivMagic dpc, const double *__restrict__ r,
                          double *__restrict__ bc, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  const int nc = DIM == 3 ? pc * pc * pc : pc * pc;
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
    const int t = dpc.div(c);
    const int I = c - t * pc;
    const int K = DIM == 3 ? dpc.div(t) : 0;
    const int J = t - K * pc;
    const bool bnd = I == 0 || I == pc - 1 || J == 0 || J == pc - 1 ||
                     (DIM == 3 && (K == 0 || K == pc - 1));
    double s = 0;
    if (!bnd) {
      const int f0 = 2 * I + pf * (2 * J + pf * 2 * K);
#pragma unroll
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            const double w = (dz ? 0.5 : 1.0) * ((dy ? 0.5 : 1.0) * (dx ? 0.5 : 1.0));
            s += w * r[f0 + dx + pf * (dy + pf * dz)];
          }
    }
    bc[c] = s;
  }
}

// x = x + P xc: fine node (i,j,k) interior adds its interior coarse parents
template <int DIM>
__global__ void __launch_bounds__(kBlock)
    mg_prolong_q1_kernel(int pf, int pc, DivMagic dpf, const double *__restrict__ xc,
                         double *__restrict__ x, const int *status) {
#pragma clang fp contract(off)
  if (stopped(status)) return;
  const int nf = DIM == 3 ? pf * pf * pf : pf * pf;
  for (int f = blockIdx.x * kBlock + threadIdx.x; f < nf; f += gridDim.x * kBlock) {
    const int t = dpf.div(f);
    const int i = f - t * pf;
    const int k = DIM == 3 ? dpf.div(t) : 0;
    const int j = t - k * pf;
    if (i == 0 || i == pf - 1 || j == 0 || j == pf - 1 || (DIM == 3 && (k == 0 || k == pf - 1)))
      continue;
    // parents per direction: g even -> g/2 (weight 1), odd -> (g-1)/2 and
    // (g+1)/2 (1/2 each); parents on the coarse boundary have no column
    const int nzp = DIM == 3 ? 1 + (k & 1) : 1, nyp = 1 + (j & 1), nxp = 1 + (i & 1);
    const int kz0 = DIM == 3 ? k >> 1 : 0, jy0 = j >> 1, ix0 = i >> 1;
    const double wz = DIM == 3 && (k & 1) ? 0.5 : 1.0, wy = (j & 1) ? 0.5 : 1.0,
                 wx = (i & 1) ? 0.5 : 1.0;
    double s = x[f];
    for (int a = 0; a < nzp; ++a) {
      const int K = kz0 + a;
      if (DIM == 3 && (K == 0 || K == pc - 1)) continue;
      for (int bq = 0; bq < nyp; ++bq) {
        const int J = jy0 + bq;
        if (J == 0 || J == pc - 1) continue;
        for (int cq = 0; cq < nxp; ++cq) {
          const int I = ix0 + cq;
          if (I == 0 || I == pc - 1) continue;
          const double w = (DIM == 3 ? wz : 1.0) * (wy * wx);
          s += w * xc[I + pc * (J + pc * K)];
        }
      }
    }
    x[f] = s;
  }
}

// ---- a dense inverse on the coarsest level ---------------------------------
// x = C b with C stored by columns (ct[j n + i] = C_ij).  The columns are
// cut into kDenseParts blocks of ceil(n / kDenseParts); wave k of a
// workgroup sums block k of 64 rows from 0 in ascending j (lane = row, so
// its loads are contiguous), and the row's sum is p_0 + p_1 + ... + p_7 in
// that order.  Eight waves per 64 rows keep enough loads in flight for a
// matrix read once per cycle.
constexpr int kDenseParts = 8;
constexpr int kDenseBlock = kDenseParts * kWave;
constexpr int kDenseMaxRows = 4096;

__global__ void __launch_bounds__(kDenseBlock)
    mg_dense_kernel(int n, const double *__restrict__ ct, const double *__restrict__ b,
                    double *__restrict__ x, const int *status) {
#pragma clang fp contract(off)
  __shared__ double part[kDenseParts][kWave];
  if (stopped(status)) return;
  const int lane = threadIdx.x & (kWave - 1), k = threadIdx.x / kWave;
  const int i = blockIdx.x * kWave + lane;
  const int cb = (n + kDenseParts - 1) / kDenseParts;
  const int j0 = k * cb, j1 = j0 + cb < n ? j0 + cb : n;
  double s = 0;
  if (i < n) {
#pragma unroll 8
    for (int j = j0; j < j1; ++j) s += ct[(size_t)j * n + i] * b[j];
  }
  part[k][lane] = s;
  __syncthreads();
  if (k == 0 && i < n) {
    double t = part[0][lane];
#pragma unroll
    for (int q = 1; q < kDenseParts; ++q) t = t + part[q][lane];
    x[i] = t;
  }
}

__global__ void __launch_bounds__(kBlock)
    mg_transpose_kernel(int n, const double *__restrict__ a, double *__restrict__ at) {
  const long long nn = (long long)n * n;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < nn;
       k += (long long)gridDim.x * kBlock) {
    const long long i = k / n, j = k - i * n;
    at[j * n + i] = a[k];
  }
}

// ---- the small coarse levels in one workgroup ------------------------------
// Levels of at most kTailMaxRows rows take microseconds of latency per
// launch and a few per kernel for little work, so the cycle from the first
// such level down to the coarsest and back runs as one launch of one
// 1024-thread workgroup, the phases separated by barriers (all its memory
// traffic stays within the workgroup).  Each phase computes exactly what
// the per-level kernels above compute.
constexpr int kTailBlock = 1024;
constexpr int kTailMaxRows = 2048;

struct MgLevelDev {
  Op A, R, P;
  const double *invd;
  double *b, *x, *x2, *r;
  int n;
};

// SellOp views in units of 256 rows: with 1024 threads, unit 4k covers rows
// 1024k .. 1024k + 1023, thread t taking row 1024k + t.
__device__ __forceinline__ void tail_first(const MgLevelDev &L, double w, const double *b,
                                           double *x) {
#pragma clang fp contract(off)
  for (int i = threadIdx.x; i < L.n; i += kTailBlock) x[i] = w * L.invd[i] * b[i];
}

__device__ __forceinline__ void tail_sweep(const MgLevelDev &L, double w, double omw,
                                           const double *b, const double *xi, double *xo) {
#pragma clang fp contract(off)
  for (int k = 0; k * kTailBlock < L.n; ++k) {
    const long long unit = 4LL * k;
    const long long r = L.A.thread_row(unit);
    const double b0 = r < L.n ? b[r] : 0.0;
    TileRow tr;
    const double t = L.A.row_sum<true, true>(unit, xi, b0, tr, nullptr);
    if (tr.has_row) xo[tr.row] = (w * L.invd[tr.row] * t) + (omw * xi[tr.row]);
  }
}

__device__ __forceinline__ void tail_residual(const MgLevelDev &L, const double *b,
                                              const double *x) {
#pragma clang fp contract(off)
  for (int k = 0; k * kTailBlock < L.n; ++k) {
    const long long unit = 4LL * k;
    const long long r = L.A.thread_row(unit);
    const double b0 = r < L.n ? b[r] : 0.0;
    TileRow tr;
    const double s = L.A.row_sum<true>(unit, x, b0, tr, nullptr);
    if (tr.has_row) L.r[tr.row] = s;
  }
}

template <bool ADD>
__device__ __forceinline__ void tail_apply(const Op &M, int nrow, const double *x, double *y) {
  for (int k = 0; k * kTailBlock < nrow; ++k) {
    const long long unit = 4LL * k;
    double s = 0;
    if (ADD) {
      const long long r = M.thread_row(unit);
      if (r < nrow) s = y[r];
    }
    TileRow tr;
    s = M.row_sum(unit, x, s, tr, nullptr);
    if (tr.has_row) y[tr.row] = s;
  }
}

// The cycle from level t (right-hand side bin) to the coarsest level L-1.
// The result of level l ends in x2 when its number of sweeps after the
// first step is odd, else in x (the host uses the same rule).
__global__ void __launch_bounds__(kTailBlock)
    mg_tail_kernel(const MgLevelDev *__restrict__ lv, int t, int L, const double *bin, int pre,
                   int post, int coarse, double w, double wc, const int *status) {
  if (stopped(status)) return;
  const double omw = 1.0 - w, omwc = 1.0 - wc;
  for (int l = t; l < L - 1; ++l) {
    const MgLevelDev Lv = lv[l];
    const double *b = l == t ? bin : Lv.b;
    double *x = Lv.x, *y = Lv.x2;
    tail_first(Lv, w, b, x);
    __syncthreads();
    for (int k = 1; k < pre; ++k) {
      tail_sweep(Lv, w, omw, b, x, y);
      __syncthreads();
      double *tmp = x;
      x = y;
      y = tmp;
    }
    tail_residual(Lv, b, x);
    __syncthreads();
    tail_apply<false>(Lv.R, lv[l + 1].n, Lv.r, lv[l + 1].b);
    __syncthreads();
  }
  {
    const MgLevelDev Lv = lv[L - 1];
    const double *b = L - 1 == t ? bin : Lv.b;
    double *x = Lv.x, *y = Lv.x2;
    tail_first(Lv, wc, b, x);
    __syncthreads();
    for (int k = 1; k < coarse; ++k) {
      tail_sweep(Lv, wc, omwc, b, x, y);
      __syncthreads();
      double *tmp = x;
      x = y;
      y = tmp;
    }
  }
  for (int l = L - 2; l >= t; --l) {
    const MgLevelDev Lv = lv[l];
    const MgLevelDev Cv = lv[l + 1];
    const int csw = l + 1 == L - 1 ? coarse - 1 : (pre - 1) + post;
    const double *xc = (csw & 1) ? Cv.x2 : Cv.x;
    const double *b = l == t ? bin : Lv.b;
    double *x = ((pre - 1) & 1) ? Lv.x2 : Lv.x;
    double *y = ((pre - 1) & 1) ? Lv.x : Lv.x2;
    tail_apply<true>(Lv.P, Lv.n, xc, x);
    __syncthreads();
    for (int k = 0; k < post; ++k) {
      tail_sweep(Lv, w, omw, b, x, y);
      __syncthreads();
      double *tmp = x;
      x = y;
      y = tmp;
    }
  }
}

struct hflmp_mg {
  int device = 0;
  int nlevels = 0;
  hflmp_mg_params prm;
  const hflmp_stencil *st = nullptr;  // level-0 operator when matrix-free
  struct Level {
    int n = 0;
    hflmp_sell *A = nullptr, *R = nullptr, *P = nullptr;
    double *invd = nullptr, *b = nullptr, *x = nullptr, *x2 = nullptr, *r = nullptr;
  } lv[kMaxLevels];
  int grid0 = 0;  // workgroups of a level-0 operator launch
  int tail = 0;   // first level of the one-workgroup tail (nlevels: none)
  int tail0 = 0;  // the tail chosen at creation
  MgLevelDev *dlv = nullptr;  // device copy of the level table
  double *cinv = nullptr;     // coarsest level's inverse by columns, or null
  int sdim = 0;               // structured Q1 transfers: dimension, or 0
  int scells[kMaxLevels] = {};  // cells per direction of each level
};

namespace {

void mg_free(hflmp_mg *m) {
  if (!m) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != m->device) (void)hipSetDevice(m->device);
  for (int l = 0; l < m->nlevels; ++l) {
    auto &L = m->lv[l];
    hflmp_sell_destroy(L.A);
    hflmp_sell_destroy(L.R);
    hflmp_sell_destroy(L.P);
    (void)hipFree(L.invd);
    (void)hipFree(L.b);
    (void)hipFree(L.x);
    (void)hipFree(L.x2);
    (void)hipFree(L.r);
  }
  (void)hipFree(m->dlv);
  (void)hipFree(m->cinv);
  if (cur >= 0 && cur != m->device) (void)hipSetDevice(cur);
  delete m;
}

// Calls fn(op) with the level-0 operator view: the stencil's or SELL-64.
template <typename Fn>
void with_fine_op(const hflmp_mg *m, Fn &&fn) {
  if (m->st)
    (void)stencil_dispatch<double>(m->st, fn);
  else
    fn(sell_view<double, int>(m->lv[0].A));
}

// The level's smoothing steps and residual.
void smooth_first(hipStream_t s, const hflmp_mg *m, int l, double w, const double *b, double *x,
                  const int *status) {
  const auto &L = m->lv[l];
  const unsigned gv = grid_for(L.n, 4);
  if (l == 0 && m->st)
    mg_st_first_kernel<<<gv, kBlock, 0, s>>>(L.n, w, L.invd, b, x, status);
  else
    mg_first_kernel<<<gv, kBlock, 0, s>>>(L.n, w, L.invd, b, x, status);
}

void smooth_sweep(hipStream_t s, const hflmp_mg *m, int l, double w, const double *b,
                  const double *xi, double *xo, const int *status) {
  const auto &L = m->lv[l];
  const unsigned ga = sell_grid(L.n);
  if (l == 0 && m->st) {
    (void)stencil_dispatch<double>(m->st, [&](auto op) {
      mg_st_sweep_kernel<<<ga, kBlock, 0, s>>>(op, w, L.invd, b, xi, xo, status);
    });
  } else {
    mg_sweep_kernel<<<ga, kBlock, 0, s>>>(sell_view<double, int>(L.A), w, 1.0 - w, L.invd, b,
                                          xi, xo, status);
  }
}

void residual(hipStream_t s, const hflmp_mg *m, int l, const double *b, const double *x,
              double *r, const int *status) {
  const auto &L = m->lv[l];
  const unsigned ga = sell_grid(L.n);
  if (l == 0 && m->st) {
    (void)stencil_dispatch<double>(m->st, [&](auto op) {
      mg_st_residual_kernel<decltype(op), false>
          <<<ga, kBlock, 0, s>>>(op, b, x, r, nullptr, status);
    });
  } else {
    mg_residual_kernel<<<ga, kBlock, 0, s>>>(sell_view<double, int>(L.A), b, x, r, status);
  }
}

// One cycle from level l with right-hand side b; returns the buffer that
// holds the result.  `out` (level 0) is the buffer the result must end in:
// the ping-pong pair starts so that the last write lands there.  Every
// launch returns at once when *status is set and not "iterate".
double *cycle(hipStream_t s, const hflmp_mg *m, int l, const double *b, double *out,
              const int *status) {
  const auto &L = m->lv[l];
  const bool coarsest = l == m->nlevels - 1;
  const double w = coarsest ? m->prm.coarse_omega : m->prm.omega;
  const unsigned ga = sell_grid(L.n);
  const bool dense = coarsest && m->cinv;
  const int first = coarsest ? m->prm.coarse_sweeps : m->prm.pre_smooth;
  const int swaps = dense ? 0 : (first - 1) + (coarsest ? 0 : m->prm.post_smooth);
  if (l > 0 && l == m->tail) {
    mg_tail_kernel<<<1, kTailBlock, 0, s>>>(m->dlv, l, m->nlevels, b, m->prm.pre_smooth,
                                            m->prm.post_smooth, m->prm.coarse_sweeps,
                                            m->prm.omega, m->prm.coarse_omega, status);
    return (swaps & 1) ? L.x2 : L.x;
  }
  double *x = L.x, *y = L.x2;
  if (out) {
    x = (swaps & 1) ? L.x2 : out;
    y = (swaps & 1) ? out : L.x2;
  }
  if (dense) {
    mg_dense_kernel<<<(L.n + kWave - 1) / kWave, kDenseBlock, 0, s>>>(L.n, m->cinv, b, x,
                                                                       status);
    return x;
  }
  smooth_first(s, m, l, w, b, x, status);
  for (int k = 1; k < first; ++k) {
    smooth_sweep(s, m, l, w, b, x, y, status);
    std::swap(x, y);
  }
  if (coarsest) return x;
  const auto &C = m->lv[l + 1];
  residual(s, m, l, b, x, L.r, status);
  const int pf = m->scells[l] + 1, pc = m->scells[l + 1] + 1;
  DivMagic dpf, dpc;
  dpf.init((unsigned)pf);
  dpc.init((unsigned)pc);
  if (m->sdim == 3)
    mg_restrict_q1_kernel<3><<<grid_for(C.n, 1), kBlock, 0, s>>>(pf, pc, dpc, L.r, C.b, status);
  else if (m->sdim == 2)
    mg_restrict_q1_kernel<2><<<grid_for(C.n, 1), kBlock, 0, s>>>(pf, pc, dpc, L.r, C.b, status);
  else
    mg_apply_kernel<false><<<sell_grid(C.n), kBlock, 0, s>>>(sell_view<double, int>(L.R), L.r,
                                                              C.b, status);
  const double *xc = cycle(s, m, l + 1, C.b, nullptr, status);
  if (m->sdim == 3)
    mg_prolong_q1_kernel<3><<<grid_for(L.n, 1), kBlock, 0, s>>>(pf, pc, dpf, xc, x, status);
  else if (m->sdim == 2)
    mg_prolong_q1_kernel<2><<<grid_for(L.n, 1), kBlock, 0, s>>>(pf, pc, dpf, xc, x, status);
  else
    mg_apply_kernel<true><<<ga, kBlock, 0, s>>>(sell_view<double, int>(L.P), xc, x, status);
  for (int k = 0; k < m->prm.post_smooth; ++k) {
    smooth_sweep(s, m, l, w, b, x, y, status);
    std::swap(x, y);
  }
  return x;
}

// z = M r: one cycle from level 0 (z must not alias r)
void apply(hipStream_t s, const hflmp_mg *m, const double *r, double *z, const int *status) {
  (void)cycle(s, m, 0, r, z, status);
}

int mg_build(hflmp_ctx *ctx, const hflmp_stencil *S, int nlevels, const hflmp_csr_f64 *A,
             const hflmp_csr_f64 *R, const hflmp_csr_f64 *P, const hflmp_mg_params *prm,
             hflmp_mg **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, A && out && prm && (nlevels == 1 || (R && P)), "mg_create: null pointer");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, nlevels >= 1 && nlevels <= kMaxLevels, "mg_create: 1 <= nlevels <= %d",
                kMaxLevels);
  HFLMP_REQUIRE(ctx, prm->pre_smooth >= 1 && prm->post_smooth >= 0 && prm->coarse_sweeps >= 1 &&
                         prm->omega > 0 && prm->coarse_omega > 0,
                "mg_create: need pre_smooth >= 1, post_smooth >= 0, coarse_sweeps >= 1, "
                "omega > 0, coarse_omega > 0");
  if (S) {
    HFLMP_REQUIRE(ctx, S->value_bytes == 8 && S->device == ctx->device,
                  "mg_create_stencil: the operator must be FP64 and on this device");
    HFLMP_REQUIRE(ctx, S->nnode < (1LL << 31), "mg_create_stencil: too many nodes");
  }
  auto rows = [&](int l) -> long long { return (l == 0 && S) ? S->nnode : A[l].nrow; };
  for (int l = 0; l < nlevels; ++l) {
    if (!(l == 0 && S))
      HFLMP_REQUIRE(ctx, A[l].nrow > 0 && A[l].ncol == A[l].nrow && A[l].row_ptr,
                    "mg_create: operator %d must be square and non-empty", l);
    if (l + 1 < nlevels) {
      HFLMP_REQUIRE(ctx, R[l].nrow == rows(l + 1) && R[l].ncol == rows(l) &&
                             P[l].nrow == rows(l) && P[l].ncol == rows(l + 1),
                    "mg_create: R[%d] must be n%d x n%d and P[%d] n%d x n%d", l, l + 1, l, l,
                    l, l + 1);
    }
  }
  auto *m = new hflmp_mg();
  m->device = ctx->device;
  m->nlevels = nlevels;
  m->prm = *prm;
  m->st = S;
  int rc = HFLMP_OK;
  auto sell_of = [&](const hflmp_csr_f64 &c, hflmp_sell **o) {
    return hflmp_sell_create_f64(ctx, c.nrow, c.ncol, c.nnz, c.row_ptr, c.col, c.val,
                                 HFLMP_SELL_ANY_PADDING, o);
  };
  for (int l = 0; l < nlevels && rc == HFLMP_OK; ++l) {
    auto &L = m->lv[l];
    const bool mf = l == 0 && S;
    L.n = (int)rows(l);
    const size_t bytes = sizeof(double) * (size_t)L.n;
    if (!mf && (rc = sell_of(A[l], &L.A)) != HFLMP_OK) break;
    if (l + 1 < nlevels && ((rc = sell_of(R[l], &L.R)) != HFLMP_OK ||
                            (rc = sell_of(P[l], &L.P)) != HFLMP_OK))
      break;
    if (hipMalloc(&L.invd, bytes) != hipSuccess || hipMalloc(&L.b, bytes) != hipSuccess ||
        hipMalloc(&L.x, bytes) != hipSuccess || hipMalloc(&L.x2, bytes) != hipSuccess ||
        hipMalloc(&L.r, bytes) != hipSuccess) {
      rc = set_error(ctx, HFLMP_ENOMEM, "mg_create: out of device memory");
      break;
    }
    int bad = 0;
    rc = mf ? hflmp_stencil_extract_invdiag_f64(ctx, S, L.invd, &bad)
            : hflmp_csr_extract_invdiag_f64(ctx, L.n, A[l].row_ptr, A[l].col, A[l].val, L.invd,
                                            &bad);
    if (rc != HFLMP_OK) break;
    if (bad) {
      rc = set_error(ctx, HFLMP_ENUMERIC, "mg_create: %d rows of operator %d have no diagonal",
                     bad, l);
      break;
    }
  }
  if (rc != HFLMP_OK) {
    mg_free(m);
    return rc;
  }
  m->grid0 = sell_grid(m->lv[0].n);
  // the tail: the longest run of small levels ending at the coarsest,
  // never level 0 (it may be matrix-free and ends in the caller's buffer)
  m->tail = nlevels;
  while (m->tail > 1 && m->lv[m->tail - 1].n <= kTailMaxRows) --m->tail;
  if (m->tail < nlevels) {
    MgLevelDev h[kMaxLevels];
    for (int l = m->tail; l < nlevels; ++l) {
      const auto &L = m->lv[l];
      MgLevelDev d = {};
      d.A = sell_view<double, int>(L.A);
      if (l + 1 < nlevels) {
        d.R = sell_view<double, int>(L.R);
        d.P = sell_view<double, int>(L.P);
      }
      d.invd = L.invd;
      d.b = L.b;
      d.x = L.x;
      d.x2 = L.x2;
      d.r = L.r;
      d.n = L.n;
      h[l] = d;
    }
    if (hipMalloc(&m->dlv, sizeof(MgLevelDev) * nlevels) != hipSuccess) {
      mg_free(m);
      return set_error(ctx, HFLMP_ENOMEM, "mg_create: out of device memory");
    }
    if (hipMemcpy(m->dlv + m->tail, h + m->tail, sizeof(MgLevelDev) * (nlevels - m->tail),
                  hipMemcpyHostToDevice) != hipSuccess) {
      mg_free(m);
      return set_error(ctx, HFLMP_EHIP, "mg_create: HIP error");
    }
  }
  m->tail0 = m->tail;
  *out = m;
  return HFLMP_OK;
}

struct MgWork {
  CgState<double> *st;
  double *partials, *v[4];
};

MgWork mg_work(const hflmp_mg *m, void *work) {
  const size_t n = (size_t)m->lv[0].n;
  const size_t g = (size_t)(m->grid0 > kMaxRedBlocks ? m->grid0 : kMaxRedBlocks);
  char *w = static_cast<char *>(work);
  MgWork k;
  k.st = reinterpret_cast<CgState<double> *>(w);
  k.partials = reinterpret_cast<double *>(w + 256);
  const size_t vec0 = 256 + align256(2 * g * sizeof(double)), vs = align256(n * sizeof(double));
  for (int i = 0; i < 4; ++i) k.v[i] = reinterpret_cast<double *>(w + vec0 + i * vs);
  return k;
}

int finish(hflmp_ctx *ctx, CgState<double> *st, std::chrono::steady_clock::time_point t0,
           hflmp_solve_info *info) {
  CgState<double> h;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  info->iters = h.iter;
  info->status = h.status == kIterate ? kMaxIts : h.status;
  info->res_init = h.res0;
  info->res_final = h.res;
  info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return HFLMP_OK;
}

int check_solve(hflmp_ctx *ctx, const hflmp_mg *m, const double *b, double *x, int maxit,
                double atol, double rtol, double dtol, void *work, size_t work_bytes,
                hflmp_solve_info *info, const char *what);

}  // namespace

extern "C" {

int hflmp_mg_create_f64(hflmp_ctx *ctx, int nlevels, const hflmp_csr_f64 *A,
                        const hflmp_csr_f64 *R, const hflmp_csr_f64 *P,
                        const hflmp_mg_params *prm, hflmp_mg **out) {
  return mg_build(ctx, nullptr, nlevels, A, R, P, prm, out);
}

int hflmp_mg_create_stencil_f64(hflmp_ctx *ctx, const hflmp_stencil *S, int nlevels,
                                const hflmp_csr_f64 *A, const hflmp_csr_f64 *R,
                                const hflmp_csr_f64 *P, const hflmp_mg_params *prm,
                                hflmp_mg **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, S != nullptr, "mg_create_stencil: null operator");
  return mg_build(ctx, S, nlevels, A, R, P, prm, out);
}

void hflmp_mg_destroy(hflmp_mg *m) { mg_free(m); }

int hflmp_mg_set_structured_q1(hflmp_ctx *ctx, hflmp_mg *m, int dim, int ncells) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m, "mg_set_structured_q1: null hierarchy");
  if (dim == 0) {
    m->sdim = 0;
    return HFLMP_OK;
  }
  HFLMP_REQUIRE(ctx, dim == 2 || dim == 3, "mg_set_structured_q1: dim must be 2 or 3");
  int nc = ncells;
  for (int l = 0; l < m->nlevels; ++l) {
    HFLMP_REQUIRE(ctx, nc >= 1 && (l + 1 == m->nlevels || nc % 2 == 0),
                  "mg_set_structured_q1: level %d has %d cells per direction", l, nc);
    long long nodes = 1;
    for (int d = 0; d < dim; ++d) nodes *= (long long)(nc + 1);
    HFLMP_REQUIRE(ctx, nodes == m->lv[l].n,
                  "mg_set_structured_q1: level %d has %d rows, not (%d+1)^%d", l, m->lv[l].n, nc,
                  dim);
    m->scells[l] = nc;
    nc /= 2;
  }
  m->sdim = dim;
  return HFLMP_OK;
}

int hflmp_mg_set_coarse_inverse_f64(hflmp_ctx *ctx, hflmp_mg *m, const double *inv) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m, "mg_set_coarse_inverse: null hierarchy");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "mg_set_coarse_inverse: hierarchy on another "
                                               "device");
  const int n = m->lv[m->nlevels - 1].n;
  if (!inv) {
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(m->cinv);
    m->cinv = nullptr;
    m->tail = m->tail0;
    return HFLMP_OK;
  }
  HFLMP_REQUIRE(ctx, n <= kDenseMaxRows, "mg_set_coarse_inverse: coarsest level has %d rows "
                                         "(at most %d)", n, kDenseMaxRows);
  if (!m->cinv &&
      hipMalloc(&m->cinv, sizeof(double) * (size_t)n * (size_t)n) != hipSuccess) {
    m->cinv = nullptr;
    return set_error(ctx, HFLMP_ENOMEM, "mg_set_coarse_inverse: out of device memory");
  }
  mg_transpose_kernel<<<grid_for((long long)n * n, 4), kBlock, 0, ctx->stream>>>(n, inv,
                                                                                m->cinv);
  HFLMP_LAUNCH_CHECK(ctx);
  // the coarsest level is no longer swept: the one-workgroup tail stops
  // above it
  m->tail = m->nlevels;
  return HFLMP_OK;
}

int hflmp_mg_apply_f64(hflmp_ctx *ctx, const hflmp_mg *m, const double *r, double *z) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && r && z, "mg_apply: null pointer");
  HFLMP_REQUIRE(ctx, r != z, "mg_apply: r and z must be different vectors");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "mg_apply: hierarchy on another device");
  apply(ctx->stream, m, r, z, nullptr);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

size_t hflmp_mg_workspace_bytes(const hflmp_mg *m) {
  if (!m) return 0;
  const size_t n = (size_t)m->lv[0].n;
  const size_t g = (size_t)(m->grid0 > kMaxRedBlocks ? m->grid0 : kMaxRedBlocks);
  return 256 + align256(2 * g * sizeof(double)) + 4 * align256(n * sizeof(double));
}

int hflmp_pcg_mg_f64(hflmp_ctx *ctx, const hflmp_mg *m, const double *b, double *x, int maxit,
                     double atol, double rtol, double dtol, void *work, size_t work_bytes,
                     hflmp_solve_info *info) {
  int rc = check_solve(ctx, m, b, x, maxit, atol, rtol, dtol, work, work_bytes, info, "pcg_mg");
  if (rc) return rc;
  const int n = m->lv[0].n;
  MgWork k = mg_work(m, work);
  CgState<double> *st = k.st;
  double *partials = k.partials, *r = k.v[0], *p = k.v[1], *z = k.v[2], *Ap = k.v[3];
  hipStream_t s = ctx->stream;
  const unsigned ga = (unsigned)m->grid0;
  const int gd = grid_for(n, 4) < kMaxRedBlocks ? (int)grid_for(n, 4) : kMaxRedBlocks;
  const int *status = &st->status;
  const auto t0 = std::chrono::steady_clock::now();
  init_state_kernel<double><<<1, kBlock, 0, s>>>(st, maxit, atol, rtol, dtol);
  with_fine_op(m, [&](auto op) { mgcg_residual_kernel<<<ga, kBlock, 0, s>>>(op, b, x, r); });
  apply(s, m, r, z, nullptr);
  mgcg_dots_kernel<<<gd, kBlock, 0, s>>>(n, r, z, partials, nullptr);
  mgcg_init_final<<<1, kFinalBlock, 0, s>>>(st, partials, gd);
  mgcg_p_kernel<<<gd, kBlock, 0, s>>>(n, z, p, st, 1);
  HFLMP_LAUNCH_CHECK(ctx);
  auto iteration = [&]() {
    with_fine_op(m, [&](auto op) {
      mgcg_spmv_kernel<<<ga, kBlock, 0, s>>>(op, p, Ap, partials, status);
    });
    mgcg_alpha_final<<<1, kFinalBlock, 0, s>>>(st, partials, (int)ga);
    mgcg_update_kernel<<<gd, kBlock, 0, s>>>(n, p, Ap, x, r, st, partials);
    mgcg_check_final<<<1, kFinalBlock, 0, s>>>(st, partials, gd);
    apply(s, m, r, z, status);
    mgcg_rz_kernel<<<gd, kBlock, 0, s>>>(n, r, z, st, partials);
    mgcg_beta_final<<<1, kFinalBlock, 0, s>>>(st, partials, gd);
    mgcg_p_kernel<<<gd, kBlock, 0, s>>>(n, z, p, st, 0);
  };
  rc = run_polled(ctx, maxit, &st->status, iteration, 2);
  if (rc) return rc;
  return finish(ctx, st, t0, info);
}

int hflmp_mg_solve_f64(hflmp_ctx *ctx, const hflmp_mg *m, const double *b, double *x, int maxit,
                       double atol, double rtol, double dtol, void *work, size_t work_bytes,
                       hflmp_solve_info *info) {
  int rc = check_solve(ctx, m, b, x, maxit, atol, rtol, dtol, work, work_bytes, info,
                       "mg_solve");
  if (rc) return rc;
  const int n = m->lv[0].n;
  MgWork k = mg_work(m, work);
  CgState<double> *st = k.st;
  double *partials = k.partials, *r = k.v[0], *z = k.v[1];
  hipStream_t s = ctx->stream;
  const unsigned ga = (unsigned)m->grid0;
  const int *status = &st->status;
  // r = b - A x with the partials of r.r: the level-0 residual of the cycle
  auto resid = [&](const int *stat) {
    if (m->st) {
      (void)stencil_dispatch<double>(m->st, [&](auto op) {
        mg_st_residual_kernel<decltype(op), true>
            <<<ga, kBlock, 0, s>>>(op, b, x, r, partials, stat);
      });
    } else {
      mgs_residual_kernel<<<ga, kBlock, 0, s>>>(sell_view<double, int>(m->lv[0].A), b, x, r,
                                                partials, stat);
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  init_state_kernel<double><<<1, kBlock, 0, s>>>(st, maxit, atol, rtol, dtol);
  resid(nullptr);
  mgs_check_final<<<1, kFinalBlock, 0, s>>>(st, partials, (int)ga, 1);
  HFLMP_LAUNCH_CHECK(ctx);
  // the reference's GMG::SolveImpl (gmg.h:979-1066): cycle, new residual, check
  auto iteration = [&]() {
    apply(s, m, r, z, status);
    mgs_add_kernel<<<grid_for(n, 4), kBlock, 0, s>>>(n, z, x, status);
    resid(status);
    mgs_check_final<<<1, kFinalBlock, 0, s>>>(st, partials, (int)ga, 0);
  };
  rc = run_polled(ctx, maxit, &st->status, iteration, 2);
  if (rc) return rc;
  return finish(ctx, st, t0, info);
}

}  // extern "C"

namespace {

int check_solve(hflmp_ctx *ctx, const hflmp_mg *m, const double *b, double *x, int maxit,
                double atol, double rtol, double dtol, void *work, size_t work_bytes,
                hflmp_solve_info *info, const char *what) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && b && x && work && info, "%s: null pointer", what);
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "%s: hierarchy on another device", what);
  HFLMP_REQUIRE(ctx, maxit > 0 && atol >= 0 && rtol >= 0 && dtol > 1,
                "%s: bad control parameters (maxit>0, atol>=0, rtol>=0, dtol>1)", what);
  HFLMP_REQUIRE(ctx, work_bytes >= hflmp_mg_workspace_bytes(m), "%s: workspace %zu < %zu", what,
                work_bytes, hflmp_mg_workspace_bytes(m));
  return HFLMP_OK;
}

}  // namespace
