// Device-resident Jacobi-preconditioned CG.
//
// Arithmetic of CG::SolvePrecond (linear_solver/cg.h:165-246):
//   r = b - A x;  z = D^-1 r;  p = z;  beta = r.z
//   loop: Ap = A p;  alpha = beta / (Ap.p);  x += alpha p;  r -= alpha Ap;
//         res = sqrt(r.r);  Check;  z = D^-1 r;  gamma = r.z;
//         p = (gamma/beta) p + z;  beta = gamma
// with the stopping rule of IterateControl::Check (common/iterate_control.cc:
// 80-107) on the recursive residual.  Every scalar stays on the device.
// One iteration is three launches (five for systems of a few hundred rows,
// whose partial area is too small for the folded form):
//   K1  SpMV Ap = A p, per-workgroup partials of p.Ap
//  (F1  one workgroup: p.Ap, alpha)
//   K2  x += alpha p, r -= alpha Ap, z = D^-1 r, partials of r.r and r.z
//  (F2  one workgroup: res, stopping test, beta)
//   K3  p = beta p + z
// In the three-launch form F1 runs at the start of every workgroup of K2
// and F2 at the start of every workgroup of K3.
// Every kernel returns at once when the state is no longer "iterate", so
// the host queues iterations ahead and reads the state every `poll`
// iterations only.
//
// The SpMV kernels take the operator as a value (CsrTileOp over the CSR
// arrays or SellOp over the SELL-64 layout, hflmp_sell.hpp, or the
// matrix-free StencilOp, hflmp_stencil.hpp); each gives every thread of a
// workgroup one row.  The CSR and SELL-64 products are identical.
#include <chrono>
#include <mutex>

#include "hflmp_cg3.hpp"
#include "hflmp_sell.hpp"
#include "hflmp_solver_common.hpp"
#include "hflmp_stencil.hpp"

namespace {

using namespace hflmp_solver;

// K1: Ap = A p, partials of p.Ap
template <typename Op, typename T = typename Op::value_type>
__global__ void __launch_bounds__(kBlock)
    cg_spmv_kernel(Op op, const T *__restrict__ p, T *__restrict__ Ap, const CgState<T> *st,
                   T *__restrict__ partials) {
  __shared__ T prod[Op::kLds];
  if (st->status != kIterate) return;
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const T s = op.row_sum(tile, p, T(0), tr, prod);
  T part = 0;
  if (tr.has_row) {
    __builtin_nontemporal_store(s, Ap + tr.row);
    part = s * p[tr.row];
  }
  block_partial1(part, partials);
}

// F1: alpha = (r.z) / (p.Ap)
template <typename T>
__global__ void __launch_bounds__(kFinalBlock)
    cg_alpha_final(CgState<T> *st, const T *__restrict__ partials, int count) {
  __shared__ T lds[kFinalBlock / kWave];
  if (st->status != kIterate) return;
  const T pAp = sum_partials(partials, count, 1, 0, lds);
  if (threadIdx.x == 0) {
    st->pAp = pAp;
    st->alpha = st->rz / pAp;
    if (pAp == T(0) || !(pAp == pAp)) st->status = kBreakdown;
  }
}

// K2: x += alpha p; r -= alpha Ap; z = D^-1 r; partials of r.r, r.z
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_update_kernel(int n, const T *__restrict__ invd, const T *__restrict__ p,
                     const T *__restrict__ Ap, T *__restrict__ x, T *__restrict__ r,
                     T *__restrict__ z, const CgState<T> *st, T *__restrict__ partials) {
  if (st->status != kIterate) return;
  const T alpha = st->alpha;
  const T malpha = -alpha;
  T rr = 0, rz = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const T ri = r[i] + malpha * Ap[i];
    const T zi = ri * invd[i];
    r[i] = ri;
    z[i] = zi;
    rr += ri * ri;
    rz += ri * zi;
  }
  block_partial2(rr, rz, partials);
}

// F2: res = sqrt(r.r), stopping test, beta = gamma / beta
template <typename T>
__global__ void __launch_bounds__(kFinalBlock)
    cg_beta_final(CgState<T> *st, const T *__restrict__ partials, int count) {
  __shared__ T lds[kFinalBlock / kWave];
  if (st->status != kIterate) return;
  const T rr = sum_partials(partials, count, 2, 0, lds);
  const T rz = sum_partials(partials, count, 2, 1, lds);
  if (threadIdx.x == 0) {
    const int iter = st->iter + 1;
    const T res = sqrt(rr);
    st->iter = iter;
    st->rr = rr;
    st->res = res;
    const int s = check_state(iter, (double)res, (double)st->res0, st->maxit, st->atol,
                              st->rtol, st->dtol);
    if (s == kIterate) {
      st->beta = rz / st->rz;
      st->rz = rz;
    }
    st->status = s;
  }
}

// K3: p = beta p + z (ScaleAdd(z, beta))
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_p_kernel(int n, const T *__restrict__ z, T *__restrict__ p, const CgState<T> *st) {
  if (st->status != kIterate) return;
  const T beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = beta * p[i] + z[i];
}

// The partial-sum area of the workspace is sized for the CSR tile grid,
// which is never smaller than the SELL grid (R <= 256 rows per tile vs 256
// rows per SELL workgroup), so one workspace size serves both operators.
// `layout` (optional) places the state, partials and vectors differently
// in `work` (the grids and so the results are the same); `events` false
// polls without creating events (run_polled_sync).
template <typename Op, typename T = typename Op::value_type>
int pcg_jacobi(hflmp_ctx *ctx, const Op &op, int nnz, const T *invd, const T *b, T *x,
               int maxit, double atol, double rtol, double dtol, void *work,
               size_t work_bytes, hflmp_solve_info *info, const SolverLayout *layout = nullptr,
               bool events = true) {
  const int nrow = op.nrow;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0, "pcg: empty system");
  HFLMP_REQUIRE(ctx, invd && b && x && work && info, "pcg: null pointer");
  HFLMP_REQUIRE(ctx, maxit > 0 && atol >= 0 && rtol >= 0 && dtol > 1,
                "pcg: bad control parameters (maxit>0, atol>=0, rtol>=0, dtol>1)");
  const SolverLayout L = layout ? *layout : solver_layout(nrow, nnz, sizeof(T), 4);
  const int grid1 = op_grid(op);
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "pcg: workspace %zu < %zu bytes", work_bytes,
                L.total);
  HFLMP_REQUIRE(ctx, grid1 <= L.grid1 || grid1 <= L.grid2,
                "pcg: operator grid %d exceeds the partial-sum area", grid1);
  char *w = static_cast<char *>(work);
  auto *st = reinterpret_cast<CgState<T> *>(w + L.state);
  T *partials = reinterpret_cast<T *>(w + L.partials);
  T *r = reinterpret_cast<T *>(w + L.vec);
  T *p = reinterpret_cast<T *>(w + L.vec + L.vstride);
  T *z = reinterpret_cast<T *>(w + L.vec + 2 * L.vstride);
  T *Ap = reinterpret_cast<T *>(w + L.vec + 3 * L.vstride);
  hipStream_t s = ctx->stream;
  const int g2 = L.grid2;
  const unsigned g3 = grid_for(nrow, 4);

  const auto t0 = std::chrono::steady_clock::now();
  init_state_kernel<T><<<1, kBlock, 0, s>>>(st, maxit, atol, rtol, dtol);
  cg_init_kernel<Op><<<grid1, kBlock, 0, s>>>(op, invd, b, x, r, z, p, partials);
  cg_init_final<T><<<1, kFinalBlock, 0, s>>>(st, partials, grid1);
  HFLMP_LAUNCH_CHECK(ctx);

  // the partial area holds 2 * max(L.grid1, L.grid2) values, at least 2 *
  // max(grid1, grid2); sizing by the latter keeps the grids, and so the
  // results, the same for every layout of the same system
  const Cg3Grids G = cg3_grids<Op>(ctx, grid1, L.grid2, grid1 > L.grid2 ? grid1 : L.grid2);
  T *part2 = partials + G.g1;
  auto iteration = [&]() {
    if (G.g1) {
      if constexpr (UnitWalk<Op>::value)
        cg3_spmv_kernel<Op><<<G.g1, kBlock, 0, s>>>(op, grid1, p, Ap, st, partials);
      else
        cg_spmv_kernel<Op><<<G.g1, kBlock, 0, s>>>(op, p, Ap, st, partials);
      cg3_update_kernel<T><<<G.g2, kBlock, 0, s>>>(nrow, G.g1, invd, p, Ap, x, r, z, st,
                                                   partials, part2);
      cg3_p_kernel<T><<<G.g2, kBlock, 0, s>>>(nrow, G.g2, z, p, st, part2);
      return;
    }
    cg_spmv_kernel<Op><<<grid1, kBlock, 0, s>>>(op, p, Ap, st, partials);
    cg_alpha_final<T><<<1, kFinalBlock, 0, s>>>(st, partials, grid1);
    cg_update_kernel<T><<<g2, kBlock, 0, s>>>(nrow, invd, p, Ap, x, r, z, st, partials);
    cg_beta_final<T><<<1, kFinalBlock, 0, s>>>(st, partials, g2);
    cg_p_kernel<T><<<g3, kBlock, 0, s>>>(nrow, z, p, st);
  };
  int rc = events ? run_polled(ctx, maxit, &st->status, iteration)
                  : run_polled_sync(ctx, maxit, &st->status, iteration);
  if (rc) return rc;
  CgState<T> h;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  info->iters = h.iter;
  info->status = h.status == kIterate ? kMaxIts : h.status;
  info->res_init = (double)h.res0;
  info->res_final = (double)h.res;
  info->seconds = std::chrono::duration<double>(t1 - t0).count();
  return HFLMP_OK;
}

// ---- FP32-inner / FP64-refined CG -----------------------------------------------
// The refinement of MPIR (linear_solver/mpir.cc:52-296) with the FP32
// Jacobi-PCG above as the error-correction solver, every step on the device:
//   r = b - A x (FP64 SpMV and residual, FP64 norm)  ->  r32 = (float) r
//   FP32 PCG on A32 c = r32 from c = 0                ->  x += (double) c
// The outer relative test uses the initial residual throughout.

struct MpState {
  double res0, res;
  int status;
  unsigned int bad16;  // entries whose column does not fit an int16 tile offset
};

// c16[j] = col[j] - r0(tile of row(j)) as int16, for the FP32 inner SpMV;
// counts the entries that do not fit.
__global__ void __launch_bounds__(kBlock)
    tile_offsets16_kernel(int nrow, int R, const int *__restrict__ rp,
                          const int *__restrict__ ci, short *__restrict__ c16,
                          unsigned int *bad) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    const int base = (i / R) * R;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      int d = ci[j] - base;
      if (d < -32768 || d > 32767) {
        atomicAdd(bad, 1u);
        d = 0;
      }
      c16[j] = (short)d;
    }
  }
}

// r = b - A x in FP64; r32 = (float) r; start of the FP32 PCG from c = 0:
// z = D32^-1 r32, p = z (keep_p: p is kept for p = z + beta p, the reliable
// update).  Partials: FP64 r.r; FP32 r32.r32 and r32.z.
template <typename Op64>
__global__ void __launch_bounds__(kBlock)
    mp_residual_kernel(Op64 op, const double *__restrict__ b, const double *__restrict__ x,
                       const float *__restrict__ invd32, float *__restrict__ r32,
                       float *__restrict__ z32, float *__restrict__ p32,
                       double *__restrict__ part64, float *__restrict__ part32, int keep_p) {
  __shared__ double prod[Op64::kLds];
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const double s = op.row_sum(tile, x, 0.0, tr, prod);
  double rr = 0;
  float rr32 = 0, rz32 = 0;
  if (tr.has_row) {
    const int i = tr.row;
    const double ri = -1.0 * s + b[i];
    const float rf = (float)ri;
    const float zf = rf * invd32[i];
    r32[i] = rf;
    z32[i] = zf;
    if (!keep_p) p32[i] = zf;
    rr = ri * ri;
    rr32 = rf * rf;
    rz32 = rf * zf;
  }
  block_partial1(rr, part64);
  block_partial2(rr32, rz32, part32);
}

// Outer test on the FP64 residual, then the FP32 PCG's initial state.  The
// inner relative tolerance is inner_rtol, relaxed to what the outer target
// still needs: max(inner_rtol, 0.5 * rtol * res_0 / res), so the last
// correction does not solve further than the outer test requires.  keep_p:
// the inner CG continues from the replaced residual with
// beta = (r32.z) / (r.z of its last iteration) instead of restarting.
__global__ void __launch_bounds__(kFinalBlock)
    mp_residual_final(MpState *ms, CgState<float> *st, const double *__restrict__ part64,
                      const float *__restrict__ part32, int count, int outer, int max_outer,
                      double atol, double rtol, double inner_rtol, int keep_p) {
  __shared__ double lds64[kFinalBlock / kWave];
  __shared__ float lds32[kFinalBlock / kWave];
  const double rr = sum_partials(part64, count, 1, 0, lds64);
  const float rr32 = sum_partials(part32, count, 2, 0, lds32);
  const float rz32 = sum_partials(part32, count, 2, 1, lds32);
  if (threadIdx.x == 0) {
    const double res = sqrt(rr);
    if (outer == 0) ms->res0 = res;
    ms->res = res;
    const double res0 = ms->res0;
    int status = kIterate;
    if (res <= atol || res < rtol * res0) status = kConverged;
    else if (!(res == res)) status = kBreakdown;
    else if (outer >= max_outer) status = kMaxIts;
    ms->status = status;
    const double need = res > 0 ? 0.5 * rtol * res0 / res : 0.0;
    const double tol = need > inner_rtol ? (need < 0.5 ? need : 0.5) : inner_rtol;
    st->rtol = tol;
    const float r0 = sqrtf(rr32);
    if (keep_p) st->beta = st->rz > 0.0f ? rz32 / st->rz : 0.0f;
    st->rr = rr32;
    st->rz = rz32;
    st->res0 = r0;
    st->res = r0;
    st->iter = 0;
    st->status = status == kIterate ? check_state(0, (double)r0, (double)r0, st->maxit,
                                                  st->atol, st->rtol, st->dtol)
                                    : kConverged;
  }
}

// x += (double) c
__global__ void __launch_bounds__(kBlock)
    mp_correct_kernel(int n, const float *__restrict__ c, double *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    x[i] += (double)c[i];
}

// Partials are sized for the larger of the grids that can run: the CSR
// tile grids (FP64 and FP32 windows), the elementwise grid, and the SELL
// grid (never larger than a CSR tile grid, and bounded by ceil(nrow/256)
// rounded to 8 when there are no CSR arrays).
struct MpLayout {
  size_t ms, st, part64, part32, vec, vstride, c16, total;
  int grid2;
};

inline MpLayout mp_layout(int nrow, int nnz, bool csr) {
  MpLayout L;
  int gmax = sell_grid(nrow);
  L.grid2 = solver_layout(nrow, 0, 4, 0).grid2;
  if (L.grid2 > gmax) gmax = L.grid2;
  if (csr) {
    const SolverLayout L64 = solver_layout(nrow, nnz, 8, 0);
    const SolverLayout L32 = solver_layout(nrow, nnz, 4, 0);
    if (L64.grid1 > gmax) gmax = L64.grid1;
    if (L32.grid1 > gmax) gmax = L32.grid1;
  }
  L.ms = 0;
  L.st = 256;
  L.part64 = 512;
  L.part32 = align256(L.part64 + (size_t)gmax * 8);
  L.vec = align256(L.part32 + 2 * (size_t)gmax * 4);
  L.vstride = align256((size_t)nrow * 4);
  L.c16 = L.vec + 5 * L.vstride;  // after r32, p32, z32, Ap32, c32
  L.total = L.c16 + (csr ? align256((size_t)nnz * 2) : 0);
  return L;
}

// The refinement loop over an FP64 operator (residual) and an FP32
// operator (inner CG); see the comment above MpState.
// reliable: the outer steps after the first replace the inner CG's residual
// and keep its search direction (no restart) unless the inner solve broke
// down.
template <typename Op64, typename Op32>
int mpcg_run(hflmp_ctx *ctx, const Op64 &op64, const Op32 &op32, const MpLayout &L,
             const float *invd32, const double *b, double *x, int max_outer, double atol,
             double rtol, int inner_maxit, double inner_rtol, bool reliable, void *work,
             hflmp_mp_info *info, std::chrono::steady_clock::time_point t0) {
  const int nrow = op64.nrow;
  const int grid64 = op_grid(op64), grid32 = op_grid(op32);
  char *w = static_cast<char *>(work);
  auto *ms = reinterpret_cast<MpState *>(w + L.ms);
  auto *st = reinterpret_cast<CgState<float> *>(w + L.st);
  double *part64 = reinterpret_cast<double *>(w + L.part64);
  float *part32 = reinterpret_cast<float *>(w + L.part32);
  float *r = reinterpret_cast<float *>(w + L.vec);
  float *p = reinterpret_cast<float *>(w + L.vec + L.vstride);
  float *z = reinterpret_cast<float *>(w + L.vec + 2 * L.vstride);
  float *Ap = reinterpret_cast<float *>(w + L.vec + 3 * L.vstride);
  float *c = reinterpret_cast<float *>(w + L.vec + 4 * L.vstride);
  hipStream_t s = ctx->stream;
  const unsigned g3 = grid_for(nrow, 4);
  // the FP32 part32 area holds 2 * gmax floats, gmax >= grid32 and grid2
  const int gmax32 = (int)((L.vec - L.part32) / (2 * sizeof(float)));
  const Cg3Grids G = cg3_grids<Op32>(ctx, grid32, L.grid2, gmax32);
  int *host_status = reinterpret_cast<int *>(static_cast<char *>(ctx->host_pinned) + 2048);
  int outer = 0, inner_total = 0;
  bool last_breakdown = false;
  for (;; ++outer) {
    const int keep_p = reliable && outer > 0 && !last_breakdown;
    if (!keep_p)
      init_state_kernel<float><<<1, kBlock, 0, s>>>(st, inner_maxit, 0.0, inner_rtol, 1e30);
    mp_residual_kernel<Op64><<<grid64, kBlock, 0, s>>>(op64, b, x, invd32, r, z, p, part64,
                                                       part32, keep_p);
    mp_residual_final<<<1, kFinalBlock, 0, s>>>(ms, st, part64, part32, grid64, outer,
                                                max_outer, atol, rtol, inner_rtol, keep_p);
    if (keep_p) cg_p_kernel<float><<<g3, kBlock, 0, s>>>(nrow, z, p, st);
    HFLMP_LAUNCH_CHECK(ctx);
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(host_status, &ms->status, sizeof(int),
                                        hipMemcpyDeviceToHost, s));
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (host_status[0] != kIterate) break;
    int rc = hflmp_fill_zero(ctx, c, sizeof(float) * (size_t)nrow);
    if (rc) return rc;
    auto iteration = [&]() {
      if (G.g1) {
        cg3_spmv_kernel<Op32><<<G.g1, kBlock, 0, s>>>(op32, grid32, p, Ap, st, part32);
        cg3_update_kernel<float><<<G.g2, kBlock, 0, s>>>(nrow, G.g1, invd32, p, Ap, c, r, z,
                                                         st, part32, part32 + G.g1);
        cg3_p_kernel<float><<<G.g2, kBlock, 0, s>>>(nrow, G.g2, z, p, st, part32 + G.g1);
        return;
      }
      cg_spmv_kernel<Op32><<<grid32, kBlock, 0, s>>>(op32, p, Ap, st, part32);
      cg_alpha_final<float><<<1, kFinalBlock, 0, s>>>(st, part32, grid32);
      cg_update_kernel<float><<<L.grid2, kBlock, 0, s>>>(nrow, invd32, p, Ap, c, r, z, st,
                                                         part32);
      cg_beta_final<float><<<1, kFinalBlock, 0, s>>>(st, part32, L.grid2);
      cg_p_kernel<float><<<g3, kBlock, 0, s>>>(nrow, z, p, st);
    };
    rc = run_polled(ctx, inner_maxit, &st->status, iteration);
    if (rc) return rc;
    mp_correct_kernel<<<grid_for(nrow, 4), kBlock, 0, s>>>(nrow, c, x);
    HFLMP_LAUNCH_CHECK(ctx);
    CgState<float> h;
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    inner_total += h.iter;
    last_breakdown = h.status == kBreakdown;
  }
  MpState hm;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&hm, ms, sizeof(hm), hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  info->outer_iters = outer;
  info->inner_iters = inner_total;
  info->status = hm.status == kIterate ? kMaxIts : hm.status;
  info->res_init = hm.res0;
  info->res_final = hm.res;
  info->seconds = std::chrono::duration<double>(t1 - t0).count();
  return HFLMP_OK;
}

#define HFLMP_MP_CHECK_CONTROL(ctx)                                                     \
  HFLMP_REQUIRE(ctx, max_outer > 0 && inner_maxit > 0 && atol >= 0 && rtol >= 0 &&      \
                         inner_rtol > 0 && inner_rtol < 1,                              \
                "mpcg: bad control parameters")

int mpcg_jacobi(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                const double *val, const float *val32, const float *invd32, const double *b,
                double *x, int max_outer, double atol, double rtol, int inner_maxit,
                double inner_rtol, int flags, void *work, size_t work_bytes,
                hflmp_mp_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0, "mpcg: empty system");
  HFLMP_REQUIRE(ctx, rp && ci && val && val32 && invd32 && b && x && work && info,
                "mpcg: null pointer");
  HFLMP_MP_CHECK_CONTROL(ctx);
  const MpLayout L = mp_layout(nrow, nnz, true);
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "mpcg: workspace %zu < %zu bytes", work_bytes,
                L.total);
  char *w = static_cast<char *>(work);
  auto *ms = reinterpret_cast<MpState *>(w + L.ms);
  short *c16 = reinterpret_cast<short *>(w + L.c16);
  hipStream_t s = ctx->stream;
  int *host_status = reinterpret_cast<int *>(static_cast<char *>(ctx->host_pinned) + 2048);

  const auto t0 = std::chrono::steady_clock::now();
  const CsrTileOp<double> op64 = csr_view<double>(nrow, nnz, rp, ci, val);
  const CsrTileOp<float> op32 = csr_view<float>(nrow, nnz, rp, ci, val32);
  // FP32 inner operator: int16 tile offsets when the band fits
  bool use16 = !(flags & HFLMP_MP_INT32_COLS);
  if (use16) {
    int rc = hflmp_fill_zero(ctx, ms, sizeof(MpState));
    if (rc) return rc;
    tile_offsets16_kernel<<<grid_for(nrow, 1), kBlock, 0, s>>>(nrow, op32.R, rp, ci, c16,
                                                               &ms->bad16);
    HFLMP_LAUNCH_CHECK(ctx);
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(host_status, &ms->bad16, sizeof(int),
                                        hipMemcpyDeviceToHost, s));
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    use16 = host_status[0] == 0;
  }
  if (use16) {
    CsrTileOp<float, short> op16;
    op16.nrow = nrow;
    op16.R = op32.R;
    op16.rp = rp;
    op16.ci = c16;
    op16.val = val32;
    return mpcg_run(ctx, op64, op16, L, invd32, b, x, max_outer, atol, rtol, inner_maxit,
                    inner_rtol, (flags & HFLMP_MP_RELIABLE) != 0, work, info, t0);
  }
  return mpcg_run(ctx, op64, op32, L, invd32, b, x, max_outer, atol, rtol, inner_maxit,
                  inner_rtol, (flags & HFLMP_MP_RELIABLE) != 0, work, info, t0);
}

int mpcg_jacobi_sell(hflmp_ctx *ctx, const hflmp_sell *A64, const hflmp_sell *A32,
                     const float *invd32, const double *b, double *x, int max_outer,
                     double atol, double rtol, int inner_maxit, double inner_rtol, int flags,
                     void *work, size_t work_bytes, hflmp_mp_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, A64 && A32 && invd32 && b && x && work && info, "mpcg: null pointer");
  HFLMP_REQUIRE(ctx, A64->value_bytes == 8 && A32->value_bytes == 4,
                "mpcg: A64 must hold FP64 values and A32 FP32 values");
  HFLMP_REQUIRE(ctx, A64->nrow == A32->nrow && A64->nrow > 0 && A64->nnz > 0,
                "mpcg: operators differ in size or are empty");
  HFLMP_REQUIRE(ctx, A64->device == ctx->device && A32->device == ctx->device,
                "mpcg: matrix on another device");
  HFLMP_MP_CHECK_CONTROL(ctx);
  const MpLayout L = mp_layout(A64->nrow, 0, false);
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "mpcg: workspace %zu < %zu bytes", work_bytes,
                L.total);
  HFLMP_REQUIRE(ctx, (flags & ~HFLMP_MP_RELIABLE) == 0, "mpcg: unknown flags %d", flags);
  const bool rel = (flags & HFLMP_MP_RELIABLE) != 0;
  const auto t0 = std::chrono::steady_clock::now();
  // four instantiations: int32 or int16 columns in each precision
  if (A64->col_bytes == 2) {
    const auto o64 = sell_view<double, short>(A64);
    if (A32->col_bytes == 2)
      return mpcg_run(ctx, o64, sell_view<float, short>(A32), L, invd32, b, x, max_outer,
                        atol, rtol, inner_maxit, inner_rtol, rel, work, info, t0);
    return mpcg_run(ctx, o64, sell_view<float, int>(A32), L, invd32, b, x, max_outer, atol,
                      rtol, inner_maxit, inner_rtol, rel, work, info, t0);
  }
  const auto o64 = sell_view<double, int>(A64);
  if (A32->col_bytes == 2)
    return mpcg_run(ctx, o64, sell_view<float, short>(A32), L, invd32, b, x, max_outer, atol,
                      rtol, inner_maxit, inner_rtol, rel, work, info, t0);
  return mpcg_run(ctx, o64, sell_view<float, int>(A32), L, invd32, b, x, max_outer, atol,
                    rtol, inner_maxit, inner_rtol, rel, work, info, t0);
}

template <typename T>
int pcg_jacobi_sell(hflmp_ctx *ctx, const hflmp_sell *A, const T *invd, const T *b, T *x,
                    int maxit, double atol, double rtol, double dtol, void *work,
                    size_t work_bytes, hflmp_solve_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, A != nullptr, "pcg: null matrix");
  HFLMP_REQUIRE(ctx, A->value_bytes == (int)sizeof(T), "pcg: value type mismatch");
  HFLMP_REQUIRE(ctx, A->device == ctx->device, "pcg: matrix on another device");
  if (A->col_bytes == 2)
    return pcg_jacobi(ctx, sell_view<T, short>(A), A->nnz, invd, b, x, maxit, atol, rtol, dtol,
                      work, work_bytes, info);
  return pcg_jacobi(ctx, sell_view<T, int>(A), A->nnz, invd, b, x, maxit, atol, rtol, dtol,
                    work, work_bytes, info);
}

// ---- the allocation-free Jacobi-PCG -----------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc): the workspace of
// pcg_jacobi over a SELL-64 operator lives in `scratch`, a static device
// array of the library sized for kNoAllocMaxRows rows.  Every GPU holds its
// own copy at its own address, resolved once per device by
// hipGetSymbolAddress; the array's name is never passed to a kernel.  All
// calls on one GPU share that GPU's copy, so calls that overlap on one GPU
// (two streams, two threads) race.  Nothing in it needs clearing: each call
// writes the state, partials and vectors before reading them.
constexpr int kNoAllocMaxRows = 1 << 18;
constexpr int kMaxDevices = 64;

SolverLayout noalloc_layout(int nrow) {
  // the grids of solver_layout / sell_grid for this row count; only the
  // partial area is sized for the SELL grid instead of the CSR tile grid
  SolverLayout L;
  L.R = 0;
  L.grid1 = sell_grid(nrow);
  long long g2 = ((long long)nrow + 4LL * kBlock - 1) / (4LL * kBlock);
  if (g2 > kMaxRedBlocks) g2 = kMaxRedBlocks;
  if (g2 < 1) g2 = 1;
  L.grid2 = (int)g2;
  const int gmax = L.grid1 > L.grid2 ? L.grid1 : L.grid2;
  L.state = 0;
  L.partials = 256;
  L.vec = align256(L.partials + 2 * (size_t)gmax * sizeof(double));
  L.vstride = align256((size_t)nrow * sizeof(double));
  L.total = L.vec + 4 * L.vstride;
  return L;
}

constexpr size_t kNoAllocBytes =
    256 + (2 * (size_t)(((kNoAllocMaxRows + kBlock - 1) / kBlock + kXcds - 1) / kXcds * kXcds) *
               sizeof(double) +
           255) / 256 * 256 +
    4 * (((size_t)kNoAllocMaxRows * sizeof(double) + 255) / 256 * 256);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the context's device (resolved on the first
// call for that device, with the device current).
int scratch_on_device(hflmp_ctx *ctx, void **out) {
  HFLMP_REQUIRE(ctx, ctx->device >= 0 && ctx->device < kMaxDevices,
                "pcg_noalloc: device %d beyond %d", ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    int cur = -1;
    HFLMP_HIP_CHECK(ctx, hipGetDevice(&cur));
    if (cur != ctx->device) HFLMP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != ctx->device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_error(ctx, HFLMP_EHIP, "pcg_noalloc: hipGetSymbolAddress(scratch): %s",
                       hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = p;
  }
  *out = g_scratch_ptr[ctx->device];
  return HFLMP_OK;
}

int pcg_jacobi_sell_noalloc(hflmp_ctx *ctx, const hflmp_sell *A, const double *invd,
                            const double *b, double *x, int maxit, double atol, double rtol,
                            double dtol, hflmp_solve_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, A != nullptr, "pcg_noalloc: null matrix");
  HFLMP_REQUIRE(ctx, A->value_bytes == 8, "pcg_noalloc: the matrix must hold FP64 values");
  HFLMP_REQUIRE(ctx, A->device == ctx->device, "pcg_noalloc: matrix on another device");
  HFLMP_REQUIRE(ctx, A->nrow <= kNoAllocMaxRows,
                "pcg_noalloc: %d rows exceed the %d rows its static workspace holds "
                "(use hflmp_pcg_jacobi_sell_f64 with a workspace)",
                A->nrow, kNoAllocMaxRows);
  const SolverLayout L = noalloc_layout(A->nrow);
  HFLMP_REQUIRE(ctx, L.total <= kNoAllocBytes, "pcg_noalloc: workspace %zu > %zu bytes",
                L.total, kNoAllocBytes);
  void *work = nullptr;
  int rc = scratch_on_device(ctx, &work);
  if (rc) return rc;
  if (A->col_bytes == 2)
    return pcg_jacobi(ctx, sell_view<double, short>(A), A->nnz, invd, b, x, maxit, atol, rtol,
                      dtol, work, kNoAllocBytes, info, &L, false);
  return pcg_jacobi(ctx, sell_view<double, int>(A), A->nnz, invd, b, x, maxit, atol, rtol,
                    dtol, work, kNoAllocBytes, info, &L, false);
}

// One-entry rows give 256-row tiles in solver_layout, the stencil grid.
inline int stencil_layout_nnz(long long nnode) { return (int)nnode; }

template <typename T>
int pcg_jacobi_stencil(hflmp_ctx *ctx, const hflmp_stencil *A, const T *invd, const T *b, T *x,
                       int maxit, double atol, double rtol, double dtol, void *work,
                       size_t work_bytes, hflmp_solve_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, A != nullptr, "pcg: null operator");
  HFLMP_REQUIRE(ctx, A->value_bytes == (int)sizeof(T), "pcg: value type mismatch");
  HFLMP_REQUIRE(ctx, A->device == ctx->device, "pcg: operator on another device");
  int rc = HFLMP_OK;
  stencil_dispatch<T>(A, [&](auto op) {
    rc = pcg_jacobi(ctx, op, stencil_layout_nnz(A->nnode), invd, b, x, maxit, atol, rtol, dtol,
                    work, work_bytes, info);
  });
  return rc;
}

template <typename T>
int pcg_jacobi_csr(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                   const T *val, const T *invd, const T *b, T *x, int maxit, double atol,
                   double rtol, double dtol, void *work, size_t work_bytes,
                   hflmp_solve_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, rp && ci && val, "pcg: null pointer");
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0, "pcg: empty system");
  return pcg_jacobi(ctx, csr_view<T>(nrow, nnz, rp, ci, val), nnz, invd, b, x, maxit, atol,
                    rtol, dtol, work, work_bytes, info);
}

}  // namespace

extern "C" {

size_t hflmp_pcg_workspace_bytes(int nrow, int nnz, int value_bytes) {
  if (nrow <= 0 || (value_bytes != 4 && value_bytes != 8)) return 0;
  return solver_layout(nrow, nnz, value_bytes, 4).total;
}

int hflmp_pcg_jacobi_f64(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                         const double *val, const double *invd, const double *b, double *x,
                         int maxit, double atol, double rtol, double dtol, void *work,
                         size_t work_bytes, hflmp_solve_info *info) {
  return pcg_jacobi_csr<double>(ctx, nrow, nnz, rp, ci, val, invd, b, x, maxit, atol, rtol,
                                dtol, work, work_bytes, info);
}

int hflmp_pcg_jacobi_sell_f64(hflmp_ctx *ctx, const hflmp_sell *A, const double *invd,
                              const double *b, double *x, int maxit, double atol, double rtol,
                              double dtol, void *work, size_t work_bytes,
                              hflmp_solve_info *info) {
  return pcg_jacobi_sell<double>(ctx, A, invd, b, x, maxit, atol, rtol, dtol, work, work_bytes,
                                 info);
}

int hflmp_pcg_jacobi_sell_noalloc_f64(hflmp_ctx *ctx, const hflmp_sell *A, const double *invd,
                                      const double *b, double *x, int maxit, double atol,
                                      double rtol, double dtol, hflmp_solve_info *info) {
  return pcg_jacobi_sell_noalloc(ctx, A, invd, b, x, maxit, atol, rtol, dtol, info);
}

int hflmp_pcg_jacobi_sell_f32(hflmp_ctx *ctx, const hflmp_sell *A, const float *invd,
                              const float *b, float *x, int maxit, double atol, double rtol,
                              double dtol, void *work, size_t work_bytes,
                              hflmp_solve_info *info) {
  return pcg_jacobi_sell<float>(ctx, A, invd, b, x, maxit, atol, rtol, dtol, work, work_bytes,
                                info);
}

size_t hflmp_pcg_stencil_workspace_bytes(long long nnode, int value_bytes) {
  if (nnode <= 0 || nnode > 0x7fffffffLL || (value_bytes != 4 && value_bytes != 8)) return 0;
  return solver_layout((int)nnode, stencil_layout_nnz(nnode), value_bytes, 4).total;
}

int hflmp_pcg_jacobi_stencil_f64(hflmp_ctx *ctx, const hflmp_stencil *A, const double *invd,
                                 const double *b, double *x, int maxit, double atol,
                                 double rtol, double dtol, void *work, size_t work_bytes,
                                 hflmp_solve_info *info) {
  return pcg_jacobi_stencil<double>(ctx, A, invd, b, x, maxit, atol, rtol, dtol, work,
                                    work_bytes, info);
}

int hflmp_pcg_jacobi_stencil_f32(hflmp_ctx *ctx, const hflmp_stencil *A, const float *invd,
                                 const float *b, float *x, int maxit, double atol, double rtol,
                                 double dtol, void *work, size_t work_bytes,
                                 hflmp_solve_info *info) {
  return pcg_jacobi_stencil<float>(ctx, A, invd, b, x, maxit, atol, rtol, dtol, work,
                                   work_bytes, info);
}

size_t hflmp_mpcg_workspace_bytes(int nrow, int nnz) {
  if (nrow <= 0) return 0;
  return mp_layout(nrow, nnz, true).total;
}

size_t hflmp_mpcg_sell_workspace_bytes(int nrow) {
  if (nrow <= 0) return 0;
  return mp_layout(nrow, 0, false).total;
}

int hflmp_mpcg_jacobi_sell(hflmp_ctx *ctx, const hflmp_sell *A64, const hflmp_sell *A32,
                           const float *invdiag32, const double *b, double *x, int max_outer,
                           double atol, double rtol, int inner_maxit, double inner_rtol,
                           int flags, void *work, size_t work_bytes, hflmp_mp_info *info) {
  return mpcg_jacobi_sell(ctx, A64, A32, invdiag32, b, x, max_outer, atol, rtol, inner_maxit,
                          inner_rtol, flags, work, work_bytes, info);
}

int hflmp_mpcg_jacobi(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                      const double *val, const float *val32, const float *invdiag32,
                      const double *b, double *x, int max_outer, double atol, double rtol,
                      int inner_maxit, double inner_rtol, int flags, void *work,
                      size_t work_bytes, hflmp_mp_info *info) {
  return mpcg_jacobi(ctx, nrow, nnz, rp, ci, val, val32, invdiag32, b, x, max_outer, atol, rtol,
                     inner_maxit, inner_rtol, flags, work, work_bytes, info);
}

int hflmp_pcg_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                         const float *val, const float *invd, const float *b, float *x,
                         int maxit, double atol, double rtol, double dtol, void *work,
                         size_t work_bytes, hflmp_solve_info *info) {
  return pcg_jacobi_csr<float>(ctx, nrow, nnz, rp, ci, val, invd, b, x, maxit, atol, rtol,
                               dtol, work, work_bytes, info);
}

}  // extern "C"
