// Device-resident Jacobi-preconditioned CG.
//
// Arithmetic of CG::SolvePrecond (linear_solver/cg.h:165-246):
//   r = b - A x;  z = D^-1 r;  p = z;  beta = r.z
//   loop: Ap = A p;  alpha = beta / (Ap.p);  x += alpha p;  r -= alpha Ap;
//         res = sqrt(r.r);  Check;  z = D^-1 r;  gamma = r.z;
//         p = (gamma/beta) p + z;  beta = gamma
// with the stopping rule of IterateControl::Check (common/iterate_control.cc:
// 80-107) on the recursive residual.  Every scalar stays on the device.
// One iteration is five launches:
//   K1  SpMV Ap = A p, per-workgroup partials of p.Ap
//   F1  one workgroup: p.Ap, alpha
//   K2  x += alpha p, r -= alpha Ap, z = D^-1 r, partials of r.r and r.z
//   F2  one workgroup: res, stopping test, beta
//   K3  p = beta p + z
// Every kernel returns at once when the state is no longer "iterate", so
// the host queues iterations ahead and reads the state every `poll`
// iterations only.
#include <chrono>

#include "hflmp_solver_common.hpp"

namespace {

using namespace hflmp_solver;

// r = b - A x (ScaleAdd(b,-1) of A x), z = D^-1 r, p = z; partials r.r, r.z
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_init_kernel(int nrow, int R, const int *__restrict__ rp, const int *__restrict__ ci,
                   const T *__restrict__ val, const T *__restrict__ invd,
                   const T *__restrict__ b, const T *__restrict__ x, T *__restrict__ r,
                   T *__restrict__ z, T *__restrict__ p, T *__restrict__ partials) {
  __shared__ T prod[TileCap<T>::value];
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const T s = tile_row_sum<T, false>(tile, R, nrow, rp, ci, val, x, T(0), tr, prod);
  T rr = 0, rz = 0;
  if (tr.has_row) {
    const int i = tr.row;
    const T ri = T(-1) * s + b[i];
    const T zi = ri * invd[i];
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    rr = ri * ri;
    rz = ri * zi;
  }
  block_partial2(rr, rz, partials);
}

template <typename T>
__global__ void __launch_bounds__(kFinalBlock)
    cg_init_final(CgState<T> *st, const T *__restrict__ partials, int count) {
  __shared__ T lds[kFinalBlock / kWave];
  const T rr = sum_partials(partials, count, 2, 0, lds);
  const T rz = sum_partials(partials, count, 2, 1, lds);
  if (threadIdx.x == 0) {
    const T res0 = sqrt(rr);
    st->rr = rr;
    st->rz = rz;
    st->res0 = res0;
    st->res = res0;
    st->iter = 0;
    st->status = check_state(0, (double)res0, (double)res0, st->maxit, st->atol, st->rtol,
                             st->dtol);
  }
}

// K1: Ap = A p, partials of p.Ap
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_spmv_kernel(int nrow, int R, const int *__restrict__ rp, const int *__restrict__ ci,
                   const T *__restrict__ val, const T *__restrict__ p, T *__restrict__ Ap,
                   const CgState<T> *st, T *__restrict__ partials) {
  __shared__ T prod[TileCap<T>::value];
  if (st->status != kIterate) return;
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const T s = tile_row_sum<T, false>(tile, R, nrow, rp, ci, val, p, T(0), tr, prod);
  T part = 0;
  if (tr.has_row) {
    Ap[tr.row] = s;
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

template <typename T>
int pcg_jacobi(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci, const T *val,
               const T *invd, const T *b, T *x, int maxit, double atol, double rtol,
               double dtol, void *work, size_t work_bytes, hflmp_solve_info *info) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0, "pcg: empty system");
  HFLMP_REQUIRE(ctx, rp && ci && val && invd && b && x && work && info, "pcg: null pointer");
  HFLMP_REQUIRE(ctx, maxit > 0 && atol >= 0 && rtol >= 0 && dtol > 1,
                "pcg: bad control parameters (maxit>0, atol>=0, rtol>=0, dtol>1)");
  const SolverLayout L = solver_layout(nrow, nnz, sizeof(T), 4);
  HFLMP_REQUIRE(ctx, work_bytes >= L.total, "pcg: workspace %zu < %zu bytes", work_bytes,
                L.total);
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
  cg_init_kernel<T><<<L.grid1, kBlock, 0, s>>>(nrow, L.R, rp, ci, val, invd, b, x, r, z, p,
                                               partials);
  cg_init_final<T><<<1, kFinalBlock, 0, s>>>(st, partials, L.grid1);
  HFLMP_LAUNCH_CHECK(ctx);

  auto iteration = [&]() {
    cg_spmv_kernel<T><<<L.grid1, kBlock, 0, s>>>(nrow, L.R, rp, ci, val, p, Ap, st, partials);
    cg_alpha_final<T><<<1, kFinalBlock, 0, s>>>(st, partials, L.grid1);
    cg_update_kernel<T><<<g2, kBlock, 0, s>>>(nrow, invd, p, Ap, x, r, z, st, partials);
    cg_beta_final<T><<<1, kFinalBlock, 0, s>>>(st, partials, g2);
    cg_p_kernel<T><<<g3, kBlock, 0, s>>>(nrow, z, p, st);
  };
  int rc = run_polled(ctx, maxit, &st->status, iteration);
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
  return pcg_jacobi<double>(ctx, nrow, nnz, rp, ci, val, invd, b, x, maxit, atol, rtol, dtol,
                            work, work_bytes, info);
}

int hflmp_pcg_jacobi_f32(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
                         const float *val, const float *invd, const float *b, float *x,
                         int maxit, double atol, double rtol, double dtol, void *work,
                         size_t work_bytes, hflmp_solve_info *info) {
  return pcg_jacobi<float>(ctx, nrow, nnz, rp, ci, val, invd, b, x, maxit, atol, rtol, dtol,
                           work, work_bytes, info);
}

}  // extern "C"
