// Kernels of the device-resident Jacobi-PCG shared by the single-system
// solver (hflmp_cg.hip) and the solve server's systems that run alone
// (hflmp_serve.hip): the start of the iteration and the three-launch
// iteration.  Arithmetic of CG::SolvePrecond (linear_solver/cg.h:165-246)
// with IterateControl::Check (common/iterate_control.cc:80-107).
#pragma once

#include <type_traits>

#include "hflmp_sell.hpp"
#include "hflmp_solver_common.hpp"
#include "hflmp_stencil.hpp"

namespace {

using namespace hflmp_solver;

// r = b - A x (ScaleAdd(b,-1) of A x), z = D^-1 r, p = z; partials r.r, r.z
template <typename Op, typename T = typename Op::value_type>
__global__ void __launch_bounds__(kBlock)
    cg_init_kernel(Op op, const T *__restrict__ invd, const T *__restrict__ b,
                   const T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                   T *__restrict__ p, T *__restrict__ partials) {
  __shared__ T prod[Op::kLds];
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  TileRow tr;
  const T s = op.row_sum(tile, x, T(0), tr, prod);
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

// ---- three launches per iteration ------------------------------------------------
// The same arithmetic with F1 and F2 folded into their consumers: every
// workgroup of K2 sums K1's partials (p.Ap) and every workgroup of K3 sums
// K2's (r.r, r.z), in the same fixed order, so all of them compute the same
// alpha, beta and stopping decision, and workgroup 0 records them in the
// state.  K1 runs a grid that fits on the device at once, each workgroup
// walking several units and keeping one partial, so the partial counts stay
// small (about 1280 and 2 x 1024 on gfx950) and every workgroup's re-read
// of them is an L2 hit.  Fields written by workgroup 0 are never read by
// the other workgroups of the same kernel: K2 leaves r.z and the iteration
// count in rz_prev / iter_prev for K3, whose workgroup 0 writes rz / iter.

// Operators whose K1 walks several units per workgroup.  The matrix-free
// stencil keeps one unit per workgroup (cg_spmv_kernel over its whole grid,
// one partial per unit): inside a walk its 3D kernel needs more than twice
// the registers.
template <typename Op>
struct UnitWalk : std::true_type {};
template <typename T, int DIM, int ORDER, bool MULTI>
struct UnitWalk<StencilOp<T, DIM, ORDER, MULTI>> : std::false_type {};

// K1: Ap = A p over all units; one partial of p.Ap per workgroup.  At least
// five waves per SIMD, so that every SELL-64 FP64 form (int16 or int32
// columns) runs the same grid and the sums, and so the iterates, do not
// depend on the column width.
template <typename Op, typename T = typename Op::value_type>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
    cg3_spmv_kernel(Op op, int vgrid, const T *__restrict__ p, T *__restrict__ Ap,
                    const CgState<T> *st, T *__restrict__ part) {
  __shared__ T prod[Op::kLds];
  if (st->status != kIterate) return;
  T acc = 0;
  // gridDim.x is a multiple of the XCD count: unit vb runs on XCD vb % 8
  for (int vb = blockIdx.x; vb < vgrid; vb += gridDim.x) {
    // operators that stage x in LDS (CSR tiles, stencil) write it before
    // their own barrier: wait until the previous unit's reads are done
    if (Op::kLds > 1 && vb != (int)blockIdx.x) __syncthreads();
    const long long tile = xcd_tile(vb, vgrid);
    TileRow tr;
    const T s = op.row_sum(tile, p, T(0), tr, prod);
    if (tr.has_row) {
      __builtin_nontemporal_store(s, Ap + tr.row);
      acc += s * p[tr.row];
    }
  }
  block_partial1(acc, part);
}

// K2: alpha from K1's partials; x += alpha p; r -= alpha Ap; z = D^-1 r;
// partials of r.r and r.z.  Four elements per thread per step, their loads
// issued together.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg3_update_kernel(int n, int g1, const T *__restrict__ invd, const T *__restrict__ p,
                      const T *__restrict__ Ap, T *__restrict__ x, T *__restrict__ r,
                      T *__restrict__ z, CgState<T> *st, const T *__restrict__ part1,
                      T *__restrict__ part2) {
  __shared__ T lds[kBlock / kWave];
  if (st->status != kIterate) return;
  const T pAp = block_total(part1, g1, 1, 0, lds);
  const T rzc = st->rz;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    st->pAp = pAp;
    st->rz_prev = rzc;
    st->iter_prev = st->iter;
  }
  if (pAp == T(0) || !(pAp == pAp)) {
    if (lead) st->status = kBreakdown;
    return;
  }
  const T alpha = rzc / pAp;
  if (lead) st->alpha = alpha;
  const T malpha = -alpha;
  T rr = 0, rz = 0;
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long i0 = (long long)blockIdx.x * kBlock + threadIdx.x; i0 < n; i0 += 4 * stride) {
    T pv[4], av[4], xv[4], rv[4], dv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = i0 + q * stride;
      if (i < n) {
        pv[q] = p[i];
        av[q] = Ap[i];
        xv[q] = x[i];
        rv[q] = r[i];
        dv[q] = invd[i];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = i0 + q * stride;
      if (i < n) {
        x[i] = xv[q] + alpha * pv[q];
        const T ri = rv[q] + malpha * av[q];
        const T zi = ri * dv[q];
        r[i] = ri;
        z[i] = zi;
        rr += ri * ri;
        rz += ri * zi;
      }
    }
  }
  block_partial2(rr, rz, part2);
}

// K3: res = sqrt(r.r), stopping test and beta from K2's partials; p = beta p + z
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg3_p_kernel(int n, int g2, const T *__restrict__ z, T *__restrict__ p, CgState<T> *st,
                 const T *__restrict__ part2) {
  __shared__ T lds[kBlock / kWave];
  if (st->status != kIterate) return;
  const T rr = block_total(part2, g2, 2, 0, lds);
  const T rz = block_total(part2, g2, 2, 1, lds);
  const int iter = st->iter_prev + 1;
  const T res = sqrt(rr);
  const int sv = check_state(iter, (double)res, (double)st->res0, st->maxit, st->atol,
                             st->rtol, st->dtol);
  const T beta = rz / st->rz_prev;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = iter;
    st->rr = rr;
    st->res = res;
    if (sv == kIterate) {
      st->beta = beta;
      st->rz = rz;
    }
    st->status = sv;
  }
  if (sv != kIterate) return;
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long i0 = (long long)blockIdx.x * kBlock + threadIdx.x; i0 < n; i0 += 4 * stride) {
    T pv[4], zv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = i0 + q * stride;
      if (i < n) {
        pv[q] = p[i];
        zv[q] = z[i];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = i0 + q * stride;
      if (i < n) p[i] = beta * pv[q] + zv[q];
    }
  }
}

// Grids of the three-launch iteration for an operator over `vgrid` units
// (a multiple of the XCD count) and a partial area of 2 * gmax values: K1
// on as many workgroups as are resident at once (a multiple of the XCD
// count), K2 and K3 on at most min(grid2, 1024).  g1 = 0: the partial area
// is too small (systems of a few hundred rows); use the five-launch form.
struct Cg3Grids {
  int g1, g2;
};

template <typename Op>
Cg3Grids cg3_grids(hflmp_ctx *ctx, int vgrid, int grid2, int gmax) {
  long long g1 = vgrid;
  if constexpr (UnitWalk<Op>::value) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, reinterpret_cast<const void *>(&cg3_spmv_kernel<Op>), kBlock, 0) !=
            hipSuccess ||
        occ < 1)
      occ = 1;
    const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
    g1 = (long long)occ * cus / kXcds * kXcds;
    if (g1 > vgrid) g1 = vgrid;
  }
  int g2 = grid2 < 1024 ? grid2 : 1024;
  const long long avail = 2LL * gmax - g1;
  if (g2 > avail / 2) g2 = (int)(avail / 2);
  if (g1 < kXcds || g2 < 1) return {0, 0};
  return {(int)g1, g2};
}

}  // namespace
