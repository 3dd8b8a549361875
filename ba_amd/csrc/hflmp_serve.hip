// Device-resident solve server: many independent Jacobi-PCG solves of
// different sizes advanced together, each in FP64 or in the FP32-inner /
// FP64-refined mode.  No reference counterpart (the reference runs one
// CG::Solve or MPIR::Solve at a time, linear_solver/cg.h:165-246,
// linear_solver/mpir.cc:52-296); each system follows that arithmetic with its
// own scalars and IterateControl test (common/iterate_control.cc:80-107).
//
// Memory.  One device arena of the HBM budget is allocated at creation.  A
// request is admitted (first come, first served) when a slot is free and
// its footprint fits in the arena; its SELL-64 layout(s), inverse diagonal,
// right-hand side, Krylov vectors and CSR arrays then occupy one arena
// block.  The CSR arrays, at the block's end, are read only by the
// admission kernels that build the SELL-64 layout: that part of the block
// returns to the arena at once; the rest stays until the result has been
// copied back.
//
// Batching.  The rows of every active system are cut into units of 256
// consecutive rows (one workgroup, four SELL-64 slices).  One launch of each
// iteration kernel covers the units of all active systems (unit -> system
// through a device table); each workgroup branches on its system's precision.
// The two scalar steps run one workgroup per system.  So an iteration of S
// systems is five launches, as for one.  A system's partial sums are
// indexed by its own units and added in a fixed order, so its iterates do
// not depend on which other systems share the launches: a result is
// bit-identical alone and in any mix.
//
// Steps.  hflmp_server_step runs one chunk of the active systems: the
// initial residuals of new FP64 systems, the FP64 outer step (x += c,
// r = b - A x, r32 = (float) r, outer test) of the mixed systems whose inner
// solve has reached its tolerance, and `chunk` iterations -- captured once
// per active set as a hipGraph and replayed while the set is unchanged.
// While the chunk runs, queued systems are admitted on a second stream
// (uploads from the caller's host arrays, layout kernels).  Then the states
// are read, the systems that stopped retire (x copied to the caller's
// buffer), and the admitted systems join from the next chunk on.
//
// Solo systems.  FP64 systems from solo_min_nnz nonzeros run the
// single-system three-launch iteration (hflmp_cg3.hpp) on a third stream,
// beside the grouped launches, with the single solver's grids: their
// iterates are that solver's bit for bit.
//
// Mixed systems use reliable updates: the outer step replaces the inner
// CG's recursive residual by the FP64 true residual and the correction c
// is folded into x, but the search direction is kept (beta formed from the
// replaced residual, p = z + beta p), so the Krylov space built so far is
// not thrown away as a restart would (the reference's MPIR restarts its
// inner solver).  Iteration counts stay close to FP64 CG's while every
// inner iteration moves FP32 bytes.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "hflmp_cg3.hpp"
#include "hflmp_sell.hpp"
#include "hflmp_solver_common.hpp"

namespace {

using namespace hflmp_solver;

constexpr int kFresh = -2;  // admitted; initial residual not yet formed
constexpr int kRefused = 4;  // a column index outside 0..nrow-1: never iterated
constexpr int kModeFp64 = 64, kModeMixed = 32;
constexpr int kMaxOuter = 50;  // refinement steps of a mixed system

struct SlotState {
  CgState<double> s64;  // FP64 systems
  CgState<float> s32;   // mixed systems: the inner FP32 CG
  double res0, res;     // mixed: FP64 true residual, initial and latest
  double atol, rtol, inner_rtol;
  int mode;
  int status;  // mixed: kFresh, kIterate or the final status
  int phase;   // mixed: 0 outer step due, 1 inner CG running, 2 its p to refresh first
  int outer, max_outer, inner_total, inner_maxit;
  int col16;  // both operators: 1 if their columns are int16 offsets from the
              // slice start (decided on the device at admission), else int32
  unsigned int bad;  // rows whose diagonal is missing or zero (counted at admission)
};

struct SysDesc {
  int nrow, nslice, nunits, unit0;
  int mode, col16;  // mixed: FP32 columns are int16 offsets from the slice start
  int col16_64;     // FP64 layout: the same (both as the device decided at admission)
  int grp0, gsz;    // K2's groups of gsz units (kGroup64 or kQuad): the system's first
  int quad0;        // K3's groups of kQuad units: the system's first
  // FP64 SELL-64 (P = 2): the operator of FP64 systems and of the residuals
  const long long *soff;
  const int *slen;
  const unsigned char *rlen;
  const void *col;  // int16 or int32 by col16_64
  const double *sval;
  // FP32 SELL-64 (P = 4): the inner operator of mixed systems
  const long long *soff32;
  const int *slen32;
  const void *col32;
  const float *sval32;
  const double *invd;
  const float *invd32;
  const double *b;
  double *x;
  double *r, *p, *z, *Ap;               // FP64 systems
  float *r32, *p32, *z32, *Ap32, *c32;  // mixed systems
  double *part;                         // 16 B per unit
  SlotState *st;
};

template <typename IT>
__host__ __device__ __forceinline__ SellOp<double, IT> op64(const SysDesc &d) {
  SellOp<double, IT> op;
  op.nrow = d.nrow;
  op.nslice = d.nslice;
  op.soff = d.soff;
  op.slen = d.slen;
  op.rlen = d.rlen;
  op.col = static_cast<const IT *>(d.col);
  op.val = d.sval;
  return op;
}

template <typename IT>
__device__ __forceinline__ SellOp<float, IT> op32(const SysDesc &d) {
  SellOp<float, IT> op;
  op.nrow = d.nrow;
  op.nslice = d.nslice;
  op.soff = d.soff32;
  op.slen = d.slen32;
  op.rlen = d.rlen;
  op.col = static_cast<const IT *>(d.col32);
  op.val = d.sval32;
  return op;
}

// sum of src[stride*u + off], u < count, over one workgroup of kBlock
// threads, eight loads in flight per thread; valid in thread 0
template <typename T>
__device__ __forceinline__ T units_sum(const T *__restrict__ src, int count, int stride,
                                       int off) {
  __shared__ T lds[kBlock / kWave];
  T a[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) a[f] = 0;
  for (int u0 = threadIdx.x; u0 < count; u0 += 8 * kBlock) {
    T v[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const int u = u0 + f * kBlock;
      v[f] = u < count ? src[(size_t)u * stride + off] : T(0);
    }
#pragma unroll
    for (int f = 0; f < 8; ++f) a[f] += v[f];
  }
  const T t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  return block_sum(t, lds);
}

// workgroup sums of a and b stored at part[2j], part[2j+1]
template <typename T>
__device__ __forceinline__ void unit_partial2(T a, T b, T *part, int j) {
  __shared__ T lds[2][kBlock / kWave];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) {
    lds[0][wid] = a;
    lds[1][wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T sa = 0, sb = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      sa += lds[0][w];
      sb += lds[1][w];
    }
    part[2 * j] = sa;
    part[2 * j + 1] = sb;
  }
}

template <typename T>
__device__ __forceinline__ void unit_partial1(T a, T *part, int j) {
  __shared__ T lds[kBlock / kWave];
  const T t = block_sum(a, lds);
  if (threadIdx.x == 0) part[j] = t;
}

// float view of a system's partials (FP32 sums of mixed systems)
__device__ __forceinline__ float *part32(const SysDesc &d) {
  return reinterpret_cast<float *>(d.part);
}

// Unit of the launch -> (system descriptor, local unit).  The launch covers
// units u0 .. u0 + U - 1 of the table, spread over the XCDs in contiguous
// bands.
__device__ __forceinline__ bool unit_of(const SysDesc *__restrict__ sys,
                                        const int *__restrict__ unit_sys, int U, SysDesc &d,
                                        int &j, int u0 = 0) {
  const long long u = xcd_tile(blockIdx.x, gridDim.x);
  if (u >= U) return false;
  d = sys[unit_sys[u0 + u]];
  j = (int)(u0 + u) - d.unit0;
  return true;
}

// The vector kernels of the iteration (K2, K3) run a workgroup over a group
// of consecutive units of one system (at most kQuad units, one row per unit
// per thread, their loads issued together).  A workgroup's life is mostly
// the latency of its dependent loads (table, descriptor, state, data), so
// more units per workgroup move more bytes per life.  K3 takes kQuad units
// for every system; K2 takes kQuad for FP32 systems and kGroup64 for FP64
// ones: their one-unit workgroups already run the update near the
// bandwidth that its mix of streams reaches, and four units ran it slower.
// Group of the launch -> (system descriptor, local group).
constexpr int kQuad = 4;
#ifndef HFLMP_GROUP64
#define HFLMP_GROUP64 1
#endif
constexpr int kGroup64 = HFLMP_GROUP64;
static_assert(kGroup64 >= 1 && kGroup64 <= kQuad, "group size");
__device__ __forceinline__ bool quad_of(const SysDesc *__restrict__ sys,
                                        const int *__restrict__ grp_sys, int G, SysDesc &d,
                                        int &k, bool quads = false) {
  const long long q = xcd_tile(blockIdx.x, gridDim.x);
  if (q >= G) return false;
  d = sys[grp_sys[q]];
  k = (int)q - (quads ? d.quad0 : d.grp0);
  return true;
}

// Per-unit sums of a[q] and b[q] over the workgroup, stored at
// part[2 (j0 + q)], part[2 (j0 + q) + 1] for the units j0 + q < jend: the
// same wave sums and the same order over the waves as unit_partial2 forms
// for one unit, so the partials do not depend on the grouping.
template <typename T>
__device__ __forceinline__ void quad_partial2(const T (&a)[kQuad], const T (&b)[kQuad],
                                              T *part, int j0, int jend) {
  __shared__ T lds[kQuad][2][kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  // only the group's own units (cross-lane sums are not free)
  const int nq = jend - j0;
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    if (q >= nq) break;
    const T sa = wave_sum(a[q]);
    const T sb = wave_sum(b[q]);
    if (lane == 0) {
      lds[q][0][wid] = sa;
      lds[q][1][wid] = sb;
    }
  }
  __syncthreads();
  if (threadIdx.x < kQuad && j0 + (int)threadIdx.x < jend) {
    const int q = threadIdx.x;
    T sa = 0, sb = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      sa += lds[q][0][w];
      sb += lds[q][1][w];
    }
    part[2 * (j0 + q)] = sa;
    part[2 * (j0 + q) + 1] = sb;
  }
}

__device__ __forceinline__ bool mixed_outer_due(const SlotState *st) {
  return st->mode == kModeMixed &&
         (st->status == kFresh || (st->status == kIterate && st->phase == 0));
}

__device__ __forceinline__ bool mixed_inner_live(const SlotState *st) {
  return st->mode == kModeMixed && st->status == kIterate && st->phase >= 1 &&
         st->s32.status == kIterate;
}

// ---- FP64 systems: start -----------------------------------------------------
// Fresh systems: x = 0, r = b, z = D^-1 b, p = z; partials r.r, r.z
__global__ void __launch_bounds__(kBlock)
    grp_init_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys, int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (d.mode != kModeFp64 || d.st->s64.status != kFresh) return;
  const long long i = (long long)j * kBlock + threadIdx.x;
  double rr = 0, rz = 0;
  if (i < d.nrow) {
    const double ri = d.b[i];
    const double zi = ri * d.invd[i];
    d.x[i] = 0.0;
    d.r[i] = ri;
    d.z[i] = zi;
    d.p[i] = zi;
    rr = ri * ri;
    rz = ri * zi;
  }
  unit_partial2(rr, rz, d.part, j);
}

__global__ void __launch_bounds__(kBlock) grp_init_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  CgState<double> *st = &d.st->s64;
  if (d.mode != kModeFp64 || st->status != kFresh) return;
  const double rr = units_sum(d.part, d.nunits, 2, 0);
  const double rz = units_sum(d.part, d.nunits, 2, 1);
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

// ---- mixed systems: the FP64 outer step --------------------------------------
// x = 0 for a fresh system, else x += (double) c
__global__ void __launch_bounds__(kBlock)
    grp_mix_correct_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys,
                           int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (!mixed_outer_due(d.st)) return;
  const bool fresh = d.st->status == kFresh;
  const long long i = (long long)j * kBlock + threadIdx.x;
  if (i < d.nrow) d.x[i] = fresh ? 0.0 : d.x[i] + (double)d.c32[i];
}

// r = b - A x in FP64; r32 = (float) r, z32 = D32^-1 r32, c = 0; a fresh
// system starts with p32 = z32 (later steps keep p32 for the refresh).
// Partials: FP64 r.r; FP32 r32.r32 and r32.z32.
__global__ void __launch_bounds__(kBlock)
    grp_mix_residual_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys,
                            int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (!mixed_outer_due(d.st)) return;
  const bool fresh = d.st->status == kFresh;
  TileRow tr;
  const double s = d.col16_64 ? op64<short>(d).row_sum(j, d.x, 0.0, tr, nullptr)
                              : op64<int>(d).row_sum(j, d.x, 0.0, tr, nullptr);
  double rr = 0;
  float rr32 = 0, rz32 = 0;
  if (tr.has_row) {
    const int i = tr.row;
    const double ri = -1.0 * s + d.b[i];
    const float rf = (float)ri;
    const float zf = rf * d.invd32[i];
    d.r32[i] = rf;
    d.z32[i] = zf;
    if (fresh) d.p32[i] = zf;
    d.c32[i] = 0.0f;
    rr = ri * ri;
    rr32 = rf * rf;
    rz32 = rf * zf;
  }
  // part[2j] = FP64 r.r; the two FP32 sums fill the unit's second 8 bytes
  // (floats 4j+2 and 4j+3)
  unit_partial1(rr, d.part, 2 * j);
  unit_partial2(rr32, rz32, part32(d), 2 * j + 1);
}

// Outer test on the FP64 residual (initial residual kept as r_0), then the
// inner CG continues from the replaced residual with tolerance
// max(inner_rtol, 0.5 rtol r_0 / r), capped at 0.5, relative to it:
// beta = (r32.z32) / (r.z of the last inner iteration), or 0 (a restart,
// p = z) for a fresh system or after an inner breakdown.
__global__ void __launch_bounds__(kBlock) grp_mix_residual_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  SlotState *st = d.st;
  if (!mixed_outer_due(st)) return;
  const double rr = units_sum(d.part, d.nunits, 2, 0);
  const float rr32 = units_sum(part32(d), d.nunits, 4, 2);
  const float rz32 = units_sum(part32(d), d.nunits, 4, 3);
  if (threadIdx.x == 0) {
    const bool fresh = st->status == kFresh;
    if (!fresh) st->outer += 1;
    const double res = sqrt(rr);
    if (fresh) st->res0 = res;
    st->res = res;
    const double res0 = st->res0;
    int status = kIterate;
    if (res <= st->atol || res < st->rtol * res0) status = kConverged;
    else if (!(res == res)) status = kBreakdown;
    else if (st->outer >= st->max_outer) status = kMaxIts;
    st->status = status;
    if (status == kIterate) {
      const double need = res > 0 ? 0.5 * st->rtol * res0 / res : 0.0;
      const double tol = need > st->inner_rtol ? (need < 0.5 ? need : 0.5) : st->inner_rtol;
      CgState<float> &s = st->s32;
      const float r0 = sqrtf(rr32);
      const bool restart = fresh || s.status == kBreakdown || !(s.rz > 0.0f);
      s.beta = restart ? 0.0f : rz32 / s.rz;
      s.rz = rz32;
      s.rr = rr32;
      s.res0 = r0;
      s.res = r0;
      s.alpha = s.pAp = 0.0f;
      s.iter = 0;
      s.maxit = st->inner_maxit;
      s.atol = 0.0;
      s.rtol = tol;
      s.dtol = 1e30;
      s.status = check_state(0, (double)r0, (double)r0, s.maxit, s.atol, s.rtol, s.dtol);
      st->phase = s.status == kIterate ? (fresh ? 1 : 2) : 0;
    }
  }
}

// p32 = beta p32 + z32 after an outer step (p32 = z32 on a restart)
__global__ void __launch_bounds__(kBlock)
    grp_mix_prefresh_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys,
                            int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  const SlotState *st = d.st;
  if (st->mode != kModeMixed || st->status != kIterate || st->phase != 2) return;
  const float beta = st->s32.beta;
  const long long i = (long long)j * kBlock + threadIdx.x;
  if (i < d.nrow) d.p32[i] = beta == 0.0f ? d.z32[i] : beta * d.p32[i] + d.z32[i];
}

// ---- the iteration, both precisions -------------------------------------------
// ---- K1: Ap = A p, partial p.Ap per unit --------------------------------------
// A grouped-SpMV unit's slices, gathered on the device when the tables are
// built, so that the SpMV's waves reach their entries one dependent load
// after the launch (the record) instead of three (unit table, system
// descriptor, slice offsets).
struct UnitRec {
  const void *col, *val;  // the unit's first slice in its form's column / value arrays
  const void *p;          // the system's search direction (p, or p32 of a mixed system)
  void *Ap;
  const unsigned char *rlen;
  void *part;             // the system's partials
  const SlotState *st;
  int row0, nrow, j, pad;
  int2 sl[kSellSlicesPerBlock];  // slice w: entry offset from slice 0, entries per lane
};

// The SpMV forms, in the order the unit table lists their units.
constexpr int kForms = 4;  // FP64 int16 columns, FP64 int32, FP32 int16, FP32 int32
__host__ __device__ inline int spmv_form(const SysDesc &d) {
  return d.mode == kModeFp64 ? (d.col16_64 ? 0 : 1) : (d.col16 ? 2 : 3);
}

__global__ void __launch_bounds__(kBlock)
    grp_unitrec_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys, int U,
                       UnitRec *__restrict__ rec) {
  for (int u = blockIdx.x * kBlock + threadIdx.x; u < U; u += gridDim.x * kBlock) {
    const SysDesc &d = sys[unit_sys[u]];
    const bool f64 = d.mode == kModeFp64;
    const long long *soff = f64 ? d.soff : d.soff32;
    const int *slen = f64 ? d.slen : d.slen32;
    const int form = spmv_form(d);
    const long long csz = (form == 0 || form == 2) ? 2 : 4;
    const int j = u - d.unit0;
    const int s0 = j * kSellSlicesPerBlock;  // < nslice: nunits = ceil(nslice / 4)
    const long long base = soff[s0];
    UnitRec r;
    r.col = static_cast<const char *>(f64 ? d.col : d.col32) + base * csz;
    r.val = f64 ? static_cast<const void *>(d.sval + base)
                : static_cast<const void *>(d.sval32 + base);
    r.p = f64 ? static_cast<const void *>(d.p) : static_cast<const void *>(d.p32);
    r.Ap = f64 ? static_cast<void *>(d.Ap) : static_cast<void *>(d.Ap32);
    r.rlen = d.rlen;
    r.part = d.part;
    r.st = d.st;
    r.row0 = s0 * kSellRows;
    r.nrow = d.nrow;
    r.j = j;
    r.pad = 0;
#pragma unroll
    for (int w = 0; w < kSellSlicesPerBlock; ++w) {
      const int sid = s0 + w;
      r.sl[w] = sid < d.nslice ? make_int2((int)(soff[sid] - base), slen[sid]) : make_int2(0, 0);
    }
    rec[u] = r;
  }
}

// One launch per form over the table's units of that form, so that each
// has its own register budget.  The loads of the state, the row's length,
// p at the row and the slice's entries go out together; the sums are those
// of SellOp::row_sum (same order, same predicates), and a stopped system's
// results are dropped.
template <typename T, typename IT>
__global__ void __launch_bounds__(kBlock)
    grp_spmv_kernel(const UnitRec *__restrict__ rec, int u0, int U) {
  const long long u = xcd_tile(blockIdx.x, gridDim.x);
  if (u >= U) return;
  const UnitRec &r = rec[u0 + u];
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & (kWave - 1);
  const int srow = r.row0 + w * kSellRows;  // the slice's first row
  const int row = srow + lane;
  const bool has_row = row < r.nrow;
  const T *__restrict__ p = static_cast<const T *>(r.p);
  const int2 sl = r.sl[w];
  const bool live = std::is_same<T, double>::value ? r.st->s64.status == kIterate
                                                   : mixed_inner_live(r.st);
  const int len = has_row ? (int)r.rlen[row] : 0;
  const T prow = has_row ? p[row] : T(0);
  // int16 columns are offsets from the slice's first row
  const T *__restrict__ x = std::is_same<IT, int>::value ? p : p + srow;
  const int dcol = std::is_same<IT, int>::value ? row : lane;
  const T s = sell_slice_sum<T, IT, false, false, true>(
      static_cast<const IT *>(r.col) + sl.x, static_cast<const T *>(r.val) + sl.x, sl.y, len,
      lane, x, dcol, T(0));
  if (!live) return;
  T part = 0;
  if (has_row) {
    static_cast<T *>(r.Ap)[row] = s;
    part = s * prow;
  }
  unit_partial1(part, static_cast<T *>(r.part), r.j);
}

// The FP32 forms run two units per workgroup: each wave sums slice w of
// both (sell_slice_sum2), so that it keeps twice the bytes in flight (the
// FP32 wave's loads are half the FP64 one's).  The units may belong to
// different systems; each keeps its own partial, summed as unit_partial1
// sums it (wave sums, then the waves in order).
template <typename T, typename IT>
__global__ void __launch_bounds__(kBlock)
    grp_spmv_pair_kernel(const UnitRec *__restrict__ rec, int u0, int U) {
  const long long pr = xcd_tile(blockIdx.x, gridDim.x);
  if (pr >= (U + 1) / 2) return;
  const int ua = u0 + 2 * (int)pr;
  const bool hasB = 2 * (int)pr + 1 < U;
  const UnitRec &ra = rec[ua];
  const UnitRec &rb = rec[hasB ? ua + 1 : ua];
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & (kWave - 1);
  const int srowA = ra.row0 + w * kSellRows, srowB = rb.row0 + w * kSellRows;
  const int rowA = srowA + lane, rowB = srowB + lane;
  const bool hA = rowA < ra.nrow, hB = hasB && rowB < rb.nrow;
  const T *__restrict__ pA = static_cast<const T *>(ra.p);
  const T *__restrict__ pB = static_cast<const T *>(rb.p);
  const int2 slA = ra.sl[w];
  const int2 slB = hasB ? rb.sl[w] : make_int2(0, 0);
  const bool liveA = mixed_inner_live(ra.st), liveB = hasB && mixed_inner_live(rb.st);
  const int lenA = hA ? (int)ra.rlen[rowA] : 0, lenB = hB ? (int)rb.rlen[rowB] : 0;
  const T prowA = hA ? pA[rowA] : T(0), prowB = hB ? pB[rowB] : T(0);
  const T *__restrict__ xA = std::is_same<IT, int>::value ? pA : pA + srowA;
  const T *__restrict__ xB = std::is_same<IT, int>::value ? pB : pB + srowB;
  T sA = 0, sB = 0;
  sell_slice_sum2<T, IT, 2>(static_cast<const IT *>(ra.col) + slA.x,
                         static_cast<const T *>(ra.val) + slA.x, slA.y, lenA, xA,
                         static_cast<const IT *>(rb.col) + slB.x,
                         static_cast<const T *>(rb.val) + slB.x, slB.y, lenB, xB, lane, sA, sB);
  T partA = 0, partB = 0;
  if (liveA && hA) {
    static_cast<T *>(ra.Ap)[rowA] = sA;
    partA = sA * prowA;
  }
  if (liveB && hB) {
    static_cast<T *>(rb.Ap)[rowB] = sB;
    partB = sB * prowB;
  }
  __shared__ T lds[2][kBlock / kWave];
  partA = wave_sum(partA);
  partB = wave_sum(partB);
  if (lane == 0) {
    lds[0][w] = partA;
    lds[1][w] = partB;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int h = threadIdx.x;
    if (h == 0 ? liveA : liveB) {
      T t = 0;
#pragma unroll
      for (int k = 0; k < kBlock / kWave; ++k) t += lds[h][k];
      const UnitRec &r = h == 0 ? ra : rb;
      static_cast<T *>(r.part)[r.j] = t;
    }
  }
}

// F1: alpha = r.z / p.Ap, one workgroup per system
__global__ void __launch_bounds__(kBlock) grp_alpha_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  if (d.mode == kModeFp64) {
    CgState<double> *st = &d.st->s64;
    if (st->status != kIterate) return;
    const double pAp = units_sum(d.part, d.nunits, 1, 0);
    if (threadIdx.x == 0) {
      st->pAp = pAp;
      st->alpha = st->rz / pAp;
      if (pAp == 0.0 || !(pAp == pAp)) st->status = kBreakdown;
    }
  } else {
    if (!mixed_inner_live(d.st)) return;
    CgState<float> *st = &d.st->s32;
    const float pAp = units_sum(part32(d), d.nunits, 1, 0);
    if (threadIdx.x == 0) {
      d.st->phase = 1;  // the refreshed p has been used
      st->pAp = pAp;
      st->alpha = st->rz / pAp;
      if (pAp == 0.0f || !(pAp == pAp)) {
        st->status = kBreakdown;
        d.st->inner_total += st->iter;
        d.st->phase = 0;
      }
    }
  }
}

// K2: x += alpha p; r -= alpha Ap; z = D^-1 r; partials r.r, r.z per unit
// (mixed systems: c in place of x, all FP32).  One workgroup per unit group;
// the loads go out before the state is read.
template <typename T>
__device__ __forceinline__ void quad_update(long long i0, int n, T alpha, const T *__restrict__ p,
                                            const T *__restrict__ Ap, T *__restrict__ x,
                                            T *__restrict__ r, T *__restrict__ z,
                                            const T (&pv)[kQuad], const T (&av)[kQuad],
                                            const T (&xv)[kQuad], const T (&rv)[kQuad],
                                            const T (&dv)[kQuad], T (&rr)[kQuad],
                                            T (&rz)[kQuad]) {
  const T malpha = -alpha;
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const long long i = i0 + q * kBlock;
    rr[q] = 0;
    rz[q] = 0;
    if (i < n) {
      x[i] = xv[q] + alpha * pv[q];
      const T ri = rv[q] + malpha * av[q];
      const T zi = ri * dv[q];
      r[i] = ri;
      z[i] = zi;
      rr[q] = ri * ri;
      rz[q] = ri * zi;
    }
  }
}

template <typename T>
__device__ __forceinline__ void quad_load5(long long i0, int n, const T *__restrict__ p,
                                           const T *__restrict__ Ap, const T *__restrict__ x,
                                           const T *__restrict__ r, const T *__restrict__ invd,
                                           T (&pv)[kQuad], T (&av)[kQuad], T (&xv)[kQuad],
                                           T (&rv)[kQuad], T (&dv)[kQuad]) {
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const long long i = i0 + q * kBlock;
    if (i < n) {
      pv[q] = p[i];
      av[q] = Ap[i];
      xv[q] = x[i];
      rv[q] = r[i];
      dv[q] = invd[i];
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    grp_update_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ grp_sys, int G) {
  SysDesc d;
  int k;
  if (!quad_of(sys, grp_sys, G, d, k)) return;
  // rows of the group: i0 + q kBlock < iend (q < gsz follows from the bound)
  const int g = d.gsz;
  const long long i0 = (long long)k * g * kBlock + threadIdx.x;
  const int iend = (int)std::min((long long)d.nrow, (long long)(k + 1) * g * kBlock);
  const int jend = std::min(d.nunits, g * (k + 1));
  if (d.mode == kModeFp64) {
    double pv[kQuad], av[kQuad], xv[kQuad], rv[kQuad], dv[kQuad], rr[kQuad], rz[kQuad];
    quad_load5(i0, iend, d.p, d.Ap, d.x, d.r, d.invd, pv, av, xv, rv, dv);
    if (d.st->s64.status != kIterate) return;
    quad_update(i0, iend, d.st->s64.alpha, d.p, d.Ap, d.x, d.r, d.z, pv, av, xv, rv, dv, rr,
                rz);
    quad_partial2(rr, rz, d.part, g * k, jend);
  } else {
    float pv[kQuad], av[kQuad], xv[kQuad], rv[kQuad], dv[kQuad], rr[kQuad], rz[kQuad];
    quad_load5(i0, iend, d.p32, d.Ap32, d.c32, d.r32, d.invd32, pv, av, xv, rv, dv);
    if (!mixed_inner_live(d.st)) return;
    quad_update(i0, iend, d.st->s32.alpha, d.p32, d.Ap32, d.c32, d.r32, d.z32, pv, av, xv,
                rv, dv, rr, rz);
    quad_partial2(rr, rz, part32(d), g * k, jend);
  }
}

// F2: res, stopping test, beta; one workgroup per system.  A mixed system
// whose inner solve stops is marked for the outer step of the next step.
__global__ void __launch_bounds__(kBlock) grp_beta_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  if (d.mode == kModeFp64) {
    CgState<double> *st = &d.st->s64;
    if (st->status != kIterate) return;
    const double rr = units_sum(d.part, d.nunits, 2, 0);
    const double rz = units_sum(d.part, d.nunits, 2, 1);
    if (threadIdx.x == 0) {
      const int iter = st->iter + 1;
      const double res = sqrt(rr);
      st->iter = iter;
      st->rr = rr;
      st->res = res;
      const int s = check_state(iter, res, st->res0, st->maxit, st->atol, st->rtol, st->dtol);
      if (s == kIterate) {
        st->beta = rz / st->rz;
        st->rz = rz;
      }
      st->status = s;
    }
  } else {
    if (!mixed_inner_live(d.st)) return;
    CgState<float> *st = &d.st->s32;
    const float rr = units_sum(part32(d), d.nunits, 2, 0);
    const float rz = units_sum(part32(d), d.nunits, 2, 1);
    if (threadIdx.x == 0) {
      const int iter = st->iter + 1;
      const float res = sqrtf(rr);
      st->iter = iter;
      st->rr = rr;
      st->res = res;
      const int s = check_state(iter, (double)res, (double)st->res0, st->maxit, st->atol,
                                st->rtol, st->dtol);
      if (s == kIterate) {
        st->beta = rz / st->rz;
        st->rz = rz;
      } else {
        d.st->inner_total += iter;
        d.st->phase = 0;
      }
      st->status = s;
    }
  }
}

// K3: p = beta p + z; one workgroup per unit group, the loads before the state
template <typename T>
__device__ __forceinline__ void quad_load2(long long i0, int n, const T *__restrict__ p,
                                           const T *__restrict__ z, T (&pv)[kQuad],
                                           T (&zv)[kQuad]) {
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const long long i = i0 + q * kBlock;
    if (i < n) {
      pv[q] = p[i];
      zv[q] = z[i];
    }
  }
}

template <typename T>
__device__ __forceinline__ void quad_p(long long i0, int n, T beta, T *__restrict__ p,
                                       const T (&pv)[kQuad], const T (&zv)[kQuad]) {
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const long long i = i0 + q * kBlock;
    if (i < n) p[i] = beta * pv[q] + zv[q];
  }
}

__global__ void __launch_bounds__(kBlock)
    grp_p_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ quad_sys, int Q) {
  SysDesc d;
  int k;
  if (!quad_of(sys, quad_sys, Q, d, k, true)) return;
  const long long i0 = (long long)k * kQuad * kBlock + threadIdx.x;
  const int iend = (int)std::min((long long)d.nrow, (long long)(k + 1) * kQuad * kBlock);
  if (d.mode == kModeFp64) {
    double pv[kQuad], zv[kQuad];
    quad_load2(i0, iend, d.p, d.z, pv, zv);
    if (d.st->s64.status != kIterate) return;
    quad_p(i0, iend, d.st->s64.beta, d.p, pv, zv);
  } else {
    float pv[kQuad], zv[kQuad];
    quad_load2(i0, iend, d.p32, d.z32, pv, zv);
    if (!mixed_inner_live(d.st)) return;
    quad_p(i0, iend, d.st->s32.beta, d.p32, pv, zv);
  }
}

// Clears st->col16 if an entry's column is not within int16 of its slice's
// first row (the flag starts at 1; every writer stores the same 0), and
// marks the system refused if a column lies outside 0..nrow-1.  It runs at
// admission before any product: every iteration kernel skips a refused
// system, so its out-of-range columns are never gathered.  (The row
// pointers were checked on the host at submission, so this kernel's own
// reads stay inside the system's arrays.)
__global__ void __launch_bounds__(kBlock)
    sys_band_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                    SlotState *st) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    const int base = (i / kSellRows) * kSellRows;
    bool fits = true, inside = true;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int c = ci[k];
      const int dlt = c - base;
      fits = fits && dlt >= -32768 && dlt <= 32767;
      inside = inside && (unsigned)c < (unsigned)nrow;
    }
    if (!fits) st->col16 = 0;
    if (!inside) {
      st->status = kRefused;
      st->s64.status = kRefused;
    }
  }
}

// CSR -> SELL-64 layouts of one system (thread per row) and its inverse
// diagonal: 1/a_ii formed in double (extract_invdiagelements,
// lmatrix_csr_cpu.cc:2419-2442); for the FP32 operator from the FP32 value,
// as the cast matrix's own extraction gives.  0 where the diagonal is
// missing or zero, counted in *bad.  Both layouts take the column width
// that sys_band_kernel decided (st->col16): int16 offsets from the slice's
// first row where every entry fits, else int32 columns.
__global__ void __launch_bounds__(kBlock)
    sys_fill_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                    const double *__restrict__ val, const long long *__restrict__ soff,
                    const int *__restrict__ slen, const SlotState *st, void *__restrict__ col,
                    double *__restrict__ sval, double *__restrict__ invd,
                    const long long *__restrict__ soff32, const int *__restrict__ slen32,
                    void *__restrict__ col32, float *__restrict__ sval32,
                    float *__restrict__ invd32, unsigned int *bad) {
  constexpr int P = SellChunk<double>::P, P32 = SellChunk<float>::P;
  const bool c16 = st->col16 != 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    const int s = i / kSellRows, lane = i % kSellRows;
    const long long off = soff[s];
    const int m = slen[s];
    const long long off32 = soff32 ? soff32[s] : 0;
    const int m32 = soff32 ? slen32[s] : 0;
    const int k0 = rp[i], len = rp[i + 1] - k0;
    double dg = 0;
    for (int j = 0; j < len; ++j) {
      const int c = ci[k0 + j];
      const double v = val[k0 + j];
      const long long pos = sell_pos(off, m, P, lane, j);
      if (c16)
        static_cast<short *>(col)[pos] = (short)(c - s * kSellRows);
      else
        static_cast<int *>(col)[pos] = c;
      sval[pos] = v;
      if (soff32) {
        const long long p32 = sell_pos(off32, m32, P32, lane, j);
        if (c16)
          static_cast<short *>(col32)[p32] = (short)(c - s * kSellRows);
        else
          static_cast<int *>(col32)[p32] = c;
        sval32[p32] = (float)v;
      }
      if (c == i && dg == 0.0) dg = v;
    }
    if (dg == 0.0) {
      if (invd) invd[i] = 0.0;
      if (invd32) invd32[i] = 0.0f;
      atomicAdd(bad, 1u);
    } else {
      if (invd) invd[i] = 1.0 / dg;
      if (invd32) {
        const float d32 = (float)dg;
        invd32[i] = d32 == 0.0f ? 0.0f : (float)(1.0 / (double)d32);
      }
    }
  }
}

// Copies nwords 4-byte words between device memory and coherent pinned host
// memory (both ends 16-byte aligned), 16 bytes per access where it can.
// The per-chunk transfers on the context stream (table uploads, the state
// copy) go through this kernel rather than the copy engine: a copy-engine
// transfer queued behind a running chunk waits for the chunk, and every
// transfer queued after it on the engine -- the retirement copies on the
// result stream -- waits with it.
__global__ void __launch_bounds__(kBlock)
    xfer_kernel(const unsigned *__restrict__ src, unsigned *__restrict__ dst, size_t nwords) {
  const size_t n4 = nwords / 4;
  const size_t stride = (size_t)gridDim.x * kBlock;
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
  for (size_t i = t; i < n4; i += stride) d4[i] = s4[i];
  for (size_t i = 4 * n4 + t; i < nwords; i += stride) dst[i] = src[i];
}

// ---- host side ------------------------------------------------------------------

// Queues xfer_kernel for `bytes` (a multiple of 4) on stream s.
int xfer(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (!bytes) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, bytes % 4 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 &&
                         reinterpret_cast<uintptr_t>(src) % 16 == 0,
                "server: transfer of %zu bytes is not word-sized and 16-byte aligned", bytes);
  const size_t nwords = bytes / 4;
  const unsigned g = (unsigned)std::min<size_t>((nwords / 4 + kBlock - 1) / kBlock + 1, 1024);
  xfer_kernel<<<g, kBlock, 0, s>>>(static_cast<const unsigned *>(src),
                                   static_cast<unsigned *>(dst), nwords);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

// First-fit allocator over the arena with coalescing of free neighbours.
class Arena {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(size_t bytes, size_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (const auto &f : free_) m = std::max(m, f.second);
    return m;
  }
  size_t total_free() const {
    size_t m = 0;
    for (const auto &f : free_) m += f.second;
    return m;
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> size
};

enum JobState { kQueued = 0, kActive = 1, kDone = 2 };

// SELL-64 slice plan (slice offsets and longest rows) from the host row_ptr
struct SlicePlan {
  std::vector<long long> soff;
  std::vector<int> slen;
  long long entries = 0;
};

void plan_slices(int nrow, const int *rp, SlicePlan &sp) {
  const int ns = (nrow + kSellRows - 1) / kSellRows;
  sp.slen.assign(ns, 0);
  sp.soff.assign((size_t)ns + 1, 0);
  long long tot = 0;
  for (int s = 0; s < ns; ++s) {
    int m = 0;
    for (int l = 0; l < kSellRows; ++l) {
      const int r = s * kSellRows + l;
      if (r >= nrow) break;
      m = std::max(m, rp[r + 1] - rp[r]);
    }
    sp.slen[s] = m;
    sp.soff[s] = tot;
    tot += (long long)kSellRows * sp.slen[s];
  }
  sp.soff[ns] = tot;
  sp.entries = tot;
}

// A solo system's partial sums: one pair per unit of its SELL-64 grid for
// the start, then the three-launch iteration's (at most 1280 of K1 and
// 2 x 1024 of K2 on gfx950).
inline size_t solo_part_values(int nrow) {
  const size_t g = 2 * (size_t)sell_grid(nrow);
  return g > 4096 ? g : 4096;
}

struct Job {
  long long ticket = 0;
  int nrow = 0, nnz = 0, nslice = 0, nunits = 0, mode = kModeFp64;
  const int *rp = nullptr, *ci = nullptr;
  const double *val = nullptr, *b = nullptr;
  double *x_out = nullptr;
  double atol = 0, rtol = 0;
  int maxit = 0;
  SlicePlan sp64, sp32;
  std::vector<unsigned char> rlen;
  size_t bytes = 0, off = 0;
  size_t kept = 0;  // bytes of the block still held (all but the CSR arrays once admitted)
  bool solo = false;  // FP64, at least solo_min_nnz nonzeros: runs alone on the solo stream
  int col16_64 = 0;   // solo: the FP64 layout's column width as the device decided it
  int slot = -1;
  int state = kQueued;
  SysDesc d{};
  // device-side block layout (offsets inside the job's arena block; arrays a
  // precision does not use keep offset 0 and are never read)
  size_t o_rp = 0, o_ci = 0, o_val = 0, o_rlen = 0, o_b = 0, o_x = 0, o_part = 0;
  size_t o_soff = 0, o_slen = 0, o_col = 0, o_sval = 0, o_invd = 0, o_r = 0, o_p = 0, o_z = 0,
         o_Ap = 0;
  size_t o_soff32 = 0, o_slen32 = 0, o_col32 = 0, o_sval32 = 0, o_invd32 = 0, o_r32 = 0,
         o_p32 = 0, o_z32 = 0, o_Ap32 = 0, o_c32 = 0;
  std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
  int iters = 0, status = 0, outer = 0, bad_diag = 0, col_bytes = 0;
  double res_init = 0, res_final = 0;
};

// Lays out the job's arrays and computes its footprint (256-byte aligned).
void plan_job(Job &jb) {
  jb.nslice = (jb.nrow + kSellRows - 1) / kSellRows;
  jb.nunits = (jb.nrow + kBlock - 1) / kBlock;
  jb.rlen.assign(jb.nrow, 0);
  for (int r = 0; r < jb.nrow; ++r) jb.rlen[r] = (unsigned char)(jb.rp[r + 1] - jb.rp[r]);
  plan_slices(jb.nrow, jb.rp, jb.sp64);
  if (jb.mode == kModeMixed) plan_slices(jb.nrow, jb.rp, jb.sp32);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + (bytes ? bytes : 1));
    return at;
  };
  const size_t n = (size_t)jb.nrow;
  const size_t e64 = (size_t)jb.sp64.entries;
  jb.o_rlen = take(n);
  jb.o_b = take(8 * n);
  jb.o_x = take(8 * n);
  jb.o_part = take(8 * (jb.solo ? solo_part_values(jb.nrow) : 2 * (size_t)jb.nunits));
  jb.o_soff = take(8 * ((size_t)jb.nslice + 1));
  jb.o_slen = take(4 * (size_t)jb.nslice);
  jb.o_col = take(4 * e64);  // int32 width; an int16 layout uses the first half
  jb.o_sval = take(8 * e64);
  if (jb.mode == kModeFp64) {
    jb.o_invd = take(8 * n);
    jb.o_r = take(8 * n);
    jb.o_p = take(8 * n);
    jb.o_z = take(8 * n);
    jb.o_Ap = take(8 * n);
  } else {
    const size_t e32 = (size_t)jb.sp32.entries;
    jb.o_soff32 = take(8 * ((size_t)jb.nslice + 1));
    jb.o_slen32 = take(4 * (size_t)jb.nslice);
    jb.o_col32 = take(4 * e32);  // as o_col
    jb.o_sval32 = take(4 * e32);
    jb.o_invd32 = take(4 * n);
    jb.o_r32 = take(4 * n);
    jb.o_p32 = take(4 * n);
    jb.o_z32 = take(4 * n);
    jb.o_Ap32 = take(4 * n);
    jb.o_c32 = take(4 * n);
  }
  // the CSR arrays last: read only at admission, then returned to the
  // arena, where they join the free space after the block
  jb.o_rp = take(4 * (n + 1));
  jb.o_ci = take(4 * (size_t)jb.nnz);
  jb.o_val = take(8 * (size_t)jb.nnz);
  jb.bytes = o;
  jb.kept = o;
}

// One of the two table sets that chunks run on in turn: the system and
// unit tables on the device (uploaded from coherent pinned staging), the
// chunk's captured graph over them, and the coherent pinned copy of all
// states taken after the chunk.  A set is rebuilt only when the chunk that
// last used it has finished.
struct TableSet {
  long long version = -1;  // the version of the active set these tables hold
  int S = 0, U = 0;
  SysDesc *d_sys = nullptr, *h_sys = nullptr;  // [max_active]
  int *d_unit = nullptr, *h_unit = nullptr;    // [unit_cap]: FP64 systems' units first
  size_t unit_cap = 0;
  int ubeg[kForms + 1] = {};  // the units of SpMV form f: ubeg[f] .. ubeg[f + 1] - 1
  UnitRec *d_rec = nullptr;   // [rec_cap]: every unit's record, built on the device
  size_t rec_cap = 0;
  int G = 0, Q = 0;                            // K2's unit groups; K3's quads
  int *d_grp = nullptr, *h_grp = nullptr;      // [grp_cap]: system of each group
  size_t grp_cap = 0;
  int *d_quad = nullptr, *h_quad = nullptr;    // [quad_cap]: system of each quad
  size_t quad_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  SlotState *h_states = nullptr;  // [max_active]
  hipEvent_t start = nullptr;     // recorded before the chunk's work (timing)
  hipEvent_t done = nullptr;      // recorded after the state copy (timing)
};

// A retired system's arena block, released once no chunk in flight lists it.
struct Deferred {
  int slot;
  size_t off, bytes;
};

}  // namespace

struct hflmp_server {
  hflmp_ctx *ctx = nullptr;
  size_t budget = 0;
  char *arena = nullptr;
  Arena alloc;
  int max_active = 0, chunk = 0, use_graph = 1, mixed_min_nnz = 0, mixed_max_nnz = 0;
  double inner_rtol = 1e-2;
  SlotState *d_states = nullptr;  // [max_active]
  SlotState *h_init = nullptr;    // pinned: each slot's initial state at admission
  hipStream_t admit_stream = nullptr;  // admission work, beside the chunk in flight
  char *stage = nullptr;               // pinned staging for admission uploads
  size_t stage_cap = 0, stage_used = 0;
  hipStream_t result_stream = nullptr; // retired systems' x to the host
  char *rstage = nullptr;              // pinned staging for those copies
  // The admission thread: the uploads and layout kernels of admitted
  // systems (admit_upload), in the order they were admitted, so that their
  // host copies overlap the calling thread's wait for the chunk.
  std::thread admitter;
  std::mutex am;
  std::condition_variable awake, aidle;
  std::deque<std::pair<const Job *, int>> apending;  // (job, slot) not yet taken
  int abusy = 0;          // taken and not finished
  bool astop = false;
  int arc = HFLMP_OK;     // the thread's first failure
  hflmp_ctx actx;         // the thread's copy of the context (its own error text)
  // per slot: an event after the slot's admission work on the admission
  // stream, and whether it has been recorded (guarded by am)
  std::vector<hipEvent_t> jev;
  std::vector<char> aready;
  size_t rstage_cap = 0;
  std::vector<int> joining;            // admitted slots that join after the chunk
  std::vector<long long> slot_job;  // slot -> ticket, -1 free
  std::vector<int> order;           // active slots in table order
  std::deque<long long> queue;
  std::map<long long, Job> jobs;
  long long next_ticket = 1;
  // Chunk c runs on table set c % 2.  A step queues the next chunk before
  // it waits for the one before, so the device always has a chunk queued
  // while the host retires, admits and returns to the caller.  The active
  // set's version changes when grouped systems join or leave.
  long long version = 0;
  TableSet sets[2];
  long long launched = 0, completed = 0;  // chunks queued / waited for
  std::vector<long long> first_chunk;     // per slot: the first chunk that lists it
  std::vector<Deferred> deferred;         // retired, still listed by the chunk in flight
  long long steps = 0, graph_captures = 0;
  int peak_active = 0;
  size_t peak_in_use = 0;
  // systems that run alone: the single-system three-launch iteration on a
  // second stream, beside the grouped launches of the others
  int solo_min_nnz = 0;
  hipStream_t solo_stream = nullptr;
  hipEvent_t ev_admitted = nullptr, ev_solo_done = nullptr;
  std::vector<int> solo;        // active solo slots
  SlotState *h_admit = nullptr; // pinned: a solo system's state read at admission
  long long solo_runs = 0;
  // host time of step(): admission (uploads, layout kernels queued), table
  // rebuild and graph capture, waiting for the step's kernels
  double admit_seconds = 0, capture_seconds = 0, wait_seconds = 0, join_seconds = 0;
  double retire_seconds = 0;
  double device_seconds = 0;  // the chunks' own device time (events)
};

namespace {

using Clock = std::chrono::steady_clock;

void drop_graph(TableSet &t) {
  if (t.graph_exec) (void)hipGraphExecDestroy(t.graph_exec);
  if (t.graph) (void)hipGraphDestroy(t.graph);
  t.graph_exec = nullptr;
  t.graph = nullptr;
}

// The work of one chunk on the stream: starts of fresh FP64 systems, outer
// steps of mixed systems, then `chunk` iterations.
void enqueue_chunk(const hflmp_server *sv, const TableSet &t, hipStream_t s) {
  auto grid = [](int n) { return (unsigned)(((n + kXcds - 1) / kXcds) * kXcds); };
  const unsigned gU = grid(t.U), gG = grid(t.G), gQ = grid(t.Q);
  const unsigned gS = (unsigned)t.S;
  const int *ub = t.ubeg;
  grp_init_kernel<<<gU, kBlock, 0, s>>>(t.d_sys, t.d_unit, t.U);
  grp_init_final<<<gS, kBlock, 0, s>>>(t.d_sys);
  grp_mix_correct_kernel<<<gU, kBlock, 0, s>>>(t.d_sys, t.d_unit, t.U);
  grp_mix_residual_kernel<<<gU, kBlock, 0, s>>>(t.d_sys, t.d_unit, t.U);
  grp_mix_residual_final<<<gS, kBlock, 0, s>>>(t.d_sys);
  grp_mix_prefresh_kernel<<<gU, kBlock, 0, s>>>(t.d_sys, t.d_unit, t.U);
  for (int it = 0; it < sv->chunk; ++it) {
    if (ub[1] > ub[0])
      grp_spmv_kernel<double, short><<<grid(ub[1] - ub[0]), kBlock, 0, s>>>(t.d_rec, ub[0],
                                                                              ub[1] - ub[0]);
    if (ub[2] > ub[1])
      grp_spmv_kernel<double, int><<<grid(ub[2] - ub[1]), kBlock, 0, s>>>(t.d_rec, ub[1],
                                                                            ub[2] - ub[1]);
    if (ub[3] > ub[2])
      grp_spmv_pair_kernel<float, short><<<grid((ub[3] - ub[2] + 1) / 2), kBlock, 0, s>>>(
          t.d_rec, ub[2], ub[3] - ub[2]);
    if (ub[4] > ub[3])
      grp_spmv_pair_kernel<float, int><<<grid((ub[4] - ub[3] + 1) / 2), kBlock, 0, s>>>(
          t.d_rec, ub[3], ub[4] - ub[3]);
    grp_alpha_final<<<gS, kBlock, 0, s>>>(t.d_sys);
    grp_update_kernel<<<gG, kBlock, 0, s>>>(t.d_sys, t.d_grp, t.G);
    grp_beta_final<<<gS, kBlock, 0, s>>>(t.d_sys);
    grp_p_kernel<<<gQ, kBlock, 0, s>>>(t.d_sys, t.d_quad, t.Q);
  }
}

// A solo system runs the single-system solver's kernels on its arena block
// (hflmp_cg3.hpp): the start from x = 0, then chunks of the three-launch
// iteration, with the grids hflmp_pcg_jacobi_sell_f64 chooses for the same
// system, so its iterates are that solver's, bit for bit.
template <typename IT>
void solo_start(const Job &jb, hipStream_t s) {
  const SellOp<double, IT> op = op64<IT>(jb.d);
  CgState<double> *st = &jb.d.st->s64;
  const int grid1 = op_grid(op);
  init_state_kernel<double><<<1, kBlock, 0, s>>>(st, jb.maxit, jb.atol, jb.rtol, 1e300);
  cg_init_kernel<SellOp<double, IT>><<<grid1, kBlock, 0, s>>>(
      op, jb.d.invd, jb.d.b, jb.d.x, jb.d.r, jb.d.z, jb.d.p, jb.d.part);
  cg_init_final<double><<<1, kFinalBlock, 0, s>>>(st, jb.d.part, grid1);
}

template <typename IT>
void solo_chunk(hflmp_ctx *ctx, const Job &jb, hipStream_t s, int iters) {
  const SellOp<double, IT> op = op64<IT>(jb.d);
  CgState<double> *st = &jb.d.st->s64;
  const int grid1 = op_grid(op);
  long long g2 = ((long long)jb.nrow + 4LL * kBlock - 1) / (4LL * kBlock);
  if (g2 > kMaxRedBlocks) g2 = kMaxRedBlocks;
  if (g2 < 1) g2 = 1;
  const Cg3Grids G = cg3_grids<SellOp<double, IT>>(
      ctx, grid1, (int)g2, (int)(solo_part_values(jb.nrow) / 2));
  double *part2 = jb.d.part + G.g1;
  for (int it = 0; it < iters; ++it) {
    cg3_spmv_kernel<SellOp<double, IT>><<<G.g1, kBlock, 0, s>>>(op, grid1, jb.d.p, jb.d.Ap, st,
                                                                 jb.d.part);
    cg3_update_kernel<double><<<G.g2, kBlock, 0, s>>>(jb.nrow, G.g1, jb.d.invd, jb.d.p,
                                                      jb.d.Ap, jb.d.x, jb.d.r, jb.d.z, st,
                                                      jb.d.part, part2);
    cg3_p_kernel<double><<<G.g2, kBlock, 0, s>>>(jb.nrow, G.g2, jb.d.z, jb.d.p, st, part2);
  }
}

// Builds table set t for the current active set (grouped systems in table
// order; the units by SpMV form, for the per-form SpMV launches, and their
// records; every system's unit groups and quads) in its pinned
// staging and queues the uploads on the context stream.  The chunk that
// last ran on t has finished, so neither its staging nor its device tables
// are in use.
int grow_table(hflmp_ctx *ctx, int **d, int **h, size_t *cap, size_t need) {
  if (need <= *cap) return HFLMP_OK;
  if (*d) (void)hipFree(*d);
  if (*h) (void)hipHostFree(*h);
  *d = nullptr;
  *h = nullptr;
  *cap = std::max(need, 2 * *cap);
  if (hipMalloc(d, sizeof(int) * *cap) != hipSuccess ||
      hipHostMalloc(h, sizeof(int) * *cap, hipHostMallocCoherent) != hipSuccess) {
    *cap = 0;
    return set_error(ctx, HFLMP_ENOMEM, "server: unit table allocation failed");
  }
  return HFLMP_OK;
}

int rebuild_tables(hflmp_server *sv, TableSet &t) {
  hflmp_ctx *ctx = sv->ctx;
  int U = 0, G = 0, Q = 0;
  for (int slot : sv->order) {
    Job &jb = sv->jobs[sv->slot_job[slot]];
    jb.d.gsz = jb.d.mode == kModeFp64 ? kGroup64 : kQuad;
    U += jb.nunits;
    G += (jb.nunits + jb.d.gsz - 1) / jb.d.gsz;
    Q += (jb.nunits + kQuad - 1) / kQuad;
  }
  if ((size_t)U > t.rec_cap) {
    if (t.d_rec) (void)hipFree(t.d_rec);
    t.d_rec = nullptr;
    t.rec_cap = std::max((size_t)U, 2 * t.rec_cap);
    if (hipMalloc(&t.d_rec, sizeof(UnitRec) * t.rec_cap) != hipSuccess) {
      t.rec_cap = 0;
      return set_error(ctx, HFLMP_ENOMEM, "server: unit record allocation failed");
    }
  }
  int rc = grow_table(ctx, &t.d_unit, &t.h_unit, &t.unit_cap, (size_t)U);
  if (!rc) rc = grow_table(ctx, &t.d_grp, &t.h_grp, &t.grp_cap, (size_t)G);
  if (!rc) rc = grow_table(ctx, &t.d_quad, &t.h_quad, &t.quad_cap, (size_t)Q);
  if (rc) return rc;
  int S = 0, unit0 = 0, grp0 = 0, quad0 = 0;
  for (int slot : sv->order) {
    Job &jb = sv->jobs[sv->slot_job[slot]];
    jb.d.grp0 = grp0;
    const int ng = (jb.nunits + jb.d.gsz - 1) / jb.d.gsz;
    for (int k = 0; k < ng; ++k) t.h_grp[grp0 + k] = S;
    grp0 += ng;
    jb.d.quad0 = quad0;
    const int nq = (jb.nunits + kQuad - 1) / kQuad;
    for (int k = 0; k < nq; ++k) t.h_quad[quad0 + k] = S;
    quad0 += nq;
    ++S;
  }
  for (int f = 0; f < kForms; ++f) {
    t.ubeg[f] = unit0;
    S = 0;
    for (int slot : sv->order) {
      Job &jb = sv->jobs[sv->slot_job[slot]];
      if (spmv_form(jb.d) == f) {
        jb.d.unit0 = unit0;
        for (int u = 0; u < jb.nunits; ++u) t.h_unit[unit0 + u] = S;
        unit0 += jb.nunits;
      }
      ++S;
    }
  }
  t.ubeg[kForms] = unit0;
  S = 0;
  for (int slot : sv->order) t.h_sys[S++] = sv->jobs[sv->slot_job[slot]].d;
  t.S = S;
  t.U = U;
  t.G = G;
  t.Q = Q;
  if (S) {
    rc = xfer(ctx, t.d_sys, t.h_sys, sizeof(SysDesc) * S, ctx->stream);
    if (!rc) rc = xfer(ctx, t.d_unit, t.h_unit, sizeof(int) * U, ctx->stream);
    if (!rc) rc = xfer(ctx, t.d_grp, t.h_grp, sizeof(int) * G, ctx->stream);
    if (!rc) rc = xfer(ctx, t.d_quad, t.h_quad, sizeof(int) * Q, ctx->stream);
    if (rc) return rc;
    const unsigned g = (unsigned)std::min<long long>(((long long)U + kBlock - 1) / kBlock, 4096);
    grp_unitrec_kernel<<<g, kBlock, 0, ctx->stream>>>(t.d_sys, t.d_unit, U, t.d_rec);
    HFLMP_LAUNCH_CHECK(ctx);
  }
  drop_graph(t);
  t.version = sv->version;
  return HFLMP_OK;
}

// Admission work on the admission stream, which runs beside the step in
// flight: that step's launches never touch this slot or block (both were
// free when its tables were built).
// Host copy into the pinned staging area, split over a few threads for
// large arrays (one thread reaches about a third of what PCIe takes).
void host_copy(char *dst, const char *src, size_t bytes) {
  constexpr size_t kSplitBytes = 8u << 20;
  constexpr int kThreads = 4;
  if (bytes < kSplitBytes) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::thread th[kThreads - 1];
  const size_t part = (bytes / kThreads + 4095) & ~(size_t)4095;
  for (int t = 1; t < kThreads; ++t) {
    const size_t a = std::min(bytes, part * t), b = std::min(bytes, part * (t + 1));
    th[t - 1] = std::thread([=] { std::memcpy(dst + a, src + a, b - a); });
  }
  std::memcpy(dst, src, std::min(bytes, part));
  for (auto &x : th) x.join();
}

// Queues an upload of the caller's (pageable) host array on the admission
// stream through the pinned staging area, so the host never waits for the
// admission kernels queued before it.  When the area is full the host
// waits for the admission stream once and starts again at its beginning.
int stage_upload(hflmp_server *sv, hflmp_ctx *ctx, void *dst, const void *src, size_t bytes) {
  constexpr size_t kMinPiece = 4u << 20;
  const char *p = static_cast<const char *>(src);
  char *d = static_cast<char *>(dst);
  while (bytes) {
    size_t room = sv->stage_cap - sv->stage_used;
    if (room < bytes && room < kMinPiece) {
      HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(sv->admit_stream));
      sv->stage_used = 0;
      room = sv->stage_cap;
    }
    const size_t piece = std::min(bytes, room);
    char *h = sv->stage + sv->stage_used;
    host_copy(h, p, piece);
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(d, h, piece, hipMemcpyHostToDevice, sv->admit_stream));
    sv->stage_used = std::min(sv->stage_cap, align256(sv->stage_used + piece));
    p += piece;
    d += piece;
    bytes -= piece;
  }
  return HFLMP_OK;
}

// The admission work of one system on the admission stream: uploads
// through the staging area, zero fills, its initial state and the layout
// kernels.  Runs on the server's admission thread (ctx: the thread's copy
// of the context, for its own error text), or on the calling thread for a
// solo system.  Touches only the system's block, the staging area and the
// slot's pinned initial state.
int admit_upload(hflmp_server *sv, hflmp_ctx *ctx, const Job &jb, int slot) {
  hipStream_t s = sv->admit_stream;
  char *base = sv->arena + jb.off;
  auto dp = [&](size_t o) { return static_cast<void *>(base + o); };
  const size_t n = (size_t)jb.nrow;
  const bool mixed = jb.mode == kModeMixed;
  int rc = HFLMP_OK;
  auto up = [&](size_t o, const void *src, size_t bytes) {
    if (!rc) rc = stage_upload(sv, ctx, dp(o), src, bytes);
  };
  up(jb.o_rp, jb.rp, 4 * (n + 1));
  up(jb.o_ci, jb.ci, 4 * (size_t)jb.nnz);
  up(jb.o_val, jb.val, 8 * (size_t)jb.nnz);
  up(jb.o_rlen, jb.rlen.data(), n);
  up(jb.o_b, jb.b, 8 * n);
  up(jb.o_soff, jb.sp64.soff.data(), 8 * jb.sp64.soff.size());
  up(jb.o_slen, jb.sp64.slen.data(), 4 * jb.sp64.slen.size());
  if (mixed) {
    up(jb.o_soff32, jb.sp32.soff.data(), 8 * jb.sp32.soff.size());
    up(jb.o_slen32, jb.sp32.slen.data(), 4 * jb.sp32.slen.size());
  }
  if (!rc) rc = fill_zero_on(ctx, dp(jb.o_col), 4 * (size_t)jb.sp64.entries, s);
  if (!rc) rc = fill_zero_on(ctx, dp(jb.o_sval), 8 * (size_t)jb.sp64.entries, s);
  if (!rc && mixed) {
    rc = fill_zero_on(ctx, dp(jb.o_col32), 4 * (size_t)jb.sp32.entries, s);
    if (!rc) rc = fill_zero_on(ctx, dp(jb.o_sval32), 4 * (size_t)jb.sp32.entries, s);
  }
  if (rc) return rc;
  SlotState st;
  std::memset(&st, 0, sizeof(st));
  st.mode = jb.mode;
  st.atol = jb.atol;
  st.rtol = jb.rtol;
  st.inner_rtol = sv->inner_rtol;
  st.inner_maxit = jb.maxit;
  st.max_outer = kMaxOuter;
  st.status = mixed ? kFresh : kIterate;
  st.s64.status = mixed ? kConverged : kFresh;
  st.s64.maxit = jb.maxit;
  st.s64.atol = jb.atol;
  st.s64.rtol = jb.rtol;
  st.s64.dtol = 1e300;
  st.s32.status = kConverged;
  st.col16 = 1;  // until the band kernel finds an entry that does not fit
  // staged in the slot's own pinned entry (the state mirror may be written
  // by the running step's state copy meanwhile)
  sv->h_init[slot] = st;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->d_states + slot, sv->h_init + slot, sizeof(st),
                                      hipMemcpyHostToDevice, s));
  // the FP64 layout's column width is decided on the device (no host pass
  // over the columns); the fill and every product read the slot's flag
  const unsigned g = grid_for(jb.nrow, 1);
  const int *rp = static_cast<const int *>(dp(jb.o_rp));
  const int *ci = static_cast<const int *>(dp(jb.o_ci));
  const double *val = static_cast<const double *>(dp(jb.o_val));
  const long long *soff = static_cast<const long long *>(dp(jb.o_soff));
  const int *slen = static_cast<const int *>(dp(jb.o_slen));
  void *col = dp(jb.o_col);
  double *sval = static_cast<double *>(dp(jb.o_sval));
  SlotState *dst = sv->d_states + slot;
  if (jb.nnz > 0) sys_band_kernel<<<g, kBlock, 0, s>>>(jb.nrow, rp, ci, dst);
  if (!mixed) {
    sys_fill_kernel<<<g, kBlock, 0, s>>>(jb.nrow, rp, ci, val, soff, slen, dst, col, sval,
                                         static_cast<double *>(dp(jb.o_invd)), nullptr,
                                         nullptr, nullptr, nullptr, nullptr,
                                         &sv->d_states[slot].bad);
  } else {
    sys_fill_kernel<<<g, kBlock, 0, s>>>(
        jb.nrow, rp, ci, val, soff, slen, dst, col, sval, nullptr,
        static_cast<const long long *>(dp(jb.o_soff32)),
        static_cast<const int *>(dp(jb.o_slen32)), dp(jb.o_col32),
        static_cast<float *>(dp(jb.o_sval32)), static_cast<float *>(dp(jb.o_invd32)),
        &sv->d_states[slot].bad);
  }
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

// The host side of an admission, on the calling thread: the system's
// descriptor over its block, its slot and state.
void admit_describe(hflmp_server *sv, Job &jb, int slot) {
  char *base = sv->arena + jb.off;
  auto dp = [&](size_t o) { return static_cast<void *>(base + o); };
  const bool mixed = jb.mode == kModeMixed;
  SysDesc &d = jb.d;
  std::memset(&d, 0, sizeof(d));
  d.nrow = jb.nrow;
  d.nslice = jb.nslice;
  d.nunits = jb.nunits;
  d.mode = jb.mode;
  d.soff = static_cast<const long long *>(dp(jb.o_soff));
  d.slen = static_cast<const int *>(dp(jb.o_slen));
  d.rlen = static_cast<const unsigned char *>(dp(jb.o_rlen));
  d.col = dp(jb.o_col);
  d.sval = static_cast<double *>(dp(jb.o_sval));
  d.b = static_cast<const double *>(dp(jb.o_b));
  d.x = static_cast<double *>(dp(jb.o_x));
  d.part = static_cast<double *>(dp(jb.o_part));
  if (!mixed) {
    d.invd = static_cast<const double *>(dp(jb.o_invd));
    d.r = static_cast<double *>(dp(jb.o_r));
    d.p = static_cast<double *>(dp(jb.o_p));
    d.z = static_cast<double *>(dp(jb.o_z));
    d.Ap = static_cast<double *>(dp(jb.o_Ap));
  } else {
    d.soff32 = static_cast<const long long *>(dp(jb.o_soff32));
    d.slen32 = static_cast<const int *>(dp(jb.o_slen32));
    d.col32 = dp(jb.o_col32);
    d.sval32 = static_cast<const float *>(dp(jb.o_sval32));
    d.invd32 = static_cast<const float *>(dp(jb.o_invd32));
    d.r32 = static_cast<float *>(dp(jb.o_r32));
    d.p32 = static_cast<float *>(dp(jb.o_p32));
    d.z32 = static_cast<float *>(dp(jb.o_z32));
    d.Ap32 = static_cast<float *>(dp(jb.o_Ap32));
    d.c32 = static_cast<float *>(dp(jb.o_c32));
  }
  d.st = sv->d_states + slot;
  jb.slot = slot;
  jb.state = kActive;
  jb.t_admit = Clock::now();
}

// The grouped systems' chunk on the context stream: a replay of the
// captured graph (captured again after the set changed), or direct launches.
int launch_group(hflmp_server *sv, TableSet &t) {
  hflmp_ctx *ctx = sv->ctx;
  hipStream_t s = ctx->stream;
  if (!sv->use_graph) {
    enqueue_chunk(sv, t, s);
    HFLMP_LAUNCH_CHECK(ctx);
    return HFLMP_OK;
  }
  if (!t.graph_exec) {
    // a graph left by a failed instantiation is freed before recapturing;
    // the capture is always ended, and a partial graph destroyed
    drop_graph(t);
    HFLMP_HIP_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_chunk(sv, t, s);
    const hipError_t launch = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &g);
    if (end != hipSuccess || launch != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return set_error(ctx, HFLMP_EHIP, "server: graph capture failed (%s)",
                       hipGetErrorString(end != hipSuccess ? end : launch));
    }
    t.graph = g;
    const hipError_t inst = hipGraphInstantiate(&t.graph_exec, t.graph, nullptr, nullptr, 0);
    if (inst != hipSuccess) {
      t.graph_exec = nullptr;
      drop_graph(t);
      return set_error(ctx, HFLMP_EHIP, "server: graph instantiation failed (%s)",
                       hipGetErrorString(inst));
    }
    ++sv->graph_captures;
  }
  HFLMP_HIP_CHECK(ctx, hipGraphLaunch(t.graph_exec, s));
  return HFLMP_OK;
}

// After a solo system's admission kernels: read its state back once (the
// FP64 layout's column width, decided by the band kernel, and whether the
// system was refused), then queue its start on the solo stream after the
// admission work (x = 0 first).
int start_solo(hflmp_server *sv, Job &jb, int slot) {
  hflmp_ctx *ctx = sv->ctx;
  hipStream_t s = sv->admit_stream;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->h_admit, sv->d_states + slot, sizeof(SlotState),
                                      hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  sv->stage_used = 0;
  jb.col16_64 = sv->h_admit->col16;
  if (sv->h_admit->s64.status == kRefused) return HFLMP_OK;  // retired at the next step
  int rc = fill_zero_on(ctx, jb.d.x, 8 * (size_t)jb.nrow, s);
  if (rc) return rc;
  HFLMP_HIP_CHECK(ctx, hipEventRecord(sv->ev_admitted, s));
  HFLMP_HIP_CHECK(ctx, hipStreamWaitEvent(sv->solo_stream, sv->ev_admitted, 0));
  if (jb.col16_64)
    solo_start<short>(jb, sv->solo_stream);
  else
    solo_start<int>(jb, sv->solo_stream);
  HFLMP_LAUNCH_CHECK(ctx);
  ++sv->solo_runs;
  return HFLMP_OK;
}

bool slot_running(const SlotState &st) {
  if (st.mode == kModeMixed) return st.status == kIterate || st.status == kFresh;
  return st.s64.status == kIterate || st.s64.status == kFresh;
}

// One step's work, in this order:
//   launch_chunk:  queue the active systems' next chunk (the grouped
//                  launches on the context stream, the solo systems' on the
//                  solo stream), then the copy of all states to the host,
//                  behind the chunk still in flight;
//   admit_queued:  admission of queued systems on the admission stream,
//                  while the chunks run (host uploads, layout kernels);
//   wait_oldest:   wait for the older chunk, retire the systems that
//                  stopped in it;
//   join_admitted: wait for the admission work; the admitted systems join
//                  the active set and run from the next chunk queued.
// A slot or block is released only when no queued chunk lists it, so an
// admission never writes memory that a chunk in flight reads.
// Queues chunk `launched` on its table set (rebuilt if the active set has
// changed since the set was last built), the solo systems' chunks, the copy
// of all states into the set's pinned buffer and the set's event.  At most
// one chunk is in flight when this is called, and it ran on the other set.
int launch_chunk(hflmp_server *sv, bool *launched) {
  hflmp_ctx *ctx = sv->ctx;
  hipStream_t s = ctx->stream;
  *launched = false;
  if (sv->order.empty() && sv->solo.empty()) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, sv->launched - sv->completed <= 1, "server: two chunks already in flight");
  TableSet &t = sv->sets[sv->launched & 1];
  const auto t0 = Clock::now();
  HFLMP_HIP_CHECK(ctx, hipEventRecord(t.start, s));
  if (t.version != sv->version) {
    int rc = rebuild_tables(sv, t);
    if (rc) return rc;
  }
  if (t.S > 0) {
    int rc = launch_group(sv, t);
    if (rc) return rc;
  }
  sv->capture_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
  if (!sv->solo.empty()) {
    // the solo systems' chunks on their stream, beside the group's; the
    // state copy waits for both
    for (int slot : sv->solo) {
      const Job &jb = sv->jobs[sv->slot_job[slot]];
      if (jb.col16_64)
        solo_chunk<short>(ctx, jb, sv->solo_stream, sv->chunk);
      else
        solo_chunk<int>(ctx, jb, sv->solo_stream, sv->chunk);
    }
    HFLMP_LAUNCH_CHECK(ctx);
    HFLMP_HIP_CHECK(ctx, hipEventRecord(sv->ev_solo_done, sv->solo_stream));
    HFLMP_HIP_CHECK(ctx, hipStreamWaitEvent(s, sv->ev_solo_done, 0));
  }
  for (const std::vector<int> *list : {&sv->order, &sv->solo})
    for (int slot : *list)
      if (sv->first_chunk[slot] < 0) sv->first_chunk[slot] = sv->launched;
  {
    const int rc = xfer(ctx, t.h_states, sv->d_states, sizeof(SlotState) * sv->max_active, s);
    if (rc) return rc;
  }
  HFLMP_HIP_CHECK(ctx, hipEventRecord(t.done, s));
  ++sv->launched;
  ++sv->steps;
  *launched = true;
  return HFLMP_OK;
}

void admitter_run(hflmp_server *sv) {
  (void)hipSetDevice(sv->actx.device);
  for (;;) {
    std::pair<const Job *, int> w;
    bool failed;
    {
      std::unique_lock<std::mutex> l(sv->am);
      sv->awake.wait(l, [sv] { return sv->astop || !sv->apending.empty(); });
      if (sv->apending.empty()) return;
      w = sv->apending.front();
      sv->apending.pop_front();
      ++sv->abusy;
      failed = sv->arc != HFLMP_OK;
    }
    // after a failure the server reports it at the next join; later
    // admissions are not started
    int rc = failed ? HFLMP_OK : admit_upload(sv, &sv->actx, *w.first, w.second);
    if (!failed && !rc && hipEventRecord(sv->jev[w.second], sv->admit_stream) != hipSuccess)
      rc = set_error(&sv->actx, HFLMP_EHIP, "server: admission event not recorded");
    std::lock_guard<std::mutex> g(sv->am);
    if (rc && sv->arc == HFLMP_OK) sv->arc = rc;
    if (!failed && !rc) sv->aready[w.second] = 1;
    if (--sv->abusy == 0 && sv->apending.empty()) sv->aidle.notify_all();
  }
}

// Waits until the admission thread has queued all admission work posted so
// far; returns its first failure (with its message in the context).
int admitter_drain(hflmp_server *sv) {
  std::unique_lock<std::mutex> l(sv->am);
  sv->aidle.wait(l, [sv] { return sv->abusy == 0 && sv->apending.empty(); });
  if (sv->arc != HFLMP_OK) {
    std::memcpy(sv->ctx->last_error, sv->actx.last_error, sizeof(sv->ctx->last_error));
    return sv->arc;
  }
  return HFLMP_OK;
}

// Admits first come, first served while a slot is free and the block fits.
int admit_queued(hflmp_server *sv) {
  const auto t0 = Clock::now();
  while (!sv->queue.empty()) {
    int slot = -1;
    for (int k = 0; k < sv->max_active; ++k)
      if (sv->slot_job[k] < 0) {
        slot = k;
        break;
      }
    if (slot < 0) break;
    Job &jb = sv->jobs[sv->queue.front()];
    if (!sv->alloc.alloc(jb.bytes, &jb.off)) break;
    admit_describe(sv, jb, slot);
    if (jb.solo) {
      // started at once after its admission work: admitted on this thread,
      // after the admission thread's work
      int rc = admitter_drain(sv);
      if (!rc) rc = admit_upload(sv, sv->ctx, jb, slot);
      if (!rc) rc = start_solo(sv, jb, slot);
      if (!rc && hipEventRecord(sv->jev[slot], sv->admit_stream) != hipSuccess)
        rc = set_error(sv->ctx, HFLMP_EHIP, "server: admission event not recorded");
      if (rc) {
        sv->alloc.release(jb.off, jb.bytes);
        return rc;
      }
      std::lock_guard<std::mutex> g(sv->am);
      sv->aready[slot] = 1;
    } else {
      {
        std::lock_guard<std::mutex> g(sv->am);
        sv->aready[slot] = 0;
        sv->apending.emplace_back(&jb, slot);
      }
      sv->awake.notify_one();
    }
    // The CSR arrays at the block's end are read only by the admission
    // kernels just queued (band check, SELL-64 fill); every later use of
    // that range (another system's uploads and kernels) is queued on the
    // same stream, after them.  (A solo system's start on the solo stream
    // reads its layout, not these arrays.)
    sv->alloc.release(jb.off + jb.o_rp, jb.bytes - jb.o_rp);
    jb.kept = jb.o_rp;
    sv->slot_job[slot] = jb.ticket;
    sv->joining.push_back(slot);
    sv->queue.pop_front();
  }
  sv->admit_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
  return HFLMP_OK;
}

// Retires systems: x to the callers' buffers (0 if refused) and the
// results from their final states.  The x of all of them go device ->
// pinned staging as one queue of asynchronous copies on the result stream
// (beside the chunk in flight, which no longer writes a stopped system's
// x), one wait, then host copies; what does not fit the staging area is
// copied directly.  Their slots and arena blocks are released by the
// caller (release_slot), when no chunk in flight lists them.
int retire_systems(hflmp_server *sv, const std::vector<int> &slots,
                   const std::vector<const SlotState *> &states, int *n_done) {
  hflmp_ctx *ctx = sv->ctx;
  const size_t none = ~(size_t)0;
  std::vector<size_t> at(slots.size(), none);
  size_t used = 0;
  for (size_t i = 0; i < slots.size(); ++i) {
    const SlotState &st = *states[i];
    const Job &jb = sv->jobs[sv->slot_job[slots[i]]];
    const bool refused = (st.mode == kModeMixed ? st.status : st.s64.status) == kRefused;
    const size_t bytes = 8 * (size_t)jb.nrow;
    if (refused || used + bytes > sv->rstage_cap) continue;
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->rstage + used, jb.d.x, bytes,
                                        hipMemcpyDeviceToHost, sv->result_stream));
    at[i] = used;
    used = align256(used + bytes);
  }
  if (used) HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(sv->result_stream));
  for (size_t i = 0; i < slots.size(); ++i) {
    const SlotState &st = *states[i];
    Job &jb = sv->jobs[sv->slot_job[slots[i]]];
    const bool refused = (st.mode == kModeMixed ? st.status : st.s64.status) == kRefused;
    const size_t bytes = 8 * (size_t)jb.nrow;
    if (refused)
      std::memset(jb.x_out, 0, bytes);
    else if (at[i] != none)
      host_copy(reinterpret_cast<char *>(jb.x_out), sv->rstage + at[i], bytes);
    else
      HFLMP_HIP_CHECK(ctx, hipMemcpy(jb.x_out, jb.d.x, bytes, hipMemcpyDeviceToHost));
    if (st.mode == kModeMixed) {
      jb.iters = st.inner_total;
      jb.status = st.status;
      jb.outer = st.outer;
      jb.res_init = st.res0;
      jb.res_final = st.res;
    } else {
      jb.iters = st.s64.iter;
      jb.status = st.s64.status;
      jb.outer = 0;
      jb.res_init = st.s64.res0;
      jb.res_final = st.s64.res;
    }
    jb.bad_diag = (int)st.bad;
    jb.col_bytes = st.col16 ? 2 : 4;
    jb.t_done = Clock::now();
    jb.state = kDone;
    ++*n_done;
  }
  return HFLMP_OK;
}

void release_slot(hflmp_server *sv, const Deferred &r) {
  sv->alloc.release(r.off, r.bytes);
  sv->slot_job[r.slot] = -1;
  sv->first_chunk[r.slot] = -1;
}

Deferred slot_block(hflmp_server *sv, int slot) {
  const Job &jb = sv->jobs[sv->slot_job[slot]];
  return Deferred{slot, jb.off, jb.kept};
}

// Waits for the oldest chunk in flight (k) and retires the systems that
// stopped in it.  Systems retired at the previous wait were still listed by
// chunk k; now that it has finished their slots and blocks are released.
// The systems retired now stay listed by the chunk queued after k (their
// kernels skip them), so their release waits for it -- or happens at once
// when no chunk is in flight.
int wait_oldest(hflmp_server *sv, int *n_done) {
  hflmp_ctx *ctx = sv->ctx;
  const long long k = sv->completed;
  TableSet &t = sv->sets[k & 1];
  const auto t_wait = Clock::now();
  HFLMP_HIP_CHECK(ctx, hipEventSynchronize(t.done));
  sv->wait_seconds += std::chrono::duration<double>(Clock::now() - t_wait).count();
  float ms = 0;
  if (hipEventElapsedTime(&ms, t.start, t.done) == hipSuccess) sv->device_seconds += 1e-3 * ms;
  ++sv->completed;
  for (const Deferred &r : sv->deferred) release_slot(sv, r);
  sv->deferred.clear();
  std::vector<int> out;
  std::vector<const SlotState *> out_st;
  for (std::vector<int> *list : {&sv->order, &sv->solo}) {
    std::vector<int> keep;
    for (int slot : *list) {
      // a system that joined after chunk k was queued is not in its states
      const SlotState &st = t.h_states[slot];
      if (sv->first_chunk[slot] < 0 || sv->first_chunk[slot] > k || slot_running(st)) {
        keep.push_back(slot);
        continue;
      }
      out.push_back(slot);
      out_st.push_back(&st);
      if (list == &sv->order) ++sv->version;
    }
    list->swap(keep);
  }
  const auto t_ret = Clock::now();
  const int rc = retire_systems(sv, out, out_st, n_done);
  if (rc) return rc;
  sv->retire_seconds += std::chrono::duration<double>(Clock::now() - t_ret).count();
  for (int slot : out) sv->deferred.push_back(slot_block(sv, slot));
  if (sv->launched == sv->completed) {
    for (const Deferred &r : sv->deferred) release_slot(sv, r);
    sv->deferred.clear();
  }
  return HFLMP_OK;
}

// A system refused at admission (a column out of range) retires here and
// never enters the tables: the grouped kernels issue their gathers before
// they read a system's state.
int join_admitted(hflmp_server *sv, int *n_done) {
  hflmp_ctx *ctx = sv->ctx;
  if (sv->joining.empty()) return HFLMP_OK;
  const auto t0 = Clock::now();
  // With systems running, only those whose admission work has finished
  // join (the others at a later step), so the next chunk is never held up
  // by admission; with nothing running, wait for all of them.
  const bool block = sv->order.empty() && sv->solo.empty() && sv->launched == sv->completed;
  if (block) {
    const int rc = admitter_drain(sv);
    if (rc) return rc;
  }
  std::vector<int> cand, ready, later;
  {
    std::lock_guard<std::mutex> g(sv->am);
    if (sv->arc != HFLMP_OK) {
      std::memcpy(ctx->last_error, sv->actx.last_error, sizeof(ctx->last_error));
      return sv->arc;
    }
    for (int slot : sv->joining) (sv->aready[slot] ? cand : later).push_back(slot);
  }
  for (int slot : cand) {
    if (block) {
      HFLMP_HIP_CHECK(ctx, hipEventSynchronize(sv->jev[slot]));
    } else {
      const hipError_t e = hipEventQuery(sv->jev[slot]);
      if (e == hipErrorNotReady) {
        later.push_back(slot);
        continue;
      }
      HFLMP_HIP_CHECK(ctx, e);
    }
    ready.push_back(slot);
  }
  sv->joining.swap(later);
  // each ready system's state as admission left it: the column width that
  // the band kernel decided (the grouped kernels take it from the system
  // table rather than from the state), refusal; read on the result stream,
  // which holds no other work now (the admission stream may hold later
  // systems' work)
  for (int slot : ready)
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->h_init + slot, sv->d_states + slot,
                                        sizeof(SlotState), hipMemcpyDeviceToHost,
                                        sv->result_stream));
  if (!ready.empty()) HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(sv->result_stream));
  sv->join_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
  for (int slot : ready) {
    const SlotState &st = sv->h_init[slot];
    if ((st.mode == kModeMixed ? st.status : st.s64.status) == kRefused) {
      const int rc = retire_systems(sv, {slot}, {&st}, n_done);
      if (rc) return rc;
      release_slot(sv, slot_block(sv, slot));  // in no chunk's tables
      continue;
    }
    // both layouts' column width, as the band kernel decided it
    sv->jobs[sv->slot_job[slot]].d.col16_64 = st.col16;
    sv->jobs[sv->slot_job[slot]].d.col16 = st.col16;
    sv->first_chunk[slot] = -1;
    if (sv->jobs[sv->slot_job[slot]].solo) {
      sv->solo.push_back(slot);
    } else {
      sv->order.push_back(slot);
      ++sv->version;
    }
  }
  sv->peak_active = std::max(sv->peak_active, (int)(sv->order.size() + sv->solo.size()));
  sv->peak_in_use = std::max(sv->peak_in_use, sv->budget - sv->alloc.total_free());
  return HFLMP_OK;
}


}  // namespace

extern "C" {

int hflmp_server_create(hflmp_ctx *ctx, const hflmp_server_config *cfg, hflmp_server **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, cfg && out, "server_create: null pointer");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, cfg->hbm_budget >= (1u << 20) && cfg->max_active >= 1 &&
                         cfg->max_active <= 4096 && cfg->chunk_iters >= 1 &&
                         cfg->chunk_iters <= 256 && cfg->mixed_min_nnz >= 0 &&
                         cfg->mixed_max_nnz >= 0 && cfg->inner_rtol >= 0 &&
                         cfg->inner_rtol < 1,
                "server_create: need hbm_budget >= 1 MiB, 1 <= max_active <= 4096, "
                "1 <= chunk_iters <= 256, mixed_min_nnz >= 0, mixed_max_nnz >= 0, "
                "0 <= inner_rtol < 1");
  auto *sv = new hflmp_server();
  sv->ctx = ctx;
  sv->budget = cfg->hbm_budget;
  sv->max_active = cfg->max_active;
  sv->chunk = cfg->chunk_iters;
  sv->use_graph = cfg->use_graphs ? 1 : 0;
  sv->mixed_min_nnz = cfg->mixed_min_nnz;
  sv->mixed_max_nnz = cfg->mixed_max_nnz;
  // FP32 CG driven far below its starting residual drifts from the true
  // residual on badly conditioned systems; frequent replacement keeps the
  // reliable updates effective
  sv->inner_rtol = cfg->inner_rtol > 0 ? cfg->inner_rtol : 1e-2;
  sv->solo_min_nnz = cfg->solo_min_nnz;
  auto fail = [&](const char *what) {
    hflmp_server_destroy(sv);
    return set_error(ctx, HFLMP_ENOMEM, "server_create: %s", what);
  };
  if (hipMalloc(&sv->arena, sv->budget) != hipSuccess) return fail("arena allocation failed");
  if (hipMalloc(&sv->d_states, sizeof(SlotState) * sv->max_active) != hipSuccess ||
      hipHostMalloc(&sv->h_init, sizeof(SlotState) * sv->max_active) != hipSuccess ||
      hipHostMalloc(&sv->h_admit, sizeof(SlotState)) != hipSuccess)
    return fail("table allocation failed");
  for (TableSet &t : sv->sets)
    if (hipMalloc(&t.d_sys, sizeof(SysDesc) * sv->max_active) != hipSuccess ||
        hipHostMalloc(&t.h_sys, sizeof(SysDesc) * sv->max_active, hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostMalloc(&t.h_states, sizeof(SlotState) * sv->max_active,
                      hipHostMallocCoherent) != hipSuccess ||
        hipEventCreate(&t.start) != hipSuccess || hipEventCreate(&t.done) != hipSuccess)
      return fail("table allocation failed");
  sv->stage_cap = std::min(std::max(sv->budget / 8, (size_t)16 << 20), (size_t)256 << 20);
  sv->rstage_cap = sv->stage_cap;
  if (hipHostMalloc(&sv->stage, sv->stage_cap) != hipSuccess ||
      hipHostMalloc(&sv->rstage, sv->rstage_cap) != hipSuccess)
    return fail("staging allocation failed");
  // the admission and solo streams keep the default priority (as every
  // stream here)
  if (hipStreamCreateWithFlags(&sv->admit_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&sv->result_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&sv->solo_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&sv->ev_admitted, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&sv->ev_solo_done, hipEventDisableTiming) != hipSuccess)
    return fail("stream or event creation failed");
  sv->alloc.reset(sv->budget);
  sv->slot_job.assign(sv->max_active, -1);
  sv->first_chunk.assign(sv->max_active, -1);
  sv->jev.assign(sv->max_active, nullptr);
  sv->aready.assign(sv->max_active, 0);
  for (hipEvent_t &e : sv->jev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail("event creation failed");
  sv->actx = *ctx;
  sv->actx.last_error[0] = 0;
  try {
    sv->admitter = std::thread(admitter_run, sv);
  } catch (...) {
    return fail("admission thread not started");
  }
  *out = sv;
  return HFLMP_OK;
}

void hflmp_server_destroy(hflmp_server *sv) {
  if (!sv) return;
  // the admission thread finishes the work already posted, then ends
  {
    std::lock_guard<std::mutex> g(sv->am);
    sv->astop = true;
  }
  sv->awake.notify_all();
  if (sv->admitter.joinable()) sv->admitter.join();
  if (sv->admit_stream) (void)hipStreamSynchronize(sv->admit_stream);
  if (sv->result_stream) (void)hipStreamSynchronize(sv->result_stream);
  if (sv->solo_stream) (void)hipStreamSynchronize(sv->solo_stream);
  if (sv->ctx) (void)hipStreamSynchronize(sv->ctx->stream);
  if (sv->admit_stream) (void)hipStreamDestroy(sv->admit_stream);
  if (sv->result_stream) (void)hipStreamDestroy(sv->result_stream);
  if (sv->solo_stream) (void)hipStreamDestroy(sv->solo_stream);
  if (sv->ev_admitted) (void)hipEventDestroy(sv->ev_admitted);
  for (hipEvent_t e : sv->jev)
    if (e) (void)hipEventDestroy(e);
  if (sv->ev_solo_done) (void)hipEventDestroy(sv->ev_solo_done);
  if (sv->h_admit) (void)hipHostFree(sv->h_admit);
  for (TableSet &t : sv->sets) {
    drop_graph(t);
    if (t.d_sys) (void)hipFree(t.d_sys);
    if (t.d_unit) (void)hipFree(t.d_unit);
    if (t.h_sys) (void)hipHostFree(t.h_sys);
    if (t.h_unit) (void)hipHostFree(t.h_unit);
    if (t.d_rec) (void)hipFree(t.d_rec);
    if (t.d_grp) (void)hipFree(t.d_grp);
    if (t.h_grp) (void)hipHostFree(t.h_grp);
    if (t.d_quad) (void)hipFree(t.d_quad);
    if (t.h_quad) (void)hipHostFree(t.h_quad);
    if (t.h_states) (void)hipHostFree(t.h_states);
    if (t.start) (void)hipEventDestroy(t.start);
    if (t.done) (void)hipEventDestroy(t.done);
  }
  (void)hipFree(sv->arena);
  (void)hipFree(sv->d_states);
  if (sv->h_init) (void)hipHostFree(sv->h_init);
  if (sv->stage) (void)hipHostFree(sv->stage);
  if (sv->rstage) (void)hipHostFree(sv->rstage);
  delete sv;
}

size_t hflmp_server_footprint(int nrow, int nnz, const int *row_ptr, const int *col,
                              int precision) {
  if (nrow <= 0 || nnz < 0 || !row_ptr || (precision != 64 && precision != 32)) return 0;
  Job jb;
  jb.nrow = nrow;
  jb.nnz = nnz;
  jb.rp = row_ptr;
  jb.ci = col;
  jb.mode = precision;
  plan_job(jb);
  return jb.bytes;
}

int hflmp_server_submit(hflmp_server *sv, int nrow, int nnz, const int *row_ptr,
                        const int *col, const double *val, const double *b, double *x,
                        double atol, double rtol, int maxit, int precision,
                        long long *ticket) {
  if (!sv) return HFLMP_EINVAL;
  hflmp_ctx *ctx = sv->ctx;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz >= 0 && row_ptr && b && x && ticket &&
                         (nnz == 0 || (col && val)),
                "server_submit: bad system");
  HFLMP_REQUIRE(ctx, maxit > 0 && atol >= 0 && rtol >= 0, "server_submit: bad tolerances");
  HFLMP_REQUIRE(ctx, precision == 0 || precision == 64 || precision == 32,
                "server_submit: precision must be 0 (by size), 64 or 32");
  HFLMP_REQUIRE(ctx, row_ptr[0] == 0 && row_ptr[nrow] == nnz,
                "server_submit: row_ptr does not span nnz");
  for (int r = 0; r < nrow; ++r)
    HFLMP_REQUIRE(ctx, row_ptr[r + 1] - row_ptr[r] >= 0 && row_ptr[r + 1] - row_ptr[r] <= 255,
                  "server_submit: row %d has %d entries (0..255 supported)", r,
                  row_ptr[r + 1] - row_ptr[r]);
  Job jb;
  jb.ticket = sv->next_ticket++;
  jb.nrow = nrow;
  jb.nnz = nnz;
  jb.rp = row_ptr;
  jb.ci = col;
  jb.val = val;
  jb.b = b;
  jb.x_out = x;
  jb.atol = atol;
  jb.rtol = rtol;
  jb.maxit = maxit;
  if (precision == 0)
    precision = (sv->mixed_min_nnz > 0 && nnz >= sv->mixed_min_nnz &&
                 (sv->mixed_max_nnz == 0 || nnz < sv->mixed_max_nnz))
                    ? 32
                    : 64;
  jb.mode = precision == 32 ? kModeMixed : kModeFp64;
  jb.solo = jb.mode == kModeFp64 && sv->solo_min_nnz > 0 && nnz >= sv->solo_min_nnz;
  plan_job(jb);
  HFLMP_REQUIRE(ctx, jb.bytes <= sv->budget,
                "server_submit: footprint %zu B exceeds the HBM budget %zu B", jb.bytes,
                sv->budget);
  jb.t_submit = Clock::now();
  *ticket = jb.ticket;
  sv->queue.push_back(jb.ticket);
  sv->jobs.emplace(jb.ticket, std::move(jb));
  return HFLMP_OK;
}

int hflmp_server_step(hflmp_server *sv, int *n_done) {
  if (!sv) return HFLMP_EINVAL;
  int done = 0;
  // A step with nothing to run admits, and runs the admitted systems at once.
  // Two chunks in flight after the launch: wait for the older one.  One
  // chunk in flight that this step did not queue (nothing left to run):
  // wait for it.  One chunk just queued and none before: leave it running.
  for (int pass = 0; pass < 2; ++pass) {
    bool launched = false;
    int rc = launch_chunk(sv, &launched);
    if (!rc) rc = admit_queued(sv);
    const long long in_flight = sv->launched - sv->completed;
    if (!rc && (in_flight >= 2 || (!launched && in_flight == 1))) rc = wait_oldest(sv, &done);
    if (!rc) rc = join_admitted(sv, &done);
    if (rc) return rc;
    if (launched || (sv->order.empty() && sv->solo.empty())) break;
  }
  if (n_done) *n_done = done;
  return HFLMP_OK;
}

int hflmp_server_poll(hflmp_server *sv, long long ticket, hflmp_server_result *res) {
  if (!sv || !res) return HFLMP_EINVAL;
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end())
    return set_error(sv->ctx, HFLMP_EINVAL, "server_poll: unknown ticket %lld", ticket);
  const Job &jb = it->second;
  res->state = jb.state;
  res->iters = jb.iters;
  res->status = jb.state == kDone ? (jb.status == kIterate ? kMaxIts : jb.status) : -1;
  res->res_init = jb.res_init;
  res->res_final = jb.res_final;
  res->zero_diagonals = jb.bad_diag;
  res->footprint = jb.bytes;
  auto sec = [](Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  res->queued_seconds = jb.state >= kActive ? sec(jb.t_submit, jb.t_admit) : -1.0;
  res->latency_seconds = jb.state == kDone ? sec(jb.t_submit, jb.t_done) : -1.0;
  res->precision = jb.mode;
  res->corrections = jb.outer;
  res->col_bytes = jb.col_bytes;
  return HFLMP_OK;
}

int hflmp_server_release(hflmp_server *sv, long long ticket) {
  if (!sv) return HFLMP_EINVAL;
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end() || it->second.state != kDone)
    return set_error(sv->ctx, HFLMP_EINVAL, "server_release: ticket %lld is not done", ticket);
  sv->jobs.erase(it);
  return HFLMP_OK;
}

int hflmp_server_stats(const hflmp_server *sv, hflmp_server_stats_t *out) {
  if (!sv || !out) return HFLMP_EINVAL;
  out->queued = (int)sv->queue.size();
  out->active = (int)(sv->order.size() + sv->solo.size());
  out->active_units = sv->launched > 0 ? sv->sets[(sv->launched - 1) & 1].U : 0;
  out->arena_free = sv->alloc.total_free();
  out->arena_largest_free = sv->alloc.largest();
  out->steps = sv->steps;
  out->graph_captures = sv->graph_captures;
  out->peak_active = sv->peak_active;
  out->peak_arena_in_use = sv->peak_in_use;
  out->solo_runs = sv->solo_runs;
  out->admit_seconds = sv->admit_seconds;
  out->capture_seconds = sv->capture_seconds;
  out->wait_seconds = sv->wait_seconds;
  out->join_seconds = sv->join_seconds;
  out->retire_seconds = sv->retire_seconds;
  out->chunk_device_seconds = sv->device_seconds;
  return HFLMP_OK;
}

}  // extern "C"
