// Matrix-free stencil operator on a structured mesh (the HIP counterpart of
// StencilOperator, linear_algebra/stencil_operator.h:43-330).
//
// The mesh has nc[0] x nc[1] (x nc[2]) cells of Q1 (2D, 3D) or Q2 (2D)
// elements; nodes are numbered lexicographically (x fastest) and so are the
// local nodes of a cell.  Each cell c carries one local matrix E_m
// (m = cell_lm[c], or 0 when there is one matrix) and a factor f_c.  The
// operator is
//     y = beta y + alpha sum_c f_c S_c^T E_m S_c x,  then  y_d = x_d for
// every fixed dof d (Dirichlet rows of the assembled matrix).
//
// The reference computes it cell by cell: gather the cell's x, tmp = E x
// (each entry summed from 0 over ascending local columns), lhs = 0 +
// (alpha f_c) tmp, scatter-add lhs into y, cells in ascending order.  Here
// one thread owns one node (row) and visits the node's cells in ascending
// cell order, forming the same tmp entry and adding the same lhs value in
// the same order, so y is bit-identical to the cell loop without atomics.
//
// Data per product: x and y (8 N B each in FP64), f (8 B per cell), a bit
// per node for the fixed dofs; no indices and no matrix values.  A
// workgroup's 256 consecutive nodes read x along (2 order + 1)^(dim - 1)
// neighbour lines; those line segments are staged in LDS with one load per
// value, and each node takes its (2 order + 1)^dim neighbourhood from there
// into registers.  Every other operand is read through the caches.
#pragma once

#include "hflmp_spmv_tile.hpp"

struct hflmp_stencil {
  int device;
  int order, dim, nloc, nmat;
  int nc[3];  // cells per direction (1 in z for 2D)
  int np[3];  // nodes per direction (1 in z for 2D)
  long long nnode, ncell;
  int nfixed;
  int value_bytes;
  void *E;                  // nmat * nloc * nloc local matrices
  void *factor;             // ncell cell factors
  unsigned char *cell_lm;   // ncell matrix indices, or null (nmat == 1)
  unsigned int *fixed_bits; // bit per node
};

// Buffer descriptor word 3 for raw (untyped) buffer access on gfx9-family
// GPUs (gfx950 included).
constexpr int kRsrcWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, unsigned off);
template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// A load that starts before the buffer returns zeros for all its words,
// one that runs past the end only for the words beyond it.  The loads to
// the left of a node (negative x offsets, the kx = 0 cells) go through this
// select, which maps an offset before the start (negative as int; ranges
// stay below 2^31 bytes) to one past the end and keeps the compiler from
// fusing such a load with its right neighbour, whose in-range words the
// fused load would lose.
__device__ __forceinline__ unsigned before_start(unsigned off) {
  return (int)off < 0 ? 0x80000000u : off;
}

// Division of n < 2^31 by a fixed d >= 1 as a multiply-high (Granlund and
// Montgomery, PLDI 1994): with s = ceil(log2 d) and M = floor(2^(32+s) / d)
// + 1 < 2^33, n / d = floor(n M / 2^(32+s)); the error n (M - 2^(32+s)/d) /
// 2^(32+s) < 2^31 / 2^(32+s) <= 1 / (2 d) never crosses an integer.
struct DivMagic {
  unsigned long long m;
  int shift;
  __host__ void init(unsigned d) {
    int s = 0;
    while ((1ull << s) < d) ++s;
    shift = 32 + s;
    m = ((1ull << shift) / d) + 1;
  }
  __device__ __forceinline__ int div(int n) const {
    return (int)(((unsigned long long)(unsigned)n * m) >> shift);
  }
};

// Operator view for the fused kernels: unit = 256 consecutive nodes.
// ORDER, DIM fixed at compile time; MULTI: per-cell matrix index.
template <typename T, int DIM, int ORDER, bool MULTI>
struct StencilOp {
  using value_type = T;
  static constexpr int Q = ORDER + 1;                        // local nodes per direction
  static constexpr int NLOC = DIM == 2 ? Q * Q : Q * Q * Q;  // local nodes per cell
  static constexpr int W = 2 * ORDER + 1;                    // neighbourhood width
  static constexpr int KZ = DIM == 3 ? 2 : 1;
  static constexpr int WZ = DIM == 3 ? W : 1;
  // staged x: one segment per (dz, dy) neighbour line of the unit's 256
  // consecutive nodes, ORDER values wider on each side
  static constexpr int kSeg = kBlock + 2 * ORDER;
  static constexpr int kLds = WZ * W * kSeg;
  int nrow;
  int ncx, ncy, ncz, px, py, pz;
  DivMagic dpx, dpy;  // division by px and py
  const T *E;
  const T *f;
  const unsigned char *lm;
  const unsigned int *fixed;

  __device__ __forceinline__ long long thread_row(long long unit) const {
    return unit * kBlock + threadIdx.x;
  }

  // The x values the unit's nodes read, for lds (kLds values): for each
  // neighbour line (dz, dy), the index range of the unit shifted by that
  // line's offset and widened by ORDER; indices outside the vector give 0,
  // as the buffer loads do.  stage_load issues the loads, stage_store
  // writes lds and waits for the workgroup.
  // One register per line (all lanes: the unit's 256 nodes) and one for
  // the 2 ORDER extra values of every line (lanes 0 .. lines * 2 ORDER - 1).
  static constexpr int kLines = WZ * W;
  static constexpr int kStageIt = kLines + 1;
  __device__ __forceinline__ int line_offset(int sg) const {
    const int dz = sg / W, dy = sg - dz * W;
    return (DIM == 3 ? (dz - ORDER) * px * py : 0) + (dy - ORDER) * px - ORDER;
  }
  // The loads go through a buffer descriptor over x: an index before the
  // start (negative, so a byte offset beyond 2^31 as unsigned) or past the
  // end reads 0 with no compare.  A lane's loads are whole lines apart, so
  // none is fused with a neighbour (the case before_start guards).
  __device__ __forceinline__ void stage_load(long long unit, const T *__restrict__ x,
                                             T (&v)[kStageIt]) const {
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T *>(x), 0, nrow * (int)sizeof(T), kRsrcWord3);
    // 32-bit index arithmetic: nodes and offsets stay below 2^31
    const int i0 = (int)unit * kBlock;
    const int t = (int)threadIdx.x;
#pragma unroll
    for (int sg = 0; sg < kLines; ++sg) {
      const int g = i0 + line_offset(sg) + t;
      v[sg] = buffer_load<T>(xr, (unsigned)g * (unsigned)sizeof(T));
    }
    const int sg = t / (2 * ORDER), q = kBlock + t - sg * (2 * ORDER);
    const int g = i0 + line_offset(sg) + q;
    v[kLines] = t < kLines * 2 * ORDER
                    ? buffer_load<T>(xr, (unsigned)g * (unsigned)sizeof(T))
                    : T(0);
  }
  __device__ __forceinline__ void stage_store(const T (&v)[kStageIt], T *lds) const {
    const int t = (int)threadIdx.x;
#pragma unroll
    for (int sg = 0; sg < kLines; ++sg) lds[sg * kSeg + t] = v[sg];
    if (t < kLines * 2 * ORDER) {
      const int sg = t / (2 * ORDER), q = kBlock + t - sg * (2 * ORDER);
      lds[sg * kSeg + q] = v[kLines];
    }
    __syncthreads();
  }

  // s + sum over the node's cells (ascending) of 0 + (alpha f_c) (E_m x_c)[local],
  // or x[row] on a fixed node.  s must not be -0 (callers start from +0 or
  // add +0 to a scaled y).  STAGED: the workgroup stages x in lds (kLds
  // values; every thread must call) and reads it from there: the same
  // values with a third of the vector loads, issued together with the
  // node's factor loads.
  template <bool STAGED = false>
  __device__ __forceinline__ T row_sum_scaled(long long unit, const T *__restrict__ x, T s,
                                              T alpha, TileRow &tr, T *lds = nullptr) const {
#pragma clang fp contract(off)
    const long long r = thread_row(unit);
    tr.has_row = r < nrow;
    tr.row = tr.has_row ? (int)r : 0;
    T sv[kStageIt];
    if (STAGED) stage_load(unit, x, sv);
    else if (!tr.has_row) return s;
    // (staged: a thread past the last row computes node 0's indices and
    // loads, and returns after the workgroup's barrier)
    const int i = tr.row;
    const int t = dpx.div(i);
    const int gx = i - t * px;
    const int gz = DIM == 3 ? dpy.div(t) : 0;
    const int gy = t - gz * py;
    // Every load is issued before the arithmetic: the x neighbourhood, the
    // factors (and matrix indices) of the node's cells, the fixed-dof word.
    const int pxy = px * py;
    const int bx = gx / ORDER - 1, by = gy / ORDER - 1, bz = DIM == 3 ? gz / ORDER - 1 : 0;
    const int rx = gx - ORDER * (bx + 1), ry = gy - ORDER * (by + 1),
              rz = DIM == 3 ? gz - ORDER * (bz + 1) : 0;
    // The loads go through buffer descriptors whose range is the vector
    // (resp. the cell array): an index past either end, reached from a
    // node on the mesh boundary, reads 0 instead of faulting, and an x row
    // (or y plane) wraps to its neighbour; both only feed cells that do not
    // exist, whose sums are dropped below.  Offsets are in bytes from one
    // per-lane base, x neighbours along a row in the immediate field.
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(x), 0, nrow * (int)sizeof(T), kRsrcWord3);
    const __amdgpu_buffer_rsrc_t fr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T *>(f), 0, ncx * ncy * ncz * (int)sizeof(T), kRsrcWord3);
    T xn[WZ][W][W];
    T fc[KZ][2][2];
    int mc[KZ][2][2];
    const unsigned ib = (unsigned)i * (unsigned)sizeof(T);
    if (!STAGED) {
#pragma unroll
      for (int dz = 0; dz < WZ; ++dz)
#pragma unroll
        for (int dy = 0; dy < W; ++dy) {
          const unsigned rb =
              ib + (unsigned)(((DIM == 3 ? (dz - ORDER) * pxy : 0) + (dy - ORDER) * px - ORDER) *
                              (int)sizeof(T));
#pragma unroll
          for (int dx = 0; dx < W; ++dx)
            xn[dz][dy][dx] = buffer_load<T>(
                xr, dx < ORDER ? before_start(rb + dx * (unsigned)sizeof(T))
                               : rb + dx * (unsigned)sizeof(T));
        }
    }
    // the upper cell (k = 1 in every direction); the others sit at fixed
    // offsets from it
    const int cu = (bx + 1) + ncx * ((by + 1) + ncy * (DIM == 3 ? bz + 1 : 0));
    const unsigned cb = (unsigned)cu * (unsigned)sizeof(T);
#pragma unroll
    for (int kz = 0; kz < KZ; ++kz)
#pragma unroll
      for (int ky = 0; ky < 2; ++ky) {
        const int roff = ncx * ((ky - 1) + (DIM == 3 ? ncy * (kz - 1) : 0)) - 1;
        const unsigned rb = cb + (unsigned)(roff * (int)sizeof(T));
#pragma unroll
        for (int kx = 0; kx < 2; ++kx) {
          fc[kz][ky][kx] =
              buffer_load<T>(fr, kx == 0 ? before_start(rb) : rb + kx * (unsigned)sizeof(T));
          if (MULTI) {
            const int c = cu + roff + kx;
            mc[kz][ky][kx] = (c >= 0 && c < ncx * ncy * ncz) ? (int)lm[c] : 0;
          } else {
            mc[kz][ky][kx] = 0;
          }
        }
      }
    const unsigned fixed_word = fixed[i >> 5];
    if (STAGED) {
      stage_store(sv, lds);
      if (!tr.has_row) return s;
      const int t0 = (int)(r - unit * kBlock);
#pragma unroll
      for (int dz = 0; dz < WZ; ++dz)
#pragma unroll
        for (int dy = 0; dy < W; ++dy)
#pragma unroll
          for (int dx = 0; dx < W; ++dx) xn[dz][dy][dx] = lds[(dz * W + dy) * kSeg + t0 + dx];
    }
    // cell k in a direction: c = g / ORDER - 1 + k, local position
    // a = g - ORDER c; it exists when 0 <= c < nc and a <= ORDER
    bool okc[KZ][2][2];
#pragma unroll
    for (int kz = 0; kz < KZ; ++kz)
#pragma unroll
      for (int ky = 0; ky < 2; ++ky)
#pragma unroll
        for (int kx = 0; kx < 2; ++kx) {
          const int cx = bx + kx, cy = by + ky, cz = bz + kz;
          const int ax = rx + ORDER * (1 - kx), ay = ry + ORDER * (1 - ky),
                    az = DIM == 3 ? rz + ORDER * (1 - kz) : 0;
          okc[kz][ky][kx] = cx >= 0 && cx < ncx && cy >= 0 && cy < ncy && cz >= 0 &&
                            cz < ncz && ax <= ORDER && ay <= ORDER && az <= ORDER;
        }
#pragma unroll
    for (int kz = 0; kz < KZ; ++kz)
#pragma unroll
      for (int ky = 0; ky < 2; ++ky)
#pragma unroll
        for (int kx = 0; kx < 2; ++kx) {
          // local position of the node in this cell
          const int ax = rx + ORDER * (1 - kx), ay = ry + ORDER * (1 - ky),
                    az = DIM == 3 ? rz + ORDER * (1 - kz) : 0;
          const bool ok = okc[kz][ky][kx];
          const T *Em = MULTI ? E + (size_t)(ok ? mc[kz][ky][kx] : 0) * (NLOC * NLOC) : E;
          // a clamped local position keeps the matrix index in range for a
          // cell that does not exist (its sum is dropped below)
          const int li = (ax > ORDER ? ORDER : ax) + Q * ((ay > ORDER ? ORDER : ay) +
                                                          Q * (az > ORDER ? ORDER : az));
          const T *Er = Em + li * NLOC;
          // local column j = o + a - ORDER for neighbour offset o: ascending o
          // is ascending j, the order of the reference's row sum.  It starts
          // from -0, the identity of addition, so tmp is the plain sum of
          // the products; the reference's start from +0 differs only in the
          // sign of an all-zero sum, which the two steps below do not see.
          T tmp = T(-0.0);
#pragma unroll
          for (int oz = 0; oz < WZ; ++oz)
#pragma unroll
            for (int oy = 0; oy < W; ++oy)
#pragma unroll
              for (int ox = 0; ox < W; ++ox) {
                const int jx = ox + ax - ORDER, jy = oy + ay - ORDER,
                          jz = DIM == 3 ? oz + az - ORDER : 0;
                if (jx < 0 || jx > ORDER || jy < 0 || jy > ORDER || jz < 0 || jz > ORDER)
                  continue;
                tmp += Er[jx + Q * (jy + Q * jz)] * xn[oz][oy][ox];
              }
          // The reference adds lhs = 0 + (alpha f_c) tmp.  Adding (alpha f_c)
          // tmp itself gives the same sum: the two differ only for a product
          // of -0, and a running sum that starts from +0 never holds -0, so
          // adding -0 or +0 leaves it unchanged.
          const T cf = alpha * fc[kz][ky][kx];
          const T sn = s + cf * tmp;
          s = ok ? sn : s;
        }
    if ((fixed_word >> (i & 31)) & 1u) s = xn[DIM == 3 ? ORDER : 0][ORDER][ORDER];
    return s;
  }

  // with x staged through lds (kLds values; every thread of the workgroup
  // must call it)
  __device__ __forceinline__ T row_sum(long long unit, const T *x, T s, TileRow &tr,
                                       T *lds) const {
    return row_sum_scaled<true>(unit, x, s, T(1), tr, lds);
  }

};

template <typename T, int DIM, int ORDER, bool MULTI>
inline StencilOp<T, DIM, ORDER, MULTI> stencil_view(const hflmp_stencil *m) {
  StencilOp<T, DIM, ORDER, MULTI> op;
  op.nrow = (int)m->nnode;
  op.ncx = m->nc[0];
  op.ncy = m->nc[1];
  op.ncz = m->nc[2];
  op.px = m->np[0];
  op.py = m->np[1];
  op.pz = m->np[2];
  op.dpx.init((unsigned)m->np[0]);
  op.dpy.init((unsigned)m->np[1]);
  op.E = static_cast<const T *>(m->E);
  op.f = static_cast<const T *>(m->factor);
  op.lm = m->cell_lm;
  op.fixed = m->fixed_bits;
  return op;
}

template <typename T, int DIM, int ORDER, bool MULTI>
inline int op_grid(const StencilOp<T, DIM, ORDER, MULTI> &op) {
  const long long units = ((long long)op.nrow + kBlock - 1) / kBlock;
  return (int)(((units + kXcds - 1) / kXcds) * kXcds);
}

// Calls fn(op) with the operator view matching the mesh; returns false for
// an element type without a kernel.
template <typename T, typename Fn>
inline bool stencil_dispatch(const hflmp_stencil *m, Fn &&fn) {
  const bool multi = m->nmat > 1;
  if (m->order == 1 && m->dim == 2) {
    if (multi) fn(stencil_view<T, 2, 1, true>(m));
    else fn(stencil_view<T, 2, 1, false>(m));
  } else if (m->order == 1 && m->dim == 3) {
    if (multi) fn(stencil_view<T, 3, 1, true>(m));
    else fn(stencil_view<T, 3, 1, false>(m));
  } else if (m->order == 2 && m->dim == 2) {
    if (multi) fn(stencil_view<T, 2, 2, true>(m));
    else fn(stencil_view<T, 2, 2, false>(m));
  } else {
    return false;
  }
  return true;
}
