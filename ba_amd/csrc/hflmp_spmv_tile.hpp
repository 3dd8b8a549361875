// CSR row-tile SpMV core shared by the SpMV entry points and the fused
// Krylov kernels.
//
// One workgroup (256 threads) owns a tile of consecutive rows [r0, r1).
// The tile's nonzeros are one contiguous range of col/val, so the whole
// workgroup streams that range with coalesced non-temporal loads, gathers
// x[col] through the cache (the banded FE stencil reuses every x entry
// from L2), and stores the rounded products val*x in LDS.  Then thread t
// adds the products of row r0+t from LDS in ascending storage order,
// starting from zero: the summation order of
// CPUsimple_CSR_lMatrix::VectorMult (CPUsimple_csr_lmatrix.cc:76-111), so
// the result is bit-identical to that CPU loop.
//
// Tiles are sized on the host from the mean row length so that a tile's
// products normally fit one 2048-entry LDS window (16 KiB in FP64, 8 KiB in
// FP32), which keeps eight workgroups (all 32 wave slots) resident per CU
// with eight loads of col and val in flight per thread.  That case runs a
// fully unrolled single-window path.  A tile whose range is longer than the
// window (long rows) is processed in consecutive windows; each row keeps
// its running sum across windows, so the order is unchanged.
//
// Column indices are int32, or int16 offsets from the tile's first row
// (IT = short, col = r0 + offset) for matrices whose band fits; the FP32
// inner solve of the mixed-precision CG uses the latter.
#pragma once

#include "hflmp_internal.hpp"

// Products staged per window.
template <typename T>
struct TileCap {
  static constexpr int value = 2048;
};

// Rows per tile: the mean row length times 1.05 must fit one window.
inline int tile_rows(long long nrow, long long nnz, int cap) {
  if (nrow <= 0) return 1;
  const double avg = nnz > 0 ? (double)nnz / (double)nrow : 1.0;
  long long r = (long long)((double)cap / (avg * 1.05));
  if (r > kBlock) r = kBlock;
  if (r < 1) r = 1;
  return (int)r;
}

inline int tile_grid(long long ntiles) {
  const long long per = (ntiles + kXcds - 1) / kXcds;
  return (int)(per * kXcds);
}

struct TileRow {
  int row;       // global row of this thread, valid if has_row
  bool has_row;
};

// Sequential sum of prod[lo..hi) onto s with eight LDS reads in flight;
// the additions stay in ascending order.
template <typename T>
__device__ __forceinline__ T lds_seq_sum(const T *__restrict__ prod, int lo, int hi, T s) {
  int k = lo;
  for (; k + 8 <= hi; k += 8) {
    const T a0 = prod[k], a1 = prod[k + 1], a2 = prod[k + 2], a3 = prod[k + 3];
    const T a4 = prod[k + 4], a5 = prod[k + 5], a6 = prod[k + 6], a7 = prod[k + 7];
    s += a0; s += a1; s += a2; s += a3; s += a4; s += a5; s += a6; s += a7;
  }
  for (; k < hi; ++k) s += prod[k];
  return s;
}

// Returns s + sum_j val[j] * x[col[j]] over this thread's row of the tile.
// SKIP_DIAG drops j with col[j] == row (y = (A - D) x).  IT = short reads
// 16-bit offsets from the tile's first row.
// All threads of the workgroup must call this (it synchronises).
template <typename T, bool SKIP_DIAG, typename IT = int>
__device__ __forceinline__ T tile_row_sum(long long tile, int rows_per_tile, int nrow,
                                          const int *__restrict__ rp,
                                          const IT *__restrict__ ci,
                                          const T *__restrict__ val,
                                          const T *__restrict__ x_in, T s, TileRow &tr,
                                          T *__restrict__ prod) {
#pragma clang fp contract(off)
  constexpr int CAP = TileCap<T>::value;
  constexpr int U = CAP / kBlock;
  const int t = threadIdx.x;
  const long long r0l = tile * (long long)rows_per_tile;
  tr.has_row = false;
  tr.row = 0;
  if (r0l >= nrow) return s;  // uniform across the workgroup
  const int r0 = (int)r0l;
  const int r1 = min(r0 + rows_per_tile, nrow);
  // x indexed by the stored column: absolute for int32, tile-relative for int16
  const T *__restrict__ x = std::is_same<IT, int>::value ? x_in : x_in + r0;
  const int cbase = std::is_same<IT, int>::value ? 0 : r0;
  const int nz0 = rp[r0];
  const int nz1 = rp[r1];
  tr.has_row = t < r1 - r0;
  tr.row = r0 + t;
  const int cnt_all = nz1 - nz0;
  if (cnt_all <= CAP) {
    // single window, fully unrolled: U loads of col and val per thread
    int c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = t + u * kBlock;
      if (k < cnt_all) {
        c[u] = ld_stream(ci + nz0 + k);
        v[u] = ld_stream(val + nz0 + k);
      }
    }
    int lo = 0, hi = 0;
    if (tr.has_row) {
      lo = rp[r0 + t] - nz0;
      hi = rp[r0 + t + 1] - nz0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = t + u * kBlock;
      if (k < cnt_all) prod[k] = v[u] * x[c[u]];
    }
    __syncthreads();
    if (tr.has_row) {
      if (SKIP_DIAG) {
        for (int k = lo; k < hi; ++k)
          if (cbase + (int)ci[nz0 + k] != tr.row) s += prod[k];
      } else {
        s = lds_seq_sum(prod, lo, hi, s);
      }
    }
    return s;
  }
  // several windows
  int my_lo = 0, my_hi = 0;
  if (tr.has_row) {
    my_lo = rp[r0 + t];
    my_hi = rp[r0 + t + 1];
  }
  for (int base = nz0; base < nz1; base += CAP) {
    const int cnt = min(CAP, nz1 - base);
    const IT *__restrict__ cb = ci + base;
    const T *__restrict__ vb = val + base;
    for (int k0 = t; k0 < cnt; k0 += kBlock * 4) {
      int c[4];
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + u * kBlock;
        if (k < cnt) {
          c[u] = ld_stream(cb + k);
          v[u] = ld_stream(vb + k);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + u * kBlock;
        if (k < cnt) prod[k] = v[u] * x[c[u]];
      }
    }
    __syncthreads();
    if (tr.has_row) {
      const int lo = max(my_lo, base) - base;
      const int hi = min(my_hi, base + cnt) - base;
      if (SKIP_DIAG) {
        for (int k = lo; k < hi; ++k)
          if (cbase + (int)cb[k] != tr.row) s += prod[k];
      } else {
        s = lds_seq_sum(prod, lo, hi, s);
      }
    }
    __syncthreads();
  }
  return s;
}
