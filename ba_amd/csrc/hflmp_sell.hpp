// Sliced-ELL execution layout of a CSR matrix (SELL-64), and the two
// "operator" views the fused solver kernels are written against.
//
// SELL-64: rows are cut into slices of 64 consecutive rows, one wave64 per
// slice and one lane per row.  Slice s stores its rows' entries padded to
// the slice's longest row (m entries), the first F = m - m mod P of them
// lane-interleaved in chunks of P consecutive entries of a row, the last
// m mod P in blocks of one entry per lane:
//     entry j < F of row 64 s + l  is at  off[s] + 64 P (j / P) + P l + (j mod P)
//     entry j >= F                 is at  off[s] + 64 j + l
// so each load instruction of a wave reads one contiguous 64 P-element
// block (FP64: P = 2, 16 B per lane; FP32: P = 4, 16 B per lane), a row of
// odd length loads no padding after its last entry, and the
// x gathers of a step j hit x at 64 neighbouring rows' j-th columns, which
// for an FE band is a nearly contiguous range.  Each lane adds its row's
// products in ascending storage order starting from zero (or from y[i] for
// y += A x), i.e. the summation order of CPUsimple_CSR_lMatrix::VectorMult
// (CPUsimple_csr_lmatrix.cc:76-111): the result is bit-identical to the
// sequential CPU loop.  A lane stops at its row's length (a byte per row),
// so only the padding inside a row's last chunk of P entries is loaded.
//
// Columns are int32, or int16 offsets from the slice's first row when the
// band fits (|col - 64 s| < 32768), which cuts a nonzero from 12 to 10 B
// (FP64) or from 8 to 6 B (FP32).
#pragma once

#include "hflmp_spmv_tile.hpp"

struct hflmp_sell {
  int device;
  int nrow, ncol, nnz, nslice;
  int chunk;        // P
  int col_bytes;    // 4 (absolute int32) or 2 (int16 offset from the slice's first row)
  int value_bytes;  // 8 or 4
  long long entries;
  long long *soff;      // nslice + 1 entry offsets
  int *slen;            // longest row of each slice (m): the first m - m mod P entries of a
                        // row sit in chunks of P, the rest in one-entry-per-lane blocks
  unsigned int *check;  // device counter of rows whose length changed (set_values)
  unsigned char *rlen;  // row lengths
  void *col;
  void *val;
};

template <typename T>
struct SellChunk;
template <>
struct SellChunk<double> {
  static constexpr int P = 2;  // entries per lane load
  static constexpr int U = 7;  // lane loads in flight per step
};
template <>
struct SellChunk<float> {
  static constexpr int P = 4;
  static constexpr int U = 3;
};

template <typename T, int P>
struct VecOf {
  typedef T type __attribute__((ext_vector_type(P)));
};

constexpr int kSellRows = 64;                         // rows per slice (one wave)
constexpr int kSellSlicesPerBlock = kBlock / kWave;   // 4 slices per workgroup

// Position of entry j of lane `lane` in a slice whose entries start at
// `off` and whose longest row has m entries.
__host__ __device__ inline long long sell_pos(long long off, int m, int P, int lane, int j) {
  const int full = m / P * P;
  return j < full ? off + (long long)kSellRows * P * (j / P) + (long long)P * lane + (j % P)
                  : off + (long long)kSellRows * j + lane;
}

// Workgroups for an operator over SELL-64: 256 consecutive rows each,
// rounded up to a multiple of the XCD count.
inline int sell_grid(int nrow) {
  const long long units = ((long long)nrow + kBlock - 1) / kBlock;
  return (int)(((units + kXcds - 1) / kXcds) * kXcds);
}

// CSR row tiles (hflmp_spmv_tile.hpp): unit = tile of R rows.
template <typename T, typename IT = int>
struct CsrTileOp {
  using value_type = T;
  static constexpr int kLds = TileCap<T>::value;
  int nrow, R;
  const int *rp;
  const IT *ci;
  const T *val;

  __device__ __forceinline__ long long thread_row(long long unit) const {
    return threadIdx.x < (unsigned)R ? unit * R + threadIdx.x : -1;
  }
  __device__ __forceinline__ T row_sum(long long unit, const T *x, T s, TileRow &tr,
                                       T *lds) const {
    return tile_row_sum<T, false, IT>(unit, R, nrow, rp, ci, val, x, s, tr, lds);
  }
};

// One lane's row of a SELL-64 slice whose entries start at col / val (the
// slice's first entry): s + sum of a_ij x_j over the row's len entries in
// ascending storage order (SUB: s - a_ij x_j; NODIAG: the entry whose
// column is dcol skipped).  smax: the slice's entries per lane.  The loads
// cover the row's own entries, or with LOAD_ALL every entry position of the
// slice (all allocated: 64 x smax entries), so that they need not wait for
// len; the sum is the same either way.
template <typename T, typename IT, bool SUB, bool NODIAG, bool LOAD_ALL>
__device__ __forceinline__ T sell_slice_sum(const IT *__restrict__ col, const T *__restrict__ val,
                                            int smax, int len, int lane,
                                            const T *__restrict__ x, int dcol, T s) {
#pragma clang fp contract(off)
  constexpr int P = SellChunk<T>::P;
  constexpr int U = SellChunk<T>::U;
  typedef typename VecOf<T, P>::type TV;
  typedef typename VecOf<IT, P>::type IV;
  const int full = smax / P * P;  // entries in P-chunks; the rest one per block
  const IV *__restrict__ c = reinterpret_cast<const IV *>(col) + lane;
  const TV *__restrict__ v = reinterpret_cast<const TV *>(val) + lane;
  // the row's entries past the last full chunk (at most P - 1), loaded
  // first so their latency overlaps the chunks'
  constexpr int TM = P > 1 ? P - 1 : 1;
  const int ntail = len - full;
  const int nload = LOAD_ALL ? smax - full : ntail;
  IT ct[TM];
  T vt[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t) {
    if (t < nload) {
      const long long q = (long long)kSellRows * (full + t) + lane;
      ct[t] = __builtin_nontemporal_load(col + q);
      vt[t] = __builtin_nontemporal_load(val + q);
    }
  }
  for (int j = 0; j < full; j += P * U) {
    IV cc[U];
    TV vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if ((LOAD_ALL || j + u * P < len) && j + u * P < full) {
        cc[u] = __builtin_nontemporal_load(c + (j / P + u) * kWave);
        vv[u] = __builtin_nontemporal_load(v + (j / P + u) * kWave);
      }
    }
    T xx[U][P];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (j + u * P + q < len && j + u * P < full) xx[u][q] = x[(int)cc[u][q]];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (j + u * P + q < len && j + u * P < full && (!NODIAG || (int)cc[u][q] != dcol)) {
          if (SUB)
            s -= vv[u][q] * xx[u][q];
          else
            s += vv[u][q] * xx[u][q];
        }
  }
  // the tail, in ascending order after the chunks
  T xt[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t)
    if (t < ntail) xt[t] = x[(int)ct[t]];
#pragma unroll
  for (int t = 0; t < TM; ++t)
    if (t < ntail && (!NODIAG || (int)ct[t] != dcol)) {
      if (SUB)
        s -= vt[t] * xt[t];
      else
        s += vt[t] * xt[t];
    }
  return s;
}

// Two slices at once, A and B, each lane one row of each: the sums that
// sell_slice_sum<T, IT, false, false, true> forms for either (the same
// order within each row), with both slices' entry loads issued before
// either's gathers, so that a wave keeps twice the bytes in flight.
// smaxB = 0: no slice B (sB stays as given).  U: chunk loads in flight
// per slice and step (the order of the sums does not depend on it).
template <typename T, typename IT, int U = SellChunk<T>::U>
__device__ __forceinline__ void sell_slice_sum2(const IT *__restrict__ colA,
                                                const T *__restrict__ valA, int smaxA, int lenA,
                                                const T *__restrict__ xA,
                                                const IT *__restrict__ colB,
                                                const T *__restrict__ valB, int smaxB, int lenB,
                                                const T *__restrict__ xB, int lane, T &sA, T &sB) {
#pragma clang fp contract(off)
  constexpr int P = SellChunk<T>::P;
  typedef typename VecOf<T, P>::type TV;
  typedef typename VecOf<IT, P>::type IV;
  const int fullA = smaxA / P * P, fullB = smaxB / P * P;
  const IV *__restrict__ cA = reinterpret_cast<const IV *>(colA) + lane;
  const TV *__restrict__ vA = reinterpret_cast<const TV *>(valA) + lane;
  const IV *__restrict__ cB = reinterpret_cast<const IV *>(colB) + lane;
  const TV *__restrict__ vB = reinterpret_cast<const TV *>(valB) + lane;
  constexpr int TM = P > 1 ? P - 1 : 1;
  const int ntA = lenA - fullA, ntB = lenB - fullB;
  const int nlA = smaxA - fullA, nlB = smaxB - fullB;
  IT ctA[TM], ctB[TM];
  T vtA[TM], vtB[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t) {
    if (t < nlA) {
      const long long q = (long long)kSellRows * (fullA + t) + lane;
      ctA[t] = __builtin_nontemporal_load(colA + q);
      vtA[t] = __builtin_nontemporal_load(valA + q);
    }
    if (t < nlB) {
      const long long q = (long long)kSellRows * (fullB + t) + lane;
      ctB[t] = __builtin_nontemporal_load(colB + q);
      vtB[t] = __builtin_nontemporal_load(valB + q);
    }
  }
  const int full = fullA > fullB ? fullA : fullB;
  for (int j = 0; j < full; j += P * U) {
    IV ccA[U], ccB[U];
    TV vvA[U], vvB[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (j + u * P < fullA) {
        ccA[u] = __builtin_nontemporal_load(cA + (j / P + u) * kWave);
        vvA[u] = __builtin_nontemporal_load(vA + (j / P + u) * kWave);
      }
      if (j + u * P < fullB) {
        ccB[u] = __builtin_nontemporal_load(cB + (j / P + u) * kWave);
        vvB[u] = __builtin_nontemporal_load(vB + (j / P + u) * kWave);
      }
    }
    T xxA[U][P], xxB[U][P];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (j + u * P + q < lenA && j + u * P < fullA) xxA[u][q] = xA[(int)ccA[u][q]];
        if (j + u * P + q < lenB && j + u * P < fullB) xxB[u][q] = xB[(int)ccB[u][q]];
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (j + u * P + q < lenA && j + u * P < fullA) sA += vvA[u][q] * xxA[u][q];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (j + u * P + q < lenB && j + u * P < fullB) sB += vvB[u][q] * xxB[u][q];
  }
  T xtA[TM], xtB[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t) {
    if (t < ntA) xtA[t] = xA[(int)ctA[t]];
    if (t < ntB) xtB[t] = xB[(int)ctB[t]];
  }
#pragma unroll
  for (int t = 0; t < TM; ++t)
    if (t < ntA) sA += vtA[t] * xtA[t];
#pragma unroll
  for (int t = 0; t < TM; ++t)
    if (t < ntB) sB += vtB[t] * xtB[t];
}

// SELL-64: unit = four slices (256 consecutive rows), wave w of the
// workgroup runs slice 4 unit + w.  No LDS and no barrier.
template <typename T, typename IT = int>
struct SellOp {
  using value_type = T;
  static constexpr int kLds = 1;
  static constexpr int P = SellChunk<T>::P;
  static constexpr int U = SellChunk<T>::U;
  int nrow, nslice;
  const long long *soff;
  const int *slen;
  const unsigned char *rlen;
  const IT *col;
  const T *val;

  __device__ __forceinline__ long long thread_row(long long unit) const {
    return unit * kBlock + threadIdx.x;
  }
  // s + sum of a_ij x_j over the row in ascending storage order; SUB:
  // s - a_ij x_j instead; NODIAG: the diagonal entry (column == row) skipped.
  template <bool SUB = false, bool NODIAG = false>
  __device__ __forceinline__ T row_sum(long long unit, const T *x_in, T s, TileRow &tr,
                                       T *) const {
    tr.has_row = false;
    tr.row = 0;
    const long long sl = unit * kSellSlicesPerBlock + (threadIdx.x >> 6);
    if (sl >= nslice) return s;  // uniform across the wave
    const int sid = __builtin_amdgcn_readfirstlane((int)sl);
    const int lane = threadIdx.x & (kWave - 1);
    const int row = sid * kSellRows + lane;
    tr.row = row;
    tr.has_row = row < nrow;
    const int len = tr.has_row ? (int)rlen[row] : 0;
    const long long base = soff[sid];
    const int smax = slen[sid];
    // int16 columns are offsets from the slice's first row
    const T *__restrict__ x = std::is_same<IT, int>::value ? x_in : x_in + sid * kSellRows;
    const int dcol = std::is_same<IT, int>::value ? row : lane;
    return sell_slice_sum<T, IT, SUB, NODIAG, false>(col + base, val + base, smax, len, lane, x,
                                                      dcol, s);
  }
};

template <typename T, typename IT>
inline SellOp<T, IT> sell_view(const hflmp_sell *m) {
  SellOp<T, IT> op;
  op.nrow = m->nrow;
  op.nslice = m->nslice;
  op.soff = m->soff;
  op.slen = m->slen;
  op.rlen = m->rlen;
  op.col = static_cast<const IT *>(m->col);
  op.val = static_cast<const T *>(m->val);
  return op;
}

template <typename T>
inline CsrTileOp<T, int> csr_view(int nrow, int nnz, const int *rp, const int *ci,
                                  const T *val) {
  CsrTileOp<T, int> op;
  op.nrow = nrow;
  op.R = tile_rows(nrow, nnz, TileCap<T>::value);
  op.rp = rp;
  op.ci = ci;
  op.val = val;
  return op;
}

template <typename T, typename IT>
inline int op_grid(const CsrTileOp<T, IT> &op) {
  return tile_grid(((long long)op.nrow + op.R - 1) / op.R);
}

template <typename T, typename IT>
inline int op_grid(const SellOp<T, IT> &op) {
  return sell_grid(op.nrow);
}
