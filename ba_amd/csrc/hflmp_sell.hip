// SELL-64 execution layout of a CSR matrix: device-side build from the CSR
// arrays, value refresh, and SpMV.  The layout is described in
// hflmp_sell.hpp.  It is an execution format of the HIP platform for a
// matrix the caller holds in CSR (lmatrix_formats.h:51-56), in the role of
// the reference's ELL format slot (la_global.h:56, MATRIX_FORMAT::ELL): the
// products y = A x and y += A x are those of CPU_CSR_lMatrix::VectorMult /
// VectorMultAdd (CPUsimple_csr_lmatrix.cc:76-111, lmatrix_csr_cpu.cc:
// 1424-1450), bit for bit.
#include <cstdio>

#include "hflmp_sell.hpp"

namespace {

enum SellMode { kSellAssign = 0, kSellAdd = 1 };

struct SellStats {
  unsigned int long_rows;  // rows longer than 255 entries
  unsigned int bad16;      // entries whose offset from the slice start needs more than 16 bits
};

// Row lengths (a byte each) and, per slice, the longest row (m; where its
// entries sit is sell_pos's business).
// Threads walk rows in whole waves, so a wave is always one slice.
__global__ void __launch_bounds__(kBlock)
    sell_lengths_kernel(int nrow, int nslice, const int *__restrict__ rp,
                        unsigned char *__restrict__ rlen, int *__restrict__ slen,
                        SellStats *stats) {
  const long long total = (long long)nslice * kSellRows;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (long long)gridDim.x * kBlock) {
    int len = 0;
    if (i < nrow) {
      len = rp[i + 1] - rp[i];
      if (len > 255) atomicAdd(&stats->long_rows, 1u);
      rlen[i] = (unsigned char)(len > 255 ? 255 : len);
    }
    int m = len;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const int other = __shfl_xor(m, o);
      m = other > m ? other : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) slen[i / kSellRows] = m;
  }
}

// soff[s] = sum over t < s of 64 slen[t]; soff[nslice] = entries.  One
// workgroup: each thread scans a contiguous run of slices.
__global__ void __launch_bounds__(kFinalBlock)
    sell_scan_kernel(int nslice, const int *__restrict__ slen, long long *__restrict__ soff) {
  __shared__ long long tot[kFinalBlock];
  const int t = threadIdx.x;
  const int per = (nslice + kFinalBlock - 1) / kFinalBlock;
  const int lo = t * per < nslice ? t * per : nslice;
  const int hi = lo + per < nslice ? lo + per : nslice;
  long long s = 0;
  for (int k = lo; k < hi; ++k) s += (long long)kSellRows * slen[k];
  tot[t] = s;
  __syncthreads();
  // inclusive Hillis-Steele scan of the per-thread totals
  for (int o = 1; o < kFinalBlock; o <<= 1) {
    const long long add = t >= o ? tot[t - o] : 0;
    __syncthreads();
    tot[t] += add;
    __syncthreads();
  }
  long long run = tot[t] - s;
  for (int k = lo; k < hi; ++k) {
    soff[k] = run;
    run += (long long)kSellRows * slen[k];
  }
  if (t == kFinalBlock - 1) soff[nslice] = tot[t];
}

// Counts entries whose column is not within int16 of its slice's first row.
__global__ void __launch_bounds__(kBlock)
    sell_band_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                     SellStats *stats) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    const int base = (i / kSellRows) * kSellRows;
    unsigned int bad = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int d = ci[k] - base;
      bad += (d < -32768 || d > 32767) ? 1u : 0u;
    }
    if (bad) atomicAdd(&stats->bad16, bad);
  }
}

// Scatter of row i's entries into its lane of its slice (thread per row).
// COLS = false refreshes the values only: a row whose length differs from
// the one stored at build time (rlen) would be placed past its lane's
// entries, so it is skipped and counted in *changed instead.
template <typename T, typename IT, bool COLS>
__global__ void __launch_bounds__(kBlock)
    sell_fill_kernel(int nrow, int P, const int *__restrict__ rp, const int *__restrict__ ci,
                     const T *__restrict__ val, const long long *__restrict__ soff,
                     const int *__restrict__ slen, IT *__restrict__ col,
                     T *__restrict__ sval, const unsigned char *__restrict__ rlen,
                     unsigned int *changed) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    const int s = i / kSellRows, lane = i % kSellRows;
    const long long off = soff[s];
    const int m = slen[s];
    const int k0 = rp[i], len = rp[i + 1] - k0;
    if (!COLS && len != (int)rlen[i]) {
      atomicAdd(changed, 1u);
      continue;
    }
    for (int j = 0; j < len; ++j) {
      const long long pos = sell_pos(off, m, P, lane, j);
      if (COLS) {
        const int c = ci[k0 + j];
        col[pos] = (IT)(std::is_same<IT, int>::value ? c : c - s * kSellRows);
      }
      sval[pos] = val[k0 + j];
    }
  }
}

template <typename Op, int MODE>
__device__ __forceinline__ void sell_spmv_rows(const Op &op,
                                               const typename Op::value_type *__restrict__ x,
                                               typename Op::value_type *__restrict__ y) {
  using T = typename Op::value_type;
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  T s = 0;
  if (MODE == kSellAdd) {
    const long long r = op.thread_row(unit);
    if (r >= 0 && r < op.nrow) s = y[r];
  }
  TileRow tr;
  s = op.row_sum(unit, x, s, tr, nullptr);
  // y is not read again by this launch: stream it past the caches
  if (tr.has_row) __builtin_nontemporal_store(s, y + tr.row);
}

template <typename Op, int MODE>
__global__ void __launch_bounds__(kBlock)
    sell_spmv_kernel(Op op, const typename Op::value_type *__restrict__ x,
                     typename Op::value_type *__restrict__ y) {
  sell_spmv_rows<Op, MODE>(op, x, y);
}

// int16 columns: the loop fits 80 VGPRs, six waves per SIMD (on 128^3 Q1
// FP64: 105 us per SpMV against 109 us at five; the int32 form would
// spill at six)
template <typename Op, int MODE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
    sell_spmv16_kernel(Op op, const typename Op::value_type *__restrict__ x,
                       typename Op::value_type *__restrict__ y) {
  sell_spmv_rows<Op, MODE>(op, x, y);
}

void sell_free(hflmp_sell *m) {
  if (!m) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != m->device) (void)hipSetDevice(m->device);
  (void)hipFree(m->soff);
  (void)hipFree(m->slen);
  (void)hipFree(m->rlen);
  (void)hipFree(m->col);
  (void)hipFree(m->val);
  (void)hipFree(m->check);
  if (cur >= 0 && cur != m->device) (void)hipSetDevice(cur);
  delete m;
}

template <typename T>
int sell_create(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *rp, const int *ci,
                const T *val, int flags, hflmp_sell **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, out != nullptr, "sell_create: null output handle");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, nrow > 0 && ncol > 0 && nnz >= 0, "sell_create: bad sizes");
  HFLMP_REQUIRE(ctx, rp && (nnz == 0 || (ci && val)), "sell_create: null pointer");
  HFLMP_REQUIRE(ctx, (flags & ~(HFLMP_SELL_COL16 | HFLMP_SELL_ANY_PADDING)) == 0,
                "sell_create: unknown flags %d", flags);
  hipStream_t s = ctx->stream;
  auto *m = new hflmp_sell();
  m->device = ctx->device;
  m->nrow = nrow;
  m->ncol = ncol;
  m->nnz = nnz;
  m->nslice = (nrow + kSellRows - 1) / kSellRows;
  m->chunk = SellChunk<T>::P;
  m->value_bytes = sizeof(T);
  m->col_bytes = 4;
  SellStats *stats = nullptr;
  int rc = HFLMP_OK;
  auto fail = [&](int code, const char *msg) {
    (void)hipFree(stats);
    sell_free(m);
    return set_error(ctx, code, "sell_create: %s", msg);
  };
  if (hipMalloc(&m->soff, sizeof(long long) * ((size_t)m->nslice + 1)) != hipSuccess ||
      hipMalloc(&m->slen, sizeof(int) * (size_t)m->nslice) != hipSuccess ||
      hipMalloc(&m->rlen, (size_t)nrow) != hipSuccess ||
      hipMalloc(&m->check, sizeof(unsigned int)) != hipSuccess ||
      hipMalloc(&stats, sizeof(SellStats)) != hipSuccess)
    return fail(HFLMP_ENOMEM, "out of device memory");
  if ((rc = hflmp_fill_zero(ctx, stats, sizeof(SellStats))) != HFLMP_OK) return fail(rc, "fill");
  const long long lanes = (long long)m->nslice * kSellRows;
  long long g = (lanes + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  sell_lengths_kernel<<<(unsigned)g, kBlock, 0, s>>>(nrow, m->nslice, rp, m->rlen, m->slen,
                                                     stats);
  sell_scan_kernel<<<1, kFinalBlock, 0, s>>>(m->nslice, m->slen, m->soff);
  if ((flags & HFLMP_SELL_COL16) && nnz > 0)
    sell_band_kernel<<<grid_for(nrow, 1), kBlock, 0, s>>>(nrow, rp, ci, stats);
  if (hipGetLastError() != hipSuccess) return fail(HFLMP_EHIP, "launch failed");
  SellStats hs;
  long long entries = 0;
  if (hipMemcpyAsync(&hs, stats, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&entries, m->soff + m->nslice, sizeof(entries), hipMemcpyDeviceToHost,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(HFLMP_EHIP, "HIP error while sizing");
  if (hs.long_rows)
    return fail(HFLMP_EFORMAT, "a row has more than 255 entries (use the CSR kernels)");
  // Padding is never loaded (a lane stops at its row's length), but it
  // costs memory and partly-used cache lines; refuse more than twice the
  // nonzeros of it.
  if (!(flags & HFLMP_SELL_ANY_PADDING) && entries > 3LL * nnz + (long long)kSellRows * m->chunk)
    return fail(HFLMP_EFORMAT, "row lengths vary too much within slices (use the CSR kernels)");
  m->entries = entries;
  if ((flags & HFLMP_SELL_COL16) && hs.bad16 == 0) m->col_bytes = 2;
  const size_t ne = entries > 0 ? (size_t)entries : 1;
  if (hipMalloc(&m->col, ne * m->col_bytes) != hipSuccess ||
      hipMalloc(&m->val, ne * sizeof(T)) != hipSuccess)
    return fail(HFLMP_ENOMEM, "out of device memory");
  if ((rc = hflmp_fill_zero(ctx, m->col, ne * m->col_bytes)) != HFLMP_OK) return fail(rc, "fill");
  if ((rc = hflmp_fill_zero(ctx, m->val, ne * sizeof(T))) != HFLMP_OK) return fail(rc, "fill");
  if (nnz > 0) {
    if (m->col_bytes == 2)
      sell_fill_kernel<T, short, true><<<grid_for(nrow, 1), kBlock, 0, s>>>(
          nrow, m->chunk, rp, ci, val, m->soff, m->slen, static_cast<short *>(m->col),
          static_cast<T *>(m->val), nullptr, nullptr);
    else
      sell_fill_kernel<T, int, true><<<grid_for(nrow, 1), kBlock, 0, s>>>(
          nrow, m->chunk, rp, ci, val, m->soff, m->slen, static_cast<int *>(m->col),
          static_cast<T *>(m->val), nullptr, nullptr);
  }
  if (hipGetLastError() != hipSuccess) return fail(HFLMP_EHIP, "launch failed");
  // the stats buffer is freed only after the stream has used it
  if (hipStreamSynchronize(s) != hipSuccess) return fail(HFLMP_EHIP, "HIP error in build");
  (void)hipFree(stats);
  *out = m;
  return HFLMP_OK;
}

template <typename T>
int sell_set_values(hflmp_ctx *ctx, hflmp_sell *m, const int *rp, const T *val) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && rp && (m->nnz == 0 || val), "sell_set_values: null pointer");
  HFLMP_REQUIRE(ctx, m->value_bytes == (int)sizeof(T), "sell_set_values: value type mismatch");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "sell_set_values: matrix on another device");
  if (m->nnz == 0) return HFLMP_OK;
  hipStream_t s = ctx->stream;
  int rc = hflmp_fill_zero(ctx, m->check, sizeof(unsigned int));
  if (rc) return rc;
  sell_fill_kernel<T, int, false><<<grid_for(m->nrow, 1), kBlock, 0, s>>>(
      m->nrow, m->chunk, rp, nullptr, val, m->soff, m->slen, nullptr, static_cast<T *>(m->val),
      m->rlen, m->check);
  HFLMP_LAUNCH_CHECK(ctx);
  unsigned int changed = 0;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(&changed, m->check, sizeof(changed),
                                      hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (changed)
    return set_error(ctx, HFLMP_EFORMAT,
                     "sell_set_values: %u rows differ in length from the pattern the layout "
                     "was built for (their values were not copied; rebuild the layout)",
                     changed);
  return HFLMP_OK;
}

template <typename T, int MODE>
int sell_spmv(hflmp_ctx *ctx, const hflmp_sell *m, const T *x, T *y) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && x && y, "sell_spmv: null pointer");
  HFLMP_REQUIRE(ctx, m->value_bytes == (int)sizeof(T), "sell_spmv: value type mismatch");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "sell_spmv: matrix on another device");
  const unsigned grid = sell_grid(m->nrow);
  hipStream_t s = ctx->stream;
  if (m->col_bytes == 2)
    sell_spmv16_kernel<SellOp<T, short>, MODE><<<grid, kBlock, 0, s>>>(sell_view<T, short>(m),
                                                                        x, y);
  else
    sell_spmv_kernel<SellOp<T, int>, MODE><<<grid, kBlock, 0, s>>>(sell_view<T, int>(m), x, y);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_sell_create_f64(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *row_ptr,
                          const int *col, const double *val, int flags, hflmp_sell **out) {
  return sell_create<double>(ctx, nrow, ncol, nnz, row_ptr, col, val, flags, out);
}
int hflmp_sell_create_f32(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *row_ptr,
                          const int *col, const float *val, int flags, hflmp_sell **out) {
  return sell_create<float>(ctx, nrow, ncol, nnz, row_ptr, col, val, flags, out);
}
int hflmp_sell_set_values_f64(hflmp_ctx *ctx, hflmp_sell *m, const int *row_ptr,
                              const double *val) {
  return sell_set_values<double>(ctx, m, row_ptr, val);
}
int hflmp_sell_set_values_f32(hflmp_ctx *ctx, hflmp_sell *m, const int *row_ptr,
                              const float *val) {
  return sell_set_values<float>(ctx, m, row_ptr, val);
}
int hflmp_sell_get_info(const hflmp_sell *m, hflmp_sell_info *info) {
  if (!m || !info) return HFLMP_EINVAL;
  info->nrow = m->nrow;
  info->ncol = m->ncol;
  info->nnz = m->nnz;
  info->chunk = m->chunk;
  info->col_bytes = m->col_bytes;
  info->value_bytes = m->value_bytes;
  info->entries = m->entries;
  info->device_bytes = (size_t)m->entries * (m->col_bytes + m->value_bytes) +
                       sizeof(long long) * ((size_t)m->nslice + 1) +
                       sizeof(int) * (size_t)m->nslice + (size_t)m->nrow;
  return HFLMP_OK;
}
void hflmp_sell_destroy(hflmp_sell *m) { sell_free(m); }

int hflmp_sell_spmv_f64(hflmp_ctx *ctx, const hflmp_sell *m, const double *x, double *y) {
  return sell_spmv<double, kSellAssign>(ctx, m, x, y);
}
int hflmp_sell_spmv_f32(hflmp_ctx *ctx, const hflmp_sell *m, const float *x, float *y) {
  return sell_spmv<float, kSellAssign>(ctx, m, x, y);
}
int hflmp_sell_spmv_add_f64(hflmp_ctx *ctx, const hflmp_sell *m, const double *x, double *y) {
  return sell_spmv<double, kSellAdd>(ctx, m, x, y);
}
int hflmp_sell_spmv_add_f32(hflmp_ctx *ctx, const hflmp_sell *m, const float *x, float *y) {
  return sell_spmv<float, kSellAdd>(ctx, m, x, y);
}

}  // extern "C"
