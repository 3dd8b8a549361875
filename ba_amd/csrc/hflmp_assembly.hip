// Finite-element assembly into CSR on the device, for any mesh given as a
// cell -> dof table.  The reference assembles on the host: the pattern
// holds every pair of dofs that share a cell, columns ascending
// (compute_sparsity_structure, assembly/assembly_utils.cc:144, then
// coo_2_csr, lmp/lmatrix_formats.h:320-456), and the values are added cell
// by cell in ascending cell order (add_values, lmp/lmatrix_csr_cpu.cc:
// 1102-1153) to a zeroed matrix.
//
// Here one thread owns one row.  A node -> cell incidence list, sorted by
// cell, gives the row's cells in ascending order; the row's columns are
// the sorted distinct dofs of those cells, found by repeated minimum
// search (no per-thread buffer); the row's values are summed from 0 in
// ascending cell order, one term f_c E[a][b] (or the cell's own matrix
// entry) per cell, so every value is bit-identical to the cell loop.  No
// atomics touch the values.
#include "hflmp_internal.hpp"

#include <climits>
#include <cstring>

struct hflmp_fe_mesh {
  int device;
  int nnode, ncell, nloc;
  int *dofs;          // ncell * nloc (own copy)
  int *inc_off;       // nnode + 1
  long long *inc;     // ncell * nloc keys: cell * 32 + local index, sorted per node
  int *row_len;       // nnode pattern row lengths
  long long *tile_sum; // scan workspace, one value per kScanTile nodes
  long long nnz;
};

namespace {

constexpr int kMaxLoc = 32;

__global__ void __launch_bounds__(kBlock)
    fe_count_kernel(long long n, int nnode, const int *__restrict__ dofs, int *__restrict__ cnt,
                    unsigned int *bad) {
  for (long long k = blockIdx.x * (long long)kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock) {
    const int d = dofs[k];
    if (d < 0 || d >= nnode) {
      atomicAdd(bad, 1u);
      continue;
    }
    atomicAdd(cnt + d, 1);
  }
}

// Exclusive scan of cnt[0..n) into off[0..n] (off[n] = total), in three
// launches: per-tile sums (kScanTile values per workgroup), a one-workgroup
// scan of the tile sums, and a rescan of each tile from its offset.
constexpr int kScanPer = 16;                    // values per thread
constexpr int kScanTile = kBlock * kScanPer;    // values per workgroup

__global__ void __launch_bounds__(kBlock)
    scan_tiles_kernel(int n, const int *__restrict__ cnt, long long *__restrict__ tile_sum) {
  __shared__ long long lds[kBlock / kWave];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) s += cnt[base + k];
  const long long t = block_sum(s, lds);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kFinalBlock)
    scan_offsets_kernel(int ntiles, long long *__restrict__ tile_sum, long long *total) {
  __shared__ long long tot[kFinalBlock];
  const int t = threadIdx.x;
  const int per = (ntiles + kFinalBlock - 1) / kFinalBlock;
  const int lo = t * per < ntiles ? t * per : ntiles;
  const int hi = lo + per < ntiles ? lo + per : ntiles;
  long long s = 0;
  for (int k = lo; k < hi; ++k) s += tile_sum[k];
  tot[t] = s;
  __syncthreads();
  for (int o = 1; o < kFinalBlock; o <<= 1) {
    const long long add = t >= o ? tot[t - o] : 0;
    __syncthreads();
    tot[t] += add;
    __syncthreads();
  }
  long long run = tot[t] - s;
  for (int k = lo; k < hi; ++k) {
    const long long v = tile_sum[k];
    tile_sum[k] = run;  // exclusive offset of tile k
    run += v;
  }
  if (t == kFinalBlock - 1 && total) *total = tot[t];
}

__global__ void __launch_bounds__(kBlock)
    scan_write_kernel(int n, const int *__restrict__ cnt, const long long *__restrict__ tile_off,
                      int *__restrict__ off) {
  __shared__ long long part[kBlock];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) s += cnt[base + k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < kBlock; o <<= 1) {
    const long long add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  long long run = tile_off[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) {
      off[base + k] = (int)run;
      run += cnt[base + k];
    }
  if (base + kScanPer >= n && base < n) off[n] = (int)run;  // the last value's thread
}

// off[0..n] = exclusive scan of cnt, *total (device, optional) = the sum;
// tile_sum holds ceil(n / kScanTile) values
void scan_exclusive(hipStream_t s, int n, const int *cnt, int *off, long long *tile_sum,
                    long long *total) {
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  scan_tiles_kernel<<<ntiles, kBlock, 0, s>>>(n, cnt, tile_sum);
  scan_offsets_kernel<<<1, kFinalBlock, 0, s>>>(ntiles, tile_sum, total);
  scan_write_kernel<<<ntiles, kBlock, 0, s>>>(n, cnt, tile_sum, off);
}

__global__ void __launch_bounds__(kBlock)
    fe_incidence_kernel(long long n, int nloc, const int *__restrict__ dofs,
                        const int *__restrict__ off, int *__restrict__ cursor,
                        long long *__restrict__ inc) {
  for (long long k = blockIdx.x * (long long)kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock) {
    const int d = dofs[k];
    const int p = off[d] + atomicAdd(cursor + d, 1);
    inc[p] = (k / nloc) * kMaxLoc + (k % nloc);
  }
}

// sort each node's incidence keys (ascending cell, then local index)
__global__ void __launch_bounds__(kBlock)
    fe_sort_kernel(int nnode, const int *__restrict__ off, long long *__restrict__ inc) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nnode; i += gridDim.x * kBlock) {
    const int lo = off[i], hi = off[i + 1];
    for (int k = lo + 1; k < hi; ++k) {
      const long long v = inc[k];
      int m = k - 1;
      while (m >= lo && inc[m] > v) {
        inc[m + 1] = inc[m];
        --m;
      }
      inc[m + 1] = v;
    }
  }
}

// Distinct dofs of row i's cells in ascending order: WRITE = false counts
// them into len[i], WRITE = true stores them at col + row_ptr[i].
template <bool WRITE>
__global__ void __launch_bounds__(kBlock)
    fe_pattern_kernel(int nnode, int nloc, const int *__restrict__ dofs,
                      const int *__restrict__ off, const long long *__restrict__ inc,
                      int *__restrict__ len, const int *__restrict__ row_ptr,
                      int *__restrict__ col) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nnode; i += gridDim.x * kBlock) {
    const int lo = off[i], hi = off[i + 1];
    int last = -1, n = 0;
    int *out = WRITE ? col + row_ptr[i] : nullptr;
    for (;;) {
      int next = INT_MAX;
      for (int k = lo; k < hi; ++k) {
        const int *cd = dofs + (inc[k] / kMaxLoc) * nloc;
        for (int b = 0; b < nloc; ++b) {
          const int d = cd[b];
          if (d > last && d < next) next = d;
        }
      }
      if (next == INT_MAX) break;
      if (WRITE) out[n] = next;
      ++n;
      last = next;
    }
    if (!WRITE) len[i] = n;
  }
}

// Row i's values: zero, then for each (cell, local index a) of the node in
// ascending order and each local column b, val[(i, dof b)] += term.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    fe_values_kernel(int nnode, int nloc, const int *__restrict__ dofs,
                     const int *__restrict__ off, const long long *__restrict__ inc,
                     const T *__restrict__ cell_mats, const T *__restrict__ E,
                     const T *__restrict__ factor, const int *__restrict__ row_ptr,
                     const int *__restrict__ col, T *__restrict__ val) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nnode; i += gridDim.x * kBlock) {
    const int r0 = row_ptr[i], r1 = row_ptr[i + 1];
    for (int p = r0; p < r1; ++p) val[p] = T(0);
    for (int k = off[i]; k < off[i + 1]; ++k) {
      const long long c = inc[k] / kMaxLoc;
      const int a = (int)(inc[k] % kMaxLoc);
      const int *cd = dofs + c * nloc;
      const T f = factor ? factor[c] : T(1);
      const T *Er = cell_mats ? cell_mats + (c * nloc + a) * nloc : E + a * nloc;
      for (int b = 0; b < nloc; ++b) {
        const int j = cd[b];
        int lo = r0, hi = r1 - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (col[mid] < j) lo = mid + 1;
          else hi = mid;
        }
        const T term = cell_mats ? Er[b] : f * Er[b];
        val[lo] += term;
      }
    }
  }
}

unsigned grid_n(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

void mesh_free(hflmp_fe_mesh *m) {
  if (!m) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != m->device) (void)hipSetDevice(m->device);
  (void)hipFree(m->dofs);
  (void)hipFree(m->inc_off);
  (void)hipFree(m->inc);
  (void)hipFree(m->row_len);
  (void)hipFree(m->tile_sum);
  if (cur >= 0 && cur != m->device) (void)hipSetDevice(cur);
  delete m;
}

template <typename T>
int fe_assemble(hflmp_ctx *ctx, const hflmp_fe_mesh *m, const T *cell_mats, const T *E,
                const T *factor, const int *row_ptr, const int *col, T *val) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && row_ptr && (m->nnz == 0 || (col && val)) && (cell_mats || E),
                "fe_assemble: null pointer (cell_mats or E is needed)");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "fe_assemble: mesh on another device");
  fe_values_kernel<T><<<grid_n(m->nnode), kBlock, 0, ctx->stream>>>(
      m->nnode, m->nloc, m->dofs, m->inc_off, m->inc, cell_mats, E, factor, row_ptr, col, val);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_fe_mesh_create(hflmp_ctx *ctx, int nnode, int ncell, int nloc, const int *cell_dofs,
                         hflmp_fe_mesh **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, out && cell_dofs, "fe_mesh_create: null pointer");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, nnode > 0 && ncell > 0 && nloc >= 1 && nloc <= kMaxLoc,
                "fe_mesh_create: need nnode > 0, ncell > 0, 1 <= nloc <= %d", kMaxLoc);
  const long long n = (long long)ncell * nloc;
  HFLMP_REQUIRE(ctx, n < 0x7fffffffLL, "fe_mesh_create: more than 2^31-1 cell dofs");
  auto *m = new hflmp_fe_mesh();
  m->device = ctx->device;
  m->nnode = nnode;
  m->ncell = ncell;
  m->nloc = nloc;
  auto fail = [&](int code, const char *msg) {
    mesh_free(m);
    return set_error(ctx, code, "fe_mesh_create: %s", msg);
  };
  int *cnt = nullptr;
  long long *total = nullptr;
  auto fail2 = [&](int code, const char *msg) {
    (void)hipFree(cnt);
    (void)hipFree(total);
    return fail(code, msg);
  };
  if (hipMalloc(&m->dofs, sizeof(int) * (size_t)n) != hipSuccess ||
      hipMalloc(&m->inc_off, sizeof(int) * ((size_t)nnode + 1)) != hipSuccess ||
      hipMalloc(&m->inc, sizeof(long long) * (size_t)n) != hipSuccess ||
      hipMalloc(&m->row_len, sizeof(int) * (size_t)nnode) != hipSuccess ||
      hipMalloc(&m->tile_sum,
                sizeof(long long) * (size_t)((nnode + kScanTile - 1) / kScanTile)) !=
          hipSuccess ||
      hipMalloc(&cnt, sizeof(int) * (size_t)nnode) != hipSuccess ||
      hipMalloc(&total, 2 * sizeof(long long)) != hipSuccess)
    return fail2(HFLMP_ENOMEM, "out of device memory");
  hipStream_t s = ctx->stream;
  unsigned int *bad = ctx->red_counter + 32;
  int rc;
  if ((rc = hflmp_fill_zero(ctx, cnt, sizeof(int) * (size_t)nnode)) != HFLMP_OK ||
      (rc = hflmp_fill_zero(ctx, bad, sizeof(unsigned int))) != HFLMP_OK)
    return fail2(rc, "fill");
  if (hipMemcpyAsync(m->dofs, cell_dofs, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    return fail2(HFLMP_EHIP, "copy failed");
  fe_count_kernel<<<grid_n(n), kBlock, 0, s>>>(n, nnode, m->dofs, cnt, bad);
  unsigned int hbad = 0;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(ctx->host_pinned, bad, sizeof(unsigned int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail2(HFLMP_EHIP, "HIP error while counting");
  std::memcpy(&hbad, ctx->host_pinned, sizeof(hbad));
  if (hbad) return fail2(HFLMP_EINVAL, "a cell dof is outside [0, nnode)");
  scan_exclusive(s, nnode, cnt, m->inc_off, m->tile_sum, nullptr);
  if ((rc = hflmp_fill_zero(ctx, cnt, sizeof(int) * (size_t)nnode)) != HFLMP_OK)
    return fail2(rc, "fill");
  fe_incidence_kernel<<<grid_n(n), kBlock, 0, s>>>(n, nloc, m->dofs, m->inc_off, cnt, m->inc);
  fe_sort_kernel<<<grid_n(nnode), kBlock, 0, s>>>(nnode, m->inc_off, m->inc);
  fe_pattern_kernel<false><<<grid_n(nnode), kBlock, 0, s>>>(
      nnode, nloc, m->dofs, m->inc_off, m->inc, m->row_len, nullptr, nullptr);
  // total pattern entries: scan of the row lengths into the (now free) cnt
  // area is not needed, only the sum
  int *scratch_off = nullptr;
  if (hipMalloc(&scratch_off, sizeof(int) * ((size_t)nnode + 1)) != hipSuccess)
    return fail2(HFLMP_ENOMEM, "out of device memory");
  scan_exclusive(s, nnode, m->row_len, scratch_off, m->tile_sum, total);
  long long nnz = 0;
  const bool ok = hipGetLastError() == hipSuccess &&
                  hipMemcpyAsync(&nnz, total, sizeof(nnz), hipMemcpyDeviceToHost, s) ==
                      hipSuccess &&
                  hipStreamSynchronize(s) == hipSuccess;
  (void)hipFree(scratch_off);
  if (!ok) return fail2(HFLMP_EHIP, "HIP error while building the pattern");
  if (nnz > 0x7fffffffLL) return fail2(HFLMP_EINVAL, "pattern has more than 2^31-1 entries");
  m->nnz = nnz;
  (void)hipFree(cnt);
  (void)hipFree(total);
  *out = m;
  return HFLMP_OK;
}

int hflmp_fe_mesh_nnz(const hflmp_fe_mesh *m, long long *nnz) {
  if (!m || !nnz) return HFLMP_EINVAL;
  *nnz = m->nnz;
  return HFLMP_OK;
}

int hflmp_fe_mesh_pattern(hflmp_ctx *ctx, const hflmp_fe_mesh *m, int *row_ptr, int *col) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && row_ptr && (m->nnz == 0 || col), "fe_mesh_pattern: null pointer");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "fe_mesh_pattern: mesh on another device");
  hipStream_t s = ctx->stream;
  scan_exclusive(s, m->nnode, m->row_len, row_ptr, m->tile_sum, nullptr);
  fe_pattern_kernel<true><<<grid_n(m->nnode), kBlock, 0, s>>>(
      m->nnode, m->nloc, m->dofs, m->inc_off, m->inc, nullptr, row_ptr, col);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

int hflmp_fe_assemble_f64(hflmp_ctx *ctx, const hflmp_fe_mesh *m, const double *cell_mats,
                          const double *E, const double *factor, const int *row_ptr,
                          const int *col, double *val) {
  return fe_assemble<double>(ctx, m, cell_mats, E, factor, row_ptr, col, val);
}

int hflmp_fe_assemble_f32(hflmp_ctx *ctx, const hflmp_fe_mesh *m, const float *cell_mats,
                          const float *E, const float *factor, const int *row_ptr,
                          const int *col, float *val) {
  return fe_assemble<float>(ctx, m, cell_mats, E, factor, row_ptr, col, val);
}

void hflmp_fe_mesh_destroy(hflmp_fe_mesh *m) { mesh_free(m); }

}  // extern "C"
