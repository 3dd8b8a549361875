// Building the row-block layout of hflmp_spmm_plan.hpp on the device, and
// Y = A X through it.
//
// Build: per block of R rows, a bitmap of the block's column window
// [cmin, cmax] (at most kPlanBits columns) is set from its entries; the
// set bits in order are the block's distinct columns, and an entry's
// position in the list is the number of set bits below its column (a
// prefix count over the bitmap words plus a popcount in its word).  A first
// pass counts each block's distinct columns and entries; the host picks the
// largest R of 64, 32, 16, 8, 4 whose largest block fits the walk's limits
// and the LDS budget, forms the offsets, and a second pass writes the lists
// and the entry positions.
#include <algorithm>
#include <climits>
#include <vector>

#include "hflmp_spmm_plan.hpp"

namespace {

constexpr int kPlanWords = 2048;  // bitmap words: a block's columns span < 65,536
constexpr int kPlanBits = 32 * kPlanWords;

__device__ __forceinline__ int block_min_max(int v, bool is_min, int *lds) {
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, kWave);
    v = is_min ? (o < v ? o : v) : (o > v ? o : v);
  }
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  int r = lds[0];
  for (int w = 1; w < kBlock / kWave; ++w) r = is_min ? (lds[w] < r ? lds[w] : r) : (lds[w] > r ? lds[w] : r);
  return r;
}

// Exclusive prefix sum of one value per thread over the workgroup (thread
// order); *total gets the sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int *total) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  int incl = v;
  for (int off = 1; off < kWave; off <<= 1) {
    const int o = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += o;
  }
  __syncthreads();
  if (lane == kWave - 1) lds[wid] = incl;
  __syncthreads();
  int before = 0, all = 0;
  for (int w = 0; w < kBlock / kWave; ++w) {
    if (w < wid) before += lds[w];
    all += lds[w];
  }
  *total = all;
  return before + incl - v;
}

struct BlockSpan {
  int j0, j1, cmin, cmax;
};

// The block's entry range and column window; every thread gets the same.
__device__ __forceinline__ BlockSpan block_span(int b, int nrow, int R, const int *__restrict__ rp,
                                                const int *__restrict__ ci, int *lds) {
  BlockSpan s;
  const int r0 = b * R;
  const int r1 = r0 + R < nrow ? r0 + R : nrow;
  s.j0 = rp[r0];
  s.j1 = rp[r1];
  int mn = INT_MAX, mx = INT_MIN;
  for (int j = s.j0 + threadIdx.x; j < s.j1; j += kBlock) {
    const int c = ci[j];
    mn = c < mn ? c : mn;
    mx = c > mx ? c : mx;
  }
  s.cmin = block_min_max(mn, true, lds);
  s.cmax = block_min_max(mx, false, lds);
  return s;
}

// The block's column bitmap over [cmin, cmin + 32 nwords).
__device__ __forceinline__ void block_bitmap(const BlockSpan &s, int nwords,
                                             const int *__restrict__ ci, unsigned *bits) {
  __syncthreads();
  for (int w = threadIdx.x; w < nwords; w += kBlock) bits[w] = 0u;
  __syncthreads();
  for (int j = s.j0 + threadIdx.x; j < s.j1; j += kBlock) {
    const int d = ci[j] - s.cmin;
    atomicOr(&bits[d >> 5], 1u << (d & 31));
  }
  __syncthreads();
}

// ucnt[b]: distinct columns of block b, -1 if its window exceeds the
// bitmap; ecnt[b]: its entries; bad: entries whose column is outside
// [0, ncol).
__global__ void __launch_bounds__(kBlock)
    plan_count_kernel(int nrow, int ncol, int R, int nblk, const int *__restrict__ rp,
                      const int *__restrict__ ci, int *__restrict__ ucnt,
                      int *__restrict__ ecnt, unsigned int *__restrict__ bad) {
  __shared__ unsigned bits[kPlanWords];
  __shared__ int lds[kBlock / kWave];
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const BlockSpan s = block_span(b, nrow, R, rp, ci, lds);
    if (threadIdx.x == 0) ecnt[b] = s.j1 - s.j0;
    if (s.j1 <= s.j0) {
      if (threadIdx.x == 0) ucnt[b] = 0;
      continue;
    }
    if (s.cmin < 0 || s.cmax >= ncol) {
      if (threadIdx.x == 0) {
        ucnt[b] = 0;
        atomicAdd(bad, 1u);
      }
      continue;
    }
    if ((long long)s.cmax - s.cmin >= kPlanBits) {
      if (threadIdx.x == 0) ucnt[b] = -1;
      continue;
    }
    const int nwords = (s.cmax - s.cmin) / 32 + 1;
    block_bitmap(s, nwords, ci, bits);
    int cnt = 0;
    for (int w = threadIdx.x; w < nwords; w += kBlock) cnt += __popc(bits[w]);
    int total;
    (void)block_exclusive_scan(cnt, lds, &total);
    if (threadIdx.x == 0) ucnt[b] = total;
  }
}

// The lists and the entry positions (blocks with a valid count only).
__global__ void __launch_bounds__(kBlock)
    plan_fill_kernel(int nrow, int R, int nblk, const int *__restrict__ rp,
                     const int *__restrict__ ci, const int *__restrict__ uoff,
                     int *__restrict__ ucol, int *__restrict__ lidx) {
  __shared__ unsigned bits[kPlanWords];
  __shared__ int pre[kPlanWords];
  __shared__ int lds[kBlock / kWave];
  constexpr int per = kPlanWords / kBlock;  // words per thread in the scan
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const BlockSpan s = block_span(b, nrow, R, rp, ci, lds);
    if (s.j1 <= s.j0) continue;
    const int nwords = (s.cmax - s.cmin) / 32 + 1;
    block_bitmap(s, nwords, ci, bits);
    // prefix counts: thread t owns words [t per, t per + per)
    int c[per], sum = 0;
#pragma unroll
    for (int q = 0; q < per; ++q) {
      const int w = threadIdx.x * per + q;
      c[q] = w < nwords ? __popc(bits[w]) : 0;
      sum += c[q];
    }
    int total;
    int run = block_exclusive_scan(sum, lds, &total);
#pragma unroll
    for (int q = 0; q < per; ++q) {
      const int w = threadIdx.x * per + q;
      if (w < nwords) pre[w] = run;
      run += c[q];
    }
    __syncthreads();
    int *__restrict__ out = ucol + uoff[b];
    for (int w = threadIdx.x; w < nwords; w += kBlock) {
      unsigned m = bits[w];
      int r = pre[w];
      while (m) {
        const int bit = __ffs(m) - 1;
        out[r++] = s.cmin + 32 * w + bit;
        m &= m - 1u;
      }
    }
    for (int j = s.j0 + threadIdx.x; j < s.j1; j += kBlock) {
      const int d = ci[j] - s.cmin;
      const int w = d >> 5;
      lidx[j] = pre[w] + __popc(bits[w] & ((1u << (d & 31)) - 1u));
    }
  }
}

template <typename T, int KC>
__global__ void __launch_bounds__(kBlock)
    plan_spmm_kernel(PlanView pv, int k, const int *__restrict__ rp, const T *__restrict__ val,
                     const T *__restrict__ X, T *__restrict__ Y) {
  extern __shared__ __align__(16) unsigned char plan_lds[];
  const int lane = threadIdx.x & (kWave - 1);
  plan_walk<T, KC>(pv, k, rp, val, X, plan_lds, [&](int i, T s) {
    if (lane < k) __builtin_nontemporal_store(s, Y + (size_t)i * k + lane);
  });
}

void free_plan(hflmp_spmm_plan *p) {
  if (!p) return;
  int cur = 0;
  const bool sw = hipGetDevice(&cur) == hipSuccess && cur != p->device;
  if (sw) (void)hipSetDevice(p->device);
  if (p->uoff) (void)hipFree(p->uoff);
  if (p->ucol) (void)hipFree(p->ucol);
  if (p->lidx) (void)hipFree(p->lidx);
  if (sw) (void)hipSetDevice(cur);
  delete p;
}

int create_plan(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *rp, const int *ci,
                int nrhs, int vb, int lds_budget, hflmp_spmm_plan **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, out != nullptr, "spmm_plan: null output");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, nrow > 0 && ncol > 0 && nnz > 0, "spmm_plan: empty matrix");
  HFLMP_REQUIRE(ctx, rp && ci, "spmm_plan: null pointer");
  HFLMP_REQUIRE(ctx, nrhs >= 1 && nrhs <= 64 && (vb == 4 || vb == 8),
                "spmm_plan: need 1 <= nrhs <= 64 and value_bytes 4 or 8");
  if (lds_budget <= 0) lds_budget = 64 * 1024;
  HFLMP_REQUIRE(ctx, lds_budget <= 160 * 1024, "spmm_plan: lds_budget above 160 KiB");
  hipStream_t s = ctx->stream;
  const int max_nblk = (nrow + 3) / 4;
  int *d_cnt = nullptr;
  unsigned int *d_bad = nullptr;
  HFLMP_HIP_CHECK(ctx, hipMalloc(&d_cnt, sizeof(int) * 2 * ((size_t)max_nblk + 1)));
  if (hipMalloc(&d_bad, sizeof(unsigned int)) != hipSuccess) {
    (void)hipFree(d_cnt);
    return set_error(ctx, HFLMP_ENOMEM, "spmm_plan: allocation failed");
  }
  int *d_ecnt = d_cnt + max_nblk + 1;
  auto release = [&]() {
    (void)hipFree(d_cnt);
    (void)hipFree(d_bad);
  };
  std::vector<int> cnt, ecnt;
  int R = 0, nblk = 0, umax = 0, emax = 0;
  for (int r : {64, 32, 16, 8, 4}) {
    const int nb = (nrow + r - 1) / r;
    if (hipMemsetAsync(d_bad, 0, sizeof(unsigned int), s) != hipSuccess) {
      release();
      return set_error(ctx, HFLMP_EHIP, "spmm_plan: memset failed");
    }
    const int g = nb < 4096 ? nb : 4096;
    plan_count_kernel<<<g, kBlock, 0, s>>>(nrow, ncol, r, nb, rp, ci, d_cnt, d_ecnt, d_bad);
    cnt.assign(nb, 0);
    ecnt.assign(nb, 0);
    unsigned int bad = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipMemcpyAsync(ecnt.data(), d_ecnt, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipMemcpyAsync(&bad, d_bad, sizeof(unsigned int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      release();
      return set_error(ctx, HFLMP_EHIP, "spmm_plan: count pass failed");
    }
    if (bad) {
      release();
      return set_error(ctx, HFLMP_EINVAL, "spmm_plan: %u row blocks hold columns outside [0, %d)",
                       bad, ncol);
    }
    const int mn = *std::min_element(cnt.begin(), cnt.end());
    const int mx = *std::max_element(cnt.begin(), cnt.end());
    const int ex = *std::max_element(ecnt.begin(), ecnt.end());
    if (mn < 0 || mx > plan_max_cols(vb) || ex > kPlanMaxEntries ||
        plan_lds_bytes(r, mx, ex, nrhs, vb) > (size_t)lds_budget)
      continue;
    R = r;
    nblk = nb;
    umax = mx;
    emax = ex;
    break;
  }
  if (!R) {
    release();
    return set_error(ctx, HFLMP_EFORMAT,
                     "spmm_plan: no block of 4..64 rows keeps to %d distinct columns within a "
                     "%d-column window, %d entries and %d bytes of LDS for %d x %d-byte rows",
                     plan_max_cols(vb), kPlanBits, kPlanMaxEntries, lds_budget, nrhs, vb);
  }
  std::vector<int> off(nblk + 1);
  long long acc = 0;
  for (int b = 0; b < nblk; ++b) {
    off[b] = (int)acc;
    acc += cnt[b];
  }
  off[nblk] = (int)acc;
  auto *p = new hflmp_spmm_plan();
  p->device = ctx->device;
  p->nrow = nrow;
  p->ncol = ncol;
  p->nnz = nnz;
  p->R = R;
  p->nblk = nblk;
  p->umax = umax;
  p->emax = emax;
  p->nrhs = nrhs;
  p->value_bytes = vb;
  p->nuniq = acc;
  if (hipMalloc(&p->uoff, sizeof(int) * ((size_t)nblk + 1)) != hipSuccess ||
      // one entry past the lists: an empty last block still loads ucol[uoff[nblk]]
      hipMalloc(&p->ucol, sizeof(int) * (size_t)(acc + 1)) != hipSuccess ||
      hipMalloc(&p->lidx, sizeof(int) * (size_t)nnz) != hipSuccess) {
    free_plan(p);
    release();
    return set_error(ctx, HFLMP_ENOMEM, "spmm_plan: allocation failed");
  }
  const int g = nblk < 4096 ? nblk : 4096;
  if (hipMemcpyAsync(p->uoff, off.data(), sizeof(int) * ((size_t)nblk + 1),
                     hipMemcpyHostToDevice, s) != hipSuccess) {
    free_plan(p);
    release();
    return set_error(ctx, HFLMP_EHIP, "spmm_plan: upload failed");
  }
  plan_fill_kernel<<<g, kBlock, 0, s>>>(nrow, R, nblk, rp, ci, p->uoff, p->ucol, p->lidx);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    free_plan(p);
    release();
    return set_error(ctx, HFLMP_EHIP, "spmm_plan: fill pass failed");
  }
  release();
  *out = p;
  return HFLMP_OK;
}

template <typename T>
int plan_spmm(hflmp_ctx *ctx, const hflmp_spmm_plan *p, const int *rp, const T *val,
              const T *X, T *Y, int k) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, p && rp && val && X && Y, "plan_spmm: null pointer");
  HFLMP_REQUIRE(ctx, k >= 1 && k <= 64, "plan_spmm: need 1 <= nrhs <= 64");
  HFLMP_REQUIRE(ctx, p->umax <= plan_max_cols((int)sizeof(T)),
                "plan_spmm: the plan's blocks have up to %d columns, above %d for %zu-byte "
                "values; build it for these values", p->umax, plan_max_cols((int)sizeof(T)),
                sizeof(T));
  const size_t lds = plan_lds_bytes(p->R, p->umax, p->emax, k, (int)sizeof(T));
  HFLMP_REQUIRE(ctx, lds <= 160 * 1024,
                "plan_spmm: %zu bytes of LDS for %d x %zu-byte rows; build the plan for "
                "these rows", lds, k, sizeof(T));
  const int g = plan_grid(ctx->num_cus, p->nblk, lds);
  // 64 vectors, the common block width, with the row stride known to the compiler
  auto kern = k == 64 ? &plan_spmm_kernel<T, 64> : &plan_spmm_kernel<T, 0>;
  if (lds > 64 * 1024)
    HFLMP_HIP_CHECK(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds));
  kern<<<g, kBlock, lds, ctx->stream>>>(plan_view(p), k, rp, val, X, Y);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_spmm_plan_create(hflmp_ctx *ctx, int nrow, int ncol, int nnz, const int *row_ptr,
                           const int *col, int nrhs, int value_bytes, int lds_budget,
                           hflmp_spmm_plan **plan) {
  return create_plan(ctx, nrow, ncol, nnz, row_ptr, col, nrhs, value_bytes, lds_budget, plan);
}

void hflmp_spmm_plan_destroy(hflmp_spmm_plan *plan) { free_plan(plan); }

int hflmp_spmm_plan_info(const hflmp_spmm_plan *plan, int *rows_per_block,
                         int *max_cols_per_block, long long *list_entries,
                         long long *device_bytes) {
  if (!plan) return HFLMP_EINVAL;
  if (rows_per_block) *rows_per_block = plan->R;
  if (max_cols_per_block) *max_cols_per_block = plan->umax;
  if (list_entries) *list_entries = plan->nuniq;
  if (device_bytes)
    *device_bytes = 4LL * (plan->nblk + 1) + 4LL * plan->nuniq + 4LL * plan->nnz;
  return HFLMP_OK;
}

int hflmp_plan_spmm_f64(hflmp_ctx *ctx, const hflmp_spmm_plan *plan, const int *row_ptr,
                        const double *val, const double *X, double *Y, int nrhs) {
  return plan_spmm<double>(ctx, plan, row_ptr, val, X, Y, nrhs);
}

int hflmp_plan_spmm_f32(hflmp_ctx *ctx, const hflmp_spmm_plan *plan, const int *row_ptr,
                        const float *val, const float *X, float *Y, int nrhs) {
  return plan_spmm<float>(ctx, plan, row_ptr, val, X, Y, nrhs);
}

}  // extern "C"
