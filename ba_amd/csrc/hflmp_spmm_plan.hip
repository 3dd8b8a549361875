// Building the block layout of hflmp_spmm_plan.hpp on the device, and
// Y = A X through it.
//
// Geometry: the line length L is read off a sample of rows in the middle
// of the matrix (the positive column offsets form a band 1..w around the
// diagonal and, on a structured mesh, the next band starts at L - w).
// Candidates are runs of 4..64 consecutive rows (one line of L = nrow) and
// a x nl patches (a = 2..16 rows of nl = 2..8 lines of L rows, a multiple
// of 4 and at most 64 rows); the host keeps the one with the fewest staged
// X rows in total that fits the walk's limits and the LDS budget (one
// counting pass each).
//
// Build: per block, a bitmap of the block's column window [cmin, cmax]
// (at most kPlanBits columns) is set from its entries; the set bits in
// order are the block's distinct columns, and an entry's position in the
// list is the number of set bits below its column (a prefix count over the
// bitmap words plus a popcount in its word).  A first pass counts each
// block's distinct columns and entries, the host forms the offsets of the
// chosen geometry, and a second pass writes the lists and the positions.
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

// Calls f(j) for the CSR entries of block b (its runs in order), each
// entry on one thread of the workgroup.
template <typename F>
__device__ __forceinline__ void for_block_entries(const PlanGeom &g, int b,
                                                  const int *__restrict__ rp, F &&f) {
  for (int t = 0; t < g.nl; ++t) {
    const PlanRun run = plan_run(g, b, t);
    const int j1 = rp[run.r0 + run.n];
    for (int j = rp[run.r0] + threadIdx.x; j < j1; j += kBlock) f(j);
  }
}

struct BlockSpan {
  int ne, cmin, cmax;
};

// The block's entry count and column window; every thread gets the same.
__device__ __forceinline__ BlockSpan block_span(const PlanGeom &g, int b,
                                                const int *__restrict__ rp,
                                                const int *__restrict__ ci, int *lds) {
  BlockSpan s;
  s.ne = 0;
  for (int t = 0; t < g.nl; ++t) {
    const PlanRun run = plan_run(g, b, t);
    s.ne += rp[run.r0 + run.n] - rp[run.r0];
  }
  int mn = INT_MAX, mx = INT_MIN;
  for_block_entries(g, b, rp, [&](int j) {
    const int c = ci[j];
    mn = c < mn ? c : mn;
    mx = c > mx ? c : mx;
  });
  s.cmin = block_min_max(mn, true, lds);
  s.cmax = block_min_max(mx, false, lds);
  return s;
}

// The block's column bitmap over [cmin, cmin + 32 nwords).
__device__ __forceinline__ void block_bitmap(const PlanGeom &g, int b, const BlockSpan &s,
                                             int nwords, const int *__restrict__ rp,
                                             const int *__restrict__ ci, unsigned *bits) {
  __syncthreads();
  for (int w = threadIdx.x; w < nwords; w += kBlock) bits[w] = 0u;
  __syncthreads();
  for_block_entries(g, b, rp, [&](int j) {
    const int d = ci[j] - s.cmin;
    atomicOr(&bits[d >> 5], 1u << (d & 31));
  });
  __syncthreads();
}

// ucnt[b]: distinct columns of block b, -1 if its window exceeds the
// bitmap; ecnt[b]: its entries; bad: entries whose column is outside
// [0, ncol).
__global__ void __launch_bounds__(kBlock)
    plan_count_kernel(PlanGeom g, int ncol, int nblk, const int *__restrict__ rp,
                      const int *__restrict__ ci, int *__restrict__ ucnt,
                      int *__restrict__ ecnt, unsigned int *__restrict__ bad) {
  __shared__ unsigned bits[kPlanWords];
  __shared__ int lds[kBlock / kWave];
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const BlockSpan s = block_span(g, b, rp, ci, lds);
    if (threadIdx.x == 0) ecnt[b] = s.ne;
    if (s.ne <= 0) {
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
    block_bitmap(g, b, s, nwords, rp, ci, bits);
    int cnt = 0;
    for (int w = threadIdx.x; w < nwords; w += kBlock) cnt += __popc(bits[w]);
    int total;
    (void)block_exclusive_scan(cnt, lds, &total);
    if (threadIdx.x == 0) ucnt[b] = total;
  }
}

// The lists and the entry positions (blocks with a valid count only).
__global__ void __launch_bounds__(kBlock)
    plan_fill_kernel(PlanGeom g, int nblk, const int *__restrict__ rp,
                     const int *__restrict__ ci, const int *__restrict__ uoff,
                     int *__restrict__ ucol, int *__restrict__ lidx) {
  __shared__ unsigned bits[kPlanWords];
  __shared__ int pre[kPlanWords];
  __shared__ int lds[kBlock / kWave];
  constexpr int per = kPlanWords / kBlock;  // words per thread in the scan
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const BlockSpan s = block_span(g, b, rp, ci, lds);
    if (s.ne <= 0) continue;
    const int nwords = (s.cmax - s.cmin) / 32 + 1;
    block_bitmap(g, b, s, nwords, rp, ci, bits);
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
    for_block_entries(g, b, rp, [&](int j) {
      const int d = ci[j] - s.cmin;
      const int w = d >> 5;
      lidx[j] = pre[w] + __popc(bits[w] & ((1u << (d & 31)) - 1u));
    });
  }
}

template <typename T, int KC>
__global__ void __launch_bounds__(kBlock)
    plan_spmm_kernel(PlanView pv, int k, const int *__restrict__ rp, const T *__restrict__ val,
                     const T *__restrict__ X, T *__restrict__ Y) {
  extern __shared__ __align__(16) unsigned char plan_lds[];
  plan_walk<T, KC>(
      pv, k, rp, val, X, plan_lds, [](int) { return T(0); },
      [&](int i, int c, T s, T) {
        if (c < k) __builtin_nontemporal_store(s, Y + (size_t)i * k + c);
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

// The line length of a structured numbering, read off up to 64 rows in the
// middle of the matrix: the positive column offsets 1..w of the band around
// the diagonal, then the first offset d1 > w + 1 of the next band, which
// is centred on L = d1 + w.  0 when there is no second band.
int detect_line(hflmp_ctx *ctx, int nrow, const int *rp, const int *ci, int *L) {
  *L = 0;
  const int r0 = nrow / 2, nr = std::min(64, nrow - r0);
  if (nr <= 0) return HFLMP_OK;
  std::vector<int> hrp(nr + 1);
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(hrp.data(), rp + r0, sizeof(int) * (nr + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const int ne = hrp[nr] - hrp[0];
  if (ne <= 0) return HFLMP_OK;
  std::vector<int> hci(ne);
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(hci.data(), ci + hrp[0], sizeof(int) * (size_t)ne,
                                      hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<long long> d;
  for (int r = 0; r < nr; ++r)
    for (int j = hrp[r] - hrp[0]; j < hrp[r + 1] - hrp[0]; ++j)
      if (hci[j] > r0 + r) d.push_back((long long)hci[j] - (r0 + r));
  std::sort(d.begin(), d.end());
  d.erase(std::unique(d.begin(), d.end()), d.end());
  long long w = 0;
  size_t i = 0;
  while (i < d.size() && d[i] == w + 1) w = d[i++];
  if (i == d.size()) return HFLMP_OK;
  const long long line = d[i] + w;
  if (line >= 2 && line <= nrow / 2) *L = (int)line;
  return HFLMP_OK;
}

struct PlanChoice {
  PlanGeom g;
  int nblk, umax, emax;
  long long nuniq;
  std::vector<int> cnt;
};

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
  int L = 0;
  int rc = detect_line(ctx, nrow, rp, ci, &L);
  if (rc) return rc;
  // candidates: runs of consecutive rows, then patches of lines of L rows
  std::vector<PlanGeom> cand;
  for (int r : {64, 32, 16, 8, 4}) cand.push_back(plan_geom(nrow, nrow, r, 1));
  // (patches of a multiple of 4 rows: the four waves take equal shares)
  if (L > 0)
    for (int nl = 2; nl <= kPlanMaxRuns; ++nl)
      for (int a = 2; a <= 16 && a <= L && a * nl <= 64; ++a)
        if ((a * nl) % 4 == 0) cand.push_back(plan_geom(nrow, L, a, nl));
  int max_nblk = 0;
  for (const PlanGeom &g : cand) max_nblk = std::max(max_nblk, plan_nblk(g));
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
  PlanChoice best;
  best.nblk = 0;
  std::vector<int> cnt, ecnt;
  for (const PlanGeom &g : cand) {
    const int nb = plan_nblk(g);
    if (hipMemsetAsync(d_bad, 0, sizeof(unsigned int), s) != hipSuccess) {
      release();
      return set_error(ctx, HFLMP_EHIP, "spmm_plan: memset failed");
    }
    const int grid = nb < 4096 ? nb : 4096;
    plan_count_kernel<<<grid, kBlock, 0, s>>>(g, ncol, nb, rp, ci, d_cnt, d_ecnt, d_bad);
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
    const int R = g.a * g.nl;
    if (mn < 0 || mx > plan_max_cols(vb) || ex > kPlanMaxEntries ||
        plan_lds_bytes(R, mx, ex, nrhs, vb) > (size_t)lds_budget)
      continue;
    long long tot = 0;
    for (int c : cnt) tot += c;
    // fewest staged X rows; on a tie, more rows per block
    if (best.nblk == 0 || tot < best.nuniq ||
        (tot == best.nuniq && R > best.g.a * best.g.nl)) {
      best.g = g;
      best.nblk = nb;
      best.umax = mx;
      best.emax = ex;
      best.nuniq = tot;
      best.cnt = cnt;
    }
  }
  if (best.nblk == 0) {
    release();
    return set_error(ctx, HFLMP_EFORMAT,
                     "spmm_plan: no block of 4..64 rows keeps to %d distinct columns within a "
                     "%d-column window, %d entries and %d bytes of LDS for %d x %d-byte rows",
                     plan_max_cols(vb), kPlanBits, kPlanMaxEntries, lds_budget, nrhs, vb);
  }
  const int nblk = best.nblk;
  std::vector<int> off(nblk + 1);
  long long acc = 0;
  for (int b = 0; b < nblk; ++b) {
    off[b] = (int)acc;
    acc += best.cnt[b];
  }
  off[nblk] = (int)acc;
  auto *p = new hflmp_spmm_plan();
  p->device = ctx->device;
  p->nrow = nrow;
  p->ncol = ncol;
  p->nnz = nnz;
  p->L = best.g.L;
  p->a = best.g.a;
  p->nl = best.g.nl;
  p->R = best.g.a * best.g.nl;
  p->nblk = nblk;
  p->umax = best.umax;
  p->emax = best.emax;
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
  const int grid = nblk < 4096 ? nblk : 4096;
  if (hipMemcpyAsync(p->uoff, off.data(), sizeof(int) * ((size_t)nblk + 1),
                     hipMemcpyHostToDevice, s) != hipSuccess) {
    free_plan(p);
    release();
    return set_error(ctx, HFLMP_EHIP, "spmm_plan: upload failed");
  }
  plan_fill_kernel<<<grid, kBlock, 0, s>>>(best.g, nblk, rp, ci, p->uoff, p->ucol, p->lidx);
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
  // 64 vectors, the common block width, with the row stride known to the compiler
  auto kern = k == 64 ? &plan_spmm_kernel<T, 64> : &plan_spmm_kernel<T, 0>;
  if (lds > 64 * 1024)
    HFLMP_HIP_CHECK(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds));
  const int g = plan_grid(ctx, reinterpret_cast<const void *>(kern), p->nblk, lds);
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

int hflmp_spmm_plan_geometry(const hflmp_spmm_plan *plan, int *line_rows, int *run_rows,
                             int *runs) {
  if (!plan) return HFLMP_EINVAL;
  if (line_rows) *line_rows = plan->L;
  if (run_rows) *run_rows = plan->a;
  if (runs) *runs = plan->nl;
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
