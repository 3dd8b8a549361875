// Symmetric Gauss-Seidel in the sequential row order of the CPU path:
// z = M^-1 r, M = (D + L) D^-1 (D + U), as CPU_CSR_lMatrix::Psgauss_seidel
// (lmatrix_csr_cpu.cc:495-565) computes it, and its block form
// BlocksPsgauss_seidel (lmatrix_csr_cpu.cc:568-641; the CUDA platform's
// GPUscalar_CSR_lMatrix::BlocksPsgauss_seidel, cuda/GPUscalar_csr_lmatrix.cu:
// 314-363, kernel_BlockPsgauss_seidel, cuda/cuda_spmv_kernel_csr_scalar.h:
// 65-143): num_blocks row ranges of nrow / num_blocks rows (the last takes
// the rest), each swept on its own over the entries whose column lies in
// the range.
//
// Per row the operations and their order are the reference's:
//   forward (rows ascending)   t = r_i; over the row in storage order up to
//                              the first diagonal entry: t -= z_c a_ij for
//                              c < i; z_i = t / a_ii
//   D                          z_i *= a_ii (first diagonal entry)
//   backward (rows descending) over the whole row: t -= z_c a_ij for c > i;
//                              z_i = t / a_ii (last diagonal entry)
// The reference runs the sweeps one row at a time (the CUDA kernel one
// thread per block).  Here the rows are grouped by dependency level: a
// row's forward level is one more than the highest level among the rows it
// reads, so every row of a level reads only values that earlier levels have
// finished.  Each level is one launch over its rows, or, for runs of levels
// of at most kSmallLevel rows, one workgroup that walks the levels with a
// barrier between them.  The values every row reads are the ones the
// sequential sweep reads, so z is bit-identical to it.  The D scaling is
// applied by each row's backward step before its first subtraction, which
// is where the sequential sweep's separate D pass leaves the value (no row
// reads another row's scaled value before that row's backward step).
#include <algorithm>
#include <vector>

#include "hflmp_internal.hpp"

namespace {

constexpr int kSmallLevel = kFinalBlock;  // rows of a level one workgroup takes

// Row range [lo, hi) of row i's block (BlockPsgauss_seidel's start_i/end_i).
struct Blocks {
  int nrow, nb, step;
  __host__ __device__ __forceinline__ void range(int i, int &lo, int &hi) const {
    if (step == 0) {  // every block but the last is empty; the last is all rows
      lo = 0;
      hi = nrow;
      return;
    }
    int b = i / step;
    if (b > nb - 1) b = nb - 1;
    lo = b * step;
    hi = b == nb - 1 ? nrow : lo + step;
  }
};

template <typename T>
struct SgsArgs {
  const int *rp, *ci;
  const T *val;
  const T *r;
  T *z;
  Blocks blk;
};

template <typename T>
__device__ __forceinline__ void forward_row(const SgsArgs<T> &a, int i) {
#pragma clang fp contract(off)
  int lo, hi;
  a.blk.range(i, lo, hi);
  T t = a.r[i];
  int dj = -1;
  for (int j = a.rp[i]; j < a.rp[i + 1]; ++j) {
    const int c = a.ci[j];
    if (c >= lo && c < hi) {
      if (c < i) t -= a.z[c] * a.val[j];
      if (c == i) {
        dj = j;
        break;
      }
    }
  }
  // the build refused rows without a diagonal entry, so dj >= 0
  a.z[i] = t / a.val[dj];
}

template <typename T>
__device__ __forceinline__ void backward_row(const SgsArgs<T> &a, int i) {
#pragma clang fp contract(off)
  int lo, hi;
  a.blk.range(i, lo, hi);
  const int j0 = a.rp[i], j1 = a.rp[i + 1];
  int dfirst = -1;
  for (int j = j0; j < j1; ++j)
    if (a.ci[j] == i) {
      dfirst = j;
      break;
    }
  T t = a.z[i] * a.val[dfirst];  // the D pass
  int dlast = dfirst;
  for (int j = j0; j < j1; ++j) {
    const int c = a.ci[j];
    if (c >= lo && c < hi) {
      if (c == i) dlast = j;
      if (c > i) t -= a.z[c] * a.val[j];
    }
  }
  a.z[i] = t / a.val[dlast];
}

// One level: rows perm[k0 .. k1), one thread each.
template <typename T, bool FWD>
__global__ void __launch_bounds__(kBlock)
    sgs_level_kernel(SgsArgs<T> a, const int *__restrict__ perm, int k0, int k1) {
  const int k = k0 + blockIdx.x * kBlock + threadIdx.x;
  if (k >= k1) return;
  const int i = perm[k];
  if (FWD)
    forward_row(a, i);
  else
    backward_row(a, i);
}

// Levels l0 .. l1 - 1 (each at most kSmallLevel rows) in one workgroup; the
// barrier between levels orders each level's stores before the next
// level's loads (all waves share the CU's caches).
template <typename T, bool FWD>
__global__ void __launch_bounds__(kFinalBlock)
    sgs_levels_kernel(SgsArgs<T> a, const int *__restrict__ perm, const int *__restrict__ loff,
                      int l0, int l1) {
  for (int l = l0; l < l1; ++l) {
    const int k = loff[l] + threadIdx.x;
    if (k < loff[l + 1]) {
      const int i = perm[k];
      if (FWD)
        forward_row(a, i);
      else
        backward_row(a, i);
    }
    __syncthreads();
  }
}

// A sweep's launches: a single level (rows k0..k1) or a run of small levels.
struct Segment {
  bool small;
  int a, b;  // small: levels [a, b); else rows [a, b) of the permutation
};

struct Sweep {
  int levels = 0;
  int *perm = nullptr;  // rows sorted by level (DEVICE)
  int *loff = nullptr;  // level offsets into perm, levels + 1 (DEVICE)
  std::vector<Segment> seg;
};

}  // namespace

struct hflmp_sgs {
  int device = 0;
  int nrow = 0, nnz = 0, value_bytes = 0;
  Blocks blk{};
  const int *rp = nullptr, *ci = nullptr;
  const void *val = nullptr;
  Sweep fwd, bwd;
};

namespace {

void sweep_free(Sweep &s) {
  (void)hipFree(s.perm);
  (void)hipFree(s.loff);
  s.perm = s.loff = nullptr;
}

void sgs_free(hflmp_sgs *m) {
  if (!m) return;
  sweep_free(m->fwd);
  sweep_free(m->bwd);
  delete m;
}

// Counting sort of the rows by level, the launch segments, and the upload.
int build_sweep(hflmp_ctx *ctx, const std::vector<int> &level, int nlev, Sweep &s) {
  const int n = (int)level.size();
  std::vector<int> off((size_t)nlev + 1, 0), perm(n);
  for (int i = 0; i < n; ++i) ++off[level[i] + 1];
  for (int l = 0; l < nlev; ++l) off[l + 1] += off[l];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int i = 0; i < n; ++i) perm[fill[level[i]]++] = i;  // ascending rows within a level
  s.levels = nlev;
  s.seg.clear();
  for (int l = 0; l < nlev;) {
    if (off[l + 1] - off[l] > kSmallLevel) {
      s.seg.push_back({false, off[l], off[l + 1]});
      ++l;
      continue;
    }
    int e = l;
    while (e < nlev && off[e + 1] - off[e] <= kSmallLevel) ++e;
    s.seg.push_back({true, l, e});
    l = e;
  }
  if (hipMalloc(&s.perm, sizeof(int) * (size_t)(n > 0 ? n : 1)) != hipSuccess ||
      hipMalloc(&s.loff, sizeof(int) * ((size_t)nlev + 1)) != hipSuccess)
    return set_error(ctx, HFLMP_ENOMEM, "sgs_create: out of device memory");
  HFLMP_HIP_CHECK(ctx, hipMemcpy(s.perm, perm.data(), sizeof(int) * (size_t)n,
                                 hipMemcpyHostToDevice));
  HFLMP_HIP_CHECK(ctx, hipMemcpy(s.loff, off.data(), sizeof(int) * ((size_t)nlev + 1),
                                 hipMemcpyHostToDevice));
  return HFLMP_OK;
}

int sgs_create(hflmp_ctx *ctx, int nrow, int nnz, const int *rp, const int *ci,
               const void *val, int value_bytes, const int *h_rp, const int *h_ci,
               int num_blocks, hflmp_sgs **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, out != nullptr, "sgs_create: null output handle");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0 && num_blocks >= 1,
                "sgs_create: need nrow > 0, nnz > 0, num_blocks >= 1");
  HFLMP_REQUIRE(ctx, rp && ci && val, "sgs_create: null device array");
  HFLMP_REQUIRE(ctx, (h_rp == nullptr) == (h_ci == nullptr),
                "sgs_create: give both host pattern arrays or neither");
  std::vector<int> hrp, hci;
  if (!h_rp) {
    hrp.resize((size_t)nrow + 1);
    hci.resize((size_t)nnz);
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    HFLMP_HIP_CHECK(ctx, hipMemcpy(hrp.data(), rp, sizeof(int) * ((size_t)nrow + 1),
                                   hipMemcpyDeviceToHost));
    HFLMP_HIP_CHECK(ctx, hipMemcpy(hci.data(), ci, sizeof(int) * (size_t)nnz,
                                   hipMemcpyDeviceToHost));
    h_rp = hrp.data();
    h_ci = hci.data();
  }
  HFLMP_REQUIRE(ctx, h_rp[0] == 0 && h_rp[nrow] == nnz, "sgs_create: row_ptr does not span nnz");
  for (int i = 0; i < nrow; ++i) {
    HFLMP_REQUIRE(ctx, h_rp[i + 1] >= h_rp[i], "sgs_create: row_ptr decreases at row %d", i);
    for (int j = h_rp[i]; j < h_rp[i + 1]; ++j)
      HFLMP_REQUIRE(ctx, h_ci[j] >= 0 && h_ci[j] < nrow,
                    "sgs_create: column %d out of range in row %d", h_ci[j], i);
  }
  Blocks blk{nrow, num_blocks, nrow / num_blocks};
  // forward levels (rows ascending): entries up to the first diagonal one
  std::vector<int> lev(nrow);
  int nf = 0;
  for (int i = 0; i < nrow; ++i) {
    int lo, hi;
    blk.range(i, lo, hi);
    int l = 0;
    bool diag = false;
    for (int j = h_rp[i]; j < h_rp[i + 1]; ++j) {
      const int c = h_ci[j];
      if (c >= lo && c < hi) {
        if (c < i) l = std::max(l, lev[c] + 1);
        if (c == i) {
          diag = true;
          break;
        }
      }
    }
    if (!diag)
      return set_error(ctx, HFLMP_EFORMAT, "sgs_create: row %d has no diagonal entry", i);
    lev[i] = l;
    nf = std::max(nf, l + 1);
  }
  auto *m = new hflmp_sgs();
  m->device = ctx->device;
  m->nrow = nrow;
  m->nnz = nnz;
  m->value_bytes = value_bytes;
  m->blk = blk;
  m->rp = rp;
  m->ci = ci;
  m->val = val;
  int rc = build_sweep(ctx, lev, nf, m->fwd);
  if (rc) {
    sgs_free(m);
    return rc;
  }
  // backward levels (rows descending): every upper entry of the range
  int nbk = 0;
  for (int i = nrow - 1; i >= 0; --i) {
    int lo, hi;
    blk.range(i, lo, hi);
    int l = 0;
    for (int j = h_rp[i]; j < h_rp[i + 1]; ++j) {
      const int c = h_ci[j];
      if (c > i && c < hi) l = std::max(l, lev[c] + 1);
    }
    lev[i] = l;  // rows above i already hold their backward levels
    nbk = std::max(nbk, l + 1);
  }
  rc = build_sweep(ctx, lev, nbk, m->bwd);
  if (rc) {
    sgs_free(m);
    return rc;
  }
  *out = m;
  return HFLMP_OK;
}

template <typename T, bool FWD>
void launch_sweep(const Sweep &s, const SgsArgs<T> &a, hipStream_t st) {
  for (const Segment &g : s.seg) {
    if (g.small) {
      sgs_levels_kernel<T, FWD><<<1, kFinalBlock, 0, st>>>(a, s.perm, s.loff, g.a, g.b);
    } else {
      const unsigned grid = (unsigned)((g.b - g.a + kBlock - 1) / kBlock);
      sgs_level_kernel<T, FWD><<<grid, kBlock, 0, st>>>(a, s.perm, g.a, g.b);
    }
  }
}

template <typename T>
int sgs_apply(hflmp_ctx *ctx, const hflmp_sgs *m, const T *r, T *z) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m && r && z, "sgs_apply: null pointer");
  HFLMP_REQUIRE(ctx, m->value_bytes == (int)sizeof(T), "sgs_apply: value type mismatch");
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "sgs_apply: object on another device");
  SgsArgs<T> a{m->rp, m->ci, static_cast<const T *>(m->val), r, z, m->blk};
  launch_sweep<T, true>(m->fwd, a, ctx->stream);
  launch_sweep<T, false>(m->bwd, a, ctx->stream);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_sgs_create_f64(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr, const int *col,
                         const double *val, const int *h_row_ptr, const int *h_col,
                         int num_blocks, hflmp_sgs **out) {
  return sgs_create(ctx, nrow, nnz, row_ptr, col, val, 8, h_row_ptr, h_col, num_blocks, out);
}
int hflmp_sgs_create_f32(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr, const int *col,
                         const float *val, const int *h_row_ptr, const int *h_col,
                         int num_blocks, hflmp_sgs **out) {
  return sgs_create(ctx, nrow, nnz, row_ptr, col, val, 4, h_row_ptr, h_col, num_blocks, out);
}
int hflmp_sgs_apply_f64(hflmp_ctx *ctx, const hflmp_sgs *m, const double *r, double *z) {
  return sgs_apply<double>(ctx, m, r, z);
}
int hflmp_sgs_apply_f32(hflmp_ctx *ctx, const hflmp_sgs *m, const float *r, float *z) {
  return sgs_apply<float>(ctx, m, r, z);
}
int hflmp_sgs_get_info(const hflmp_sgs *m, hflmp_sgs_info *info) {
  if (!m || !info) return HFLMP_EINVAL;
  info->nrow = m->nrow;
  info->num_blocks = m->blk.nb;
  info->levels_forward = m->fwd.levels;
  info->levels_backward = m->bwd.levels;
  info->launches = (int)(m->fwd.seg.size() + m->bwd.seg.size());
  return HFLMP_OK;
}
void hflmp_sgs_destroy(hflmp_sgs *m) {
  if (!m) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != m->device) (void)hipSetDevice(m->device);
  const int dev = m->device;
  sgs_free(m);
  if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
}

}  // extern "C"
