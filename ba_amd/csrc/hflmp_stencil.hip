// Matrix-free stencil operator: build, products, inverse diagonal.
// Replaces StencilOperator::Init / VectorMult / VectorMultAdd
// (linear_algebra/stencil_operator.h:60-100,131-330) with the cell loop
// turned into one thread per node (hflmp_stencil.hpp).
#include "hflmp_stencil.hpp"

#include <cstring>

namespace {

enum StencilMode { kStAssign = 0, kStScaleAdd = 1 };

// MODE kStAssign: y = A x (VectorMult).  kStScaleAdd: y = beta y + alpha A x
// (VectorMultAdd: Scale(beta), then the cell contributions).
template <typename Op, int MODE>
__global__ void __launch_bounds__(kBlock)
    stencil_kernel(Op op, typename Op::value_type alpha,
                   const typename Op::value_type *__restrict__ x,
                   typename Op::value_type beta, typename Op::value_type *__restrict__ y) {
  using T = typename Op::value_type;
  __shared__ T xs[Op::kLds];
  const long long unit = xcd_tile(blockIdx.x, gridDim.x);
  T s = 0;
  if (MODE == kStScaleAdd) {
    const long long r = op.thread_row(unit);
    // + 0 turns a -0 into +0: the running sum never holds -0 (the
    // reference's sum differs from it only in the sign of that zero,
    // which the first cell's term overwrites either way)
    if (r < op.nrow) s = y[r] * beta + T(0);
  }
  TileRow tr;
  s = op.template row_sum_scaled<true>(unit, x, s, MODE == kStScaleAdd ? alpha : T(1), tr, xs);
  if (tr.has_row) __builtin_nontemporal_store(s, y + tr.row);
}

// d_i = sum over the node's cells (ascending) of f_c E_m[l][l], the value
// cell-by-cell assembly puts on the diagonal; 1 on fixed nodes.  invd =
// 1/d formed in double; a zero diagonal gives 0 and is counted.
template <typename Op>
__global__ void __launch_bounds__(kBlock)
    stencil_invdiag_kernel(Op op, typename Op::value_type *__restrict__ invd,
                           unsigned int *bad) {
#pragma clang fp contract(off)
  using T = typename Op::value_type;
  const long long r = op.thread_row(xcd_tile(blockIdx.x, gridDim.x));
  if (r >= op.nrow) return;
  const int i = (int)r;
  constexpr int ORDER = Op::Q - 1;
  const int gx = i % op.px;
  const int t = i / op.px;
  const int gy = t % op.py;
  const int gz = Op::KZ == 2 ? t / op.py : 0;
  T d = 0;
  for (int kz = 0; kz < Op::KZ; ++kz)
    for (int ky = 0; ky < 2; ++ky)
      for (int kx = 0; kx < 2; ++kx) {
        const int cx = gx / ORDER - 1 + kx, cy = gy / ORDER - 1 + ky,
                  cz = Op::KZ == 2 ? gz / ORDER - 1 + kz : 0;
        const int ax = gx - ORDER * cx, ay = gy - ORDER * cy,
                  az = Op::KZ == 2 ? gz - ORDER * cz : 0;
        if (cx < 0 || cx >= op.ncx || cy < 0 || cy >= op.ncy || cz < 0 || cz >= op.ncz ||
            ax > ORDER || ay > ORDER || az > ORDER)
          continue;
        const long long c = cx + (long long)op.ncx * (cy + (long long)op.ncy * cz);
        const T *Em = op.lm ? op.E + (size_t)op.lm[c] * (Op::NLOC * Op::NLOC) : op.E;
        const int l = ax + Op::Q * (ay + Op::Q * az);
        d += op.f[c] * Em[l * Op::NLOC + l];
      }
  if ((op.fixed[i >> 5] >> (i & 31)) & 1u) d = T(1);
  if (d == T(0)) {
    invd[i] = T(0);
    atomicAdd(bad, 1u);
  } else {
    invd[i] = static_cast<T>(1.0 / static_cast<double>(d));
  }
}

// fixed dof list -> bit per node; out-of-range entries are counted
__global__ void __launch_bounds__(kBlock)
    stencil_fixed_kernel(long long nnode, int nfixed, const int *__restrict__ fixed,
                         unsigned int *__restrict__ bits, unsigned int *bad) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nfixed; k += gridDim.x * kBlock) {
    const int d = fixed[k];
    if (d < 0 || d >= nnode) {
      atomicAdd(bad, 1u);
      continue;
    }
    atomicOr(bits + (d >> 5), 1u << (d & 31));
  }
}

// counts cell matrix indices >= nmat
__global__ void __launch_bounds__(kBlock)
    stencil_lm_check_kernel(long long ncell, int nmat, const unsigned char *__restrict__ lm,
                            unsigned int *bad) {
  for (long long c = blockIdx.x * (long long)kBlock + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * kBlock)
    if (lm[c] >= nmat) atomicAdd(bad, 1u);
}

void stencil_free(hflmp_stencil *m) {
  if (!m) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != m->device) (void)hipSetDevice(m->device);
  (void)hipFree(m->E);
  (void)hipFree(m->factor);
  (void)hipFree(m->cell_lm);
  (void)hipFree(m->fixed_bits);
  if (cur >= 0 && cur != m->device) (void)hipSetDevice(cur);
  delete m;
}

unsigned grid_cap(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

template <typename T>
int stencil_create(hflmp_ctx *ctx, const hflmp_stencil_desc *desc, const T *lm_host,
                   const unsigned char *cell_lm, const T *cell_factor, int nfixed,
                   const int *fixed, hflmp_stencil **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, desc && out && lm_host && cell_factor, "stencil_create: null pointer");
  *out = nullptr;
  const int order = desc->order, dim = desc->dim;
  HFLMP_REQUIRE(ctx, (order == 1 && (dim == 2 || dim == 3)) || (order == 2 && dim == 2),
                "stencil_create: elements Q1 (2D, 3D) and Q2 (2D) only, got Q%d in %dD", order,
                dim);
  HFLMP_REQUIRE(ctx, desc->nmat >= 1 && desc->nmat <= 256 && (desc->nmat == 1 || cell_lm),
                "stencil_create: need 1 <= nmat <= 256 and cell_lm when nmat > 1");
  HFLMP_REQUIRE(ctx, nfixed >= 0 && (nfixed == 0 || fixed), "stencil_create: bad fixed list");
  long long nnode = 1, ncell = 1;
  int nc[3] = {1, 1, 1}, np[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) {
    HFLMP_REQUIRE(ctx, desc->ncell[d] >= 1, "stencil_create: ncell[%d] < 1", d);
    nc[d] = desc->ncell[d];
    np[d] = order * nc[d] + 1;
    nnode *= np[d];
    ncell *= nc[d];
  }
  // byte offsets of the vector and cell arrays stay below 2^31
  HFLMP_REQUIRE(ctx, nnode * (long long)sizeof(T) < 0x80000000LL &&
                         ncell * (long long)sizeof(T) < 0x80000000LL,
                "stencil_create: %lld nodes / %lld cells exceed the 2 GiB vector limit", nnode,
                ncell);
  const int q = order + 1;
  const int nloc = dim == 2 ? q * q : q * q * q;
  auto *m = new hflmp_stencil();
  m->device = ctx->device;
  m->order = order;
  m->dim = dim;
  m->nloc = nloc;
  m->nmat = desc->nmat;
  for (int d = 0; d < 3; ++d) {
    m->nc[d] = nc[d];
    m->np[d] = np[d];
  }
  m->nnode = nnode;
  m->ncell = ncell;
  m->nfixed = nfixed;
  m->value_bytes = sizeof(T);
  auto fail = [&](int code, const char *msg) {
    stencil_free(m);
    return set_error(ctx, code, "stencil_create: %s", msg);
  };
  const size_t ebytes = sizeof(T) * (size_t)m->nmat * nloc * nloc;
  const size_t words = (size_t)((nnode + 31) / 32);
  if (hipMalloc(&m->E, ebytes) != hipSuccess ||
      hipMalloc(&m->factor, sizeof(T) * (size_t)ncell) != hipSuccess ||
      hipMalloc(&m->fixed_bits, sizeof(unsigned int) * words) != hipSuccess ||
      (m->nmat > 1 && hipMalloc(&m->cell_lm, (size_t)ncell) != hipSuccess))
    return fail(HFLMP_ENOMEM, "out of device memory");
  hipStream_t s = ctx->stream;
  int rc;
  unsigned int *bad = ctx->red_counter + 32;
  if ((rc = hflmp_fill_zero(ctx, bad, 2 * sizeof(unsigned int))) != HFLMP_OK ||
      (rc = hflmp_fill_zero(ctx, m->fixed_bits, sizeof(unsigned int) * words)) != HFLMP_OK)
    return fail(rc, "fill");
  if (hipMemcpyAsync(m->E, lm_host, ebytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(m->factor, cell_factor, sizeof(T) * (size_t)ncell, hipMemcpyDeviceToDevice,
                     s) != hipSuccess ||
      (m->nmat > 1 && hipMemcpyAsync(m->cell_lm, cell_lm, (size_t)ncell,
                                     hipMemcpyDeviceToDevice, s) != hipSuccess))
    return fail(HFLMP_EHIP, "copy failed");
  if (nfixed > 0)
    stencil_fixed_kernel<<<grid_cap(nfixed), kBlock, 0, s>>>(nnode, nfixed, fixed, m->fixed_bits,
                                                             bad);
  if (m->nmat > 1)
    stencil_lm_check_kernel<<<grid_cap(ncell), kBlock, 0, s>>>(ncell, m->nmat, m->cell_lm,
                                                               bad + 1);
  if (hipGetLastError() != hipSuccess) return fail(HFLMP_EHIP, "launch failed");
  unsigned int hb[2];
  if (hipMemcpyAsync(ctx->host_pinned, bad, sizeof(hb), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(HFLMP_EHIP, "HIP error in build");
  std::memcpy(hb, ctx->host_pinned, sizeof(hb));
  if (hb[0]) return fail(HFLMP_EINVAL, "fixed dof index out of range");
  if (hb[1]) return fail(HFLMP_EINVAL, "cell matrix index >= nmat");
  *out = m;
  return HFLMP_OK;
}

template <typename T>
int stencil_check(hflmp_ctx *ctx, const hflmp_stencil *m, const char *what) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, m, "%s: null operator", what);
  HFLMP_REQUIRE(ctx, m->value_bytes == (int)sizeof(T), "%s: value type mismatch", what);
  HFLMP_REQUIRE(ctx, m->device == ctx->device, "%s: operator on another device", what);
  return HFLMP_OK;
}

template <typename T, int MODE>
int stencil_apply(hflmp_ctx *ctx, const hflmp_stencil *m, T alpha, const T *x, T beta, T *y) {
  int rc = stencil_check<T>(ctx, m, "stencil_apply");
  if (rc) return rc;
  HFLMP_REQUIRE(ctx, x && y, "stencil_apply: null vector");
  HFLMP_REQUIRE(ctx, (const void *)x != (const void *)y, "stencil_apply: x and y must differ");
  hipStream_t s = ctx->stream;
  stencil_dispatch<T>(m, [&](auto op) {
    stencil_kernel<decltype(op), MODE><<<op_grid(op), kBlock, 0, s>>>(op, alpha, x, beta, y);
  });
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

template <typename T>
int stencil_set_factors(hflmp_ctx *ctx, hflmp_stencil *m, const T *f) {
  int rc = stencil_check<T>(ctx, m, "stencil_set_factors");
  if (rc) return rc;
  HFLMP_REQUIRE(ctx, f, "stencil_set_factors: null pointer");
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(m->factor, f, sizeof(T) * (size_t)m->ncell,
                                      hipMemcpyDeviceToDevice, ctx->stream));
  return HFLMP_OK;
}

template <typename T>
int stencil_invdiag(hflmp_ctx *ctx, const hflmp_stencil *m, T *invd, int *n_bad) {
  int rc = stencil_check<T>(ctx, m, "stencil_extract_invdiag");
  if (rc) return rc;
  HFLMP_REQUIRE(ctx, invd, "stencil_extract_invdiag: null pointer");
  unsigned int *bad = ctx->red_counter + 32;
  if ((rc = hflmp_fill_zero(ctx, bad, sizeof(unsigned int))) != HFLMP_OK) return rc;
  hipStream_t s = ctx->stream;
  stencil_dispatch<T>(m, [&](auto op) {
    stencil_invdiag_kernel<decltype(op)><<<op_grid(op), kBlock, 0, s>>>(op, invd, bad);
  });
  HFLMP_LAUNCH_CHECK(ctx);
  unsigned int h = 0;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(ctx->host_pinned, bad, sizeof(unsigned int),
                                      hipMemcpyDeviceToHost, s));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  std::memcpy(&h, ctx->host_pinned, sizeof(h));
  if (n_bad) *n_bad = (int)h;
  if (h)
    return set_error(ctx, HFLMP_ENUMERIC, "stencil_extract_invdiag: %u zero diagonals", h);
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_stencil_create_f64(hflmp_ctx *ctx, const hflmp_stencil_desc *desc,
                             const double *local_matrices, const unsigned char *cell_lm,
                             const double *cell_factor, int nfixed, const int *fixed,
                             hflmp_stencil **out) {
  return stencil_create<double>(ctx, desc, local_matrices, cell_lm, cell_factor, nfixed, fixed,
                                out);
}
int hflmp_stencil_create_f32(hflmp_ctx *ctx, const hflmp_stencil_desc *desc,
                             const float *local_matrices, const unsigned char *cell_lm,
                             const float *cell_factor, int nfixed, const int *fixed,
                             hflmp_stencil **out) {
  return stencil_create<float>(ctx, desc, local_matrices, cell_lm, cell_factor, nfixed, fixed,
                               out);
}
int hflmp_stencil_set_factors_f64(hflmp_ctx *ctx, hflmp_stencil *m, const double *f) {
  return stencil_set_factors<double>(ctx, m, f);
}
int hflmp_stencil_set_factors_f32(hflmp_ctx *ctx, hflmp_stencil *m, const float *f) {
  return stencil_set_factors<float>(ctx, m, f);
}
int hflmp_stencil_get_info(const hflmp_stencil *m, hflmp_stencil_info *info) {
  if (!m || !info) return HFLMP_EINVAL;
  info->order = m->order;
  info->dim = m->dim;
  info->nloc = m->nloc;
  info->nmat = m->nmat;
  for (int d = 0; d < 3; ++d) info->ncell[d] = m->nc[d];
  info->nnode = m->nnode;
  info->ncell_total = m->ncell;
  info->nfixed = m->nfixed;
  info->value_bytes = m->value_bytes;
  info->device_bytes = (size_t)m->value_bytes * ((size_t)m->nmat * m->nloc * m->nloc +
                                                 (size_t)m->ncell) +
                       sizeof(unsigned int) * (size_t)((m->nnode + 31) / 32) +
                       (m->nmat > 1 ? (size_t)m->ncell : 0);
  return HFLMP_OK;
}
void hflmp_stencil_destroy(hflmp_stencil *m) { stencil_free(m); }
int hflmp_stencil_apply_f64(hflmp_ctx *ctx, const hflmp_stencil *m, const double *x, double *y) {
  return stencil_apply<double, kStAssign>(ctx, m, 1.0, x, 0.0, y);
}
int hflmp_stencil_apply_f32(hflmp_ctx *ctx, const hflmp_stencil *m, const float *x, float *y) {
  return stencil_apply<float, kStAssign>(ctx, m, 1.0f, x, 0.0f, y);
}
int hflmp_stencil_apply_add_f64(hflmp_ctx *ctx, const hflmp_stencil *m, double alpha,
                                const double *x, double beta, double *y) {
  return stencil_apply<double, kStScaleAdd>(ctx, m, alpha, x, beta, y);
}
int hflmp_stencil_apply_add_f32(hflmp_ctx *ctx, const hflmp_stencil *m, float alpha,
                                const float *x, float beta, float *y) {
  return stencil_apply<float, kStScaleAdd>(ctx, m, alpha, x, beta, y);
}
int hflmp_stencil_extract_invdiag_f64(hflmp_ctx *ctx, const hflmp_stencil *m, double *invd,
                                      int *n_bad) {
  return stencil_invdiag<double>(ctx, m, invd, n_bad);
}
int hflmp_stencil_extract_invdiag_f32(hflmp_ctx *ctx, const hflmp_stencil *m, float *invd,
                                      int *n_bad) {
  return stencil_invdiag<float>(ctx, m, invd, n_bad);
}

}  // extern "C"
