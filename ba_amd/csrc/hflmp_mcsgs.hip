// Multicoloured symmetric Gauss-Seidel on the device
// (lPreconditioner_MultiColoring_SymmetricGaussSeidel,
//  linear_algebra/lmp/lpreconditioner_mc.cc:147-723,
//  lpreconditioner_mc_sgs.cc:44-90, lpreconditioner_mc_gs.cc:83-102).
//
// With the rows grouped by colour (same-coloured rows are not coupled), the
// reference permutes the operator and splits it into per-colour diagonal,
// lower (L) and upper (R) blocks scaled by omega, then applies
//   forward, colours ascending:  z_c = -(r_c - sum_{c'<c} w L z_c') / -d
//   diagonal:                    z   = z * (-d / (w (2 - w)))
//   backward, colours descending: z_c = (z_c + sum_{c'>c} w R z_c') * (-1/d)
// Here the matrix is not permuted: create() builds, per colour, the list of
// its rows and, for every such row, the entries of the original CSR row that
// couple to lower colours and to higher colours (two split CSR arrays with
// positions into val, so value updates need no rebuild).  One kernel per
// colour and sweep processes that colour's rows, one thread per row, summing
// the row's entries in ascending column order.
#include <vector>

#include "hflmp_internal.hpp"

struct hflmp_mcsgs {
  int nrow = 0, ncolors = 0;
  double omega = 1.0;
  std::vector<int> color_off;  // ncolors + 1 offsets into rows
  int *rows = nullptr;         // rows grouped by colour
  int *lo_ptr = nullptr, *lo_pos = nullptr;  // lower-colour entries per listed row
  int *up_ptr = nullptr, *up_pos = nullptr;  // higher-colour entries
  int *diag_pos = nullptr;                   // position of a_ii per listed row
  const int *col = nullptr;
  const double *val = nullptr;
};

namespace {

// forward colour c: z_i = (-r_i + sum_lower w a_ij z_j) * (-1/a_ii)
__global__ void __launch_bounds__(kBlock)
    mcsgs_forward(int m, const int *__restrict__ rows, const int *__restrict__ lp,
                  const int *__restrict__ lpos, const int *__restrict__ dpos,
                  const int *__restrict__ col, const double *__restrict__ val, double omega,
                  const double *__restrict__ r, double *__restrict__ z) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m) return;
  const int i = rows[k];
  double acc = -r[i];
  for (int q = lp[k]; q < lp[k + 1]; ++q) {
    const int p = lpos[q];
    acc += (omega * val[p]) * z[col[p]];
  }
  z[i] = acc * (-(1.0 / val[dpos[k]]));
}

// diagonal step over all rows: z_i *= -a_ii / (w (2 - w))
__global__ void __launch_bounds__(kBlock)
    mcsgs_diag(int n, const int *__restrict__ rows, const int *__restrict__ dpos,
               const double *__restrict__ val, double dscale, double *__restrict__ z) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const int i = rows[k];
  z[i] *= (-val[dpos[k]]) * dscale;
}

// backward colour c: z_i = (z_i + sum_upper w a_ij z_j) * (-1/a_ii)
__global__ void __launch_bounds__(kBlock)
    mcsgs_backward(int m, const int *__restrict__ rows, const int *__restrict__ up,
                   const int *__restrict__ upos, const int *__restrict__ dpos,
                   const int *__restrict__ col, const double *__restrict__ val, double omega,
                   double *__restrict__ z) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m) return;
  const int i = rows[k];
  double acc = z[i];
  for (int q = up[k]; q < up[k + 1]; ++q) {
    const int p = upos[q];
    acc += (omega * val[p]) * z[col[p]];
  }
  z[i] = acc * (-(1.0 / val[dpos[k]]));
}

void free_mcsgs(hflmp_mcsgs *m) {
  if (!m) return;
  for (int *p : {m->rows, m->lo_ptr, m->lo_pos, m->up_ptr, m->up_pos, m->diag_pos})
    if (p) (void)hipFree(p);
  delete m;
}

template <typename V>
int upload(hflmp_ctx *ctx, const std::vector<V> &h, V **d) {
  *d = nullptr;
  if (h.empty()) return HFLMP_OK;
  HFLMP_HIP_CHECK(ctx, hipMalloc(reinterpret_cast<void **>(d), h.size() * sizeof(V)));
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(*d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice,
                                      ctx->stream));
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_mcsgs_create_f64(hflmp_ctx *ctx, int nrow, int nnz, const int *row_ptr, const int *col,
                           const double *val, int ncolors, const int *h_color_sizes,
                           const int *h_perm, double omega, hflmp_mcsgs **out) {
  if (!ctx || !out) return HFLMP_EINVAL;
  *out = nullptr;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz > 0 && ncolors > 0, "mcsgs: empty system or colouring");
  HFLMP_REQUIRE(ctx, row_ptr && col && val && h_color_sizes && h_perm, "mcsgs: null pointer");
  HFLMP_REQUIRE(ctx, omega > 0.0 && omega < 2.0, "mcsgs: omega must lie in (0, 2)");
  // host copy of the structure
  std::vector<int> rp(nrow + 1), ci(nnz);
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(rp.data(), row_ptr, sizeof(int) * (nrow + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(ci.data(), col, sizeof(int) * nnz, hipMemcpyDeviceToHost,
                                      ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  HFLMP_REQUIRE(ctx, rp[0] == 0 && rp[nrow] == nnz, "mcsgs: row_ptr does not span nnz");
  std::vector<int> color(nrow, -1);
  std::vector<int> off(ncolors + 1, 0);
  for (int c = 0; c < ncolors; ++c) {
    HFLMP_REQUIRE(ctx, h_color_sizes[c] >= 0, "mcsgs: negative colour size");
    off[c + 1] = off[c] + h_color_sizes[c];
  }
  HFLMP_REQUIRE(ctx, off[ncolors] == nrow, "mcsgs: colour sizes do not sum to nrow");
  for (int c = 0; c < ncolors; ++c)
    for (int k = off[c]; k < off[c + 1]; ++k) {
      const int i = h_perm[k];
      HFLMP_REQUIRE(ctx, i >= 0 && i < nrow && color[i] < 0, "mcsgs: perm is not a permutation");
      color[i] = c;
    }
  std::vector<int> lp(nrow + 1, 0), up(nrow + 1, 0), dpos(nrow, -1), lpos, upos;
  lpos.reserve(nnz / 2 + 1);
  upos.reserve(nnz / 2 + 1);
  for (int k = 0; k < nrow; ++k) {
    const int i = h_perm[k];
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
      const int j = ci[p];
      HFLMP_REQUIRE(ctx, j >= 0 && j < nrow, "mcsgs: column %d out of range", j);
      if (j == i) {
        dpos[k] = p;
      } else if (color[j] == color[i]) {
        return set_error(ctx, HFLMP_EINVAL, "mcsgs: rows %d and %d share a colour but couple", i, j);
      } else if (color[j] < color[i]) {
        lpos.push_back(p);
      } else {
        upos.push_back(p);
      }
    }
    HFLMP_REQUIRE(ctx, dpos[k] >= 0, "mcsgs: row %d has no diagonal entry", i);
    lp[k + 1] = (int)lpos.size();
    up[k + 1] = (int)upos.size();
  }
  hflmp_mcsgs *m = new hflmp_mcsgs();
  m->nrow = nrow;
  m->ncolors = ncolors;
  m->omega = omega;
  m->color_off = off;
  m->col = col;
  m->val = val;
  std::vector<int> rows(h_perm, h_perm + nrow);
  int rc = upload(ctx, rows, &m->rows);
  if (!rc) rc = upload(ctx, lp, &m->lo_ptr);
  if (!rc) rc = upload(ctx, lpos, &m->lo_pos);
  if (!rc) rc = upload(ctx, up, &m->up_ptr);
  if (!rc) rc = upload(ctx, upos, &m->up_pos);
  if (!rc) rc = upload(ctx, dpos, &m->diag_pos);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = HFLMP_EHIP;
  if (rc) {
    free_mcsgs(m);
    return rc;
  }
  *out = m;
  return HFLMP_OK;
}

int hflmp_mcsgs_apply_f64(hflmp_ctx *ctx, const hflmp_mcsgs *m, const double *r, double *z) {
  if (!ctx || !m) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, r && z && r != z, "mcsgs_apply: null or aliased vectors");
  const double dscale = 1.0 / (m->omega * (2.0 - m->omega));
  for (int c = 0; c < m->ncolors; ++c) {
    const int a = m->color_off[c], cnt = m->color_off[c + 1] - a;
    if (!cnt) continue;
    mcsgs_forward<<<grid_for(cnt, 1), kBlock, 0, ctx->stream>>>(
        cnt, m->rows + a, m->lo_ptr + a, m->lo_pos, m->diag_pos + a, m->col, m->val, m->omega, r,
        z);
  }
  mcsgs_diag<<<grid_for(m->nrow, 1), kBlock, 0, ctx->stream>>>(m->nrow, m->rows, m->diag_pos,
                                                               m->val, dscale, z);
  for (int c = m->ncolors - 1; c >= 0; --c) {
    const int a = m->color_off[c], cnt = m->color_off[c + 1] - a;
    if (!cnt) continue;
    mcsgs_backward<<<grid_for(cnt, 1), kBlock, 0, ctx->stream>>>(
        cnt, m->rows + a, m->up_ptr + a, m->up_pos, m->diag_pos + a, m->col, m->val, m->omega,
        z);
  }
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

int hflmp_mcsgs_destroy(hflmp_ctx *ctx, hflmp_mcsgs *m) {
  if (!ctx) return HFLMP_EINVAL;
  if (m) {
    (void)hipStreamSynchronize(ctx->stream);
    free_mcsgs(m);
  }
  return HFLMP_OK;
}

}  // extern "C"
