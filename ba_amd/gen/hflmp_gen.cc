// Structured FE system generator (host, OpenMP).  Rows are produced
// independently: for row i the coupled nodes are enumerated in ascending
// index and each value is summed over the shared cells in ascending cell
// index, which is the value cell-by-cell assembly into a sorted CSR
// pattern produces.  Built with -ffp-contract=off so each term is a
// rounded product added to the running sum.
#include "hflmp_gen.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace {

int g_threads = 0;

// 1D stiffness/mass numerators of Q2 (K = Kp / 3h, M = h Mp / 30).
constexpr int kKp[3][3] = {{7, -8, 1}, {-8, 16, -8}, {1, -8, 7}};
constexpr int kMp[3][3] = {{4, 2, -1}, {2, 16, 2}, {-1, 2, 4}};

struct Element {
  int order, dim, q, nloc;
  double E[27 * 27];
};

// Exact element stiffness, local nodes lexicographic (x fastest).
void make_element(int order, int dim, int n, Element &el) {
  el.order = order;
  el.dim = dim;
  el.q = order + 1;
  el.nloc = dim == 2 ? el.q * el.q : el.q * el.q * el.q;
  const double h = 1.0 / (double)n;
  if (order == 1 && dim == 2) {
    // [[4,-1,-2,-1],...]/6 in tensor order: 0 differing coords 2/3, 1: -1/6, 2: -1/3
    const double c[3] = {2.0 / 3.0, -1.0 / 6.0, -1.0 / 3.0};
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) el.E[a * 4 + b] = c[__builtin_popcount(a ^ b)];
  } else if (order == 1 && dim == 3) {
    const double c[4] = {h * (1.0 / 3.0), h * 0.0, h * (-1.0 / 12.0), h * (-1.0 / 12.0)};
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) el.E[a * 8 + b] = c[__builtin_popcount(a ^ b)];
  } else {
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) {
        const int ax = a % 3, ay = a / 3, bx = b % 3, by = b / 3;
        el.E[a * 9 + b] =
            (double)(kKp[ax][bx] * kMp[ay][by] + kMp[ax][bx] * kKp[ay][by]) / 90.0;
      }
  }
}

inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// cells c in [0, n) with order*c <= g <= order*c + order
inline void cell_range(int64_t g, int order, int n, int64_t &lo, int64_t &hi) {
  lo = std::max<int64_t>(0, (g + order - 1) / order - 1);
  hi = std::min<int64_t>(n - 1, g / order);
}

inline bool is_boundary(const int64_t *g, int dim, int64_t p) {
  for (int d = 0; d < dim; ++d)
    if (g[d] == 0 || g[d] == p - 1) return true;
  return false;
}

void set_threads() {
  if (g_threads > 0) omp_set_num_threads(g_threads);
}

}  // namespace

extern "C" {

void hflmp_gen_set_threads(int nthreads) { g_threads = nthreads; }

int hflmp_gen_fe_sizes(int order, int dim, int n, int64_t *nrow, int64_t *nnz) {
  if (!((order == 1 && (dim == 2 || dim == 3)) || (order == 2 && dim == 2)) || n < 1)
    return -1;
  const int64_t p = (int64_t)order * n + 1;
  const int64_t r = order == 1 ? 3 * (int64_t)n + 1 : 8 * (int64_t)n + 1;
  int64_t N = 1, Z = 1;
  for (int d = 0; d < dim; ++d) {
    N *= p;
    Z *= r;
  }
  if (N > INT32_MAX || Z > INT32_MAX) return -1;
  *nrow = N;
  *nnz = Z;
  return 0;
}

int hflmp_gen_fe_csr(int order, int dim, int n, const double *kappa, int *row_ptr, int *col,
                     double *val) {
  int64_t N, Z;
  if (hflmp_gen_fe_sizes(order, dim, n, &N, &Z)) return -1;
  Element el;
  make_element(order, dim, n, el);
  const int64_t p = (int64_t)order * n + 1;
  const int q = el.q;
  // 1D coupling width per node: vertices of Q2 reach 2 nodes each side,
  // Q2 midpoints and Q1 nodes 1.
  auto reach = [&](int64_t g, int64_t &lo, int64_t &hi) {
    int64_t c0, c1;
    cell_range(g, order, n, c0, c1);
    lo = order * c0;
    hi = order * c1 + order;
  };
  // 1D row lengths -> row_ptr by prefix sums over the tensor product.
  auto len1 = [&](int64_t g) {
    int64_t lo, hi;
    reach(g, lo, hi);
    return hi - lo + 1;
  };
  set_threads();
  // row lengths
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t g[3] = {i % p, (i / p) % p, dim == 3 ? i / (p * p) : 0};
    int64_t L = 1;
    for (int d = 0; d < dim; ++d) L *= len1(g[d]);
    row_ptr[i + 1] = (int)L;
  }
  row_ptr[0] = 0;
  for (int64_t i = 0; i < N; ++i) row_ptr[i + 1] += row_ptr[i];
  if (row_ptr[N] != Z) return -2;

#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t g[3] = {i % p, (i / p) % p, dim == 3 ? i / (p * p) : 0};
    int64_t jlo[3] = {0, 0, 0}, jhi[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) reach(g[d], jlo[d], jhi[d]);
    const bool bnd = is_boundary(g, dim, p);
    int64_t pos = row_ptr[i];
    for (int64_t jz = jlo[2]; jz <= jhi[2]; ++jz)
      for (int64_t jy = jlo[1]; jy <= jhi[1]; ++jy)
        for (int64_t jx = jlo[0]; jx <= jhi[0]; ++jx, ++pos) {
          const int64_t j = jx + p * (jy + p * jz);
          col[pos] = (int)j;
          if (bnd) {
            val[pos] = (j == i) ? 1.0 : 0.0;
            continue;
          }
          const int64_t h[3] = {jx, jy, jz};
          // shared cells per dimension: both nodes inside [order*c, order*c+order]
          int64_t c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
          for (int d = 0; d < dim; ++d) {
            int64_t a0, a1, b0, b1;
            cell_range(g[d], order, n, a0, a1);
            cell_range(h[d], order, n, b0, b1);
            c0[d] = std::max(a0, b0);
            c1[d] = std::min(a1, b1);
          }
          double acc = 0.0;
          for (int64_t cz = c0[2]; cz <= c1[2]; ++cz)
            for (int64_t cy = c0[1]; cy <= c1[1]; ++cy)
              for (int64_t cx = c0[0]; cx <= c1[0]; ++cx) {
                const int64_t cell = cx + (int64_t)n * (cy + (int64_t)n * cz);
                const int64_t cc[3] = {cx, cy, cz};
                int la = 0, lb = 0, mul = 1;
                for (int d = 0; d < dim; ++d) {
                  la += (int)(g[d] - order * cc[d]) * mul;
                  lb += (int)(h[d] - order * cc[d]) * mul;
                  mul *= q;
                }
                const double k = kappa ? kappa[cell] : 1.0;
                acc += k * el.E[la * el.nloc + lb];
              }
          val[pos] = acc;
        }
  }
  return 0;
}

int hflmp_gen_fe_element(int order, int dim, int n, double *E) {
  int64_t N, Z;
  if (hflmp_gen_fe_sizes(order, dim, n, &N, &Z)) return -1;
  Element el;
  make_element(order, dim, n, el);
  for (int k = 0; k < el.nloc * el.nloc; ++k) E[k] = el.E[k];
  return el.nloc;
}

void hflmp_gen_kappa(int dim, int n, double sigma, uint64_t seed, double *kappa) {
  int64_t nc = 1;
  for (int d = 0; d < dim; ++d) nc *= n;
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    const double u1 =
        (double)((splitmix64_at(seed, 2u * (uint64_t)c) >> 11) + 1u) * 0x1.0p-53;
    const double u2 = (double)(splitmix64_at(seed, 2u * (uint64_t)c + 1u) >> 11) * 0x1.0p-53;
    const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    kappa[c] = std::exp(sigma * g);
  }
}

void hflmp_gen_rhs(int order, int dim, int n, uint64_t seed, double *b) {
  int64_t N, Z;
  if (hflmp_gen_fe_sizes(order, dim, n, &N, &Z)) return;
  const int64_t p = (int64_t)order * n + 1;
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int64_t g[3] = {i % p, (i / p) % p, dim == 3 ? i / (p * p) : 0};
    b[i] = is_boundary(g, dim, p)
               ? 0.0
               : 2.0 * ((double)(splitmix64_at(seed, (uint64_t)i) >> 11) * 0x1.0p-53) - 1.0;
  }
}

}  // extern "C"
