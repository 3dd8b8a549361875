/*
 * oracle_core.c -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
 * Random streams, the structured FE system generator and CSR utilities.
 * See hflmp_oracle.h for the reference file each routine follows.
 */
#include "hflmp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* SplitMix64: the i-th output of the generator whose state starts at seed
 * (Steele, Lea, Flood 2014).  Output i = mix(seed + (i+1)*gamma).          */
uint64_t or_splitmix64_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* uniform on [-1, 1): 53 high bits -> [0,1) -> 2u - 1 */
double or_uniform_pm1(uint64_t seed, uint64_t i) {
  const double u = (double)(or_splitmix64_at(seed, i) >> 11) * 0x1.0p-53;
  return 2.0 * u - 1.0;
}

/* standard normal for cell c by Box-Muller from outputs 2c and 2c+1 */
double or_gauss(uint64_t seed, uint64_t c) {
  const double u1 =
      (double)((or_splitmix64_at(seed, 2u * c) >> 11) + 1u) * 0x1.0p-53;
  const double u2 = (double)(or_splitmix64_at(seed, 2u * c + 1u) >> 11) *
                    0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ---------------------------------------------------------------------- */
/* Element matrices.  Local nodes are numbered lexicographically (x fastest).
 *  Q1 2D: same node 2/3, nodes differing in one coordinate -1/6, in two -1/3
 *  Q1 3D: h * {1/3, 0, -1/12, -1/12} for 0,1,2,3 differing coordinates
 *  Q2 2D: (Kp[ax][bx] Mp[ay][by] + Mp[ax][bx] Kp[ay][by]) / 90 with the 1D
 *         Q2 matrices K = Kp/(3h), M = h Mp/30.                             */
static const int KP[3][3] = {{7, -8, 1}, {-8, 16, -8}, {1, -8, 7}};
static const int MP[3][3] = {{4, 2, -1}, {2, 16, 2}, {-1, 2, 4}};

static void element_matrix(int order, int dim, int n, double *E) {
  const double h = 1.0 / (double)n;
  if (order == 1 && dim == 2) {
    const double c[3] = {2.0 / 3.0, -1.0 / 6.0, -1.0 / 3.0};
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        const int d = ((a ^ b) & 1) + (((a ^ b) >> 1) & 1);
        E[a * 4 + b] = c[d];
      }
  } else if (order == 1 && dim == 3) {
    const double c[4] = {h * (1.0 / 3.0), h * 0.0, h * (-1.0 / 12.0),
                         h * (-1.0 / 12.0)};
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        const int x = a ^ b;
        const int d = (x & 1) + ((x >> 1) & 1) + ((x >> 2) & 1);
        E[a * 8 + b] = c[d];
      }
  } else { /* Q2 2D */
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) {
        const int ax = a % 3, ay = a / 3, bx = b % 3, by = b / 3;
        const int num = KP[ax][bx] * MP[ay][by] + MP[ax][bx] * KP[ay][by];
        E[a * 9 + b] = (double)num / 90.0;
      }
  }
}

int or_fe_sizes(int order, int dim, int n, int64_t *nrow, int64_t *nnz) {
  if (!((order == 1 && (dim == 2 || dim == 3)) || (order == 2 && dim == 2)) ||
      n < 1)
    return -1;
  const int64_t p = (int64_t)order * n + 1;        /* nodes per side */
  const int64_t r = (order == 1) ? 3 * (int64_t)n + 1 : 8 * (int64_t)n + 1;
  int64_t N = 1, Z = 1;
  for (int d = 0; d < dim; ++d) {
    N *= p;
    Z *= r;
  }
  *nrow = N;
  *nnz = Z;
  return 0;
}

/* global node of local node l in cell (cx,cy,cz) */
static inline int64_t cell_node(int order, int dim, int64_t p, int64_t cx,
                                int64_t cy, int64_t cz, int l) {
  const int q = order + 1;
  const int lx = l % q, ly = (l / q) % q, lz = (dim == 3) ? l / (q * q) : 0;
  const int64_t gx = order * cx + lx, gy = order * cy + ly,
                gz = order * cz + lz;
  return gx + p * (gy + p * gz);
}

static int cmp_int(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* Pattern = the set of (i,j) with i,j in a common cell, each row sorted
 * ascending (compute_sparsity_structure + coo_2_csr in the reference:
 * assembly/assembly_utils.cc:144, lmatrix_formats.h:320-456).  Values are
 * accumulated cell by cell in lexicographic cell order (add_values,
 * lmatrix_csr_cpu.cc:1102-1153), then the Dirichlet rows are replaced by
 * unit rows with ZeroRows (lmatrix_csr_cpu.cc:908-927). */
int or_fe_assemble(int order, int dim, int n, const double *kappa,
                   int *rowptr, int *col, double *val) {
  int64_t N, Z;
  if (or_fe_sizes(order, dim, n, &N, &Z)) return -1;
  const int64_t p = (int64_t)order * n + 1;
  const int q = order + 1;
  const int nloc = (dim == 2) ? q * q : q * q * q;
  double E[81 * 9];
  element_matrix(order, dim, n, E);

  /* 1. pattern, row by row */
  int64_t pos = 0;
  int buf[256];
  for (int64_t i = 0; i < N; ++i) {
    const int64_t ix = i % p, iy = (i / p) % p, iz = (dim == 3) ? i / (p * p) : 0;
    int cnt = 0;
    /* cells containing node i: cell index c with order*c <= g <= order*c+order */
    int64_t c0[3], c1[3];
    const int64_t g[3] = {ix, iy, iz};
    for (int d = 0; d < 3; ++d) {
      if (d >= dim) {
        c0[d] = c1[d] = 0;
        continue;
      }
      /* order*c <= g <= order*c + order  <=>  ceil(g/order)-1 <= c <= g/order */
      int64_t lo = (g[d] + order - 1) / order - 1;
      int64_t hi = g[d] / order;
      if (hi > n - 1) hi = n - 1;
      if (lo < 0) lo = 0;
      c0[d] = lo;
      c1[d] = hi;
    }
    for (int64_t cz = c0[2]; cz <= c1[2]; ++cz)
      for (int64_t cy = c0[1]; cy <= c1[1]; ++cy)
        for (int64_t cx = c0[0]; cx <= c1[0]; ++cx)
          for (int l = 0; l < nloc; ++l)
            buf[cnt++] = (int)cell_node(order, dim, p, cx, cy, cz, l);
    qsort(buf, (size_t)cnt, sizeof(int), cmp_int);
    rowptr[i] = (int)pos;
    for (int k = 0; k < cnt; ++k)
      if (k == 0 || buf[k] != buf[k - 1]) col[pos++] = buf[k];
  }
  rowptr[N] = (int)pos;
  if (pos != Z) return -2;
  memset(val, 0, sizeof(double) * (size_t)Z);

  /* 2. values, cell by cell in lexicographic order */
  const int64_t ncz = (dim == 3) ? n : 1;
  int64_t gl[27];
  for (int64_t cz = 0; cz < ncz; ++cz)
    for (int64_t cy = 0; cy < n; ++cy)
      for (int64_t cx = 0; cx < n; ++cx) {
        const int64_t cell = cx + (int64_t)n * (cy + (int64_t)n * cz);
        const double k = kappa ? kappa[cell] : 1.0;
        for (int l = 0; l < nloc; ++l)
          gl[l] = cell_node(order, dim, p, cx, cy, cz, l);
        for (int a = 0; a < nloc; ++a) {
          const int64_t r = gl[a];
          for (int b = 0; b < nloc; ++b) {
            const int c = (int)gl[b];
            /* binary search in the sorted row */
            int lo = rowptr[r], hi = rowptr[r + 1] - 1;
            while (lo < hi) {
              const int mid = (lo + hi) / 2;
              if (col[mid] < c) lo = mid + 1;
              else hi = mid;
            }
            val[lo] += k * E[a * nloc + b];
          }
        }
      }

  /* 3. Dirichlet rows (all boundary nodes) */
  for (int64_t i = 0; i < N; ++i) {
    const int64_t ix = i % p, iy = (i / p) % p, iz = (dim == 3) ? i / (p * p) : 0;
    int bnd = ix == 0 || ix == p - 1 || iy == 0 || iy == p - 1;
    if (dim == 3) bnd = bnd || iz == 0 || iz == p - 1;
    if (bnd) {
      const int ri = (int)i;
      or_csr_zero_rows_d(&ri, 1, rowptr, col, val, 1.0);
    }
  }
  return 0;
}

void or_fe_element(int order, int dim, int n, double *E) {
  element_matrix(order, dim, n, E);
}

void or_fe_boundary_mask(int order, int dim, int n, uint8_t *mask) {
  int64_t N, Z;
  or_fe_sizes(order, dim, n, &N, &Z);
  const int64_t p = (int64_t)order * n + 1;
  for (int64_t i = 0; i < N; ++i) {
    const int64_t ix = i % p, iy = (i / p) % p, iz = (dim == 3) ? i / (p * p) : 0;
    int bnd = ix == 0 || ix == p - 1 || iy == 0 || iy == p - 1;
    if (dim == 3) bnd = bnd || iz == 0 || iz == p - 1;
    mask[i] = (uint8_t)bnd;
  }
}

void or_fe_kappa(int dim, int n, double sigma, uint64_t seed, double *kappa) {
  int64_t nc = 1;
  for (int d = 0; d < dim; ++d) nc *= n;
  for (int64_t c = 0; c < nc; ++c) kappa[c] = exp(sigma * or_gauss(seed, (uint64_t)c));
}

void or_fe_rhs(int order, int dim, int n, uint64_t seed, double *b) {
  int64_t N, Z;
  or_fe_sizes(order, dim, n, &N, &Z);
  const int64_t p = (int64_t)order * n + 1;
  for (int64_t i = 0; i < N; ++i) {
    const int64_t ix = i % p, iy = (i / p) % p, iz = (dim == 3) ? i / (p * p) : 0;
    int bnd = ix == 0 || ix == p - 1 || iy == 0 || iy == p - 1;
    if (dim == 3) bnd = bnd || iz == 0 || iz == p - 1;
    b[i] = bnd ? 0.0 : or_uniform_pm1(seed, (uint64_t)i);
  }
}

/* ---------------------------------------------------------------------- */
/* coo_2_csr (lmatrix_formats.h:320-456): count, exclusive prefix sum,
 * scatter in input order, then a per-row selection sort on the columns. */
void or_coo_to_csr_d(int nrow, int nnz, const int *ri, const int *ci,
                     const double *v, int *rowptr, int *col, double *val) {
  int *next = (int *)malloc(sizeof(int) * (size_t)(nrow > 0 ? nrow : 1));
  memset(rowptr, 0, sizeof(int) * (size_t)(nrow + 1));
  for (int i = 0; i < nnz; ++i) ++rowptr[ri[i]];
  int acc = 0;
  for (int i = 0; i < nrow; ++i) {
    const int c = rowptr[i];
    rowptr[i] = acc;
    next[i] = acc;
    acc += c;
  }
  rowptr[nrow] = nnz;
  for (int i = 0; i < nnz; ++i) {
    const int t = next[ri[i]]++;
    col[t] = ci[i];
    val[t] = v ? v[i] : 0.0;
  }
  for (int k = 0; k < nrow; ++k)
    for (int i = rowptr[k]; i < rowptr[k + 1]; ++i) {
      int mi = i;
      for (int j = i + 1; j < rowptr[k + 1]; ++j)
        if (col[j] < col[mi]) mi = j;
      const int tc = col[i];
      col[i] = col[mi];
      col[mi] = tc;
      const double tv = val[i];
      val[i] = val[mi];
      val[mi] = tv;
    }
  free(next);
}

/* get_diag_ptr (lmatrix_formats.h:489-532): position of a_ii in each row.
 * Returns the number of rows without a stored diagonal. */
int or_csr_diag_ptr(int nrow, const int *rowptr, const int *col, int *diag) {
  int missing = 0;
  for (int i = 0; i < nrow; ++i) {
    diag[i] = -1;
    for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
      if (col[j] == i) {
        diag[i] = j;
        break;
      }
    if (diag[i] < 0) ++missing;
  }
  return missing;
}

/* ZeroRows (lmatrix_csr_cpu.cc:908-927): row values -> 0, diagonal ->
 * alpha; the columns are left untouched. */
void or_csr_zero_rows_d(const int *rows, int nrows_zero, const int *rowptr,
                        const int *col, double *val, double alpha) {
  for (int k = 0; k < nrows_zero; ++k) {
    const int r = rows[k];
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j)
      val[j] = (col[j] == r) ? alpha : 0.0;
  }
}

/* Multicoloring (lmatrix_csr_cpu.cc:2195-2300): natural-order first-fit.
 * Row j starts at colour 1 and steps up while any stored neighbour (other
 * than j) already holds the same colour; the reference repeats the sweep
 * over the row once per entry, which ends at the smallest colour >= 1 not
 * held by a neighbour.  perm[position] = node, nodes grouped by colour and
 * ascending inside a colour.  color[] receives the 0-based colour.
 * color_sizes must hold at least (max row length + 1) entries. */
int or_multicoloring(int nrow, const int *rowptr, const int *col,
                     int *ncolors, int *color_sizes, int *perm, int *color) {
  int *nc = (int *)calloc((size_t)(nrow > 0 ? nrow : 1), sizeof(int));
  int hicol = 0;
  for (int j = 0; j < nrow; ++j) {
    nc[j]++;
    const int m = rowptr[j], m1 = rowptr[j + 1];
    for (int k = m; k < m1; ++k)
      for (int i = m; i < m1; ++i) {
        const int dep = col[i];
        if (dep != j) {
          if (nc[dep] == nc[j]) ++nc[j];
          if (nc[j] > hicol) hicol = nc[j];
        }
      }
    if (nc[j] > hicol) hicol = nc[j];
  }
  *ncolors = hicol;
  for (int k = 0; k < hicol; ++k) color_sizes[k] = 0;
  for (int i = 0; i < nrow; ++i) ++color_sizes[nc[i] - 1];
  int *start = (int *)calloc((size_t)hicol + 1, sizeof(int));
  for (int k = 1; k <= hicol; ++k) start[k] = start[k - 1] + color_sizes[k - 1];
  for (int j = 0; j < nrow; ++j) {
    perm[start[nc[j] - 1]++] = j;
    if (color) color[j] = nc[j] - 1;
  }
  free(start);
  free(nc);
  return 0;
}

void or_cast_d2s(int n, const double *x, float *y) {
  for (int i = 0; i < n; ++i) y[i] = (float)x[i];
}
void or_cast_s2d(int n, const float *x, double *y) {
  for (int i = 0; i < n; ++i) y[i] = (double)x[i];
}
