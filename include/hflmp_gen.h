/*
 * hflmp_gen.h -- structured finite-element test systems (host side).
 *
 * Stands in for HiFlow3's mesh -> DoF -> assembly chain for the structured
 * meshes the benchmarks use (assembly/assembly_utils.cc:144
 * compute_sparsity_structure, :1077 InitStructure, then
 * CoupledMatrix::diagonalize_rows, coupled_matrix.cc:452-481).
 *
 * Meshes: n cells per side on [0,1]^dim, Q1 (dim 2 or 3) or Q2 (dim 2)
 * Lagrange elements, nodes numbered lexicographically (x fastest), every
 * boundary node Dirichlet.  The CSR pattern holds every pair of nodes that
 * share a cell, columns ascending; entry (i,j) is the sum over the shared
 * cells, in ascending cell index, of kappa_cell * E[l_i][l_j] with the
 * exact element stiffness E; Dirichlet rows are unit rows with the pattern
 * kept (ZeroRows, lmatrix_csr_cpu.cc:908-927).
 *
 * Random streams: SplitMix64 (Steele, Lea, Flood 2014), output i of the
 * stream with state `seed` is mix(seed + (i+1) * 0x9E3779B97F4A7C15).
 *   rhs:   b_i = 2 * (out_i >> 11) * 2^-53 - 1 on interior nodes, 0 on the
 *          boundary
 *   kappa: kappa_c = exp(sigma * g_c), g_c Box-Muller from outputs 2c, 2c+1
 */
#ifndef HFLMP_GEN_H
#define HFLMP_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* order 1 (dim 2, 3) or 2 (dim 2); returns 0 or -1 if unsupported */
int hflmp_gen_fe_sizes(int order, int dim, int n, int64_t *nrow, int64_t *nnz);
/* fills row_ptr[nrow+1], col[nnz], val[nnz]; kappa per cell or NULL */
int hflmp_gen_fe_csr(int order, int dim, int n, const double *kappa,
                     int *row_ptr, int *col, double *val);
void hflmp_gen_kappa(int dim, int n, double sigma, uint64_t seed,
                     double *kappa);
void hflmp_gen_rhs(int order, int dim, int n, uint64_t seed, double *b);
/* the element stiffness E (nloc x nloc, row-major, local nodes
 * lexicographic, h included) that cell c enters with factor kappa_c: the
 * local matrix of the matrix-free stencil operator; returns nloc, or -1 */
int hflmp_gen_fe_element(int order, int dim, int n, double *E);
/* number of host threads the generator uses (0 = all) */
void hflmp_gen_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
