// lMatrix<ValueType>: the local matrix interface (linear_algebra/lmp/
// lmatrix.h:39-468), reduced to the part the Coupled layer, the solvers and
// the preconditioners use.  Host-side arrays passed to init_structure,
// add_values, ZeroRows, ... are host memory, as in the reference.
#pragma once

#include <iostream>
#include <string>

#include <vector>

#include "la_global.h"
#include "lmp_log.h"
#include "lmp_mmio.h"
#include "lvector.h"

namespace hiflow {
namespace la {

template <typename ValueType>
class lMatrix {
 public:
  lMatrix() = default;
  virtual ~lMatrix() = default;
  lMatrix(const lMatrix &) = delete;
  lMatrix &operator=(const lMatrix &) = delete;

  int get_num_row() const { return num_row_; }
  int get_num_col() const { return num_col_; }
  int get_nnz() const { return nnz_; }
  std::string get_name() const { return name_; }
  PLATFORM get_platform() const { return platform_id_; }
  IMPLEMENTATION get_implementation() const { return implementation_id_; }
  MATRIX_FORMAT get_matrix_format() const { return matrix_format_id_; }
  std::string get_platform_name() const { return platform_name(platform_id_); }
  std::string get_implementation_name() const { return implementation_name(implementation_id_); }
  std::string get_matrix_format_name() const { return format_name(matrix_format_id_); }
  virtual void print(std::ostream &out = std::cout) const;

  virtual void Init(const int init_nnz, const int init_num_row, const int init_num_col,
                    const std::string init_name) = 0;
  virtual void Clear() = 0;
  virtual void Zeros() = 0;
  // MatrixMarket coordinate file (lmatrix.h:344-348; CSR through COO,
  // lmatrix_csr_cpu.cc:1781-1810): read replaces structure and values
  // (duplicate entries are added), write stores the matrix in CSR order.
  virtual void ReadFile(const char *filename);
  virtual void WriteFile(const char *filename) const;

  // structure from (row, col) pairs; duplicates are merged, columns sorted
  // ascending per row (init_structure / coo_2_csr, lmatrix_formats.h:320-456)
  virtual void init_structure(const int *rows, const int *cols) = 0;
  // structure and values from COO triplets (TransformFromCOO,
  // lmatrix_csr_cpu.cc:883-905); duplicate entries are summed
  virtual void TransformFromCOO(const int *rows, const int *cols, const ValueType *data,
                                const int num_rows, const int num_cols, const int num_nonzeros) = 0;

  // assembly into the existing pattern (entries outside it are an error)
  virtual void add_value(const int row, const int col, const ValueType val) = 0;
  virtual void add_values(const int *rows, int num_rows, const int *cols, int num_cols,
                          const ValueType *values) = 0;
  virtual void set_values(const int *rows, int num_rows, const int *cols, int num_cols,
                          const ValueType *values) = 0;
  virtual void get_value(const int row, const int col, ValueType *val) const = 0;
  // Dirichlet rows: row -> 0, diagonal -> alpha, columns kept
  virtual void ZeroRows(const int *index_set, const int size, const ValueType alpha) = 0;

  virtual void CopyFrom(const lMatrix<ValueType> &mat) = 0;
  virtual void CopyStructureFrom(const lMatrix<ValueType> &mat) = 0;
  virtual void CloneFrom(const lMatrix<ValueType> &mat);
  virtual void CastFrom(const lMatrix<double> &mat) = 0;
  virtual void CastFrom(const lMatrix<float> &mat) = 0;
  virtual lMatrix<ValueType> *CloneWithoutContent() const = 0;

  // out = A in
  virtual void VectorMult(const lVector<ValueType> &invec, lVector<ValueType> *outvec) const = 0;
  // out += A in
  virtual void VectorMultAdd(const lVector<ValueType> &invec,
                             lVector<ValueType> *outvec) const = 0;
  // out = (A - D) in
  virtual void VectorMultNoDiag(const lVector<ValueType> &invec,
                                lVector<ValueType> *outvec) const = 0;
  // vec[i - start_i] = 1 / a_ii for start_i <= i < end_i
  virtual void extract_invdiagelements(const int start_i, const int end_i,
                                       lVector<ValueType> *vec) const = 0;
  // greedy natural-order colouring of the pattern
  // (CPU_CSR_lMatrix::Multicoloring, lmatrix_csr_cpu.cc:2195-2300)
  virtual void Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const = 0;
  // out = M^-1 in, M = (D + L) D^-1 (D + U), rows in the sequential order
  // (Psgauss_seidel, lmatrix.h:419; CPU_CSR_lMatrix, lmatrix_csr_cpu.cc:495-565)
  virtual void Psgauss_seidel(const lVector<ValueType> &invec,
                              lVector<ValueType> *outvec) const;
  // the same on num_blocks row ranges swept independently (BlocksPsgauss_seidel,
  // lmatrix.h:437; lmatrix_csr_cpu.cc:568-641)
  virtual void BlocksPsgauss_seidel(const lVector<ValueType> &invec, lVector<ValueType> *outvec,
                                    const int num_blocks) const;

 protected:
  int num_row_ = 0, num_col_ = 0, nnz_ = 0;
  std::string name_;
  IMPLEMENTATION implementation_id_ = HIPKERNEL;
  PLATFORM platform_id_ = HIP;
  MATRIX_FORMAT matrix_format_id_ = CSR;
};

template <typename ValueType>
void lMatrix<ValueType>::print(std::ostream &out) const {
  out << "lMatrix name=" << name_ << " nnz=" << nnz_ << " rows=" << num_row_
      << " cols=" << num_col_ << " precision=" << 8 * sizeof(ValueType)
      << "bit format=" << get_matrix_format_name() << " platform=" << get_platform_name()
      << " implementation=" << get_implementation_name() << std::endl;
}

template <typename ValueType>
void lMatrix<ValueType>::ReadFile(const char *filename) {
  MmCoo coo;
  const std::string err = mm_read_coo(filename, coo);
  if (!err.empty()) LMP_FATAL("lMatrix::ReadFile(" << filename << "): " << err);
  std::vector<ValueType> v(coo.val.begin(), coo.val.end());
  this->Clear();
  this->Init((int)v.size(), coo.nrow, coo.ncol, filename);
  this->TransformFromCOO(coo.row.data(), coo.col.data(), v.data(), coo.nrow, coo.ncol,
                         (int)v.size());
}

template <typename ValueType>
void lMatrix<ValueType>::WriteFile(const char *filename) const {
  LMP_FATAL("lMatrix::WriteFile() does not support the format " << get_matrix_format_name());
}

template <typename ValueType>
void lMatrix<ValueType>::Psgauss_seidel(const lVector<ValueType> &, lVector<ValueType> *) const {
  LMP_FATAL("lMatrix::Psgauss_seidel() is not implemented for " << get_matrix_format_name());
}

template <typename ValueType>
void lMatrix<ValueType>::BlocksPsgauss_seidel(const lVector<ValueType> &, lVector<ValueType> *,
                                              const int) const {
  LMP_FATAL("lMatrix::BlocksPsgauss_seidel() is not implemented for "
            << get_matrix_format_name());
}

template <typename ValueType>
void lMatrix<ValueType>::CloneFrom(const lMatrix<ValueType> &mat) {
  this->Clear();
  this->CopyStructureFrom(mat);
  this->CopyFrom(mat);
}

}  // namespace la
}  // namespace hiflow
