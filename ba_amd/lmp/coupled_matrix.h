// CoupledMatrix<DataType> on one process (linear_algebra/coupled_matrix.h:47;
// coupled_matrix.cc:65-481).  The reference splits the local rows into a
// diagonal block (owned columns) and an off-diagonal block (ghost columns)
// and runs ReceiveGhost / diagonal VectorMult / WaitForRecv / off-diagonal
// VectorMultAdd (coupled_matrix.cc:328-364).  On one device the
// off-diagonal block is empty, so VectorMult is the diagonal block's.
#pragma once

#include <memory>
#include <vector>

#include "coupled_vector.h"
#include "init_vec_mat.h"
#include "la_couplings.h"
#include "lmatrix_csr_hip.h"
#include "lmp_log.h"

namespace hiflow {
namespace la {

template <typename DataType>
class CoupledMatrix {
 public:
  CoupledMatrix() = default;
  CoupledMatrix(const CoupledMatrix &) = delete;
  CoupledMatrix &operator=(const CoupledMatrix &) = delete;

  void Init(const LaCouplings &row_cp, const LaCouplings &col_cp, PLATFORM plat = HIP,
            IMPLEMENTATION impl = HIPKERNEL, MATRIX_FORMAT format = CSR) {
    row_cp_ = row_cp;
    col_cp_ = col_cp;
    platform_ = plat;
    impl_ = impl;
    format_ = format;
    diagonal_.reset(init_matrix<DataType>(0, row_cp.size_global(), col_cp.size_global(),
                                          "diagonal", plat, impl, format));
    offdiagonal_.reset(init_matrix<DataType>(0, row_cp.size_global(), 0, "offdiagonal", plat,
                                             impl, format));
  }
  void Init(const LaCouplings &cp, PLATFORM plat = HIP, IMPLEMENTATION impl = HIPKERNEL,
            MATRIX_FORMAT format = CSR) {
    Init(cp, cp, plat, impl, format);
  }
  void Clear() {
    diagonal_.reset();
    offdiagonal_.reset();
  }
  bool is_initialized() const { return (bool)diagonal_; }

  // Pattern from global (row, col) pairs; on one process every column is
  // owned, so the off-diagonal lists must be empty.
  void InitStructure(const int *rows_diag, const int *cols_diag, const int num_nonzeros_diag,
                     const int *rows_offdiag, const int *cols_offdiag,
                     const int num_nonzeros_offdiag) {
    (void)rows_offdiag;
    (void)cols_offdiag;
    if (num_nonzeros_offdiag != 0)
      LMP_FATAL("CoupledMatrix::InitStructure: off-diagonal entries need more than one process");
    diagonal().Init(num_nonzeros_diag, row_cp_.size_global(), col_cp_.size_global(), "diagonal");
    diagonal().init_structure(rows_diag, cols_diag);
  }
  // Pattern and values from CSR arrays (host), columns ascending per row.
  void InitCSR(const int *row_ptr, const int *col, const DataType *val) {
    const int nrow = row_cp_.size_global();
    auto *h = dynamic_cast<HIP_CSR_lMatrix<DataType> *>(diagonal_.get());
    if (!h) LMP_FATAL("CoupledMatrix::InitCSR: the diagonal block is not a HIP CSR matrix");
    h->Init(row_ptr[nrow], nrow, col_cp_.size_global(), "diagonal");
    h->SetCSR(row_ptr, col, val);
  }

  void Zeros() { diagonal().Zeros(); }
  void Add(int i, int j, DataType val) { diagonal().add_value(i, j, val); }
  void Add(const int *rows, int num_rows, const int *cols, int num_cols, const DataType *values) {
    diagonal().add_values(rows, num_rows, cols, num_cols, values);
  }
  void SetValues(const int *rows, int num_rows, const int *cols, int num_cols,
                 const DataType *values) {
    diagonal().set_values(rows, num_rows, cols, num_cols, values);
  }
  void GetValues(const int *rows, int num_rows, const int *cols, int num_cols,
                 DataType *values) const {
    for (int i = 0; i < num_rows; ++i)
      for (int j = 0; j < num_cols; ++j)
        diagonal().get_value(rows[i], cols[j], &values[i * num_cols + j]);
  }
  // Dirichlet rows: row values 0, diagonal = diagonal_value, columns kept
  // (coupled_matrix.cc:452-481 -> ZeroRows)
  void diagonalize_rows(const int *row_indices, int num_rows, DataType diagonal_value) {
    diagonal().ZeroRows(row_indices, num_rows, diagonal_value);
  }

  // out = A in
  void VectorMult(CoupledVector<DataType> &in, CoupledVector<DataType> *out) const {
    in.ReceiveGhost();
    in.SendBorder();
    diagonal().VectorMult(in.interior(), &out->interior());
    in.WaitForRecv();
    offdiagonal().VectorMultAdd(in.ghost(), &out->interior());
    in.WaitForSend();
  }
  // out = beta out + alpha A in
  void VectorMultAdd(DataType alpha, CoupledVector<DataType> &in, DataType beta,
                     CoupledVector<DataType> *out) const {
    CoupledVector<DataType> tmp;
    tmp.CloneFromWithoutContent(*out);
    VectorMult(in, &tmp);
    out->Scale(beta);
    out->Axpy(tmp, alpha);
  }

  void CloneFrom(const CoupledMatrix<DataType> &mat) {
    CopyStructureFrom(mat);
    CopyFrom(mat);
  }
  void CopyStructureFrom(const CoupledMatrix<DataType> &mat) {
    Init(mat.row_cp_, mat.col_cp_, mat.platform_, mat.impl_, mat.format_);
    diagonal().CopyStructureFrom(mat.diagonal());
  }
  void CopyFrom(const CoupledMatrix<DataType> &mat) { diagonal().CopyFrom(mat.diagonal()); }
  // structure and values of a matrix of another precision
  template <typename Other>
  void CastFrom(const CoupledMatrix<Other> &mat) {
    Init(mat.row_couplings(), mat.col_couplings(), mat.platform(), HIPKERNEL, CSR);
    diagonal().CastFrom(mat.diagonal());
  }

  lMatrix<DataType> &diagonal() {
    if (!diagonal_) LMP_FATAL("CoupledMatrix: not initialized");
    return *diagonal_;
  }
  const lMatrix<DataType> &diagonal() const {
    if (!diagonal_) LMP_FATAL("CoupledMatrix: not initialized");
    return *diagonal_;
  }
  const lMatrix<DataType> &offdiagonal() const { return *offdiagonal_; }

  int num_rows_local() const { return row_cp_.size_global(); }
  int num_rows_global() const { return row_cp_.size_global(); }
  int num_cols_global() const { return col_cp_.size_global(); }
  int nnz_local_diag() const { return diagonal().get_nnz(); }
  int nnz_global() const { return diagonal().get_nnz(); }
  const LaCouplings &row_couplings() const { return row_cp_; }
  const LaCouplings &col_couplings() const { return col_cp_; }
  PLATFORM platform() const { return platform_; }

 private:
  LaCouplings row_cp_, col_cp_;
  PLATFORM platform_ = HIP;
  IMPLEMENTATION impl_ = HIPKERNEL;
  MATRIX_FORMAT format_ = CSR;
  std::unique_ptr<lMatrix<DataType>> diagonal_, offdiagonal_;
};

}  // namespace la
}  // namespace hiflow
