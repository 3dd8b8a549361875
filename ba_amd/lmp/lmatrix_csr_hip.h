// HIP_CSR_lMatrix<ValueType>: a device CSR matrix of the HIP platform
// (the role of GPU_CSR_lMatrix / GPUscalar_CSR_lMatrix,
// cuda/lmatrix_csr_gpu.h:39-160).  row_ptr/col/val live on the device;
// a host copy of the pattern serves assembly (add_values finds positions
// on the host, then scatters on the device) and colouring.
//
// VectorMult and VectorMultAdd run through a SELL-64 copy of the matrix
// (hflmp_sell_*: slices of 64 rows, one wave each, lane per row), built on
// the device at the first product and refreshed after any value change.
// Its columns are int16 offsets from each slice's first row when every
// column fits (10 B per FP64 nonzero instead of 12), else int32.
// Its products are bit-identical to the CSR kernels'; a pattern that does
// not suit it (a row longer than 255 entries, heavy padding) stays on the
// CSR kernels, and set_use_sell(false) forces them.  The copy costs about
// the matrix's own device memory again.
#pragma once

#include <map>
#include <vector>

#include "hflmp_hip.h"
#include "lmatrix.h"

namespace hiflow {
namespace la {

template <typename ValueType>
class HIP_CSR_lMatrix : public lMatrix<ValueType> {
 public:
  HIP_CSR_lMatrix();
  HIP_CSR_lMatrix(int init_nnz, int init_num_row, int init_num_col, std::string init_name);
  ~HIP_CSR_lMatrix() override;

  void Init(const int init_nnz, const int init_num_row, const int init_num_col,
            const std::string init_name) override;
  void Clear() override;
  void Zeros() override;

  void init_structure(const int *rows, const int *cols) override;
  void TransformFromCOO(const int *rows, const int *cols, const ValueType *data,
                        const int num_rows, const int num_cols,
                        const int num_nonzeros) override;
  // CSR arrays given directly (host), columns sorted ascending in each row
  void SetCSR(const int *row_ptr, const int *col, const ValueType *val);
  // MatrixMarket "coordinate real general", entries in CSR order
  void WriteFile(const char *filename) const override;

  void add_value(const int row, const int col, const ValueType val) override;
  void add_values(const int *rows, int num_rows, const int *cols, int num_cols,
                  const ValueType *values) override;
  void set_values(const int *rows, int num_rows, const int *cols, int num_cols,
                  const ValueType *values) override;
  void get_value(const int row, const int col, ValueType *val) const override;
  void ZeroRows(const int *index_set, const int size, const ValueType alpha) override;

  void CopyFrom(const lMatrix<ValueType> &mat) override;
  void CopyStructureFrom(const lMatrix<ValueType> &mat) override;
  void CastFrom(const lMatrix<double> &mat) override;
  void CastFrom(const lMatrix<float> &mat) override;
  lMatrix<ValueType> *CloneWithoutContent() const override;

  void VectorMult(const lVector<ValueType> &invec, lVector<ValueType> *outvec) const override;
  void VectorMultAdd(const lVector<ValueType> &invec, lVector<ValueType> *outvec) const override;
  void VectorMultNoDiag(const lVector<ValueType> &invec,
                        lVector<ValueType> *outvec) const override;
  void extract_invdiagelements(const int start_i, const int end_i,
                               lVector<ValueType> *vec) const override;
  void Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const override;
  // Sequential and block symmetric Gauss-Seidel on the device (hflmp_sgs_*),
  // bit-identical to the CPU sweeps; the level schedule of each block count
  // is built at the first call and kept until the pattern changes.
  void Psgauss_seidel(const lVector<ValueType> &invec, lVector<ValueType> *outvec) const override;
  void BlocksPsgauss_seidel(const lVector<ValueType> &invec, lVector<ValueType> *outvec,
                            const int num_blocks) const override;

  // device arrays
  const int *row_ptr() const { return row_; }
  const int *col_ind() const { return col_; }
  const ValueType *values() const { return val_; }
  // host copy of the pattern
  const std::vector<int> &host_row_ptr() const { return h_row_; }
  const std::vector<int> &host_col_ind() const { return h_col_; }

  // SELL-64 execution layout for the products (on by default)
  void set_use_sell(bool on);
  // the layout the next product will use: nullptr means the CSR kernels
  const hflmp_sell *sell_layout() const;

 private:
  void values_changed() { sell_stale_ = true; }
  void drop_layout();
  void allocate(int nnz, int nrow, int ncol);
  void upload_structure();
  int position(int row, int col) const;  // -1 if (row,col) is not in the pattern
  void scatter_values(const std::vector<int> &pos, const std::vector<ValueType> &v, bool add);
  void check_vectors(const lVector<ValueType> &in, const lVector<ValueType> *out,
                     const char *op) const;

  int *row_ = nullptr;
  int *col_ = nullptr;
  ValueType *val_ = nullptr;
  std::vector<int> h_row_, h_col_;
  const hflmp_sgs *sgs_schedule(int num_blocks) const;

  mutable std::map<int, hflmp_sgs *> sgs_;  // block count -> SGS level schedule
  mutable hflmp_sell *sell_ = nullptr;
  mutable bool sell_refused_ = false;  // the pattern does not suit SELL-64
  mutable bool sell_stale_ = false;    // values changed since the copy was made
  bool use_sell_ = true;
};

}  // namespace la
}  // namespace hiflow
