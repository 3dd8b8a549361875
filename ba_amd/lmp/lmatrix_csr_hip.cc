// HIP_CSR_lMatrix: CSR matrices of the HIP platform.  Structure building
// follows the reference's COO -> CSR conversion (count, prefix sum,
// scatter, ascending columns per row: lmatrix_formats.h:320-456, here with
// duplicate pairs merged); products and the diagonal extraction run on the
// device through libhflmp_hip.
#include "lmatrix_csr_hip.h"

#include <algorithm>
#include <numeric>
#include <type_traits>

#include "hip_ops.h"
#include "lmp_log.h"
#include "lvector_hip.h"
#include "platform_management.h"

namespace hiflow {
namespace la {

namespace {

template <typename V>
V *dev_alloc(size_t n, const char *what) {
  V *p = nullptr;
  if (n) hip_check(hflmp_malloc(hip_ctx(), n * sizeof(V), reinterpret_cast<void **>(&p)), what);
  return p;
}

template <typename V>
void dev_free(V *&p) {
  if (p) hflmp_free(hip_ctx(), p);
  p = nullptr;
}

}  // namespace

template <typename T>
HIP_CSR_lMatrix<T>::HIP_CSR_lMatrix() {
  this->platform_id_ = HIP;
  this->implementation_id_ = HIPKERNEL;
  this->matrix_format_id_ = CSR;
}

template <typename T>
HIP_CSR_lMatrix<T>::HIP_CSR_lMatrix(int init_nnz, int init_num_row, int init_num_col,
                                    std::string init_name)
    : HIP_CSR_lMatrix() {
  Init(init_nnz, init_num_row, init_num_col, init_name);
}

template <typename T>
HIP_CSR_lMatrix<T>::~HIP_CSR_lMatrix() {
  Clear();
}

template <typename T>
void HIP_CSR_lMatrix<T>::drop_layout() {
  for (auto &e : sgs_) hflmp_sgs_destroy(e.second);
  sgs_.clear();
  if (sell_) hflmp_sell_destroy(sell_);
  sell_ = nullptr;
  sell_refused_ = false;
  sell_stale_ = false;
}

template <typename T>
void HIP_CSR_lMatrix<T>::set_use_sell(bool on) {
  use_sell_ = on;
  if (!on) drop_layout();
}

template <typename T>
const hflmp_sell *HIP_CSR_lMatrix<T>::sell_layout() const {
  if (!use_sell_ || sell_refused_ || this->nnz_ == 0 || this->num_row_ == 0 ||
      this->num_col_ == 0)
    return nullptr;
  if (!sell_) {
    const int rc = HipOps<T>::sell_create(hip_ctx(), this->num_row_, this->num_col_,
                                          this->nnz_, row_, col_, val_, HFLMP_SELL_COL16,
                                          &sell_);
    if (rc == HFLMP_EFORMAT) {
      sell_ = nullptr;
      sell_refused_ = true;
      return nullptr;
    }
    hip_check(rc, "SELL-64 layout");
    sell_stale_ = false;
  } else if (sell_stale_) {
    hip_check(HipOps<T>::sell_set_values(hip_ctx(), sell_, row_, val_), "SELL-64 values");
    sell_stale_ = false;
  }
  return sell_;
}

template <typename T>
void HIP_CSR_lMatrix<T>::allocate(int nnz, int nrow, int ncol) {
  drop_layout();
  dev_free(row_);
  dev_free(col_);
  dev_free(val_);
  this->nnz_ = nnz;
  this->num_row_ = nrow;
  this->num_col_ = ncol;
  row_ = dev_alloc<int>((size_t)nrow + 1, "HIP_CSR_lMatrix row_ptr");
  col_ = dev_alloc<int>((size_t)nnz, "HIP_CSR_lMatrix col");
  val_ = dev_alloc<T>((size_t)nnz, "HIP_CSR_lMatrix val");
  h_row_.assign((size_t)nrow + 1, 0);
  h_col_.assign((size_t)nnz, 0);
  if (nnz) hip_check(hflmp_fill_zero(hip_ctx(), val_, sizeof(T) * (size_t)nnz), "Zeros");
}

template <typename T>
void HIP_CSR_lMatrix<T>::Init(const int init_nnz, const int init_num_row, const int init_num_col,
                              const std::string init_name) {
  if (init_nnz < 0 || init_num_row < 0 || init_num_col < 0)
    LMP_FATAL("HIP_CSR_lMatrix::Init: negative size");
  this->name_ = init_name;
  allocate(init_nnz, init_num_row, init_num_col);
  upload_structure();  // empty rows until a structure is set
}

template <typename T>
void HIP_CSR_lMatrix<T>::Clear() {
  drop_layout();
  dev_free(row_);
  dev_free(col_);
  dev_free(val_);
  h_row_.clear();
  h_col_.clear();
  this->nnz_ = this->num_row_ = this->num_col_ = 0;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Zeros() {
  values_changed();
  if (this->nnz_) hip_check(hflmp_fill_zero(hip_ctx(), val_, sizeof(T) * (size_t)this->nnz_), "Zeros");
}

template <typename T>
void HIP_CSR_lMatrix<T>::upload_structure() {
  drop_layout();
  hip_check(hflmp_memcpy_h2d(hip_ctx(), row_, h_row_.data(), sizeof(int) * h_row_.size()),
            "row_ptr upload");
  if (this->nnz_)
    hip_check(hflmp_memcpy_h2d(hip_ctx(), col_, h_col_.data(), sizeof(int) * h_col_.size()),
              "col upload");
}

// COO pairs -> sorted, duplicate-free CSR pattern; `vals` (optional) are
// summed into the merged entries.
template <typename T>
static void coo_to_csr(int nrow, int ncol, int n, const int *rows, const int *cols,
                       const T *vals, std::vector<int> &rp, std::vector<int> &ci,
                       std::vector<T> &v) {
  std::vector<int> cnt((size_t)nrow + 1, 0);
  for (int k = 0; k < n; ++k) {
    if (rows[k] < 0 || rows[k] >= nrow || cols[k] < 0 || cols[k] >= ncol)
      LMP_FATAL("COO entry (" << rows[k] << "," << cols[k] << ") outside " << nrow << "x" << ncol);
    ++cnt[rows[k] + 1];
  }
  std::partial_sum(cnt.begin(), cnt.end(), cnt.begin());
  std::vector<int> order((size_t)n), next(cnt.begin(), cnt.end() - 1);
  for (int k = 0; k < n; ++k) order[next[rows[k]]++] = k;
  rp.assign((size_t)nrow + 1, 0);
  ci.clear();
  v.clear();
  ci.reserve(n);
  if (vals) v.reserve(n);
  for (int r = 0; r < nrow; ++r) {
    std::stable_sort(order.begin() + cnt[r], order.begin() + cnt[r + 1],
                     [&](int a, int b) { return cols[a] < cols[b]; });
    for (int q = cnt[r]; q < cnt[r + 1]; ++q) {
      const int k = order[q];
      if (q > cnt[r] && cols[order[q - 1]] == cols[k]) {
        if (vals) v.back() += vals[k];
        continue;
      }
      ci.push_back(cols[k]);
      if (vals) v.push_back(vals[k]);
    }
    rp[r + 1] = (int)ci.size();
  }
}

template <typename T>
void HIP_CSR_lMatrix<T>::init_structure(const int *rows, const int *cols) {
  std::vector<int> rp, ci;
  std::vector<T> dummy;
  coo_to_csr<T>(this->num_row_, this->num_col_, this->nnz_, rows, cols, nullptr, rp, ci, dummy);
  const int nrow = this->num_row_, ncol = this->num_col_;
  allocate((int)ci.size(), nrow, ncol);
  h_row_ = rp;
  h_col_ = ci;
  upload_structure();
}

template <typename T>
void HIP_CSR_lMatrix<T>::TransformFromCOO(const int *rows, const int *cols, const T *data,
                                          const int num_rows, const int num_cols,
                                          const int num_nonzeros) {
  std::vector<int> rp, ci;
  std::vector<T> v;
  coo_to_csr<T>(num_rows, num_cols, num_nonzeros, rows, cols, data, rp, ci, v);
  allocate((int)ci.size(), num_rows, num_cols);
  h_row_ = rp;
  h_col_ = ci;
  upload_structure();
  if (!v.empty())
    hip_check(hflmp_memcpy_h2d(hip_ctx(), val_, v.data(), sizeof(T) * v.size()), "val upload");
}

template <typename T>
void HIP_CSR_lMatrix<T>::SetCSR(const int *row_ptr, const int *col, const T *val) {
  if (row_ptr[0] != 0 || row_ptr[this->num_row_] != this->nnz_)
    LMP_FATAL("HIP_CSR_lMatrix::SetCSR: row_ptr does not span nnz = " << this->nnz_);
  for (int r = 0; r < this->num_row_; ++r) {
    if (row_ptr[r + 1] < row_ptr[r]) LMP_FATAL("SetCSR: row_ptr decreases at row " << r);
    for (int j = row_ptr[r]; j < row_ptr[r + 1]; ++j) {
      if (col[j] < 0 || col[j] >= this->num_col_) LMP_FATAL("SetCSR: column out of range");
      if (j > row_ptr[r] && col[j] <= col[j - 1])
        LMP_FATAL("SetCSR: columns of row " << r << " are not strictly ascending");
    }
  }
  h_row_.assign(row_ptr, row_ptr + this->num_row_ + 1);
  h_col_.assign(col, col + this->nnz_);
  upload_structure();
  if (val && this->nnz_)
    hip_check(hflmp_memcpy_h2d(hip_ctx(), val_, val, sizeof(T) * (size_t)this->nnz_), "val upload");
}

template <typename T>
void HIP_CSR_lMatrix<T>::WriteFile(const char *filename) const {
  std::vector<T> v((size_t)this->nnz_);
  if (this->nnz_)
    hip_check(hflmp_memcpy_d2h(hip_ctx(), v.data(), val_, sizeof(T) * v.size()), "val download");
  std::vector<double> d(v.begin(), v.end());
  const std::vector<int> rp = h_row_.empty() ? std::vector<int>((size_t)this->num_row_ + 1, 0)
                                             : h_row_;
  const std::string err = mm_write_csr(filename, this->num_row_, this->num_col_, rp.data(),
                                       h_col_.data(), d.data());
  if (!err.empty()) LMP_FATAL("HIP_CSR_lMatrix::WriteFile(" << filename << "): " << err);
}

template <typename T>
int HIP_CSR_lMatrix<T>::position(int row, int col) const {
  if (row < 0 || row >= this->num_row_) return -1;
  const auto b = h_col_.begin() + h_row_[row], e = h_col_.begin() + h_row_[row + 1];
  const auto it = std::lower_bound(b, e, col);
  return (it != e && *it == col) ? (int)(it - h_col_.begin()) : -1;
}

template <typename T>
void HIP_CSR_lMatrix<T>::scatter_values(const std::vector<int> &pos, const std::vector<T> &v,
                                        bool add) {
  if (pos.empty()) return;
  values_changed();
  int *dp = dev_alloc<int>(pos.size(), "scatter positions");
  T *dv = dev_alloc<T>(v.size(), "scatter values");
  hip_check(hflmp_memcpy_h2d(hip_ctx(), dp, pos.data(), sizeof(int) * pos.size()), "upload");
  hip_check(hflmp_memcpy_h2d(hip_ctx(), dv, v.data(), sizeof(T) * v.size()), "upload");
  if (add)
    hip_check(HipOps<T>::scatter_add(hip_ctx(), (int)pos.size(), dp, dv, val_), "add_values");
  else
    hip_check(HipOps<T>::scatter(hip_ctx(), (int)pos.size(), dp, dv, val_), "set_values");
  dev_free(dp);
  dev_free(dv);
}

template <typename T>
void HIP_CSR_lMatrix<T>::add_value(const int row, const int col, const T val) {
  add_values(&row, 1, &col, 1, &val);
}

template <typename T>
void HIP_CSR_lMatrix<T>::add_values(const int *rows, int num_rows, const int *cols, int num_cols,
                                    const T *values) {
  std::vector<int> pos;
  std::vector<T> v;
  for (int i = 0; i < num_rows; ++i)
    for (int j = 0; j < num_cols; ++j) {
      const int p = position(rows[i], cols[j]);
      if (p < 0)
        LMP_FATAL("HIP_CSR_lMatrix::add_values: (" << rows[i] << "," << cols[j]
                                                   << ") is not in the sparsity pattern");
      pos.push_back(p);
      v.push_back(values[i * num_cols + j]);
    }
  scatter_values(pos, v, true);
}

template <typename T>
void HIP_CSR_lMatrix<T>::set_values(const int *rows, int num_rows, const int *cols, int num_cols,
                                    const T *values) {
  std::vector<int> pos;
  std::vector<T> v;
  for (int i = 0; i < num_rows; ++i)
    for (int j = 0; j < num_cols; ++j) {
      const int p = position(rows[i], cols[j]);
      if (p < 0)
        LMP_FATAL("HIP_CSR_lMatrix::set_values: (" << rows[i] << "," << cols[j]
                                                   << ") is not in the sparsity pattern");
      pos.push_back(p);
      v.push_back(values[i * num_cols + j]);
    }
  scatter_values(pos, v, false);
}

template <typename T>
void HIP_CSR_lMatrix<T>::get_value(const int row, const int col, T *val) const {
  const int p = position(row, col);
  if (p < 0) {
    *val = T(0);
    return;
  }
  hip_check(hflmp_memcpy_d2h(hip_ctx(), val, val_ + p, sizeof(T)), "get_value");
}

template <typename T>
void HIP_CSR_lMatrix<T>::ZeroRows(const int *index_set, const int size, const T alpha) {
  if (size <= 0) return;
  for (int k = 0; k < size; ++k)
    if (index_set[k] < 0 || index_set[k] >= this->num_row_)
      LMP_FATAL("HIP_CSR_lMatrix::ZeroRows: row " << index_set[k] << " out of range");
  values_changed();
  int *d = dev_alloc<int>((size_t)size, "ZeroRows");
  hip_check(hflmp_memcpy_h2d(hip_ctx(), d, index_set, sizeof(int) * (size_t)size), "upload");
  hip_check(HipOps<T>::zero_rows(hip_ctx(), size, d, row_, col_, val_, alpha), "ZeroRows");
  dev_free(d);
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyStructureFrom(const lMatrix<T> &mat) {
  const auto *h = dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat);
  if (!h) LMP_FATAL("HIP_CSR_lMatrix::CopyStructureFrom: source is not a HIP CSR matrix");
  if (h == this) return;
  this->name_ = h->name_;
  allocate(h->nnz_, h->num_row_, h->num_col_);
  h_row_ = h->h_row_;
  h_col_ = h->h_col_;
  upload_structure();
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyFrom(const lMatrix<T> &mat) {
  const auto *h = dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat);
  if (!h) LMP_FATAL("HIP_CSR_lMatrix::CopyFrom: source is not a HIP CSR matrix");
  if (h == this) return;
  if (h->nnz_ != this->nnz_ || h->num_row_ != this->num_row_)
    LMP_FATAL("HIP_CSR_lMatrix::CopyFrom: structure differs (call CopyStructureFrom first)");
  values_changed();
  if (this->nnz_)
    hip_check(hflmp_memcpy_d2d(hip_ctx(), val_, h->val_, sizeof(T) * (size_t)this->nnz_),
              "CopyFrom");
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastFrom(const lMatrix<double> &mat) {
  const auto *h = dynamic_cast<const HIP_CSR_lMatrix<double> *>(&mat);
  if (!h) LMP_FATAL("HIP_CSR_lMatrix::CastFrom: source is not a HIP CSR matrix");
  this->name_ = h->get_name();
  allocate(h->get_nnz(), h->get_num_row(), h->get_num_col());
  h_row_ = h->host_row_ptr();
  h_col_ = h->host_col_ind();
  upload_structure();
  if (!this->nnz_) return;
  if constexpr (std::is_same<T, double>::value)
    hip_check(hflmp_memcpy_d2d(hip_ctx(), val_, h->values(), sizeof(T) * (size_t)this->nnz_),
              "CastFrom");
  else
    hip_check(hflmp_cast_csr_f64_to_f32(hip_ctx(), this->nnz_, h->values(), val_), "CastFrom");
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastFrom(const lMatrix<float> &mat) {
  const auto *h = dynamic_cast<const HIP_CSR_lMatrix<float> *>(&mat);
  if (!h) LMP_FATAL("HIP_CSR_lMatrix::CastFrom: source is not a HIP CSR matrix");
  this->name_ = h->get_name();
  allocate(h->get_nnz(), h->get_num_row(), h->get_num_col());
  h_row_ = h->host_row_ptr();
  h_col_ = h->host_col_ind();
  upload_structure();
  if (!this->nnz_) return;
  if constexpr (std::is_same<T, float>::value)
    hip_check(hflmp_memcpy_d2d(hip_ctx(), val_, h->values(), sizeof(T) * (size_t)this->nnz_),
              "CastFrom");
  else
    hip_check(hflmp_cast_f32_to_f64(hip_ctx(), this->nnz_, h->values(), val_), "CastFrom");
}

template <typename T>
lMatrix<T> *HIP_CSR_lMatrix<T>::CloneWithoutContent() const {
  auto *m = new HIP_CSR_lMatrix<T>();
  m->CopyStructureFrom(*this);
  return m;
}

template <typename T>
void HIP_CSR_lMatrix<T>::check_vectors(const lVector<T> &in, const lVector<T> *out,
                                       const char *op) const {
  if (!dynamic_cast<const HIP_lVector<T> *>(&in) || !dynamic_cast<const HIP_lVector<T> *>(out))
    LMP_FATAL("HIP_CSR_lMatrix::" << op << ": vectors must be HIP vectors");
  if (in.get_size() != this->num_col_ || out->get_size() != this->num_row_)
    LMP_FATAL("HIP_CSR_lMatrix::" << op << ": size mismatch (matrix " << this->num_row_ << "x"
                                  << this->num_col_ << ", in " << in.get_size() << ", out "
                                  << out->get_size() << ")");
}

template <typename T>
void HIP_CSR_lMatrix<T>::VectorMult(const lVector<T> &invec, lVector<T> *outvec) const {
  check_vectors(invec, outvec, "VectorMult");
  if (const hflmp_sell *s = sell_layout()) {
    hip_check(HipOps<T>::sell_spmv(hip_ctx(), s, invec.GetBuffer(), outvec->GetBuffer()),
              "VectorMult");
    return;
  }
  hip_check(HipOps<T>::spmv(hip_ctx(), this->num_row_, this->num_col_, this->nnz_, row_, col_,
                            val_, invec.GetBuffer(), outvec->GetBuffer()),
            "VectorMult");
}

template <typename T>
void HIP_CSR_lMatrix<T>::VectorMultAdd(const lVector<T> &invec, lVector<T> *outvec) const {
  check_vectors(invec, outvec, "VectorMultAdd");
  if (this->nnz_ == 0) return;  // empty off-diagonal block: nothing to add
  if (const hflmp_sell *s = sell_layout()) {
    hip_check(HipOps<T>::sell_spmv_add(hip_ctx(), s, invec.GetBuffer(), outvec->GetBuffer()),
              "VectorMultAdd");
    return;
  }
  hip_check(HipOps<T>::spmv_add(hip_ctx(), this->num_row_, this->num_col_, this->nnz_, row_, col_,
                                val_, invec.GetBuffer(), outvec->GetBuffer()),
            "VectorMultAdd");
}

template <typename T>
void HIP_CSR_lMatrix<T>::VectorMultNoDiag(const lVector<T> &invec, lVector<T> *outvec) const {
  check_vectors(invec, outvec, "VectorMultNoDiag");
  hip_check(HipOps<T>::spmv_nodiag(hip_ctx(), this->num_row_, this->num_col_, this->nnz_, row_,
                                   col_, val_, invec.GetBuffer(), outvec->GetBuffer()),
            "VectorMultNoDiag");
}

template <typename T>
void HIP_CSR_lMatrix<T>::extract_invdiagelements(const int start_i, const int end_i,
                                                 lVector<T> *vec) const {
  auto *h = dynamic_cast<HIP_lVector<T> *>(vec);
  if (!h) LMP_FATAL("extract_invdiagelements: target is not a HIP vector");
  if (start_i < 0 || end_i > this->num_row_ || start_i > end_i || h->get_size() < end_i - start_i)
    LMP_FATAL("extract_invdiagelements: bad range [" << start_i << "," << end_i << ")");
  int bad = 0;
  if (start_i == 0 && end_i == this->num_row_) {
    hip_check(HipOps<T>::invdiag(hip_ctx(), this->num_row_, row_, col_, val_, h->GetBuffer(), &bad),
              "extract_invdiagelements");
    return;
  }
  HIP_lVector<T> all(this->num_row_, "invdiag");
  hip_check(HipOps<T>::invdiag(hip_ctx(), this->num_row_, row_, col_, val_, all.GetBuffer(), &bad),
            "extract_invdiagelements");
  h->partial_replace_subvector(0, start_i, end_i - start_i, all);
}

template <typename T>
const hflmp_sgs *HIP_CSR_lMatrix<T>::sgs_schedule(int num_blocks) const {
  auto it = sgs_.find(num_blocks);
  if (it != sgs_.end()) return it->second;
  hflmp_sgs *m = nullptr;
  hip_check(HipOps<T>::sgs_create(hip_ctx(), this->num_row_, this->nnz_, row_, col_, val_,
                                  h_row_.data(), h_col_.data(), num_blocks, &m),
            "symmetric Gauss-Seidel schedule");
  sgs_[num_blocks] = m;
  return m;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Psgauss_seidel(const lVector<T> &invec, lVector<T> *outvec) const {
  BlocksPsgauss_seidel(invec, outvec, 1);
}

template <typename T>
void HIP_CSR_lMatrix<T>::BlocksPsgauss_seidel(const lVector<T> &invec, lVector<T> *outvec,
                                              const int num_blocks) const {
  check_vectors(invec, outvec, "BlocksPsgauss_seidel");
  if (this->num_row_ != this->num_col_)
    LMP_FATAL("HIP_CSR_lMatrix::BlocksPsgauss_seidel: the matrix is not square");
  if (num_blocks < 1) LMP_FATAL("HIP_CSR_lMatrix::BlocksPsgauss_seidel: num_blocks < 1");
  if (this->num_row_ == 0) return;
  hip_check(HipOps<T>::sgs_apply(hip_ctx(), sgs_schedule(num_blocks), invec.GetBuffer(),
                                 outvec->GetBuffer()),
            "BlocksPsgauss_seidel");
}

// Natural-order first fit: row j takes the smallest colour >= 1 that no
// stored neighbour of lower index holds (the colour the reference's repeated
// sweep arrives at, lmatrix_csr_cpu.cc:2195-2300).  permut_index[position]
// = row, rows grouped by colour and ascending inside a colour.
template <typename T>
void HIP_CSR_lMatrix<T>::Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const {
  const int n = this->num_row_;
  std::vector<int> color((size_t)n, 0), mark;
  int hicol = 0;
  for (int j = 0; j < n; ++j) {
    mark.assign((size_t)(h_row_[j + 1] - h_row_[j] + 2), 0);
    for (int p = h_row_[j]; p < h_row_[j + 1]; ++p) {
      const int d = h_col_[p];
      if (d < j) {
        const int c = color[d];
        if (c < (int)mark.size()) mark[c] = 1;
      }
    }
    int c = 1;
    while (c < (int)mark.size() && mark[c]) ++c;
    color[j] = c;
    hicol = std::max(hicol, c);
  }
  ncolors = hicol;
  *color_sizes = new int[std::max(hicol, 1)]();
  *permut_index = new int[std::max(n, 1)];
  for (int j = 0; j < n; ++j) ++(*color_sizes)[color[j] - 1];
  std::vector<int> start((size_t)hicol + 1, 0);
  for (int c = 1; c <= hicol; ++c) start[c] = start[c - 1] + (*color_sizes)[c - 1];
  for (int j = 0; j < n; ++j) (*permut_index)[start[color[j] - 1]++] = j;
}

template class HIP_CSR_lMatrix<double>;
template class HIP_CSR_lMatrix<float>;

}  // namespace la
}  // namespace hiflow
