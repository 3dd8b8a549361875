// Row ownership of a distributed system (linear_algebra/la_couplings.h:
// 65-198).  On one process every row is owned locally: the global offsets
// are {0, n} and there are no off-diagonal (ghost) columns.
#pragma once

#include <vector>

namespace hiflow {
namespace la {

class LaCouplings {
 public:
  LaCouplings() = default;
  explicit LaCouplings(int n) : global_offsets_{0, n} {}
  // single-process form of InitializeCouplings(global_offsets,
  // offdiag_cols, offdiag_offsets): offdiag_cols must be empty
  void InitializeCouplings(const std::vector<int> &global_offsets,
                           const std::vector<int> &offdiag_cols,
                           const std::vector<int> & /*offdiag_offsets*/) {
    global_offsets_ = global_offsets;
    num_offdiag_ = (int)offdiag_cols.size();
  }
  int size_global() const { return global_offsets_.empty() ? 0 : global_offsets_.back(); }
  int dof_offset(int rank) const { return global_offsets_[rank]; }
  int nb_dofs(int rank) const { return global_offsets_[rank + 1] - global_offsets_[rank]; }
  int size_ghost() const { return num_offdiag_; }
  const std::vector<int> &global_offsets() const { return global_offsets_; }

 private:
  std::vector<int> global_offsets_{0, 0};
  int num_offdiag_ = 0;
};

}  // namespace la
}  // namespace hiflow
