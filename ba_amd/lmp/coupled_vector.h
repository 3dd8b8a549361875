// CoupledVector<DataType> on one process (linear_algebra/coupled_vector.h:46;
// coupled_vector.cc:122-395).  The reference couples an interior lVector
// with a ghost lVector and MPI; on a single device the ghost part is empty,
// the exchange calls are no-ops, and every operation is the interior
// lVector's.  Norm2 is sqrt(Dot(*this)) as in coupled_vector.cc:392-395.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "init_vec_mat.h"
#include "la_couplings.h"
#include "lmp_log.h"

namespace hiflow {
namespace la {

template <typename DataType>
class CoupledVector {
 public:
  CoupledVector() = default;
  CoupledVector(const CoupledVector &) = delete;
  CoupledVector &operator=(const CoupledVector &) = delete;

  void Init(const LaCouplings &cp, PLATFORM plat = HIP, IMPLEMENTATION impl = HIPKERNEL) {
    Init(cp.size_global(), plat, impl);
    cp_ = cp;
  }
  void Init(int size, PLATFORM plat = HIP, IMPLEMENTATION impl = HIPKERNEL) {
    platform_ = plat;
    impl_ = impl;
    cp_ = LaCouplings(size);
    interior_.reset(init_vector<DataType>(size, "interior", plat, impl));
    ghost_.reset(init_vector<DataType>(0, "ghost", plat, impl));
  }
  void Clear() {
    interior_.reset();
    ghost_.reset();
  }
  bool is_initialized() const { return (bool)interior_; }

  // structure (and platform) of vec, zero content
  void CloneFromWithoutContent(const CoupledVector<DataType> &vec) {
    Init(vec.cp_, vec.platform_, vec.impl_);
  }
  void CloneFrom(const CoupledVector<DataType> &vec) {
    CloneFromWithoutContent(vec);
    CopyFrom(vec);
  }
  void CopyStructureFrom(const CoupledVector<DataType> &vec) { CloneFromWithoutContent(vec); }
  void CopyFrom(const CoupledVector<DataType> &vec) {
    if (&vec != this) interior().CopyFrom(vec.interior());
  }
  template <typename Other>
  void CastInteriorFrom(const CoupledVector<Other> &vec) {
    interior().CastFrom(vec.interior());
  }
  template <typename Other>
  void CastInteriorTo(CoupledVector<Other> &vec) const {
    interior().CastTo(vec.interior());
  }

  void Zeros() { interior().Zeros(); }
  void Add(int global_id, DataType val) { interior().add_value(global_id, val); }
  void Add(const int *ids, int n, const DataType *vals) { interior().add_values(ids, n, vals); }
  void SetValue(int global_id, DataType val) { interior().SetValue(global_id, val); }
  void SetValues(const int *ids, int n, const DataType *vals) { interior().SetValues(ids, n, vals); }
  DataType GetValue(int global_id) const { return interior().GetValue(global_id); }
  void GetValues(const int *ids, int n, DataType *vals) const { interior().GetValues(ids, n, vals); }

  DataType Dot(const CoupledVector<DataType> &vec) const { return interior().Dot(vec.interior()); }
  DataType Norm1() const { return interior().Norm1(); }
  DataType Norm2() const { return std::sqrt(Dot(*this)); }
  DataType NormMax() const { return interior().NormMax(); }
  // this = this + alpha * vec
  void Axpy(const CoupledVector<DataType> &vec, DataType alpha) {
    interior().Axpy(vec.interior(), alpha);
  }
  // this = alpha * this + vec
  void ScaleAdd(const CoupledVector<DataType> &vec, DataType alpha) {
    interior().ScaleAdd(alpha, vec.interior());
  }
  void Scale(DataType alpha) { interior().Scale(alpha); }

  // ghost exchange: nothing to send or receive on one process
  void Update() {}
  void UpdateGhost() {}
  void SendBorder() {}
  void ReceiveGhost() {}
  void WaitForSend() {}
  void WaitForRecv() {}

  int size_local() const { return interior().get_size(); }
  int size_global() const { return cp_.size_global(); }
  int ownership_begin() const { return 0; }
  int ownership_end() const { return size_local(); }
  const LaCouplings &la_couplings() const { return cp_; }
  PLATFORM platform() const { return platform_; }

  lVector<DataType> &interior() {
    if (!interior_) LMP_FATAL("CoupledVector: not initialized");
    return *interior_;
  }
  const lVector<DataType> &interior() const {
    if (!interior_) LMP_FATAL("CoupledVector: not initialized");
    return *interior_;
  }
  lVector<DataType> &ghost() { return *ghost_; }
  const lVector<DataType> &ghost() const { return *ghost_; }

 private:
  LaCouplings cp_;
  PLATFORM platform_ = HIP;
  IMPLEMENTATION impl_ = HIPKERNEL;
  std::unique_ptr<lVector<DataType>> interior_, ghost_;
};

}  // namespace la
}  // namespace hiflow
