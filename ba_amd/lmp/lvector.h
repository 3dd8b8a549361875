// lVector<ValueType>: the local vector interface the Coupled layer and the
// Krylov solvers program against (linear_algebra/lmp/lvector.h:38-321).
// Method names and argument meanings follow that interface so code written
// for it compiles unchanged; the set is the part exercised by
// CoupledVector, CoupledMatrix, CG, GMRES, FGMRES, MPIR and the Jacobi and
// multicoloured Gauss-Seidel preconditioners.  Host-side arrays passed to
// the Set/Get/Add methods are host memory, as in the reference.
#pragma once

#include <iostream>
#include <string>
#include <vector>

#include "la_global.h"
#include "lmp_log.h"
#include "lmp_mmio.h"

namespace hiflow {
namespace la {

template <typename ValueType>
class lVector {
 public:
  lVector() = default;
  virtual ~lVector() = default;
  lVector(const lVector &) = delete;
  lVector &operator=(const lVector &) = delete;

  int get_size() const { return size_; }
  std::string get_name() const { return name_; }
  PLATFORM get_platform() const { return platform_id_; }
  IMPLEMENTATION get_implementation() const { return implementation_id_; }
  std::string get_platform_name() const { return platform_name(platform_id_); }
  std::string get_implementation_name() const { return implementation_name(implementation_id_); }
  void print(std::ostream &out = std::cout) const;

  // allocation
  virtual void Init(const int size, const std::string name) = 0;
  virtual void Clear() = 0;
  virtual void Zeros() = 0;
  // plain text, one value per line (lvector_cpu.cc:496-565): read sizes
  // the vector to the values in the file; write stores 17 significant
  // digits, so a round trip is exact
  virtual void ReadFile(const char *filename);
  virtual void WriteFile(const char *filename);

  // element access (host arrays)
  virtual void add_value(const int i, ValueType val) = 0;
  virtual void add_values(const int *indices, int length, const ValueType *values) = 0;
  virtual void SetValues(const int *index, const int size, const ValueType *values) = 0;
  virtual void SetValue(const int index, const ValueType value) = 0;
  virtual void GetValues(const int *index, const int size, ValueType *values) const = 0;
  virtual ValueType GetValue(const int index) const = 0;
  virtual void SetBlockValues(const int start_i, const int end_i, const ValueType *values) = 0;
  virtual void GetBlockValues(const int start_i, const int end_i, ValueType *values) const = 0;

  // index set for ghost exchange (set_indexset / GetIndexedValues)
  virtual void set_indexset(const int *indexset, const int size) = 0;
  virtual int get_indexset_size() const = 0;
  virtual void GetIndexedValues(ValueType *values) const = 0;
  virtual void SetIndexedValues(const ValueType *values) = 0;

  // this[start_i + i] = sub_vec[start_sub_vec + i], i < size
  virtual void partial_replace_subvector(const int start_i, const int start_sub_vec,
                                         const int size, const lVector<ValueType> &sub_vec) = 0;
  // this[start_i + i] += weight * sub_vec[start_sub_vec + i]
  virtual void partial_add_subvector(const int start_i, const int start_sub_vec, const int size,
                                     const ValueType weight,
                                     const lVector<ValueType> &sub_vec) = 0;

  // copies within and across vectors of this platform
  virtual void CopyFrom(const lVector<ValueType> &vec) = 0;
  virtual void CopyTo(lVector<ValueType> &vec) const = 0;
  virtual void CopyStructureFrom(const lVector<ValueType> &vec);
  virtual void CastFrom(const lVector<double> &vec) = 0;
  virtual void CastFrom(const lVector<float> &vec) = 0;
  virtual void CastTo(lVector<double> &vec) const = 0;
  virtual void CastTo(lVector<float> &vec) const = 0;
  // Clear, Init, CopyStructureFrom, CopyFrom (lvector.h:244-246)
  virtual void CloneFrom(const lVector<ValueType> &vec);
  // a new vector of the same class and size, zero filled (lvector.cc:47-57)
  virtual lVector<ValueType> *CloneWithoutContent() const = 0;
  virtual void Sync() const {}

  // BLAS-1
  virtual void ElementWiseMult(const lVector<ValueType> &vec) const = 0;
  virtual int ArgMin() const = 0;
  virtual int ArgMax() const = 0;
  virtual ValueType Norm1() const = 0;
  virtual ValueType Norm2() const = 0;
  virtual ValueType NormMax() const = 0;
  virtual ValueType Dot(const lVector<ValueType> &vec) const = 0;
  virtual void Scale(const ValueType scalar) = 0;
  // this = this + scalar * vec
  virtual void Axpy(const lVector<ValueType> &vec, const ValueType scalar) = 0;
  // this = scalar * this + vec
  virtual void ScaleAdd(const ValueType scalar, const lVector<ValueType> &vec) = 0;
  // x' = c x + s y, y' = c y - s x with x = this, y = vec
  virtual void Rot(lVector<ValueType> *vec, const ValueType &sc, const ValueType &ss) = 0;
  virtual bool ContainsNaN() const = 0;

  // raw storage (device memory for the HIP platform)
  virtual ValueType *GetBuffer() const { return nullptr; }

 protected:
  int size_ = 0;
  std::string name_;
  IMPLEMENTATION implementation_id_ = HIPKERNEL;
  PLATFORM platform_id_ = HIP;
};

template <typename ValueType>
void lVector<ValueType>::ReadFile(const char *filename) {
  std::vector<double> v;
  const std::string err = vec_read(filename, v);
  if (!err.empty()) LMP_FATAL("lVector::ReadFile(" << filename << "): " << err);
  std::vector<ValueType> t(v.begin(), v.end());
  this->Clear();
  this->Init((int)t.size(), filename);
  if (!t.empty()) this->SetBlockValues(0, (int)t.size(), t.data());
}

template <typename ValueType>
void lVector<ValueType>::WriteFile(const char *filename) {
  std::vector<ValueType> t((size_t)size_);
  if (size_ > 0) this->GetBlockValues(0, size_, t.data());
  std::vector<double> v(t.begin(), t.end());
  const std::string err = vec_write(filename, v.data(), (int)v.size());
  if (!err.empty()) LMP_FATAL("lVector::WriteFile(" << filename << "): " << err);
}

template <typename ValueType>
void lVector<ValueType>::print(std::ostream &out) const {
  out << "lVector name=" << name_ << " size=" << size_ << " precision=" << 8 * sizeof(ValueType)
      << "bit platform=" << get_platform_name() << " implementation="
      << get_implementation_name() << std::endl;
}

template <typename ValueType>
void lVector<ValueType>::CopyStructureFrom(const lVector<ValueType> &vec) {
  if (this->get_size() != vec.get_size()) {
    this->Clear();
    this->Init(vec.get_size(), vec.get_name());
  }
}

template <typename ValueType>
void lVector<ValueType>::CloneFrom(const lVector<ValueType> &vec) {
  this->Clear();
  this->Init(vec.get_size(), vec.get_name());
  this->CopyStructureFrom(vec);
  this->CopyFrom(vec);
}

}  // namespace la
}  // namespace hiflow
