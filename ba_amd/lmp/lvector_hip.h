// HIP_lVector<ValueType>: a device vector of the HIP platform.  Storage is
// one device buffer (the role of GPU_lVector::buffer, cuda/lvector_gpu.h:118);
// every operation is one call into libhflmp_hip (include/hflmp_hip.h).
// Scalar results (Dot, Norm*) are returned by value, which synchronises the
// context's stream as lVector::Dot does on the CUDA platform.
#pragma once

#include "hflmp_hip.h"
#include "lvector.h"

namespace hiflow {
namespace la {

template <typename ValueType>
class HIP_lVector : public lVector<ValueType> {
 public:
  HIP_lVector();
  HIP_lVector(const int size, const std::string name);
  ~HIP_lVector() override;

  void Init(const int size, const std::string name) override;
  void Clear() override;
  void Zeros() override;

  void add_value(const int i, ValueType val) override;
  void add_values(const int *indices, int length, const ValueType *values) override;
  void SetValues(const int *index, const int size, const ValueType *values) override;
  void SetValue(const int index, const ValueType value) override;
  void GetValues(const int *index, const int size, ValueType *values) const override;
  ValueType GetValue(const int index) const override;
  void SetBlockValues(const int start_i, const int end_i, const ValueType *values) override;
  void GetBlockValues(const int start_i, const int end_i, ValueType *values) const override;

  void set_indexset(const int *indexset, const int size) override;
  int get_indexset_size() const override { return index_size_; }
  void GetIndexedValues(ValueType *values) const override;
  void SetIndexedValues(const ValueType *values) override;

  void partial_replace_subvector(const int start_i, const int start_sub_vec, const int size,
                                 const lVector<ValueType> &sub_vec) override;
  void partial_add_subvector(const int start_i, const int start_sub_vec, const int size,
                             const ValueType weight, const lVector<ValueType> &sub_vec) override;

  void CopyFrom(const lVector<ValueType> &vec) override;
  void CopyTo(lVector<ValueType> &vec) const override;
  void CastFrom(const lVector<double> &vec) override;
  void CastFrom(const lVector<float> &vec) override;
  void CastTo(lVector<double> &vec) const override;
  void CastTo(lVector<float> &vec) const override;
  lVector<ValueType> *CloneWithoutContent() const override;
  void Sync() const override;

  void ElementWiseMult(const lVector<ValueType> &vec) const override;
  int ArgMin() const override;
  int ArgMax() const override;
  ValueType Norm1() const override;
  ValueType Norm2() const override;
  ValueType NormMax() const override;
  ValueType Dot(const lVector<ValueType> &vec) const override;
  void Scale(const ValueType scalar) override;
  void Axpy(const lVector<ValueType> &vec, const ValueType scalar) override;
  void ScaleAdd(const ValueType scalar, const lVector<ValueType> &vec) override;
  void Rot(lVector<ValueType> *vec, const ValueType &sc, const ValueType &ss) override;
  bool ContainsNaN() const override;

  ValueType *GetBuffer() const override { return buffer_; }

 private:
  const HIP_lVector<ValueType> &peer(const lVector<ValueType> &v, const char *op) const;
  ValueType *buffer_ = nullptr;  // device
  int *index_ = nullptr;         // device index set
  int index_size_ = 0;
};

}  // namespace la
}  // namespace hiflow
