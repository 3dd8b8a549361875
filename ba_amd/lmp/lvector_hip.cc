// HIP_lVector: device vectors of the HIP platform.  The CUDA platform's
// GPU_lVector / GPUblas_lVector (cuda/lvector_gpu.cc, cuda/GPUblas_lvector.cu)
// have the same responsibilities: host<->device transfers for the
// Set/Get methods, device-to-device copies and casts, and BLAS-1.
#include "lvector_hip.h"

#include <type_traits>
#include <vector>

#include "hip_ops.h"
#include "lmp_log.h"
#include "platform_management.h"

namespace hiflow {
namespace la {

namespace {

// Device scratch for index/value lists handed over from the host.
template <typename V>
struct DevTemp {
  V *p = nullptr;
  explicit DevTemp(size_t n, const V *host = nullptr) {
    if (!n) return;
    hip_check(hflmp_malloc(hip_ctx(), n * sizeof(V), reinterpret_cast<void **>(&p)),
              "hflmp_malloc");
    if (host) hip_check(hflmp_memcpy_h2d(hip_ctx(), p, host, n * sizeof(V)), "memcpy_h2d");
  }
  ~DevTemp() {
    if (p) hflmp_free(hip_ctx(), p);
  }
  DevTemp(const DevTemp &) = delete;
  DevTemp &operator=(const DevTemp &) = delete;
};

}  // namespace

template <typename T>
HIP_lVector<T>::HIP_lVector() {
  this->platform_id_ = HIP;
  this->implementation_id_ = HIPKERNEL;
}

template <typename T>
HIP_lVector<T>::HIP_lVector(const int size, const std::string name) : HIP_lVector() {
  Init(size, name);
}

template <typename T>
HIP_lVector<T>::~HIP_lVector() {
  Clear();
}

template <typename T>
void HIP_lVector<T>::Init(const int size, const std::string name) {
  if (size < 0) LMP_FATAL("HIP_lVector::Init: negative size " << size);
  Clear();
  this->size_ = size;
  this->name_ = name;
  if (size > 0) {
    hip_check(hflmp_malloc(hip_ctx(), sizeof(T) * (size_t)size, reinterpret_cast<void **>(&buffer_)),
              "HIP_lVector::Init");
    Zeros();
  }
}

template <typename T>
void HIP_lVector<T>::Clear() {
  if (buffer_) hflmp_free(hip_ctx(), buffer_);
  if (index_) hflmp_free(hip_ctx(), index_);
  buffer_ = nullptr;
  index_ = nullptr;
  index_size_ = 0;
  this->size_ = 0;
}

template <typename T>
void HIP_lVector<T>::Zeros() {
  if (this->size_ > 0)
    hip_check(hflmp_fill_zero(hip_ctx(), buffer_, sizeof(T) * (size_t)this->size_),
              "HIP_lVector::Zeros");
}

template <typename T>
const HIP_lVector<T> &HIP_lVector<T>::peer(const lVector<T> &v, const char *op) const {
  const auto *h = dynamic_cast<const HIP_lVector<T> *>(&v);
  if (!h)
    LMP_FATAL("HIP_lVector::" << op << ": the other vector is on platform "
                              << v.get_platform_name() << ", not HIP");
  if (h->get_size() != this->get_size())
    LMP_FATAL("HIP_lVector::" << op << ": sizes differ (" << this->get_size() << " vs "
                              << h->get_size() << ")");
  return *h;
}

template <typename T>
void HIP_lVector<T>::add_value(const int i, T val) {
  add_values(&i, 1, &val);
}

template <typename T>
void HIP_lVector<T>::add_values(const int *indices, int length, const T *values) {
  if (length <= 0) return;
  for (int k = 0; k < length; ++k)
    if (indices[k] < 0 || indices[k] >= this->size_)
      LMP_FATAL("HIP_lVector::add_values: index " << indices[k] << " out of range");
  DevTemp<int> di(length, indices);
  DevTemp<T> dv(length, values);
  hip_check(HipOps<T>::scatter_add(hip_ctx(), length, di.p, dv.p, buffer_), "add_values");
}

template <typename T>
void HIP_lVector<T>::SetValues(const int *index, const int size, const T *values) {
  if (size <= 0) return;
  for (int k = 0; k < size; ++k)
    if (index[k] < 0 || index[k] >= this->size_)
      LMP_FATAL("HIP_lVector::SetValues: index " << index[k] << " out of range");
  DevTemp<int> di(size, index);
  DevTemp<T> dv(size, values);
  hip_check(HipOps<T>::scatter(hip_ctx(), size, di.p, dv.p, buffer_), "SetValues");
}

template <typename T>
void HIP_lVector<T>::SetValue(const int index, const T value) {
  if (index < 0 || index >= this->size_) LMP_FATAL("HIP_lVector::SetValue: index out of range");
  hip_check(hflmp_memcpy_h2d(hip_ctx(), buffer_ + index, &value, sizeof(T)), "SetValue");
}

template <typename T>
void HIP_lVector<T>::GetValues(const int *index, const int size, T *values) const {
  if (size <= 0) return;
  for (int k = 0; k < size; ++k)
    if (index[k] < 0 || index[k] >= this->size_)
      LMP_FATAL("HIP_lVector::GetValues: index " << index[k] << " out of range");
  DevTemp<int> di(size, index);
  DevTemp<T> dv(size);
  hip_check(HipOps<T>::gather(hip_ctx(), size, di.p, buffer_, dv.p), "GetValues");
  hip_check(hflmp_memcpy_d2h(hip_ctx(), values, dv.p, sizeof(T) * (size_t)size), "memcpy_d2h");
}

template <typename T>
T HIP_lVector<T>::GetValue(const int index) const {
  if (index < 0 || index >= this->size_) LMP_FATAL("HIP_lVector::GetValue: index out of range");
  T v;
  hip_check(hflmp_memcpy_d2h(hip_ctx(), &v, buffer_ + index, sizeof(T)), "GetValue");
  return v;
}

template <typename T>
void HIP_lVector<T>::SetBlockValues(const int start_i, const int end_i, const T *values) {
  if (start_i < 0 || end_i > this->size_ || start_i > end_i)
    LMP_FATAL("HIP_lVector::SetBlockValues: bad range [" << start_i << "," << end_i << ")");
  if (end_i > start_i)
    hip_check(hflmp_memcpy_h2d(hip_ctx(), buffer_ + start_i, values,
                               sizeof(T) * (size_t)(end_i - start_i)),
              "SetBlockValues");
}

template <typename T>
void HIP_lVector<T>::GetBlockValues(const int start_i, const int end_i, T *values) const {
  if (start_i < 0 || end_i > this->size_ || start_i > end_i)
    LMP_FATAL("HIP_lVector::GetBlockValues: bad range [" << start_i << "," << end_i << ")");
  if (end_i > start_i)
    hip_check(hflmp_memcpy_d2h(hip_ctx(), values, buffer_ + start_i,
                               sizeof(T) * (size_t)(end_i - start_i)),
              "GetBlockValues");
}

template <typename T>
void HIP_lVector<T>::set_indexset(const int *indexset, const int size) {
  if (index_) hflmp_free(hip_ctx(), index_);
  index_ = nullptr;
  index_size_ = size;
  if (size <= 0) return;
  for (int k = 0; k < size; ++k)
    if (indexset[k] < 0 || indexset[k] >= this->size_)
      LMP_FATAL("HIP_lVector::set_indexset: index " << indexset[k] << " out of range");
  hip_check(hflmp_malloc(hip_ctx(), sizeof(int) * (size_t)size, reinterpret_cast<void **>(&index_)),
            "set_indexset");
  hip_check(hflmp_memcpy_h2d(hip_ctx(), index_, indexset, sizeof(int) * (size_t)size),
            "set_indexset");
}

template <typename T>
void HIP_lVector<T>::GetIndexedValues(T *values) const {
  if (index_size_ <= 0) return;
  DevTemp<T> dv(index_size_);
  hip_check(HipOps<T>::gather(hip_ctx(), index_size_, index_, buffer_, dv.p), "GetIndexedValues");
  hip_check(hflmp_memcpy_d2h(hip_ctx(), values, dv.p, sizeof(T) * (size_t)index_size_),
            "GetIndexedValues");
}

template <typename T>
void HIP_lVector<T>::SetIndexedValues(const T *values) {
  if (index_size_ <= 0) return;
  DevTemp<T> dv(index_size_, values);
  hip_check(HipOps<T>::scatter(hip_ctx(), index_size_, index_, dv.p, buffer_), "SetIndexedValues");
}

template <typename T>
void HIP_lVector<T>::partial_replace_subvector(const int start_i, const int start_sub_vec,
                                               const int size, const lVector<T> &sub_vec) {
  const auto *h = dynamic_cast<const HIP_lVector<T> *>(&sub_vec);
  if (!h) LMP_FATAL("partial_replace_subvector: the sub vector is not a HIP vector");
  if (size < 0 || start_i < 0 || start_sub_vec < 0 || start_i + size > this->size_ ||
      start_sub_vec + size > h->get_size())
    LMP_FATAL("partial_replace_subvector: range out of bounds");
  if (size)
    hip_check(HipOps<T>::copy(hip_ctx(), size, h->buffer_ + start_sub_vec, buffer_ + start_i),
              "partial_replace_subvector");
}

template <typename T>
void HIP_lVector<T>::partial_add_subvector(const int start_i, const int start_sub_vec,
                                           const int size, const T weight,
                                           const lVector<T> &sub_vec) {
  const auto *h = dynamic_cast<const HIP_lVector<T> *>(&sub_vec);
  if (!h) LMP_FATAL("partial_add_subvector: the sub vector is not a HIP vector");
  if (size < 0 || start_i < 0 || start_sub_vec < 0 || start_i + size > this->size_ ||
      start_sub_vec + size > h->get_size())
    LMP_FATAL("partial_add_subvector: range out of bounds");
  if (size)
    hip_check(HipOps<T>::axpy(hip_ctx(), size, weight, h->buffer_ + start_sub_vec,
                              buffer_ + start_i),
              "partial_add_subvector");
}

template <typename T>
void HIP_lVector<T>::CopyFrom(const lVector<T> &vec) {
  if (&vec == this) return;
  const auto &h = peer(vec, "CopyFrom");
  if (this->size_) hip_check(HipOps<T>::copy(hip_ctx(), this->size_, h.buffer_, buffer_), "CopyFrom");
}

template <typename T>
void HIP_lVector<T>::CopyTo(lVector<T> &vec) const {
  vec.CopyFrom(*this);
}

template <typename T>
void HIP_lVector<T>::CastFrom(const lVector<double> &vec) {
  const auto *h = dynamic_cast<const HIP_lVector<double> *>(&vec);
  if (!h || h->get_size() != this->size_) LMP_FATAL("HIP_lVector::CastFrom: incompatible vector");
  if (!this->size_) return;
  if constexpr (std::is_same<T, double>::value)
    hip_check(hflmp_copy_f64(hip_ctx(), this->size_, h->GetBuffer(), buffer_), "CastFrom");
  else
    hip_check(hflmp_cast_f64_to_f32(hip_ctx(), this->size_, h->GetBuffer(), buffer_), "CastFrom");
}

template <typename T>
void HIP_lVector<T>::CastFrom(const lVector<float> &vec) {
  const auto *h = dynamic_cast<const HIP_lVector<float> *>(&vec);
  if (!h || h->get_size() != this->size_) LMP_FATAL("HIP_lVector::CastFrom: incompatible vector");
  if (!this->size_) return;
  if constexpr (std::is_same<T, float>::value)
    hip_check(hflmp_copy_f32(hip_ctx(), this->size_, h->GetBuffer(), buffer_), "CastFrom");
  else
    hip_check(hflmp_cast_f32_to_f64(hip_ctx(), this->size_, h->GetBuffer(), buffer_), "CastFrom");
}

template <typename T>
void HIP_lVector<T>::CastTo(lVector<double> &vec) const {
  vec.CastFrom(*this);
}

template <typename T>
void HIP_lVector<T>::CastTo(lVector<float> &vec) const {
  vec.CastFrom(*this);
}

template <typename T>
lVector<T> *HIP_lVector<T>::CloneWithoutContent() const {
  return new HIP_lVector<T>(this->size_, this->name_);
}

template <typename T>
void HIP_lVector<T>::Sync() const {
  hip_check(hflmp_ctx_sync(hip_ctx()), "Sync");
}

template <typename T>
void HIP_lVector<T>::ElementWiseMult(const lVector<T> &vec) const {
  const auto &h = peer(vec, "ElementWiseMult");
  if (this->size_)
    hip_check(HipOps<T>::ewmul(hip_ctx(), this->size_, h.buffer_, buffer_), "ElementWiseMult");
}

template <typename T>
int HIP_lVector<T>::ArgMin() const {
  int i = -1;
  hip_check(HipOps<T>::argmin(hip_ctx(), this->size_, buffer_, &i), "ArgMin");
  return i;
}

template <typename T>
int HIP_lVector<T>::ArgMax() const {
  int i = -1;
  hip_check(HipOps<T>::argmax(hip_ctx(), this->size_, buffer_, &i), "ArgMax");
  return i;
}

template <typename T>
T HIP_lVector<T>::Norm1() const {
  T r = 0;
  hip_check(HipOps<T>::norm1(hip_ctx(), this->size_, buffer_, &r), "Norm1");
  return r;
}

template <typename T>
T HIP_lVector<T>::Norm2() const {
  T r = 0;
  hip_check(HipOps<T>::nrm2(hip_ctx(), this->size_, buffer_, &r), "Norm2");
  return r;
}

template <typename T>
T HIP_lVector<T>::NormMax() const {
  T r = 0;
  hip_check(HipOps<T>::normmax(hip_ctx(), this->size_, buffer_, &r), "NormMax");
  return r;
}

template <typename T>
T HIP_lVector<T>::Dot(const lVector<T> &vec) const {
  const auto &h = peer(vec, "Dot");
  T r = 0;
  hip_check(HipOps<T>::dot(hip_ctx(), this->size_, buffer_, h.buffer_, &r), "Dot");
  return r;
}

template <typename T>
void HIP_lVector<T>::Scale(const T scalar) {
  if (this->size_) hip_check(HipOps<T>::scal(hip_ctx(), this->size_, scalar, buffer_), "Scale");
}

template <typename T>
void HIP_lVector<T>::Axpy(const lVector<T> &vec, const T scalar) {
  const auto &h = peer(vec, "Axpy");
  if (this->size_)
    hip_check(HipOps<T>::axpy(hip_ctx(), this->size_, scalar, h.buffer_, buffer_), "Axpy");
}

template <typename T>
void HIP_lVector<T>::ScaleAdd(const T scalar, const lVector<T> &vec) {
  const auto &h = peer(vec, "ScaleAdd");
  if (this->size_)
    hip_check(HipOps<T>::scaleadd(hip_ctx(), this->size_, scalar, h.buffer_, buffer_), "ScaleAdd");
}

template <typename T>
void HIP_lVector<T>::Rot(lVector<T> *vec, const T &sc, const T &ss) {
  auto *h = dynamic_cast<HIP_lVector<T> *>(vec);
  if (!h || h->get_size() != this->size_) LMP_FATAL("HIP_lVector::Rot: incompatible vector");
  if (this->size_) hip_check(HipOps<T>::rot(hip_ctx(), this->size_, buffer_, h->buffer_, sc, ss), "Rot");
}

template <typename T>
bool HIP_lVector<T>::ContainsNaN() const {
  int k = 0;
  if (this->size_) hip_check(HipOps<T>::count_nan(hip_ctx(), this->size_, buffer_, &k), "ContainsNaN");
  return k > 0;
}

template class HIP_lVector<double>;
template class HIP_lVector<float>;

}  // namespace la
}  // namespace hiflow
