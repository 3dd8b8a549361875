#include "lpreconditioner.h"

#include <type_traits>
#include <vector>

#include "hip_ops.h"
#include "lmatrix_csr_hip.h"
#include "lmp_log.h"
#include "lvector_hip.h"
#include "platform_management.h"

namespace hiflow {
namespace la {

template <typename T>
lPreconditioner_Jacobi<T>::lPreconditioner_Jacobi() {
  this->precond_name_ = "Jacobi";
}

template <typename T>
lPreconditioner_Jacobi<T>::~lPreconditioner_Jacobi() = default;

template <typename T>
void lPreconditioner_Jacobi<T>::Build() {
  if (!this->Operator_) LMP_FATAL("lPreconditioner_Jacobi::Build: no operator");
  const int n = this->Operator_->get_num_row();
  inv_D_.reset(new HIP_lVector<T>(n, "inv_D"));
  this->Operator_->extract_invdiagelements(0, n, inv_D_.get());
  this->SetState(true);
}

template <typename T>
void lPreconditioner_Jacobi<T>::ApplylPreconditioner(const lVector<T> &input, lVector<T> *output) {
  if (!this->GetState()) Build();
  if (input.get_size() != output->get_size() || output->get_size() != inv_D_->get_size())
    LMP_FATAL("lPreconditioner_Jacobi::Apply: size mismatch");
  if (input.get_size() == 0) return;
  hip_check(HipOps<T>::jacobi(hip_ctx(), input.get_size(), inv_D_->GetBuffer(),
                              input.GetBuffer(), output->GetBuffer()),
            "Jacobi apply");
}

template <typename T>
void lPreconditioner_Jacobi<T>::Clear() {
  inv_D_.reset();
  lPreconditioner<T>::Clear();
}

template <typename T>
lPreconditioner_SymmetricGaussSeidel<T>::lPreconditioner_SymmetricGaussSeidel() {
  this->precond_name_ = "Symmetric Gauss-Seidel";
}

template <typename T>
lPreconditioner_SymmetricGaussSeidel<T>::~lPreconditioner_SymmetricGaussSeidel() = default;

template <typename T>
void lPreconditioner_SymmetricGaussSeidel<T>::Build() {
  if (!this->Operator_) LMP_FATAL("lPreconditioner_SymmetricGaussSeidel::Build: no operator");
  this->SetState(true);
}

template <typename T>
void lPreconditioner_SymmetricGaussSeidel<T>::ApplylPreconditioner(const lVector<T> &input,
                                                                   lVector<T> *output) {
  if (!this->GetState()) Build();
  this->Operator_->Psgauss_seidel(input, output);
}

template <typename T>
lPreconditioner_BlocksSymmetricGaussSeidel<T>::lPreconditioner_BlocksSymmetricGaussSeidel() {
  this->precond_name_ = "Block-wise Symmetric Gauss-Seidel";
}

template <typename T>
lPreconditioner_BlocksSymmetricGaussSeidel<T>::~lPreconditioner_BlocksSymmetricGaussSeidel() =
    default;

template <typename T>
void lPreconditioner_BlocksSymmetricGaussSeidel<T>::Init(const int num_blocks) {
  if (num_blocks < 1) LMP_FATAL("lPreconditioner_BlocksSymmetricGaussSeidel: num_blocks < 1");
  num_blocks_ = num_blocks;
}

template <typename T>
void lPreconditioner_BlocksSymmetricGaussSeidel<T>::Build() {
  if (!this->Operator_)
    LMP_FATAL("lPreconditioner_BlocksSymmetricGaussSeidel::Build: no operator");
  this->SetState(true);
}

template <typename T>
void lPreconditioner_BlocksSymmetricGaussSeidel<T>::ApplylPreconditioner(
    const lVector<T> &input, lVector<T> *output) {
  if (!this->GetState()) Build();
  this->Operator_->BlocksPsgauss_seidel(input, output, num_blocks_);
}

template <typename T>
lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>::
    lPreconditioner_MultiColoring_SymmetricGaussSeidel() {
  this->precond_name_ = "Multi-Coloring / Symmetric Gauss-Seidel";
}

template <typename T>
lPreconditioner_MultiColoring_SymmetricGaussSeidel<
    T>::~lPreconditioner_MultiColoring_SymmetricGaussSeidel() {
  Clear();
}

template <typename T>
void lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>::SetRelax(T omega_relax) {
  if (!(omega_relax > T(0) && omega_relax < T(2)))
    LMP_FATAL("MC-SGS: relaxation must lie in (0, 2)");
  omega_ = omega_relax;
  this->SetState(false);
}

template <typename T>
void lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>::Build() {
  if constexpr (!std::is_same<T, double>::value) {
    LMP_FATAL("MC-SGS on the HIP platform is provided for double only");
  } else {
    const auto *A = dynamic_cast<const HIP_CSR_lMatrix<double> *>(this->Operator_);
    if (!A) LMP_FATAL("MC-SGS: the operator must be a HIP CSR matrix");
    if (mc_) hflmp_mcsgs_destroy(hip_ctx(), mc_);
    mc_ = nullptr;
    int *sizes = nullptr, *perm = nullptr;
    A->Multicoloring(ncolors_, &sizes, &perm);
    const int rc = hflmp_mcsgs_create_f64(hip_ctx(), A->get_num_row(), A->get_nnz(),
                                          A->row_ptr(), A->col_ind(), A->values(), ncolors_,
                                          sizes, perm, (double)omega_, &mc_);
    delete[] sizes;
    delete[] perm;
    hip_check(rc, "MC-SGS build");
    this->SetState(true);
  }
}

template <typename T>
void lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>::ApplylPreconditioner(
    const lVector<T> &input, lVector<T> *output) {
  if constexpr (!std::is_same<T, double>::value) {
    LMP_FATAL("MC-SGS on the HIP platform is provided for double only");
  } else {
    if (!this->GetState()) Build();
    if (input.get_size() != output->get_size() || input.GetBuffer() == output->GetBuffer())
      LMP_FATAL("MC-SGS apply: sizes differ or input aliases output");
    hip_check(hflmp_mcsgs_apply_f64(hip_ctx(), mc_, input.GetBuffer(), output->GetBuffer()),
              "MC-SGS apply");
  }
}

template <typename T>
void lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>::Clear() {
  if (mc_) hflmp_mcsgs_destroy(hip_ctx(), mc_);
  mc_ = nullptr;
  ncolors_ = 0;
  lPreconditioner<T>::Clear();
}

template class lPreconditioner_Jacobi<double>;
template class lPreconditioner_Jacobi<float>;
template class lPreconditioner_SymmetricGaussSeidel<double>;
template class lPreconditioner_SymmetricGaussSeidel<float>;
template class lPreconditioner_BlocksSymmetricGaussSeidel<double>;
template class lPreconditioner_BlocksSymmetricGaussSeidel<float>;
template class lPreconditioner_MultiColoring_SymmetricGaussSeidel<double>;
template class lPreconditioner_MultiColoring_SymmetricGaussSeidel<float>;

}  // namespace la
}  // namespace hiflow
