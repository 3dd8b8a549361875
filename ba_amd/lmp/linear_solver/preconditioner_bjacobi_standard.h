// PreconditionerBlockJacobiStand<LAD> (linear_solver/
// preconditioner_bjacobi_standard.cc:41-196): a local lPreconditioner on
// each process's diagonal block.  Jacobi; symmetric Gauss-Seidel in the
// sequential row order (Init_SymmetricGaussSeidel, :95-101, the operator's
// Psgauss_seidel, level-scheduled on the device with the CPU sweep's
// results); its block form (the lPreconditioner_BlocksSymmetricGaussSeidel
// of lpreconditioner.h:179-200, which the reference's adapter has no Init
// for); and the multicoloured SGS of lpreconditioner_mc_sgs, a different
// preconditioner (colour-permuted sweeps) under its own name.
#pragma once

#include <memory>

#include "../lpreconditioner.h"
#include "preconditioner.h"

namespace hiflow {
namespace la {

template <class LAD>
class PreconditionerBlockJacobiStand : public Preconditioner<LAD> {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  void Init_Jacobi() {
    local_.reset(new lPreconditioner_Jacobi<DataType>());
    this->state_ = false;
  }
  void Init_SymmetricGaussSeidel() {
    local_.reset(new lPreconditioner_SymmetricGaussSeidel<DataType>());
    this->state_ = false;
  }
  void Init_BlocksSymmetricGaussSeidel(int num_blocks) {
    auto *p = new lPreconditioner_BlocksSymmetricGaussSeidel<DataType>();
    p->Init(num_blocks);
    local_.reset(p);
    this->state_ = false;
  }
  void Init_MultiColoring_SymmetricGaussSeidel(DataType omega) {
    auto *p = new lPreconditioner_MultiColoring_SymmetricGaussSeidel<DataType>();
    p->SetRelax(omega);
    local_.reset(p);
    this->state_ = false;
  }
  lPreconditioner<DataType> *local() { return local_.get(); }

 protected:
  void BuildImpl() override {
    if (!local_) LMP_FATAL("PreconditionerBlockJacobiStand: no local preconditioner chosen");
    if (!this->op_) LMP_FATAL("PreconditionerBlockJacobiStand: no operator");
    local_->SetupOperator(this->op_->diagonal());
    local_->Build();
  }
  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    local_->ApplylPreconditioner(b.interior(), &x->interior());
    return kSolverSuccess;
  }

 private:
  std::unique_ptr<lPreconditioner<DataType>> local_;
};

}  // namespace la
}  // namespace hiflow
