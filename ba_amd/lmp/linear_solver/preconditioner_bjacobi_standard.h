// PreconditionerBlockJacobiStand<LAD> (linear_solver/
// preconditioner_bjacobi_standard.cc:41-196): a local lPreconditioner on
// each process's diagonal block.  Jacobi, and symmetric Gauss-Seidel in its
// parallel multicoloured form (the sequential row sweep of the CPU path
// has no parallel counterpart on the device; the colour ordering is the
// reference's lpreconditioner_mc_sgs).
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
  void Init_SymmetricGaussSeidel() { Init_MultiColoring_SymmetricGaussSeidel(DataType(1)); }
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
