// LinearSolver<LAD> (linear_solver/linear_solver.h:38-371): a Preconditioner
// that iterates under an IterateControl, with an optional preconditioner
// applied through ApplyPreconditioner.
#pragma once

#include <string>

#include "iterate_control.h"
#include "preconditioner.h"

namespace hiflow {
namespace la {

template <class LAD, class PreLAD = LAD>
class LinearSolver : virtual public Preconditioner<LAD> {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  ~LinearSolver() override = default;

  virtual void InitControl(int maxits, double abstol, double reltol, double divtol) {
    control_.Init(maxits, abstol, reltol, divtol);
  }
  virtual void InitControl(int maxits) {
    control_.Init(maxits, control_.absolute_tol(), control_.relative_tol(),
                  control_.divergence_tol());
  }
  virtual void InitParameter(std::string const &method) { method_ = method; }
  virtual void SetupPreconditioner(Preconditioner<PreLAD> &precond) { precond_ = &precond; }
  void SetupOperator(OperatorType &op) override {
    Preconditioner<LAD>::SetupOperator(op);
    if (precond_) precond_->SetupOperator(op);
  }
  IterateControl &control() { return control_; }
  const std::string &Method() const { return method_; }
  virtual DataType res() const { return res_; }
  DataType res_init() const { return res_init_; }
  DataType res_rel() const { return res_rel_; }

 protected:
  // z = M^-1 r
  virtual LinearSolverState ApplyPreconditioner(const VectorType &r, VectorType *z) {
    if (!precond_) {
      z->CopyFrom(r);
      return kSolverSuccess;
    }
    return precond_->Solve(r, z);
  }

  IterateControl control_;
  std::string method_ = "NoPreconditioning";
  Preconditioner<PreLAD> *precond_ = nullptr;
  DataType res_ = 0, res_init_ = 0, res_rel_ = 0;
};

}  // namespace la
}  // namespace hiflow
