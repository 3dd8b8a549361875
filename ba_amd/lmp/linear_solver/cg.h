// CG<LAD>: conjugate gradients with the arithmetic of linear_solver/cg.h:
// 87-246 (SolveNoPrecond / SolvePrecond).  Convergence is tested on the
// recursive residual ||r_k|| = sqrt(r.r) with IterateControl.  Every vector
// operation is a CoupledVector call, so on the HIP platform each one is a
// device kernel and each scalar product returns to the host (three per
// iteration), as in the reference's GPU runs.
#pragma once

#include <cmath>

#include "linear_solver.h"

namespace hiflow {
namespace la {

template <class LAD>
class CG : public LinearSolver<LAD> {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  CG() { this->SetName("CG"); }

 protected:
  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    const std::string &m = this->Method();
    if (m == "NoPreconditioning") return SolveNoPrecond(b, x);
    if (m == "Preconditioning" || m == "RightPreconditioning" || m == "LeftPreconditioning")
      return SolvePrecond(b, x);
    return kSolverError;
  }

 private:
  LinearSolverState finish(IterateControl::State conv) {
    LMP_INFO(this->name_, "iterations " << this->iter_ << ", residual " << this->res_
                                        << " (relative " << this->res_rel_ << ")");
    if (conv == IterateControl::kFailureDivergenceTol ||
        conv == IterateControl::kFailureMaxitsExceeded)
      return kSolverExceeded;
    return kSolverSuccess;
  }

  LinearSolverState SolveNoPrecond(const VectorType &b, VectorType *x) {
    if (!this->op_) return kSolverError;
    VectorType r, p, Ap;
    r.CloneFromWithoutContent(b);
    p.CloneFromWithoutContent(b);
    Ap.CloneFromWithoutContent(b);
    this->iter_ = 0;
    this->op_->VectorMult(*x, &r);
    r.ScaleAdd(b, DataType(-1));  // r = b - A x
    DataType rr = r.Dot(r);
    this->res_init_ = this->res_ = std::sqrt(rr);
    this->res_rel_ = 1;
    IterateControl::State conv = this->control().Check(this->iter_, this->res_);
    p.CopyFrom(r);
    DataType beta = rr;
    while (conv == IterateControl::kIterate) {
      ++this->iter_;
      this->op_->VectorMult(p, &Ap);
      const DataType alpha = beta / Ap.Dot(p);
      x->Axpy(p, alpha);
      r.Axpy(Ap, -alpha);
      rr = r.Dot(r);
      this->res_ = std::sqrt(rr);
      this->res_rel_ = this->res_ / this->res_init_;
      conv = this->control().Check(this->iter_, this->res_);
      if (conv != IterateControl::kIterate) break;
      p.ScaleAdd(r, rr / beta);  // p = (r.r / beta) p + r
      beta = rr;
    }
    return finish(conv);
  }

  LinearSolverState SolvePrecond(const VectorType &b, VectorType *x) {
    if (!this->op_) return kSolverError;
    VectorType r, p, z, Ap;
    r.CloneFromWithoutContent(b);
    p.CloneFromWithoutContent(b);
    z.CloneFromWithoutContent(b);
    Ap.CloneFromWithoutContent(b);
    this->iter_ = 0;
    this->op_->VectorMult(*x, &r);
    r.ScaleAdd(b, DataType(-1));
    DataType rr = r.Dot(r);
    this->res_init_ = this->res_ = std::sqrt(rr);
    this->res_rel_ = 1;
    IterateControl::State conv = this->control().Check(this->iter_, this->res_);
    z.Zeros();
    this->ApplyPreconditioner(r, &z);
    p.CopyFrom(z);
    DataType beta = r.Dot(z);
    while (conv == IterateControl::kIterate) {
      ++this->iter_;
      this->op_->VectorMult(p, &Ap);
      const DataType alpha = beta / Ap.Dot(p);
      x->Axpy(p, alpha);
      r.Axpy(Ap, -alpha);
      rr = r.Dot(r);
      this->res_ = std::sqrt(rr);
      this->res_rel_ = this->res_ / this->res_init_;
      conv = this->control().Check(this->iter_, this->res_);
      if (conv != IterateControl::kIterate) break;
      z.Zeros();
      this->ApplyPreconditioner(r, &z);
      const DataType gamma = r.Dot(z);
      p.ScaleAdd(z, gamma / beta);  // p = (gamma / beta) p + z
      beta = gamma;
    }
    return finish(conv);
  }
};

}  // namespace la
}  // namespace hiflow
