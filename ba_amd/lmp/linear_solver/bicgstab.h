// BiCGSTAB<LAD>: stabilised bi-conjugate gradients (van der Vorst, SIAM J.
// Sci. Stat. Comput. 13, 1992) with the arithmetic of linear_solver/
// bicgstab.h:127-483 (SolveNoPrecond, SolvePrecondRight): shadow residual
// r0 = b - A x, rho = (r, r0), p = r + beta (p - omega Ap); a restart from
// the true residual when |(Ap, r0)| <= tol ||Ap|| ||r0||, and an update with
// alpha alone when ||s|| falls below the same tolerance (1e-15 without,
// 5e-16 with preconditioning).  The iteration count is not reset by a
// restart.  Convergence is tested on the recursive residual ||r|| with
// IterateControl.  Left preconditioning is refused, as in the reference.
#pragma once

#include <cmath>

#include "linear_solver.h"

namespace hiflow {
namespace la {

template <class LAD>
class BiCGSTAB : public LinearSolver<LAD> {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  BiCGSTAB() { this->SetName("BiCGSTAB"); }

  void InitParameter(std::string const &method) override {
    if (method != "NoPreconditioning" && method != "RightPreconditioning" &&
        method != "LeftPreconditioning")
      LMP_FATAL("BiCGSTAB: unknown method " << method);
    LinearSolver<LAD>::InitParameter(method);
  }

 protected:
  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    const std::string &m = this->Method();
    if (m == "NoPreconditioning") return Solve_(b, x, false);
    if (m == "RightPreconditioning") return Solve_(b, x, true);
    if (m == "LeftPreconditioning") LMP_FATAL("BiCGSTAB: left preconditioning is not implemented");
    return kSolverError;
  }

 private:
  // h = M^-1 v (right preconditioning) or a copy of v
  void apply(bool prec, const VectorType &v, VectorType *h) {
    h->Zeros();
    if (prec)
      this->ApplyPreconditioner(v, h);
    else
      h->CopyFrom(v);
  }

  void residual_norm(const VectorType &r) {
    this->res_ = std::sqrt(r.Dot(r));
    this->res_rel_ = this->res_ / this->res_init_;
  }

  LinearSolverState Solve_(const VectorType &b, VectorType *x, bool prec) {
    if (!this->op_) return kSolverError;
    const DataType tol = prec ? DataType(5e-16) : DataType(1e-15);
    VectorType r, r0, p, Ap, s, t, h, h1;
    for (VectorType *v : {&r, &r0, &p, &Ap, &s, &t, &h, &h1}) v->CloneFromWithoutContent(b);
    IterateControl::State conv = IterateControl::kIterate;
    this->iter_ = 0;
    bool restart = true;
    while (restart) {
      restart = false;
      this->op_->VectorMult(*x, &r0);
      r0.ScaleAdd(b, DataType(-1));  // r0 = b - A x
      r.CopyFrom(r0);
      const DataType rr = r.Dot(r);
      DataType rho = rr, rho_old = rr;
      this->res_init_ = this->res_ = std::sqrt(rr);
      this->res_rel_ = 1;
      conv = this->control().Check(this->iter_, this->res_);
      p.CopyFrom(r);
      while (conv == IterateControl::kIterate && !restart) {
        ++this->iter_;
        // Ap = A M^-1 p
        if (prec) {
          apply(true, p, &h);
          this->op_->VectorMult(h, &Ap);
        } else {
          this->op_->VectorMult(p, &Ap);
        }
        const DataType sigma = Ap.Dot(r0);
        const DataType tol1 = tol * Ap.Norm2() * r0.Norm2();
        if (!(std::abs(sigma) > tol1)) {
          restart = true;
          break;
        }
        const DataType alpha = rho / sigma;
        s.CopyFrom(r);
        s.Axpy(Ap, -alpha);  // s = r - alpha Ap
        if (s.Norm2() > tol) {
          // t = A M^-1 s
          if (prec) {
            apply(true, s, &h);
            this->op_->VectorMult(h, &t);
          } else {
            this->op_->VectorMult(s, &t);
          }
          const DataType omega = t.Dot(s) / t.Dot(t);
          // x += M^-1 (alpha p + omega s)
          if (prec) {
            h1.Zeros();
            h1.Axpy(p, alpha);
            h1.Axpy(s, omega);
            apply(true, h1, &h);
            x->Axpy(h, DataType(1));
          } else {
            x->Axpy(p, alpha);
            x->Axpy(s, omega);
          }
          r.CopyFrom(s);
          r.Axpy(t, -omega);  // r = s - omega t
          residual_norm(r);
          conv = this->control().Check(this->iter_, this->res_);
          if (conv != IterateControl::kIterate) break;
          rho = r.Dot(r0);
          const DataType beta = (alpha * rho) / (omega * rho_old);
          rho_old = rho;
          p.ScaleAdd(r, beta);         // p = beta p + r
          p.Axpy(Ap, -omega * beta);   // p -= omega beta Ap
        } else {
          // x += M^-1 (alpha p), r = s
          if (prec) {
            h1.Zeros();
            h1.Axpy(p, alpha);
            apply(true, h1, &h);
            x->Axpy(h, DataType(1));
          } else {
            x->Axpy(p, alpha);
          }
          r.CopyFrom(s);
          residual_norm(r);
          conv = this->control().Check(this->iter_, this->res_);
          if (conv != IterateControl::kIterate) break;
        }
      }
    }
    LMP_INFO(this->name_, "iterations " << this->iter_ << ", residual " << this->res_);
    if (conv == IterateControl::kFailureDivergenceTol ||
        conv == IterateControl::kFailureMaxitsExceeded)
      return kSolverExceeded;
    return kSolverSuccess;
  }
};

}  // namespace la
}  // namespace hiflow
