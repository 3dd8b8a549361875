// MPIR<LAD_high, LAD_low>: mixed-precision iterative refinement
// (linear_solver/mpir.cc:52-296) with the low-precision error correction:
//   r_h = b - A_h x;  loop: r_l = (low) r_h;  c_l = 0;  solve A_l c_l = r_l
//   with any LinearSolver<LAD_low>;  x += (high) c_l;  r_h = b - A_h x.
// The outer relative test uses the initial residual for the whole run; the
// reference resets res0_ after each correction (mpir.cc:240), which makes
// its tolerance relative to the previous step instead.
#pragma once

#include <memory>

#include "linear_solver.h"

namespace hiflow {
namespace la {

template <class LADh, class LADl>
class MPIR : public LinearSolver<LADh> {
 public:
  typedef typename LADh::MatrixType OperatorType;
  typedef typename LADh::VectorType VectorType;
  typedef typename LADh::DataType DataType;
  typedef typename LADl::MatrixType OperatorTypeLow;
  typedef typename LADl::VectorType VectorTypeLow;
  typedef typename LADl::DataType DataTypeLow;

  MPIR() { this->SetName("MPIR"); }

  // The low-precision operator is the high one rounded entrywise.
  void SetupOperator(OperatorType &op) override {
    Preconditioner<LADh>::SetupOperator(op);
    Al_.reset(new OperatorTypeLow());
    Al_->CastFrom(op);
    if (ecl_) ecl_->SetupOperator(*Al_);
  }
  void SetupErrorCorrectionSolverLow(LinearSolver<LADl> &solver) {
    ecl_ = &solver;
    if (Al_) ecl_->SetupOperator(*Al_);
  }
  // outer loop: at most max_ecl corrections, stop when res <= abs_tol,
  // res < rel_tol * res_0, or res >= div_tol
  void SetMpirParameters(int max_ecl, double abs_tol, double rel_tol, double div_tol) {
    max_ecl_ = max_ecl;
    abs_tol_ = abs_tol;
    rel_tol_ = rel_tol;
    div_tol_ = div_tol;
  }
  int num_corrections() const { return n_ecl_; }
  int inner_iterations() const { return accu_iter_ecl_; }
  OperatorTypeLow *low_operator() { return Al_.get(); }

 protected:
  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    if (!this->op_ || !ecl_ || !Al_) return kSolverError;
    VectorType resh, corh;
    VectorTypeLow resl, corl;
    resh.CloneFromWithoutContent(b);
    corh.CloneFromWithoutContent(b);
    resl.Init(b.la_couplings(), b.platform(), HIPKERNEL);
    corl.Init(b.la_couplings(), b.platform(), HIPKERNEL);
    n_ecl_ = 0;
    accu_iter_ecl_ = 0;
    this->op_->VectorMult(*x, &resh);
    resh.ScaleAdd(b, DataType(-1));
    this->res_ = resh.Norm2();
    this->res_init_ = this->res_;
    const DataType res0 = this->res_;
    while (this->res_ > abs_tol_ && this->res_ > rel_tol_ * res0 && this->res_ < div_tol_ &&
           n_ecl_ < max_ecl_) {
      ++n_ecl_;
      resl.CastInteriorFrom(resh);
      corl.Zeros();
      ecl_->Solve(resl, &corl);
      accu_iter_ecl_ += ecl_->iter();
      corh.CastInteriorFrom(corl);
      x->Axpy(corh, DataType(1));
      this->op_->VectorMult(*x, &resh);
      resh.ScaleAdd(b, DataType(-1));
      this->res_ = resh.Norm2();
      LMP_INFO(this->name_, "correction " << n_ecl_ << ": residual " << this->res_);
    }
    this->res_rel_ = res0 > 0 ? this->res_ / res0 : DataType(0);
    this->iter_ = accu_iter_ecl_;
    if (this->res_ <= abs_tol_ || this->res_ < rel_tol_ * res0) return kSolverSuccess;
    return kSolverExceeded;
  }

 private:
  std::unique_ptr<OperatorTypeLow> Al_;
  LinearSolver<LADl> *ecl_ = nullptr;
  int max_ecl_ = 100;
  double abs_tol_ = 1e-14, rel_tol_ = 1e-10, div_tol_ = 1e6;
  int n_ecl_ = 0, accu_iter_ecl_ = 0;
};

}  // namespace la
}  // namespace hiflow
