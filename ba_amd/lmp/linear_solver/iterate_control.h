// Convergence control of the iterative solvers (common/iterate_control.h,
// iterate_control.cc:48-107): success when res <= atol or res < rtol *
// res_0, failure when res > dtol * res_0 or the step count reaches maxits.
#pragma once

#include "../lmp_log.h"

namespace hiflow {

class IterateControl {
 public:
  enum State {
    kIterate = 0,
    kSuccessAbsoluteTol,
    kSuccessRelativeTol,
    kFailureDivergenceTol,
    kFailureMaxitsExceeded
  };

  IterateControl() = default;
  IterateControl(int maxits, double atol, double rtol, double dtol) {
    Init(maxits, atol, rtol, dtol);
  }
  void Init(int maxits, double atol, double rtol, double dtol) {
    if (!(maxits > 0 && atol >= 0. && rtol >= 0. && rtol < 1. && dtol > 1.))
      LMP_FATAL("IterateControl::Init: need maxits > 0, atol >= 0, 0 <= rtol < 1, dtol > 1");
    maxits_ = maxits;
    atol_ = atol;
    rtol_ = rtol;
    dtol_ = dtol;
  }
  void set_first(int first_step, double first_value) {
    first_step_ = first_step;
    first_value_ = first_value;
  }
  State Check(int current_step, double current_value) {
    if (current_step == first_step_) first_value_ = current_value;
    if (current_value <= atol_) status_ = kSuccessAbsoluteTol;
    else if (current_value < rtol_ * first_value_) status_ = kSuccessRelativeTol;
    else if (current_value > dtol_ * first_value_) status_ = kFailureDivergenceTol;
    else if (current_step >= maxits_) status_ = kFailureMaxitsExceeded;
    else status_ = kIterate;
    return status_;
  }
  State status() const { return status_; }
  int maxits() const { return maxits_; }
  double absolute_tol() const { return atol_; }
  double relative_tol() const { return rtol_; }
  double divergence_tol() const { return dtol_; }
  int first_step() const { return first_step_; }
  double first_value() const { return first_value_; }

 private:
  int maxits_ = 1000;
  double atol_ = 1e-12, rtol_ = 1e-10, dtol_ = 1e6;
  int first_step_ = 0;
  double first_value_ = 0.;
  State status_ = kIterate;
};

}  // namespace hiflow
