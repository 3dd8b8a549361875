// Preconditioner<LAD> (linear_solver/preconditioner.h:33-259): Build on the
// first Solve after the operator changed, then SolveImpl; counts builds,
// solves and accumulated iterations.
#pragma once

#include <chrono>
#include <string>

#include "../la_descriptor.h"

namespace hiflow {
namespace la {

enum LinearSolverState { kSolverSuccess = 0, kSolverExceeded, kSolverError };

template <class LAD>
class Preconditioner {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  Preconditioner() = default;
  virtual ~Preconditioner() = default;

  virtual void SetupOperator(OperatorType &op) {
    op_ = &op;
    state_ = false;
  }
  virtual void Build() {
    const auto t0 = std::chrono::steady_clock::now();
    BuildImpl();
    state_ = true;
    ++num_build_;
    time_build_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  virtual LinearSolverState Solve(const VectorType &b, VectorType *x) {
    if (!state_) Build();
    const auto t0 = std::chrono::steady_clock::now();
    const LinearSolverState s = SolveImpl(b, x);
    ++num_solve_;
    time_solve_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    acc_iter_ += iter_;
    return s;
  }
  virtual void Clear() {
    op_ = nullptr;
    state_ = false;
    iter_ = 0;
  }

  bool GetState() const { return state_; }
  void SetState(bool s) { state_ = s; }
  int iter() const { return iter_; }
  void SetPrintLevel(int l) { print_level_ = l; }
  void SetName(const std::string &n) { name_ = n; }
  const std::string &name() const { return name_; }
  int num_build() const { return num_build_; }
  int num_solve() const { return num_solve_; }
  int acc_iter() const { return acc_iter_; }
  double time_build() const { return time_build_; }
  double time_solve() const { return time_solve_; }

 protected:
  virtual void BuildImpl() {}
  virtual LinearSolverState SolveImpl(const VectorType &b, VectorType *x) = 0;

  OperatorType *op_ = nullptr;
  bool state_ = false;
  int iter_ = 0;
  int print_level_ = 0;
  std::string name_ = "Preconditioner";
  int num_build_ = 0, num_solve_ = 0, acc_iter_ = 0;
  double time_build_ = 0., time_solve_ = 0.;
};

}  // namespace la
}  // namespace hiflow
