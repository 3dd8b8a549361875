// The lmp-local conjugate gradient entry (linear_algebra/lmp/solvers/cg.h:
// 30-209, cg.cc:35-53): cg() over the lVector / lMatrix / lPreconditioner
// interface, so it runs unchanged on any platform the factory provides,
// here the HIP one.
//
// Arithmetic, in the reference's order (CGFunctor::operator(), cg.h:
// 111-205): r = A x; r = b - r (ScaleAdd(-1, b)); rho = r.r; norm =
// sqrt(rho); conv = max(rel_eps norm, abs_eps); then per iteration
//   p = beta p + r (ScaleAdd; p starts at zero and beta at 0)
//   q = A p;  alpha = rho / (p.q);  x += alpha p;  r -= alpha q
//   rho_old = rho;  rho = r.r;  beta = rho / rho_old;  norm = sqrt(rho)
//   stop when norm <= conv
// max_iter returns the index of the last iteration (the loop counter at
// the break, or max_iter when none broke), and the function returns 0.
//
// The reference's quirk is kept on purpose: when a preconditioner is given
// it is applied to r (once before the loop, then at the start of every
// iteration) into a work vector z that the iteration never reads, so the
// iterates are those of unpreconditioned CG.  Callers that rely on the
// reference's results get the same ones; the preconditioner's own work
// (its Build on first use) still happens.  For preconditioned CG use
// CG<LAD> (linear_solver/cg.h).  Deviation: nothing is printed unless
// print > -1 (the reference always prints the preconditioner's name).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>
#include <memory>

#include "../lmatrix.h"
#include "../lmp_log.h"
#include "../lpreconditioner.h"
#include "../lvector.h"

namespace hiflow {
namespace la {

template <typename ValueType>
int cg(lVector<ValueType> *x, const lVector<ValueType> *b, const lMatrix<ValueType> *matrix,
       const ValueType rel_eps, const ValueType abs_eps, int &max_iter, const int print,
       lPreconditioner<ValueType> *lPrecond = nullptr);

// Timer = true accumulates the wall time of each kind of operation (the
// reference's ScopedTimer members, cg.h:49-91); read them with seconds().
template <typename ValueType, bool Timer>
struct CGFunctor {
  enum Part { kNrm2, kDot, kAxpy, kScaleAdd, kVectorMult, kTotal, kPreconditioner, kParts };

  int operator()(lVector<ValueType> *x, const lVector<ValueType> *b,
                 const lMatrix<ValueType> *matrix, const ValueType rel_eps,
                 const ValueType abs_eps, int &max_iter, const int print,
                 lPreconditioner<ValueType> *lPrecond = nullptr) const;

  double seconds(Part p) const { return t_[p]; }

 private:
  template <typename F>
  void timed(Part p, F &&f) const {
    if (!Timer) {
      f();
      return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    f();
    x_sync_();
    t_[p] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void x_sync_() const {
    if (sync_) sync_->Sync();
  }

  mutable double t_[kParts] = {0, 0, 0, 0, 0, 0, 0};
  mutable const lVector<ValueType> *sync_ = nullptr;
};

template <typename ValueType, bool Timer>
int CGFunctor<ValueType, Timer>::operator()(lVector<ValueType> *x, const lVector<ValueType> *b,
                                            const lMatrix<ValueType> *matrix,
                                            const ValueType rel_eps, const ValueType abs_eps,
                                            int &max_iter, const int print,
                                            lPreconditioner<ValueType> *lPrecond) const {
  for (double &t : t_) t = 0;
  if (!x || !b || !matrix) LMP_FATAL("cg: null argument");
  sync_ = x;
  const auto t_total = std::chrono::steady_clock::now();
  if (print > 0) {
    x->print();
    b->print();
    matrix->print();
  }
  std::unique_ptr<lVector<ValueType>> p(x->CloneWithoutContent());
  std::unique_ptr<lVector<ValueType>> q(x->CloneWithoutContent());
  std::unique_ptr<lVector<ValueType>> r(x->CloneWithoutContent());
  timed(kVectorMult, [&] { matrix->VectorMult(*x, r.get()); });
  timed(kScaleAdd, [&] { r->ScaleAdd(ValueType(-1), *b); });  // r = b - A x
  std::unique_ptr<lVector<ValueType>> z;
  if (lPrecond) {
    z.reset(x->CloneWithoutContent());
    timed(kPreconditioner, [&] { lPrecond->ApplylPreconditioner(*r, z.get()); });
    if (print > -1) lPrecond->print();
  }
  ValueType rho = 0, norm = 0;
  timed(kDot, [&] { rho = r->Dot(*r); });
  timed(kNrm2, [&] { norm = std::sqrt(rho); });
  if (print > -1)
    std::cout << "cg: CG with " << (lPrecond ? "" : "No") << " Preconditioner\n"
              << "cg: Init, rho = " << rho << ", norm = " << norm << std::endl;
  int iter = 0;
  ValueType beta = 0, rho_old = 0;
  const ValueType conv_crit = std::max(rel_eps * norm, abs_eps);
  for (; iter < max_iter; ++iter) {
    if (lPrecond)
      timed(kPreconditioner, [&] { lPrecond->ApplylPreconditioner(*r, z.get()); });
    timed(kScaleAdd, [&] { p->ScaleAdd(beta, *r); });  // p = beta p + r
    timed(kVectorMult, [&] { matrix->VectorMult(*p, q.get()); });
    ValueType alpha = 0;
    timed(kDot, [&] { alpha = rho / p->Dot(*q); });
    timed(kAxpy, [&] { x->Axpy(*p, alpha); });
    timed(kAxpy, [&] { r->Axpy(*q, -alpha); });
    rho_old = rho;
    timed(kDot, [&] { rho = r->Dot(*r); });
    beta = rho / rho_old;
    timed(kNrm2, [&] { norm = std::sqrt(rho); });
    if (print > 0)
      std::cout << "cg: iter = " << iter << ", rho = " << rho << ", norm = " << norm
                << std::endl;
    if (norm <= conv_crit) break;
  }
  if (print > -1)
    std::cout << "cg: End after: iterations=" << iter << ", rho = " << rho
              << ", norm = " << norm << std::endl;
  max_iter = iter;
  if (Timer)
    t_[kTotal] =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_total).count();
  return 0;
}

}  // namespace la
}  // namespace hiflow
