// Local preconditioners (linear_algebra/lmp/lpreconditioner.h:35-312,
// lpreconditioner_mc.h, lpreconditioner_mc_sgs.h): the lPreconditioner
// interface, point Jacobi, and multicoloured symmetric Gauss-Seidel, all
// applied on the device for HIP operators.
#pragma once

#include <iostream>
#include <memory>
#include <string>

#include "hflmp_hip.h"
#include "lmatrix.h"
#include "lvector.h"

namespace hiflow {
namespace la {

template <typename ValueType>
class lPreconditioner {
 public:
  lPreconditioner() = default;
  virtual ~lPreconditioner() = default;
  virtual void Init() {}
  // the operator must outlive the preconditioner
  virtual void SetupOperator(const lMatrix<ValueType> &op) {
    Operator_ = &op;
    state_ = false;
  }
  virtual void Build() = 0;
  virtual void ApplylPreconditioner(const lVector<ValueType> &input,
                                    lVector<ValueType> *output) = 0;
  virtual void Clear() {
    Operator_ = nullptr;
    state_ = false;
  }
  bool GetState() const { return state_; }
  void SetState(bool s) { state_ = s; }
  virtual void print(std::ostream &out = std::cout) const {
    out << "lPreconditioner " << precond_name_ << std::endl;
  }

 protected:
  const lMatrix<ValueType> *Operator_ = nullptr;
  bool state_ = false;
  std::string precond_name_;
};

// z = D^-1 r (lPreconditioner_Jacobi, lpreconditioner.cc:125-200); on the
// HIP platform the copy and the multiplication are one kernel.
template <typename ValueType>
class lPreconditioner_Jacobi : public lPreconditioner<ValueType> {
 public:
  lPreconditioner_Jacobi();
  ~lPreconditioner_Jacobi() override;
  void Build() override;
  void ApplylPreconditioner(const lVector<ValueType> &input,
                            lVector<ValueType> *output) override;
  void Clear() override;
  const lVector<ValueType> *inv_diagonal() const { return inv_D_.get(); }

 private:
  std::unique_ptr<lVector<ValueType>> inv_D_;
};

// Symmetric Gauss-Seidel in the sequential row order: the operator's
// Psgauss_seidel (lPreconditioner_SymmetricGaussSeidel, lpreconditioner.h:
// 165-177, lpreconditioner.cc:230-257).  Matrix-free: Build only records the
// operator.
template <typename ValueType>
class lPreconditioner_SymmetricGaussSeidel : public lPreconditioner<ValueType> {
 public:
  lPreconditioner_SymmetricGaussSeidel();
  ~lPreconditioner_SymmetricGaussSeidel() override;
  void Init() override {}
  void Build() override;
  void ApplylPreconditioner(const lVector<ValueType> &input,
                            lVector<ValueType> *output) override;
};

// The same on num_blocks row ranges: the operator's BlocksPsgauss_seidel
// (lPreconditioner_BlocksSymmetricGaussSeidel, lpreconditioner.h:179-200,
// lpreconditioner.cc:259-293); Init() means one block.
template <typename ValueType>
class lPreconditioner_BlocksSymmetricGaussSeidel : public lPreconditioner<ValueType> {
 public:
  lPreconditioner_BlocksSymmetricGaussSeidel();
  ~lPreconditioner_BlocksSymmetricGaussSeidel() override;
  void Init() override { Init(1); }
  virtual void Init(const int num_blocks);
  void Build() override;
  void ApplylPreconditioner(const lVector<ValueType> &input,
                            lVector<ValueType> *output) override;
  int num_blocks() const { return num_blocks_; }

 private:
  int num_blocks_ = 1;
};

// Multicoloured SGS with relaxation omega: colouring of the operator's
// pattern, then the colour-ordered forward sweep, diagonal scaling and
// backward sweep on the device (lpreconditioner_mc.cc:147-723,
// lpreconditioner_mc_sgs.cc:44-90).  Provided for double.
template <typename ValueType>
class lPreconditioner_MultiColoring_SymmetricGaussSeidel : public lPreconditioner<ValueType> {
 public:
  lPreconditioner_MultiColoring_SymmetricGaussSeidel();
  ~lPreconditioner_MultiColoring_SymmetricGaussSeidel() override;
  void SetRelax(ValueType omega_relax);
  void Build() override;
  void ApplylPreconditioner(const lVector<ValueType> &input,
                            lVector<ValueType> *output) override;
  void Clear() override;
  int get_ncolors() const { return ncolors_; }

 private:
  ValueType omega_ = ValueType(1);
  int ncolors_ = 0;
  hflmp_mcsgs *mc_ = nullptr;
};

}  // namespace la
}  // namespace hiflow
