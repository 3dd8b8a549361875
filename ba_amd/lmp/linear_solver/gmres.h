// GMRES<LAD> and FGMRES<LAD>: restarted GMRES with modified Gram-Schmidt
// Arnoldi, Givens rotations and back substitution on the host, the basis
// vectors on the device.  Arithmetic of linear_solver/gmres.h:97-116,
// 464-718 (NoPreconditioning and RightPreconditioning) and
// linear_solver/fgmres.h:119-320 (flexible: the preconditioned directions
// Z_j are stored and x += Z y).
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "linear_solver.h"

namespace hiflow {
namespace la {

template <class LAD>
class GMRESBase : public LinearSolver<LAD> {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  void InitParameter(int size_basis, std::string const &method) {
    size_basis_ = size_basis;
    this->method_ = method;
  }
  using LinearSolver<LAD>::InitParameter;
  int size_basis() const { return size_basis_; }

 protected:
  explicit GMRESBase(bool flexible) : flexible_(flexible) {}

  // H(j, i) = H_[j * ldh + i]: column j of the Hessenberg matrix
  static void rot_apply(DataType cs, DataType sn, DataType *dx, DataType *dy) {
    const DataType t = cs * (*dx) + sn * (*dy);
    *dy = -sn * (*dx) + cs * (*dy);
    *dx = t;
  }
  static void rot_gen(DataType dx, DataType dy, DataType *cs, DataType *sn) {
    const DataType beta = std::sqrt(dx * dx + dy * dy);
    *cs = dx / beta;
    *sn = dy / beta;
  }
  // x += sum_{j<=k} y_j B_j with H y = g (upper triangle of the first k+1
  // columns)
  void update_solution(std::vector<std::unique_ptr<VectorType>> &B, const std::vector<DataType> &H,
                       int ldh, const std::vector<DataType> &g, int k, VectorType *x) const {
    std::vector<DataType> y(g.begin(), g.begin() + k + 1);
    for (int i = k; i >= 0; --i) {
      y[i] /= H[i * ldh + i];
      const DataType t = y[i];
      for (int j = 0; j < i; ++j) y[j] -= H[i * ldh + j] * t;
    }
    for (int j = 0; j <= k; ++j) x->Axpy(*B[j], y[j]);
  }

  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    if (!this->op_ || size_basis_ <= 0) return kSolverError;
    const bool prec = this->method_ != "NoPreconditioning";
    if (flexible_ && !prec) return kSolverError;
    if (prec && this->method_ != "RightPreconditioning" && this->method_ != "Preconditioning")
      return kSolverError;
    const int m = std::min(size_basis_, this->control().maxits());
    const int ldh = m + 1;
    std::vector<std::unique_ptr<VectorType>> V(m + 1), Z(flexible_ ? m + 1 : 0);
    for (auto &v : V) {
      v.reset(new VectorType());
      v->CloneFromWithoutContent(b);
    }
    for (auto &z : Z) {
      z.reset(new VectorType());
      z->CloneFromWithoutContent(b);
    }
    VectorType w, zz;
    w.CloneFromWithoutContent(b);
    zz.CloneFromWithoutContent(b);
    std::vector<DataType> H((size_t)m * ldh), g(m + 1), cs(m + 1), sn(m + 1);

    int iter = 0;
    this->op_->VectorMult(*x, V[0].get());
    V[0]->ScaleAdd(b, DataType(-1));  // V0 = b - A x
    this->res_init_ = this->res_ = V[0]->Norm2();
    this->res_rel_ = 1;
    IterateControl::State conv = this->control().Check(iter, this->res_);
    while (conv == IterateControl::kIterate) {
      std::fill(g.begin(), g.end(), DataType(0));
      std::fill(H.begin(), H.end(), DataType(0));
      V[0]->Scale(DataType(1) / this->res_);
      g[0] = this->res_;
      bool done = false;
      for (int j = 0; j < m; ++j) {
        ++iter;
        const VectorType *dir = V[j].get();
        if (prec) {
          VectorType *dst = flexible_ ? Z[j].get() : &zz;
          dst->Zeros();
          this->ApplyPreconditioner(*V[j], dst);
          dir = dst;
        }
        this->op_->VectorMult(const_cast<VectorType &>(*dir), &w);
        for (int i = 0; i <= j; ++i) {
          H[j * ldh + i] = w.Dot(*V[i]);
          w.Axpy(*V[i], -H[j * ldh + i]);
        }
        H[j * ldh + j + 1] = w.Norm2();
        w.Scale(DataType(1) / H[j * ldh + j + 1]);
        V[j + 1]->CopyFrom(w);
        for (int k = 0; k < j; ++k) rot_apply(cs[k], sn[k], &H[j * ldh + k], &H[j * ldh + k + 1]);
        rot_gen(H[j * ldh + j], H[j * ldh + j + 1], &cs[j], &sn[j]);
        rot_apply(cs[j], sn[j], &H[j * ldh + j], &H[j * ldh + j + 1]);
        rot_apply(cs[j], sn[j], &g[j], &g[j + 1]);
        this->res_ = std::abs(g[j + 1]);
        this->res_rel_ = this->res_ / this->res_init_;
        conv = this->control().Check(iter, this->res_);
        if (conv != IterateControl::kIterate) {
          finish_cycle(V, Z, H, ldh, g, j, prec, x, zz, w);
          done = true;
          break;
        }
      }
      if (!done) {
        finish_cycle(V, Z, H, ldh, g, m - 1, prec, x, zz, w);
        // restart from the true residual
        this->op_->VectorMult(*x, V[0].get());
        V[0]->ScaleAdd(b, DataType(-1));
      }
    }
    this->iter_ = iter;
    LMP_INFO(this->name_, "iterations " << iter << ", residual " << this->res_);
    if (conv == IterateControl::kFailureDivergenceTol ||
        conv == IterateControl::kFailureMaxitsExceeded)
      return kSolverExceeded;
    return kSolverSuccess;
  }

 private:
  void finish_cycle(std::vector<std::unique_ptr<VectorType>> &V,
                    std::vector<std::unique_ptr<VectorType>> &Z, const std::vector<DataType> &H,
                    int ldh, const std::vector<DataType> &g, int k, bool prec, VectorType *x,
                    VectorType &zz, VectorType &w) {
    if (flexible_) {
      update_solution(Z, H, ldh, g, k, x);  // x += Z y
    } else if (prec) {
      zz.Zeros();
      update_solution(V, H, ldh, g, k, &zz);  // x += M^-1 V y
      w.Zeros();
      this->ApplyPreconditioner(zz, &w);
      x->Axpy(w, DataType(1));
    } else {
      update_solution(V, H, ldh, g, k, x);
    }
  }

  bool flexible_;
  int size_basis_ = 100;
};

template <class LAD>
class GMRES : public GMRESBase<LAD> {
 public:
  GMRES() : GMRESBase<LAD>(false) { this->SetName("GMRES"); }
};

template <class LAD>
class FGMRES : public GMRESBase<LAD> {
 public:
  FGMRES() : GMRESBase<LAD>(true) {
    this->SetName("FGMRES");
    this->method_ = "RightPreconditioning";
  }
};

}  // namespace la
}  // namespace hiflow
