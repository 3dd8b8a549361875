// PreconditionerGMG<LAD>: one V-cycle of the device geometric multigrid
// (hflmp_mg_*, the GMG kernels of lmp/cuda/cuda_gmg_kernel.h driven as
// GMGBase::SolveImplLevel does, linear_solver/gmg.h:1070-1187) as a
// Preconditioner, so the unchanged CG/GMRES/FGMRES templates use it
// through SetupPreconditioner like any other.  Level 0 is the operator
// given to SetupOperator (the HIP matrix's device CSR arrays); the coarser
// operators, restrictions and prolongations are host CSR arrays given by
// SetHierarchy, uploaded at Build.  FP64 only.
#pragma once

#include <memory>
#include <vector>

#include "../lmatrix_csr_hip.h"
#include "../lmp_log.h"
#include "../lvector_hip.h"
#include "../platform_management.h"
#include "preconditioner.h"

namespace hiflow {
namespace la {

template <class LAD>
class PreconditionerGMG : public Preconditioner<LAD> {
 public:
  typedef typename LAD::MatrixType OperatorType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;
  static_assert(std::is_same<DataType, double>::value, "PreconditionerGMG is FP64");

  // A host CSR matrix
  struct Csr {
    int nrow = 0, ncol = 0;
    std::vector<int> row_ptr, col;
    std::vector<double> val;
  };

  PreconditionerGMG() { this->SetName("GMG"); }
  ~PreconditionerGMG() override { release(); }

  // coarse[l] is level l + 1's operator; R[l], P[l] connect levels l and
  // l + 1.  coarse_inverse (optional, n x n row-major): the dense inverse
  // of the coarsest operator, which then replaces its sweeps.
  void SetHierarchy(std::vector<Csr> coarse, std::vector<Csr> R, std::vector<Csr> P,
                    const hflmp_mg_params &prm, std::vector<double> coarse_inverse = {}) {
    if (R.size() != coarse.size() || P.size() != coarse.size())
      LMP_FATAL("PreconditionerGMG: need one R and one P per coarse level");
    coarse_ = std::move(coarse);
    R_ = std::move(R);
    P_ = std::move(P);
    prm_ = prm;
    cinv_ = std::move(coarse_inverse);
    this->state_ = false;
  }
  int num_levels() const { return (int)coarse_.size() + 1; }

  // frees the device hierarchy; the host levels stay for the next Build
  void Clear() override {
    release();
    Preconditioner<LAD>::Clear();
  }

 protected:
  void BuildImpl() override {
    if (!this->op_) LMP_FATAL("PreconditionerGMG: no operator");
    auto *A0 = dynamic_cast<const HIP_CSR_lMatrix<double> *>(&this->op_->diagonal());
    if (!A0) LMP_FATAL("PreconditionerGMG: the operator is not a HIP CSR matrix");
    release();
    hflmp_ctx *ctx = hip_ctx();
    const int L = num_levels();
    std::vector<hflmp_csr_f64> A(L), R(L > 1 ? L - 1 : 1), P(L > 1 ? L - 1 : 1);
    A[0] = {A0->get_num_row(), A0->get_num_col(), A0->get_nnz(), A0->row_ptr(), A0->col_ind(),
            A0->values()};
    // the hierarchy converts every level to its own storage at creation:
    // the uploaded host arrays are staging only
    Staging st;
    for (int l = 1; l < L; ++l) {
      A[l] = st.upload(coarse_[l - 1]);
      R[l - 1] = st.upload(R_[l - 1]);
      P[l - 1] = st.upload(P_[l - 1]);
    }
    hip_check(hflmp_mg_create_f64(ctx, L, A.data(), R.data(), P.data(), &prm_, &mg_),
              "PreconditionerGMG: hflmp_mg_create_f64");
    if (!cinv_.empty())
      hip_check(hflmp_mg_set_coarse_inverse_f64(
                    ctx, mg_, static_cast<const double *>(
                                  st.put(cinv_.data(), sizeof(double) * cinv_.size()))),
                "PreconditionerGMG: hflmp_mg_set_coarse_inverse_f64");
    hip_check(hflmp_ctx_sync(ctx), "PreconditionerGMG: sync");
  }

  // x = M b: one V-cycle from zero
  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    hip_check(hflmp_mg_apply_f64(hip_ctx(), mg_, b.interior().GetBuffer(),
                                 x->interior().GetBuffer()),
              "PreconditionerGMG: hflmp_mg_apply_f64");
    this->iter_ = 1;
    return kSolverSuccess;
  }

 private:
  // device copies of host arrays, freed when it goes out of scope
  struct Staging {
    std::vector<void *> bufs;
    ~Staging() {
      for (void *p : bufs) (void)hflmp_free(hip_ctx(), p);
    }
    void *put(const void *src, size_t bytes) {
      void *d = nullptr;
      hip_check(hflmp_malloc(hip_ctx(), bytes > 0 ? bytes : 1, &d), "PreconditionerGMG: malloc");
      bufs.push_back(d);
      if (bytes) hip_check(hflmp_memcpy_h2d(hip_ctx(), d, src, bytes), "PreconditionerGMG: upload");
      return d;
    }
    hflmp_csr_f64 upload(const Csr &m) {
      const int nnz = m.row_ptr.empty() ? 0 : m.row_ptr.back();
      hflmp_csr_f64 c;
      c.nrow = m.nrow;
      c.ncol = m.ncol;
      c.nnz = nnz;
      c.row_ptr = static_cast<const int *>(put(m.row_ptr.data(), sizeof(int) * m.row_ptr.size()));
      c.col = static_cast<const int *>(put(m.col.data(), sizeof(int) * (size_t)nnz));
      c.val = static_cast<const double *>(put(m.val.data(), sizeof(double) * (size_t)nnz));
      return c;
    }
  };

  void release() {
    if (mg_) {
      hflmp_mg_destroy(mg_);
      mg_ = nullptr;
    }
  }

  std::vector<Csr> coarse_, R_, P_;
  std::vector<double> cinv_;
  hflmp_mg_params prm_{1, 1, 4, 0.7, 0.7};
  hflmp_mg *mg_ = nullptr;
};

}  // namespace la
}  // namespace hiflow
