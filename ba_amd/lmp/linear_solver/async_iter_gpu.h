// AsynchronousIterationGPU<LAD> (linear_solver/gpu-based/async_iter_gpu.h:
// 32-157, async_iter_gpu.cc:106-874) on the HIP platform: block Jacobi
// iterations on the operator's diagonal block through
// hflmp_async_jacobi_{f64,f32}, with the reference's control surface
// (SetNumIter, SetSolveMode, SetDampingParameter, Prepare, Solve).
//
// One process and one device: the reference's split of rows over devices
// (use_gpus) and its BlockUpdate / GlobalUpdate exchanges have nothing to
// exchange here, so Prepare accepts use_gpus == 1 only.  The reference's
// in-place updates inside a device read whatever neighbouring rows hold at
// the moment; on the HIP platform rows are grouped in blocks of
// block_rows (one workgroup each), every block runs n_inner synchronous
// sweeps per launch with the couplings outside it taken from the launch's
// input, and results do not depend on scheduling.  SetCudaBlockSize sets
// that block size (the reference's thread-block size is the unit its
// kernels are launched in, as the block is here).  The residual is always
// formed on the device (compute_residual_on_dev), in the reference's device
// form b - (A - D) x - D x.
#pragma once

#include <string>
#include <type_traits>

#include "../lmatrix_csr_hip.h"
#include "../lmp_log.h"
#include "../platform_management.h"
#include "linear_solver.h"

namespace hiflow {
namespace la {

template <class LAD>
class AsynchronousIterationGPU : public LinearSolver<LAD> {
 public:
  typedef typename LAD::MatrixType MatrixType;
  typedef typename LAD::VectorType VectorType;
  typedef typename LAD::DataType DataType;

  AsynchronousIterationGPU() { this->SetName("AsynchronousIterationGPU"); }
  ~AsynchronousIterationGPU() override { Clear(); }

  void Clear() override {
    if (work_) hflmp_free(hip_ctx(), work_);
    work_ = nullptr;
    work_bytes_ = 0;
    LinearSolver<LAD>::Clear();
  }

  // The reference copies A, b and x to its devices here; on the HIP
  // platform they already are device objects, so Prepare records the
  // operator and checks the arguments.
  void Prepare(MatrixType &A, const VectorType &b, VectorType &x,
               const bool compute_residual_on_dev, const int use_gpus,
               const int procs_per_node) {
    if (use_gpus != 1 || procs_per_node < 1)
      LMP_FATAL("AsynchronousIterationGPU::Prepare: one device per process (use_gpus 1)");
    if (A.diagonal().get_num_row() != b.interior().get_size() ||
        A.diagonal().get_num_row() != x.interior().get_size())
      LMP_FATAL("AsynchronousIterationGPU::Prepare: sizes differ");
    (void)compute_residual_on_dev;
    this->SetupOperator(A);
  }

  void SetNumIter(const int n_global, const int n_block = 1, const int n_inner = 5) {
    if (n_global < 1 || n_block < 1 || n_inner < 1)
      LMP_FATAL("AsynchronousIterationGPU::SetNumIter: counts must be positive");
    prm_.n_global = n_global;
    prm_.n_block = n_block;
    prm_.n_inner = n_inner;
  }
  // 0 solve normal, 1 solve damped, 2 smooth normal, 3 smooth damped
  void SetSolveMode(const int mode) {
    if (mode < 0 || mode > 3) LMP_FATAL("AsynchronousIterationGPU: no solve mode " << mode);
    prm_.mode = mode;
  }
  void SetDampingParameter(const DataType w) {
    if (!(w > DataType(0))) LMP_FATAL("AsynchronousIterationGPU: damping must be positive");
    prm_.w = (double)w;
  }
  void SetCudaBlockSize(const int block_size) { SetBlockRows(block_size); }
  void SetBlockRows(const int rows) {
    if (rows < 1) LMP_FATAL("AsynchronousIterationGPU: block size must be positive");
    prm_.block_rows = rows;
  }
  void SetRelativeTolerance(double reltol) {
    this->control_.Init(this->control_.maxits(), this->control_.absolute_tol(), reltol,
                        this->control_.divergence_tol());
  }
  const hflmp_async_params &params() const { return prm_; }

 protected:
  LinearSolverState SolveImpl(const VectorType &b, VectorType *x) override {
    if (!this->op_) return kSolverError;
    const auto *A = dynamic_cast<const HIP_CSR_lMatrix<DataType> *>(&this->op_->diagonal());
    if (!A) LMP_FATAL("AsynchronousIterationGPU: the operator must be a HIP CSR matrix");
    const int n = A->get_num_row();
    const size_t need = hflmp_async_workspace_bytes(n, (int)sizeof(DataType));
    if (need > work_bytes_) {
      if (work_) hflmp_free(hip_ctx(), work_);
      work_ = nullptr;
      hip_check(hflmp_malloc(hip_ctx(), need, &work_), "AsynchronousIterationGPU workspace");
      work_bytes_ = need;
    }
    hflmp_solve_info info;
    const IterateControl &c = this->control_;
    int rc;
    if constexpr (std::is_same<DataType, double>::value)
      rc = hflmp_async_jacobi_f64(hip_ctx(), n, A->get_nnz(), A->row_ptr(), A->col_ind(),
                                  A->values(), b.interior().GetBuffer(),
                                  x->interior().GetBuffer(), &prm_, c.maxits(),
                                  c.absolute_tol(), c.relative_tol(), c.divergence_tol(),
                                  work_, work_bytes_, &info);
    else
      rc = hflmp_async_jacobi_f32(hip_ctx(), n, A->get_nnz(), A->row_ptr(), A->col_ind(),
                                  A->values(), b.interior().GetBuffer(),
                                  x->interior().GetBuffer(), &prm_, c.maxits(),
                                  c.absolute_tol(), c.relative_tol(), c.divergence_tol(),
                                  work_, work_bytes_, &info);
    hip_check(rc, "AsynchronousIterationGPU::Solve");
    this->iter_ = info.iters;
    this->res_ = (DataType)info.res_final;
    this->res_init_ = (DataType)info.res_init;
    this->res_rel_ = info.res_init > 0 ? (DataType)(info.res_final / info.res_init) : 0;
    // SolveNormal / SolveDamped / Smooth* return success whatever the
    // residual (async_iter_gpu.cc:775,837,851,873); callers read res() and
    // iter(), as the reference's do
    return kSolverSuccess;
  }

 private:
  hflmp_async_params prm_{128, 1, 1, 5, 0, 1.0};
  void *work_ = nullptr;
  size_t work_bytes_ = 0;
};

}  // namespace la
}  // namespace hiflow
