// C entry points of the lmp layer: set up CoupledMatrix / CoupledVector on
// the HIP platform from host CSR arrays and run the solver templates, as an
// application would (tutorial pattern: assemble, diagonalize_rows, Solve).
#include <chrono>
#include <cstring>
#include <memory>
#include <string>

#include "hflmp_lmp.h"
#include "la_descriptor.h"
#include "lmatrix_csr_hip.h"
#include "lmp_mmio.h"
#include "lvector_hip.h"
#include "linear_solver/bicgstab.h"
#include "linear_solver/cg.h"
#include "linear_solver/gmres.h"
#include "linear_solver/mpir.h"
#include "linear_solver/preconditioner_bjacobi_standard.h"
#include "linear_solver/preconditioner_gmg.h"
#include "platform_management.h"

using namespace hiflow;
using namespace hiflow::la;

namespace {

thread_local std::string g_last_error;

struct ThrowOnFatal {
  ThrowOnFatal() {
    set_fatal_handler([](const std::string &m) { g_last_error = m; });
  }
  ~ThrowOnFatal() { set_fatal_handler(nullptr); }
};

template <class LAD>
std::unique_ptr<LinearSolver<LAD>> make_solver(int kind, const hflmp_lmp_options &o) {
  std::unique_ptr<LinearSolver<LAD>> s;
  if (kind == HFLMP_LMP_CG) {
    s.reset(new CG<LAD>());
    s->InitParameter(o.precond == HFLMP_LMP_PREC_NONE ? "NoPreconditioning" : "Preconditioning");
  } else if (kind == HFLMP_LMP_GMRES) {
    auto *g = new GMRES<LAD>();
    g->InitParameter(o.basis, o.precond == HFLMP_LMP_PREC_NONE ? "NoPreconditioning"
                                                                : "RightPreconditioning");
    s.reset(g);
  } else if (kind == HFLMP_LMP_BICGSTAB) {
    auto *g = new BiCGSTAB<LAD>();
    g->InitParameter(o.precond == HFLMP_LMP_PREC_NONE ? "NoPreconditioning"
                                                      : "RightPreconditioning");
    s.reset(g);
  } else if (kind == HFLMP_LMP_FGMRES) {
    auto *g = new FGMRES<LAD>();
    g->InitParameter(o.basis, "RightPreconditioning");
    s.reset(g);
  } else {
    LMP_FATAL("unknown solver kind " << kind);
  }
  s->InitControl(o.maxit, o.atol, o.rtol, o.dtol);
  return s;
}

template <class LAD>
std::unique_ptr<PreconditionerBlockJacobiStand<LAD>> make_precond(const hflmp_lmp_options &o) {
  std::unique_ptr<PreconditionerBlockJacobiStand<LAD>> p;
  if (o.precond == HFLMP_LMP_PREC_JACOBI) {
    p.reset(new PreconditionerBlockJacobiStand<LAD>());
    p->Init_Jacobi();
  } else if (o.precond == HFLMP_LMP_PREC_MCSGS) {
    p.reset(new PreconditionerBlockJacobiStand<LAD>());
    p->Init_MultiColoring_SymmetricGaussSeidel(typename LAD::DataType(o.omega));
  } else if (o.precond == HFLMP_LMP_PREC_SGS) {
    p.reset(new PreconditionerBlockJacobiStand<LAD>());
    p->Init_SymmetricGaussSeidel();
  } else if (o.precond == HFLMP_LMP_PREC_BSGS) {
    p.reset(new PreconditionerBlockJacobiStand<LAD>());
    p->Init_BlocksSymmetricGaussSeidel(o.sgs_blocks);
  } else if (o.precond != HFLMP_LMP_PREC_NONE) {
    LMP_FATAL("unknown preconditioner " << o.precond);
  }
  return p;
}

// CG/GMRES/FGMRES in precision LAD on the FP64 system (A, b, x).
template <class LAD>
void run_plain(CoupledMatrix<double> &A, CoupledVector<double> &b, CoupledVector<double> &x,
               const hflmp_lmp_options &o, hflmp_lmp_result *r) {
  typedef typename LAD::MatrixType M;
  typedef typename LAD::VectorType V;
  M Al;
  V bl, xl;
  M *op;
  V *bp, *xp;
  if constexpr (std::is_same<typename LAD::DataType, double>::value) {
    op = &A;
    bp = &b;
    xp = &x;
  } else {
    Al.CastFrom(A);
    bl.Init(b.la_couplings());
    xl.Init(x.la_couplings());
    bl.CastInteriorFrom(b);
    xl.CastInteriorFrom(x);
    op = &Al;
    bp = &bl;
    xp = &xl;
  }
  auto solver = make_solver<LAD>(o.solver, o);
  auto pre = make_precond<LAD>(o);
  if (pre) solver->SetupPreconditioner(*pre);
  solver->SetupOperator(*op);
  hip_check(hflmp_ctx_sync(hip_ctx()), "sync");
  const auto t0 = std::chrono::steady_clock::now();
  const LinearSolverState st = solver->Solve(*bp, xp);
  hip_check(hflmp_ctx_sync(hip_ctx()), "sync");
  r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if constexpr (!std::is_same<typename LAD::DataType, double>::value) x.CastInteriorFrom(xl);
  r->iters = solver->iter();
  r->status = (int)st;
  r->res_init = (double)solver->res_init();
  r->res_final = (double)solver->res();
}

void run_mpir(CoupledMatrix<double> &A, CoupledVector<double> &b, CoupledVector<double> &x,
              const hflmp_lmp_options &o, hflmp_lmp_result *r) {
  MPIR<LADescriptorCoupledD, LADescriptorCoupledS> mpir;
  mpir.SetMpirParameters(o.mpir_max_outer, o.mpir_atol, o.mpir_rtol, 1e300);
  auto inner = make_solver<LADescriptorCoupledS>(o.solver_inner, o);
  auto pre = make_precond<LADescriptorCoupledS>(o);
  mpir.SetupOperator(A);
  if (pre) inner->SetupPreconditioner(*pre);
  mpir.SetupErrorCorrectionSolverLow(*inner);
  hip_check(hflmp_ctx_sync(hip_ctx()), "sync");
  const auto t0 = std::chrono::steady_clock::now();
  const LinearSolverState st = mpir.Solve(b, &x);
  hip_check(hflmp_ctx_sync(hip_ctx()), "sync");
  r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  r->iters = mpir.inner_iterations();
  r->outer = mpir.num_corrections();
  r->status = (int)st;
  r->res_init = (double)mpir.res_init();
  r->res_final = (double)mpir.res();
}

void setup_system(int device, int nrow, const int *rp, const int *ci, const double *val,
                  CoupledMatrix<double> &A) {
  SYSTEM sys;
  sys.Platform = HIP;
  sys.device = device;
  init_platform(sys);
  LaCouplings cp(nrow);
  A.Init(cp);
  A.InitCSR(rp, ci, val);
}

// host CSR arrays of one level of the hierarchy
PreconditionerGMG<LADescriptorCoupledD>::Csr host_csr(const hflmp_csr_f64 &m, const char *what) {
  if (m.nrow <= 0 || m.ncol <= 0 || !m.row_ptr || m.row_ptr[0] != 0 ||
      m.row_ptr[m.nrow] != m.nnz || (m.nnz > 0 && (!m.col || !m.val)))
    LMP_FATAL("hflmp_lmp_solve_gmg: bad " << what << " level");
  // the device gathers at these columns: check them here
  for (int r = 0; r < m.nrow; ++r)
    if (m.row_ptr[r + 1] < m.row_ptr[r])
      LMP_FATAL("hflmp_lmp_solve_gmg: " << what << " row_ptr decreases at row " << r);
  for (int k = 0; k < m.nnz; ++k)
    if (m.col[k] < 0 || m.col[k] >= m.ncol)
      LMP_FATAL("hflmp_lmp_solve_gmg: " << what << " column " << m.col[k] << " out of range");
  PreconditionerGMG<LADescriptorCoupledD>::Csr c;
  c.nrow = m.nrow;
  c.ncol = m.ncol;
  c.row_ptr.assign(m.row_ptr, m.row_ptr + m.nrow + 1);
  c.col.assign(m.col, m.col + m.nnz);
  c.val.assign(m.val, m.val + m.nnz);
  return c;
}

}  // namespace

extern "C" {

void hflmp_lmp_default_options(hflmp_lmp_options *o) {
  std::memset(o, 0, sizeof(*o));
  o->solver = HFLMP_LMP_CG;
  o->precond = HFLMP_LMP_PREC_JACOBI;
  o->precision = 64;
  o->maxit = 10000;
  o->atol = 0.0;
  o->rtol = 1e-8;
  o->dtol = 1e6;
  o->basis = 100;
  o->omega = 1.0;
  o->mpir_max_outer = 50;
  o->mpir_atol = 0.0;
  o->mpir_rtol = 1e-10;
  o->solver_inner = HFLMP_LMP_CG;
  o->sgs_blocks = 1;
}

int hflmp_lmp_solve(int device, int nrow, int nnz, const int *rp, const int *ci,
                    const double *val, const double *b, double *x, const hflmp_lmp_options *opt,
                    hflmp_lmp_result *res) {
  ThrowOnFatal guard;
  if (!rp || !ci || !val || !b || !x || !opt || !res || nrow <= 0 || rp[nrow] != nnz) {
    g_last_error = "hflmp_lmp_solve: bad arguments";
    return -1;
  }
  if (opt->precond == HFLMP_LMP_PREC_GMG) {
    g_last_error = "hflmp_lmp_solve: multigrid needs its hierarchy (hflmp_lmp_solve_gmg)";
    return -1;
  }
  std::memset(res, 0, sizeof(*res));
  try {
    CoupledMatrix<double> A;
    setup_system(device, nrow, rp, ci, val, A);
    CoupledVector<double> vb, vx;
    vb.Init(nrow);
    vx.Init(nrow);
    vb.interior().SetBlockValues(0, nrow, b);
    vx.interior().SetBlockValues(0, nrow, x);
    if (opt->solver == HFLMP_LMP_MPIR)
      run_mpir(A, vb, vx, *opt, res);
    else if (opt->precision == 32)
      run_plain<LADescriptorCoupledS>(A, vb, vx, *opt, res);
    else
      run_plain<LADescriptorCoupledD>(A, vb, vx, *opt, res);
    vx.interior().GetBlockValues(0, nrow, x);
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
  return 0;
}

int hflmp_lmp_solve_gmg(int device, int nrow, int nnz, const int *rp, const int *ci,
                        const double *val, const double *b, double *x, int nlevels,
                        const hflmp_csr_f64 *coarse, const hflmp_csr_f64 *R,
                        const hflmp_csr_f64 *P, const hflmp_mg_params *prm,
                        const double *coarse_inverse, const hflmp_lmp_options *opt,
                        hflmp_lmp_result *res) {
  ThrowOnFatal guard;
  if (!rp || !ci || !val || !b || !x || !opt || !res || !prm || nrow <= 0 || rp[nrow] != nnz ||
      nlevels < 1 || (nlevels > 1 && (!coarse || !R || !P)) ||
      (coarse_inverse && nlevels < 2) ||
      (opt->solver != HFLMP_LMP_CG && opt->solver != HFLMP_LMP_GMRES &&
       opt->solver != HFLMP_LMP_FGMRES && opt->solver != HFLMP_LMP_BICGSTAB)) {
    g_last_error = "hflmp_lmp_solve_gmg: bad arguments";
    return -1;
  }
  std::memset(res, 0, sizeof(*res));
  try {
    typedef LADescriptorCoupledD LAD;
    std::vector<PreconditionerGMG<LAD>::Csr> Ac, Rc, Pc;
    for (int l = 0; l + 1 < nlevels; ++l) {
      Ac.push_back(host_csr(coarse[l], "coarse operator"));
      Rc.push_back(host_csr(R[l], "restriction"));
      Pc.push_back(host_csr(P[l], "prolongation"));
    }
    std::vector<double> cinv;
    if (coarse_inverse) {
      const size_t n = (size_t)coarse[nlevels - 2].nrow;
      cinv.assign(coarse_inverse, coarse_inverse + n * n);
    }
    CoupledMatrix<double> A;
    setup_system(device, nrow, rp, ci, val, A);
    CoupledVector<double> vb, vx;
    vb.Init(nrow);
    vx.Init(nrow);
    vb.interior().SetBlockValues(0, nrow, b);
    vx.interior().SetBlockValues(0, nrow, x);
    hflmp_lmp_options o = *opt;
    o.precond = HFLMP_LMP_PREC_GMG;
    auto solver = make_solver<LAD>(o.solver, o);
    PreconditionerGMG<LAD> gmg;
    gmg.SetHierarchy(std::move(Ac), std::move(Rc), std::move(Pc), *prm, std::move(cinv));
    solver->SetupPreconditioner(gmg);
    solver->SetupOperator(A);
    hip_check(hflmp_ctx_sync(hip_ctx()), "sync");
    const auto t0 = std::chrono::steady_clock::now();
    const LinearSolverState st = solver->Solve(vb, &vx);
    hip_check(hflmp_ctx_sync(hip_ctx()), "sync");
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->iters = solver->iter();
    res->status = (int)st;
    res->res_init = solver->res_init();
    res->res_final = solver->res();
    vx.interior().GetBlockValues(0, nrow, x);
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
  return 0;
}

int hflmp_lmp_vector_mult(int device, int nrow, int nnz, const int *rp, const int *ci,
                          const double *val, const double *x, double *y) {
  ThrowOnFatal guard;
  if (!rp || !ci || !val || !x || !y || nrow <= 0 || rp[nrow] != nnz) {
    g_last_error = "hflmp_lmp_vector_mult: bad arguments";
    return -1;
  }
  try {
    CoupledMatrix<double> A;
    setup_system(device, nrow, rp, ci, val, A);
    CoupledVector<double> vx, vy;
    vx.Init(nrow);
    vy.Init(nrow);
    vx.interior().SetBlockValues(0, nrow, x);
    A.VectorMult(vx, &vy);
    vy.interior().GetBlockValues(0, nrow, y);
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
  return 0;
}

int hflmp_lmp_mm_read_coo(const char *path, int *nrow, int *ncol, int *nentries, int *rows,
                          int *cols, double *vals, int capacity) {
  if (!path || !nrow || !ncol || !nentries) {
    g_last_error = "hflmp_lmp_mm_read_coo: null argument";
    return -1;
  }
  MmCoo coo;
  const std::string err = mm_read_coo(path, coo);
  if (!err.empty()) {
    g_last_error = err;
    return -1;
  }
  *nrow = coo.nrow;
  *ncol = coo.ncol;
  *nentries = (int)coo.val.size();
  if (!rows || !cols || !vals) return 0;
  if (capacity < (int)coo.val.size()) {
    g_last_error = "hflmp_lmp_mm_read_coo: capacity below the number of entries";
    return -1;
  }
  std::memcpy(rows, coo.row.data(), sizeof(int) * coo.row.size());
  std::memcpy(cols, coo.col.data(), sizeof(int) * coo.col.size());
  std::memcpy(vals, coo.val.data(), sizeof(double) * coo.val.size());
  return 0;
}

int hflmp_lmp_mm_write_csr(const char *path, int nrow, int ncol, const int *rp, const int *ci,
                           const double *val) {
  if (!path || !rp || nrow < 0 || ncol < 0 || (rp[nrow] > 0 && (!ci || !val))) {
    g_last_error = "hflmp_lmp_mm_write_csr: bad arguments";
    return -1;
  }
  const std::string err = mm_write_csr(path, nrow, ncol, rp, ci, val);
  if (!err.empty()) {
    g_last_error = err;
    return -1;
  }
  return 0;
}

int hflmp_lmp_file_spmv(int device, const char *matrix_file, const char *x_file,
                        const char *y_file, const char *matrix_out, int *nrow, int *nnz) {
  ThrowOnFatal guard;
  if (!matrix_file || !x_file || !y_file) {
    g_last_error = "hflmp_lmp_file_spmv: null file name";
    return -1;
  }
  try {
    SYSTEM sys;
    sys.Platform = HIP;
    sys.device = device;
    init_platform(sys);
    HIP_CSR_lMatrix<double> A;
    A.ReadFile(matrix_file);
    HIP_lVector<double> x, y;
    x.ReadFile(x_file);
    if (x.get_size() != A.get_num_col())
      LMP_FATAL("x has " << x.get_size() << " entries, the matrix " << A.get_num_col()
                         << " columns");
    y.Init(A.get_num_row(), "y");
    A.VectorMult(x, &y);
    y.WriteFile(y_file);
    if (matrix_out) A.WriteFile(matrix_out);
    if (nrow) *nrow = A.get_num_row();
    if (nnz) *nnz = A.get_nnz();
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
  return 0;
}

const char *hflmp_lmp_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
