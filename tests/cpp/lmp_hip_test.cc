// The reference's lmp known-answer tests, run on the HIP platform through
// the factory, the Coupled layer and the unchanged solver templates:
//   tests/lmp_test.cc:42-90, tests/cuda_test.cc:29-108, tests/cg_test.cc:
//   30-139 (through CG<LAD> and through the lmp-local cg() itself),
//   tests/parallel_vector_test.cc:402-660,
//   tests/parallel_matrix_test.cc:922-1048,1187+ (one process).
// Prints one line per case and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "init_vec_mat.h"
#include "la_descriptor.h"
#include "linear_solver/async_iter_gpu.h"
#include "linear_solver/cg.h"
#include "linear_solver/gmres.h"
#include "linear_solver/mpir.h"
#include "linear_solver/preconditioner_bjacobi_standard.h"
#include "lmatrix_csr_hip.h"
#include "lmp_log.h"
#include "platform_management.h"
#include "solvers/cg.h"

using namespace hiflow;
using namespace hiflow::la;

static int g_fail = 0;
#define CHECK_CLOSE(a, b, tol)                                                         \
  do {                                                                                 \
    const double a_ = (a), b_ = (b);                                                   \
    if (!(std::fabs(a_ - b_) <= (tol))) {                                              \
      std::printf("  FAIL %s:%d: %s = %.17g, expected %.17g\n", __FILE__, __LINE__, #a, \
                  a_, b_);                                                             \
      ++g_fail;                                                                        \
    }                                                                                  \
  } while (0)
#define CHECK(c)                                                                \
  do {                                                                          \
    if (!(c)) {                                                                 \
      std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);                \
      ++g_fail;                                                                 \
    }                                                                           \
  } while (0)

static void run(const char *name, void (*f)()) {
  const int before = g_fail;
  f();
  std::printf("%s %s\n", g_fail == before ? "ok  " : "FAIL", name);
  std::fflush(stdout);
}

// lmp_test.cc: TransformFromCOO of [[2,-1,0],[-1,2,-1],[0,-1,2]] times [1,-2,0]
static void test_lmp_vectormult() {
  std::unique_ptr<lMatrix<double>> A(init_matrix<double>(7, 3, 3, "sm", HIP, HIPKERNEL, CSR));
  const int r[] = {0, 0, 1, 1, 1, 2, 2}, c[] = {0, 1, 0, 1, 2, 1, 2};
  const double v[] = {2, -1, -1, 2, -1, -1, 2};
  A->TransformFromCOO(r, c, v, 3, 3, 7);
  std::unique_ptr<lVector<double>> x(init_vector<double>(3, "svi", HIP, HIPKERNEL));
  std::unique_ptr<lVector<double>> y(init_vector<double>(3, "svo", HIP, HIPKERNEL));
  x->add_value(0, 1.0);
  x->add_value(1, -2.0);
  x->add_value(2, 0.0);
  A->VectorMult(*x, y.get());
  CHECK(y->GetValue(0) == 4.0);
  CHECK(y->GetValue(1) == -5.0);
  CHECK(y->GetValue(2) == 2.0);
}

// cuda_test.cc: dense 9-entry CSR (explicit zeros) times [1.2,-2.3,0.7]
static void test_dense_csr_vectormult() {
  std::unique_ptr<lMatrix<double>> A(init_matrix<double>(9, 3, 3, "A", HIP, HIPKERNEL, CSR));
  std::vector<int> r, c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      r.push_back(i);
      c.push_back(j);
    }
  A->init_structure(r.data(), c.data());
  const double vals[3][3] = {{2, -1, 0}, {-1, 2, -1}, {0, -1, 2}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A->add_value(i, j, vals[i][j]);
  CHECK(A->get_nnz() == 9);
  std::unique_ptr<lVector<double>> x(init_vector<double>(3, "x", HIP, HIPKERNEL));
  std::unique_ptr<lVector<double>> y(x->CloneWithoutContent());
  const int idx[] = {0, 1, 2};
  const double xv[] = {1.2, -2.3, 0.7};
  x->SetValues(idx, 3, xv);
  y->Zeros();
  A->VectorMult(*x, y.get());
  double out[3];
  y->GetValues(idx, 3, out);
  CHECK_CLOSE(out[0], 4.7, 1e-12);
  CHECK_CLOSE(out[1], -6.5, 1e-12);
  CHECK_CLOSE(out[2], 3.7, 1e-12);
}

static void make_3x3(CoupledMatrix<double> &A) {
  LaCouplings cp(3);
  A.Init(cp);
  const int r[] = {0, 0, 1, 1, 1, 2, 2}, c[] = {0, 1, 0, 1, 2, 1, 2};
  A.InitStructure(r, c, 7, nullptr, nullptr, 0);
  const double v[] = {2, -1, -1, 2, -1, -1, 2};
  for (int k = 0; k < 7; ++k) A.Add(r[k], c[k], v[k]);
}

// cg_test.cc: A x = [1,-2,0] -> x = [-0.25,-1.5,-0.75] to 1e-12
template <class LAD>
static void cg_3x3(bool precond) {
  CoupledMatrix<double> A;
  make_3x3(A);
  CoupledVector<double> b, x;
  b.Init(3);
  x.Init(3);
  const int idx[] = {0, 1, 2};
  const double bv[] = {1, -2, 0};
  b.SetValues(idx, 3, bv);
  CG<LAD> cg;
  cg.InitControl(100, 1e-15, 1e-14, 1e6);
  PreconditionerBlockJacobiStand<LAD> pre;
  if (precond) {
    pre.Init_Jacobi();
    cg.SetupPreconditioner(pre);
    cg.InitParameter("Preconditioning");
  } else {
    cg.InitParameter("NoPreconditioning");
  }
  cg.SetupOperator(A);
  CHECK(cg.Solve(b, &x) == kSolverSuccess);
  CHECK_CLOSE(x.GetValue(0), -0.25, 1e-12);
  CHECK_CLOSE(x.GetValue(1), -1.5, 1e-12);
  CHECK_CLOSE(x.GetValue(2), -0.75, 1e-12);
  CHECK(cg.iter() <= 3);
}
static void test_cg_noprecond() { cg_3x3<LADescriptorCoupledD>(false); }
static void test_cg_jacobi() { cg_3x3<LADescriptorCoupledD>(true); }

// cg_test.cc:30-139 through the lmp-local cg() (solvers/cg.cc:35-53): the
// full 3x3 CSR pattern from init_structure, values by add_value (explicit
// zeros kept), b = [1,-2,0], x0 = 0, rel_eps = abs_eps = 0, max_iter = 100,
// print -1.  The reference's three CSR cases: NAIVE, OPENMP (the same
// computation on another implementation: here the matrix's SELL-64 layout
// is switched off so the CSR kernels run) and JACOBI, which cg_test builds
// but never passes to cg() (cg_test.cc:83); a fourth case passes a Jacobi
// preconditioner, whose output cg() does not use, so x is bit-identical.
template <typename T>
static void cg_local_3x3(int variant, T tol, std::vector<T> *xout) {
  std::unique_ptr<lVector<T>> b(init_vector<T>(3, "b", HIP, HIPKERNEL));
  std::unique_ptr<lVector<T>> x(init_vector<T>(3, "x", HIP, HIPKERNEL));
  std::unique_ptr<lMatrix<T>> A(init_matrix<T>(9, 3, 3, "A", HIP, HIPKERNEL, CSR));
  b->Zeros();
  b->add_value(0, T(1));
  b->add_value(1, T(-2));
  b->add_value(2, T(0));
  x->Zeros();
  int rows[9], cols[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      rows[3 * i + j] = i;
      cols[3 * i + j] = j;
    }
  A->init_structure(rows, cols);
  const T av[3][3] = {{2, -1, 0}, {-1, 2, -1}, {0, -1, 2}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A->add_value(i, j, av[i][j]);
  if (variant == 1) dynamic_cast<HIP_CSR_lMatrix<T> *>(A.get())->set_use_sell(false);
  lPreconditioner_Jacobi<T> jac;
  lPreconditioner<T> *pre = nullptr;
  if (variant == 3) {
    jac.SetupOperator(*A);
    pre = &jac;
  }
  int max_iter = 100;
  const int rv = cg(x.get(), b.get(), A.get(), T(0), T(0), max_iter, -1, pre);
  CHECK(rv == 0);
  CHECK(max_iter >= 2 && max_iter <= 100);
  if (variant == 3) CHECK(jac.GetState());  // applied (and built), output unused
  const int idx[3] = {0, 1, 2};
  T xv[3];
  x->GetValues(idx, 3, xv);
  CHECK_CLOSE(xv[0], -0.25, tol);
  CHECK_CLOSE(xv[1], -1.50, tol);
  CHECK_CLOSE(xv[2], -0.75, tol);
  if (xout) xout->assign(xv, xv + 3);
}
static void test_cg_local_naive() { cg_local_3x3<double>(0, 1e-12, nullptr); }
static void test_cg_local_openmp() { cg_local_3x3<double>(1, 1e-12, nullptr); }
static void test_cg_local_jacobi() {
  std::vector<double> a, b;
  cg_local_3x3<double>(2, 1e-12, &a);
  cg_local_3x3<double>(3, 1e-12, &b);
  CHECK(a == b);
  cg_local_3x3<float>(0, 1e-5f, nullptr);
}

// Init_SymmetricGaussSeidel (preconditioner_bjacobi_standard.cc:95-101)
// builds the sequential SGS: on the 3x3 system z = M^-1 r with
// M = (D+L) D^-1 (D+U) in closed form, and CG<LAD> with it converges.
static void test_sgs_preconditioner() {
  std::unique_ptr<lMatrix<double>> A(init_matrix<double>(7, 3, 3, "A", HIP, HIPKERNEL, CSR));
  const int r[] = {0, 0, 1, 1, 1, 2, 2}, c[] = {0, 1, 0, 1, 2, 1, 2};
  A->init_structure(r, c);
  const double v[] = {2, -1, -1, 2, -1, -1, 2};
  for (int k = 0; k < 7; ++k) A->add_value(r[k], c[k], v[k]);
  std::unique_ptr<lVector<double>> rv(init_vector<double>(3, "r", HIP, HIPKERNEL));
  std::unique_ptr<lVector<double>> z(init_vector<double>(3, "z", HIP, HIPKERNEL));
  const int idx[3] = {0, 1, 2};
  const double rr[3] = {1, -2, 0.5};
  rv->SetValues(idx, 3, rr);
  lPreconditioner_SymmetricGaussSeidel<double> sgs;
  sgs.SetupOperator(*A);
  sgs.ApplylPreconditioner(*rv, z.get());
  // forward: y0 = 1/2, y1 = (-2 + y0)/2, y2 = (0.5 + y1)/2; D: w = 2 y;
  // backward: z2 = w2/2, z1 = (w1 + z2)/2, z0 = (w0 + z1)/2
  const double y0 = 0.5, y1 = (-2 + y0) / 2, y2 = (0.5 + y1) / 2;
  const double z2 = 2 * y2 / 2, z1 = (2 * y1 + z2) / 2, z0 = (2 * y0 + z1) / 2;
  double zv[3];
  z->GetValues(idx, 3, zv);
  CHECK(zv[0] == z0 && zv[1] == z1 && zv[2] == z2);
  CoupledMatrix<double> Ac;
  make_3x3(Ac);
  CoupledVector<double> b, x;
  b.Init(3);
  x.Init(3);
  const double bv[] = {1, -2, 0};
  b.SetValues(idx, 3, bv);
  CG<LADescriptorCoupledD> solver;
  solver.InitControl(100, 1e-15, 1e-14, 1e6);
  PreconditionerBlockJacobiStand<LADescriptorCoupledD> pre;
  pre.Init_SymmetricGaussSeidel();
  CHECK(dynamic_cast<lPreconditioner_SymmetricGaussSeidel<double> *>(pre.local()) != nullptr);
  solver.SetupPreconditioner(pre);
  solver.InitParameter("Preconditioning");
  solver.SetupOperator(Ac);
  CHECK(solver.Solve(b, &x) == kSolverSuccess);
  CHECK_CLOSE(x.GetValue(0), -0.25, 1e-12);
  CHECK_CLOSE(x.GetValue(1), -1.5, 1e-12);
  CHECK_CLOSE(x.GetValue(2), -0.75, 1e-12);
}

// AsynchronousIterationGPU<LAD> (async_iter_gpu.cc:682-874) on a 1D Laplace
// system of 200 rows (diagonal 2.5, off-diagonal -1: Jacobi converges):
// damped solve mode to 1e-8, checked against the exact solution; then one
// smoothing run (mode 2) leaves the residual smaller and counts its sweeps.
static void test_async_iteration() {
  const int n = 200;
  std::vector<int> r, c;
  for (int i = 0; i < n; ++i)
    for (int j = i - 1; j <= i + 1; ++j)
      if (j >= 0 && j < n) {
        r.push_back(i);
        c.push_back(j);
      }
  CoupledMatrix<double> A;
  LaCouplings cp(n);
  A.Init(cp);
  A.InitStructure(r.data(), c.data(), (int)r.size(), nullptr, nullptr, 0);
  for (size_t k = 0; k < r.size(); ++k) A.Add(r[k], c[k], r[k] == c[k] ? 2.5 : -1.0);
  CoupledVector<double> b, x, ax;
  b.Init(n);
  x.Init(n);
  ax.Init(n);
  std::vector<double> xe(n), bv(n);
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) {
    idx[i] = i;
    xe[i] = std::sin(0.05 * i);
  }
  for (int i = 0; i < n; ++i)
    bv[i] = 2.5 * xe[i] - (i > 0 ? xe[i - 1] : 0.0) - (i + 1 < n ? xe[i + 1] : 0.0);
  b.SetValues(idx.data(), n, bv.data());
  AsynchronousIterationGPU<LADescriptorCoupledD> as;
  as.Prepare(A, b, x, true, 1, 1);
  as.SetNumIter(1, 1, 4);
  as.SetSolveMode(1);
  as.SetDampingParameter(0.9);
  as.SetCudaBlockSize(32);
  as.InitControl(100000, 0.0, 1e-8, 1e6);
  CHECK(as.Solve(b, &x) == kSolverSuccess);
  CHECK(as.iter() > 0 && as.iter() % 4 == 0);
  CHECK(as.res() < 1e-8 * as.res_init());
  std::vector<double> xv(n);
  x.GetValues(idx.data(), n, xv.data());
  double err = 0;
  for (int i = 0; i < n; ++i) err = std::max(err, std::fabs(xv[i] - xe[i]));
  CHECK(err < 1e-6);
  CoupledVector<double> y;
  y.Init(n);
  as.SetSolveMode(2);
  as.SetNumIter(2, 1, 3);
  CHECK(as.Solve(b, &y) == kSolverSuccess);
  CHECK(as.iter() == 6);
  A.VectorMult(y, &ax);
  ax.ScaleAdd(b, -1.0);
  CHECK(ax.Norm2() < b.Norm2());
}

// parallel_vector_test.cc: v_i = i, w_i = i + 2, i < 10
static void test_vector_ops() {
  CoupledVector<double> v, w;
  v.Init(10);
  w.CloneFromWithoutContent(v);
  CHECK(w.size_global() == 10 && w.ghost().get_size() == 0);
  for (int i = 0; i < 10; ++i) {
    v.Add(i, double(i));
    w.Add(i, double(i + 2));
  }
  const double N = 9.0;
  CHECK_CLOSE(v.Dot(w), N * (N + 1) * (2 * N + 1) / 6 + N * (N + 1), 1e-10);
  CoupledVector<double> t;
  t.CloneFrom(v);
  t.Axpy(w, 4.0);  // axpy1
  for (int i = 0; i < 10; ++i) CHECK_CLOSE(t.GetValue(i), i + 4.0 * (i + 2), 1e-12);
  t.CopyFrom(v);
  t.Scale(0.5);  // scale
  for (int i = 0; i < 10; ++i) CHECK_CLOSE(t.GetValue(i), 0.5 * i, 1e-12);
  t.CopyFrom(v);
  t.ScaleAdd(w, 4.0);  // scale_add: t = 4 t + w
  for (int i = 0; i < 10; ++i) CHECK_CLOSE(t.GetValue(i), 4.0 * i + (i + 2), 1e-12);
  CHECK_CLOSE(v.Norm2(), std::sqrt(N * (N + 1) * (2 * N + 1) / 6), 1e-12);
  CHECK_CLOSE(v.NormMax(), 9.0, 0);
  CHECK_CLOSE(v.Norm1(), 45.0, 1e-12);
  CHECK(v.interior().ArgMax() == 9 && v.interior().ArgMin() == 0);
  CoupledVector<float> f;
  f.Init(10);
  f.CastInteriorFrom(w);
  CHECK_CLOSE(f.GetValue(7), 9.0, 0);
  CHECK(!v.interior().ContainsNaN());
}

// parallel_matrix_test.cc: tridiagonal block times ones -> [1,0,...,0,1];
// diagonalize_rows keeps the columns
static void test_matrix_tridiagonal() {
  const int n = 10;
  std::vector<int> r, c;
  for (int i = 0; i < n; ++i)
    for (int j = i - 1; j <= i + 1; ++j)
      if (j >= 0 && j < n) {
        r.push_back(i);
        c.push_back(j);
      }
  CoupledMatrix<double> A;
  A.Init(LaCouplings(n));
  A.InitStructure(r.data(), c.data(), (int)r.size(), nullptr, nullptr, 0);
  std::vector<int> rows(n);
  for (int i = 0; i < n; ++i) rows[i] = i;
  std::vector<double> val(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = i - 1; j <= i + 1; ++j)
      if (j >= 0 && j < n) val[i * n + j] = (i == j) ? 2.0 : -1.0;
  // full row x column block: entries outside the pattern are skipped by the
  // reference's SetValues only when zero; set the pattern entries one by one
  for (int k = 0; k < (int)r.size(); ++k) A.SetValues(&r[k], 1, &c[k], 1, &val[r[k] * n + c[k]]);
  CoupledVector<double> in, out;
  in.Init(n);
  out.Init(n);
  std::vector<double> ones(n, 1.0);
  in.Add(rows.data(), n, ones.data());
  A.VectorMult(in, &out);
  CHECK_CLOSE(out.GetValue(0), 1.0, 1e-14);
  for (int i = 1; i < n - 1; ++i) CHECK_CLOSE(out.GetValue(i), 0.0, 1e-14);
  CHECK_CLOSE(out.GetValue(n - 1), 1.0, 1e-14);
  const int drow[] = {0, 4};
  A.diagonalize_rows(drow, 2, 3.0);
  double g;
  A.diagonal().get_value(4, 4, &g);
  CHECK(g == 3.0);
  A.diagonal().get_value(4, 3, &g);
  CHECK(g == 0.0);
  A.diagonal().get_value(3, 4, &g);
  CHECK(g == -1.0);
}

// GMRES / FGMRES / MPIR on a 1D Poisson system through the same templates
static void poisson_1d(int n, CoupledMatrix<double> &A, CoupledVector<double> &b) {
  std::vector<int> r, c;
  std::vector<double> v;
  for (int i = 0; i < n; ++i)
    for (int j = i - 1; j <= i + 1; ++j)
      if (j >= 0 && j < n) {
        r.push_back(i);
        c.push_back(j);
        v.push_back(i == j ? 2.0 : -1.0);
      }
  A.Init(LaCouplings(n));
  A.InitStructure(r.data(), c.data(), (int)r.size(), nullptr, nullptr, 0);
  for (size_t k = 0; k < r.size(); ++k) A.Add(r[k], c[k], v[k]);
  b.Init(n);
  for (int i = 0; i < n; ++i) b.Add(i, std::sin(0.37 * i) + 0.1);
}

static double residual(CoupledMatrix<double> &A, CoupledVector<double> &b,
                       CoupledVector<double> &x) {
  CoupledVector<double> r;
  r.CloneFromWithoutContent(b);
  A.VectorMult(x, &r);
  r.ScaleAdd(b, -1.0);
  return r.Norm2() / b.Norm2();
}

static void test_gmres_fgmres_mpir() {
  CoupledMatrix<double> A;
  CoupledVector<double> b;
  poisson_1d(60, A, b);
  {
    CoupledVector<double> x;
    x.CloneFromWithoutContent(b);
    GMRES<LADescriptorCoupledD> gm;
    gm.InitControl(1000, 0.0, 1e-10, 1e6);
    gm.InitParameter(60, "NoPreconditioning");
    gm.SetupOperator(A);
    CHECK(gm.Solve(b, &x) == kSolverSuccess);
    CHECK(residual(A, b, x) < 1e-9);
  }
  {
    CoupledVector<double> x;
    x.CloneFromWithoutContent(b);
    FGMRES<LADescriptorCoupledD> fg;
    PreconditionerBlockJacobiStand<LADescriptorCoupledD> pre;
    pre.Init_MultiColoring_SymmetricGaussSeidel(1.0);
    fg.InitControl(1000, 0.0, 1e-10, 1e6);
    fg.InitParameter(20, "RightPreconditioning");
    fg.SetupPreconditioner(pre);
    fg.SetupOperator(A);
    CHECK(fg.Solve(b, &x) == kSolverSuccess);
    CHECK(residual(A, b, x) < 1e-9);
  }
  {
    CoupledVector<double> x;
    x.CloneFromWithoutContent(b);
    MPIR<LADescriptorCoupledD, LADescriptorCoupledS> mpir;
    CG<LADescriptorCoupledS> inner;
    PreconditionerBlockJacobiStand<LADescriptorCoupledS> pre;
    pre.Init_Jacobi();
    inner.InitControl(5000, 0.0, 1e-4, 1e6);
    inner.InitParameter("Preconditioning");
    inner.SetupPreconditioner(pre);
    mpir.SetMpirParameters(30, 0.0, 1e-11, 1e300);
    mpir.SetupOperator(A);
    mpir.SetupErrorCorrectionSolverLow(inner);
    CHECK(mpir.Solve(b, &x) == kSolverSuccess);
    CHECK(mpir.num_corrections() >= 2);
    CHECK(residual(A, b, x) < 1e-11);
  }
}

// The SELL-64 copy behind VectorMult / VectorMultAdd against the CSR
// kernels (set_use_sell(false)) bit for bit, through value changes that
// must refresh the copy: ZeroRows, add_values, Zeros + set_values, CopyFrom.
static void sell_products_equal(const HIP_CSR_lMatrix<double> &A,
                                const HIP_CSR_lMatrix<double> &B, int n) {
  std::unique_ptr<lVector<double>> x(init_vector<double>(n, "x", HIP, HIPKERNEL));
  std::unique_ptr<lVector<double>> ya(init_vector<double>(n, "ya", HIP, HIPKERNEL));
  std::unique_ptr<lVector<double>> yb(init_vector<double>(n, "yb", HIP, HIPKERNEL));
  std::vector<double> hx(n), a(n), b(n);
  for (int i = 0; i < n; ++i) hx[i] = std::sin(0.37 * i + 0.1) * (1.0 + (i % 7));
  x->SetBlockValues(0, n, hx.data());
  A.VectorMult(*x, ya.get());
  B.VectorMult(*x, yb.get());
  ya->GetBlockValues(0, n, a.data());
  yb->GetBlockValues(0, n, b.data());
  int diff = 0;
  for (int i = 0; i < n; ++i) diff += a[i] != b[i];
  CHECK(diff == 0);
  A.VectorMultAdd(*x, ya.get());
  B.VectorMultAdd(*x, yb.get());
  ya->GetBlockValues(0, n, a.data());
  yb->GetBlockValues(0, n, b.data());
  diff = 0;
  for (int i = 0; i < n; ++i) diff += a[i] != b[i];
  CHECK(diff == 0);
}

static void test_sell_layout() {
  // 2D 5-point pattern on a 37 x 41 grid (a partial last slice)
  const int nx = 37, ny = 41, n = nx * ny;
  std::vector<int> rp(1, 0), ci;
  std::vector<double> v;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int r = j * nx + i;
      const int nb[5] = {r - nx, r - 1, r, r + 1, r + nx};
      const bool ok[5] = {j > 0, i > 0, true, i < nx - 1, j < ny - 1};
      for (int k = 0; k < 5; ++k)
        if (ok[k]) {
          ci.push_back(nb[k]);
          v.push_back(nb[k] == r ? 4.0 + 0.01 * (r % 13) : -1.0 - 0.001 * (r % 5));
        }
      rp.push_back((int)ci.size());
    }
  const int nnz = (int)ci.size();
  HIP_CSR_lMatrix<double> A(nnz, n, n, "A"), B(nnz, n, n, "B");
  A.SetCSR(rp.data(), ci.data(), v.data());
  B.SetCSR(rp.data(), ci.data(), v.data());
  B.set_use_sell(false);
  CHECK(B.sell_layout() == nullptr);
  sell_products_equal(A, B, n);
  CHECK(A.sell_layout() != nullptr);
  const int zr[] = {0, 5, 700, n - 1};
  A.ZeroRows(zr, 4, 2.5);
  B.ZeroRows(zr, 4, 2.5);
  sell_products_equal(A, B, n);
  const int ar[] = {100, 101}, ac[] = {100, 101};
  const double av[] = {0.5, -0.25, 0.125, 3.0};
  A.add_values(ar, 2, ac, 2, av);
  B.add_values(ar, 2, ac, 2, av);
  sell_products_equal(A, B, n);
  A.Zeros();
  B.Zeros();
  const int sr[] = {7}, sc[] = {7};
  const double sv[] = {9.0};
  A.set_values(sr, 1, sc, 1, sv);
  B.set_values(sr, 1, sc, 1, sv);
  sell_products_equal(A, B, n);
  HIP_CSR_lMatrix<double> C(nnz, n, n, "C");
  C.SetCSR(rp.data(), ci.data(), v.data());
  sell_products_equal(C, C, n);  // builds C's copy
  C.CopyFrom(B);                  // B's values (CSR kernels) into C
  sell_products_equal(C, B, n);
  // a row of 300 entries: the layout is refused and the CSR kernels run
  const int m = 400;
  std::vector<int> rp2(1, 0), ci2;
  std::vector<double> v2;
  for (int r = 0; r < m; ++r) {
    if (r == 3)
      for (int c = 0; c < 300; ++c) ci2.push_back(c);
    else
      ci2.push_back(r);
    rp2.push_back((int)ci2.size());
  }
  for (size_t k = 0; k < ci2.size(); ++k) v2.push_back(1.0 + 0.5 * (double)(k % 11));
  HIP_CSR_lMatrix<double> D((int)ci2.size(), m, m, "D"), E((int)ci2.size(), m, m, "E");
  D.SetCSR(rp2.data(), ci2.data(), v2.data());
  E.SetCSR(rp2.data(), ci2.data(), v2.data());
  E.set_use_sell(false);
  sell_products_equal(D, E, m);
  CHECK(D.sell_layout() == nullptr);
}

static void test_factory_errors() {
  set_fatal_handler([](const std::string &) {});
  bool threw = false;
  try {
    std::unique_ptr<lVector<double>> v(init_vector<double>(4, "cpu", CPU, NAIVE));
  } catch (const LmpError &) {
    threw = true;
  }
  CHECK(threw);
  threw = false;
  try {
    CoupledMatrix<double> A;
    make_3x3(A);
    A.Add(0, 2, 1.0);  // (0,2) is outside the pattern
  } catch (const LmpError &) {
    threw = true;
  }
  CHECK(threw);
  set_fatal_handler(nullptr);
}

int main() {
  SYSTEM sys;
  sys.Platform = HIP;
  sys.device = 0;
  init_platform(sys);
  run("lmp_test VectorMult (TransformFromCOO)", test_lmp_vectormult);
  run("cuda_test VectorMult (dense CSR)", test_dense_csr_vectormult);
  run("cg_test NoPreconditioning", test_cg_noprecond);
  run("cg_test Jacobi", test_cg_jacobi);
  run("cg_test cg() NAIVE CSR", test_cg_local_naive);
  run("cg_test cg() OPENMP CSR (CSR kernels)", test_cg_local_openmp);
  run("cg_test cg() JACOBI (preconditioner output unused)", test_cg_local_jacobi);
  run("Init_SymmetricGaussSeidel: sequential SGS", test_sgs_preconditioner);
  run("AsynchronousIterationGPU solve and smooth", test_async_iteration);
  run("parallel_vector_test dot/axpy/scale/scale_add", test_vector_ops);
  run("parallel_matrix_test vector_mult_diag + diagonalize_rows", test_matrix_tridiagonal);
  run("GMRES / FGMRES(MC-SGS) / MPIR(FP32 CG)", test_gmres_fgmres_mpir);
  run("SELL-64 products vs CSR kernels through value changes", test_sell_layout);
  run("factory and pattern errors", test_factory_errors);
  stop_platform(sys);
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}
