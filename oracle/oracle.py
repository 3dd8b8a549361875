"""ctypes front end of the CPU oracle (oracle/lib/libhflmp_oracle.so).

TEST INFRASTRUCTURE ONLY: the correctness checker and the timed CPU
baseline.  Only the tests, the smoke check and the CPU-baseline leg of the
benchmark import this module; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libhflmp_oracle.so")
_lib = None

PREC_NONE, PREC_JACOBI, PREC_SGS, PREC_MCSGS, PREC_BSGS = 0, 1, 2, 3, 4

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_splitmix64_at.restype = C.c_uint64
        L.or_splitmix64_at.argtypes = [C.c_uint64, C.c_uint64]
        L.or_uniform_pm1.restype = C.c_double
        L.or_uniform_pm1.argtypes = [C.c_uint64, C.c_uint64]
        L.or_gauss.restype = C.c_double
        L.or_gauss.argtypes = [C.c_uint64, C.c_uint64]
        L.or_fe_sizes.argtypes = [C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.or_fe_assemble.argtypes = [C.c_int, C.c_int, C.c_int, _vp,
                                     _i32p, _i32p, _f64p]
        L.or_fe_kappa.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64, _f64p]
        L.or_fe_rhs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, _f64p]
        L.or_fe_boundary_mask.argtypes = [C.c_int, C.c_int, C.c_int, _u8p]
        L.or_fe_element.argtypes = [C.c_int, C.c_int, C.c_int, _f64p]
        L.or_coo_to_csr_d.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p,
                                      _i32p, _i32p, _f64p]
        L.or_csr_diag_ptr.argtypes = [C.c_int, _i32p, _i32p, _i32p]
        L.or_csr_zero_rows_d.argtypes = [_i32p, C.c_int, _i32p, _i32p, _f64p,
                                         C.c_double]
        L.or_multicoloring.argtypes = [C.c_int, _i32p, _i32p, C.POINTER(C.c_int),
                                       _i32p, _i32p, _i32p]
        L.or_set_solver_threads.argtypes = [C.c_int]
        L.or_get_solver_threads.restype = C.c_int
        L.or_set_blas_threads.argtypes = [C.c_int]
        L.or_get_blas_threads.restype = C.c_int
        L.or_cast_d2s.argtypes = [C.c_int, _f64p, _f32p]
        L.or_cast_s2d.argtypes = [C.c_int, _f32p, _f64p]
        for s, fp, ct in (("d", _f64p, C.c_double), ("s", _f32p, C.c_float)):
            g = lambda n: getattr(L, "or_%s_%s" % (n, s))
            for n in ("spmv", "spmv_add", "spmv_nodiag"):
                g(n).argtypes = [C.c_int, _i32p, _i32p, fp, fp, fp]
            g("spmv_omp").argtypes = [C.c_int, _i32p, _i32p, fp, fp, fp, C.c_int]
            g("residual_sub").argtypes = [C.c_int, _i32p, _i32p, fp, fp, fp, fp]
            g("jacobi_damped").argtypes = [C.c_int, _i32p, _i32p, fp, fp, ct, fp, fp, fp]
            g("dot").argtypes = [C.c_int, fp, fp]
            g("dot").restype = ct
            for n in ("nrm2", "norm1", "normmax"):
                g(n).argtypes = [C.c_int, fp]
                g(n).restype = ct
            for n in ("axpy", "scaleadd"):
                g(n).argtypes = [C.c_int, ct, fp, fp]
            g("scale").argtypes = [C.c_int, ct, fp]
            g("ewmult").argtypes = [C.c_int, fp, fp]
            g("invdiag").argtypes = [C.c_int, _i32p, _i32p, fp, fp]
            g("jacobi_apply").argtypes = [C.c_int, fp, fp, fp]
            g("sgs_apply").argtypes = [C.c_int, _i32p, _i32p, fp, fp, fp]
            g("block_sgs_apply").argtypes = [C.c_int, _i32p, _i32p, fp, C.c_int, fp, fp]
            g("mcsgs_apply").argtypes = [C.c_int, _i32p, _i32p, fp, C.c_int,
                                         _i32p, _i32p, _i32p, ct, fp, fp]
            g("cg").argtypes = [C.c_int, _i32p, _i32p, fp, C.c_int, _vp,
                                C.c_int, _vp, _vp, _vp, fp, fp, C.c_int,
                                C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_int), C.POINTER(C.c_double),
                                C.POINTER(C.c_double), _vp]
            g("async_jacobi").argtypes = [C.c_int, _i32p, _i32p, fp, fp, fp, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, ct, C.c_int, C.c_double,
                                          C.c_double, C.c_double, C.POINTER(C.c_int),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
            g("bicgstab").argtypes = [C.c_int, _i32p, _i32p, fp, C.c_int, _vp, C.c_int, _vp,
                                      _vp, _vp, fp, fp, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
            g("fe_assemble_cells").argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _vp, _vp,
                                               _vp, _i32p, _vp, _vp]
            g("stencil_vmadd").argtypes = [C.c_int, C.c_int, _i32p, fp, _vp, fp, C.c_int,
                                           _vp, ct, fp, ct, fp]
            g("gmres").argtypes = [C.c_int, _i32p, _i32p, fp, C.c_int, _vp,
                                   C.c_int, _vp, _vp, _vp, C.c_int, C.c_int,
                                   fp, fp, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_mpir_ds.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f32p, _f64p,
                                 _f64p, C.c_int, C.c_double, C.c_double,
                                 C.c_int, C.c_double, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _sfx(a):
    return "d" if a.dtype == np.float64 else "s"


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- generator ------------------------------------------------------------
def fe_sizes(order, dim, n):
    N, Z = C.c_int64(), C.c_int64()
    if lib().or_fe_sizes(order, dim, n, C.byref(N), C.byref(Z)):
        raise ValueError("unsupported element")
    return N.value, Z.value


def fe_kappa(dim, n, sigma, seed):
    k = np.empty(n ** dim, np.float64)
    lib().or_fe_kappa(dim, n, sigma, seed, k)
    return k


def fe_system(order, dim, n, kappa=None):
    N, Z = fe_sizes(order, dim, n)
    rp = np.empty(N + 1, np.int32)
    ci = np.empty(Z, np.int32)
    v = np.empty(Z, np.float64)
    kap = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
    rc = lib().or_fe_assemble(order, dim, n, _ptr(kap), rp, ci, v)
    if rc:
        raise RuntimeError("or_fe_assemble failed %d" % rc)
    return rp, ci, v


def fe_rhs(order, dim, n, seed):
    N, _ = fe_sizes(order, dim, n)
    b = np.empty(N, np.float64)
    lib().or_fe_rhs(order, dim, n, seed, b)
    return b


def fe_boundary_mask(order, dim, n):
    N, _ = fe_sizes(order, dim, n)
    m = np.empty(N, np.uint8)
    lib().or_fe_boundary_mask(order, dim, n, m)
    return m.astype(bool)


def fe_element(order, dim, n):
    """The generator's element stiffness (h included; the cell factor is
    kappa), local nodes lexicographic, x fastest."""
    q = order + 1
    nloc = q ** dim
    E = np.empty(nloc * nloc, np.float64)
    lib().or_fe_element(order, dim, n, E)
    return E.reshape(nloc, nloc)


def stencil_vmadd(order, dim, nc, E, cell_factor, fixed, x, y=None, alpha=1.0, beta=0.0,
                  cell_lm=None):
    """out = beta out + alpha A x by the matrix-free cell loop of
    StencilOperator::VectorMultAdd (stencil_operator.h:153-330); E holds one
    or more nloc x nloc local matrices, cell_lm picks one per cell."""
    dt = np.asarray(x).dtype
    nc3 = np.array(list(nc) + [1] * (3 - len(nc)), np.int32)
    E = np.ascontiguousarray(E, dt).ravel()
    f = np.ascontiguousarray(cell_factor, dt)
    fx = np.ascontiguousarray(fixed, np.int32)
    lm = None if cell_lm is None else np.ascontiguousarray(cell_lm, np.uint8)
    x = np.ascontiguousarray(x, dt)
    out = np.zeros_like(x) if y is None else np.array(y, dt, copy=True)
    getattr(lib(), "or_stencil_vmadd_" + _sfx(x))(order, dim, nc3, E, _ptr(lm), f, len(fx),
                                                  _ptr(fx), alpha, x, beta, out)
    return out


def splitmix64_at(seed, i):
    return lib().or_splitmix64_at(seed, i)


def uniform_pm1(seed, i):
    return lib().or_uniform_pm1(seed, i)


# ---- CSR utilities ----------------------------------------------------------
def coo_to_csr(nrow, ri, ci, v):
    ri = np.ascontiguousarray(ri, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    rp = np.empty(nrow + 1, np.int32)
    col = np.empty(len(ci), np.int32)
    val = np.empty(len(ci), np.float64)
    lib().or_coo_to_csr_d(nrow, len(ci), ri, ci, v, rp, col, val)
    return rp, col, val


def diag_ptr(rp, ci):
    d = np.empty(len(rp) - 1, np.int32)
    missing = lib().or_csr_diag_ptr(len(rp) - 1, rp, ci, d)
    return d, missing


def zero_rows(rows, rp, ci, v, alpha=1.0):
    rows = np.ascontiguousarray(rows, np.int32)
    lib().or_csr_zero_rows_d(rows, len(rows), rp, ci, v, alpha)


def multicoloring(rp, ci):
    n = len(rp) - 1
    maxlen = int(np.max(np.diff(rp))) if n else 0
    csz = np.zeros(maxlen + 2, np.int32)
    perm = np.empty(n, np.int32)
    color = np.empty(n, np.int32)
    nc = C.c_int()
    lib().or_multicoloring(n, rp, ci, C.byref(nc), csz, perm, color)
    return nc.value, csz[: nc.value].copy(), perm, color


# ---- kernels -----------------------------------------------------------------
def spmv(rp, ci, v, x):
    y = np.empty(len(rp) - 1, v.dtype)
    getattr(lib(), "or_spmv_" + _sfx(v))(len(rp) - 1, rp, ci, v, x, y)
    return y


def spmv_add(rp, ci, v, x, y):
    getattr(lib(), "or_spmv_add_" + _sfx(v))(len(rp) - 1, rp, ci, v, x, y)
    return y


def spmv_nodiag(rp, ci, v, x):
    y = np.empty(len(rp) - 1, v.dtype)
    getattr(lib(), "or_spmv_nodiag_" + _sfx(v))(len(rp) - 1, rp, ci, v, x, y)
    return y


def spmv_omp(rp, ci, v, x, y, nthreads):
    getattr(lib(), "or_spmv_omp_" + _sfx(v))(len(rp) - 1, rp, ci, v, x, y,
                                             nthreads)
    return y


def dot(x, y):
    return getattr(lib(), "or_dot_" + _sfx(x))(len(x), x, y)


def nrm2(x):
    return getattr(lib(), "or_nrm2_" + _sfx(x))(len(x), x)


def norm1(x):
    return getattr(lib(), "or_norm1_" + _sfx(x))(len(x), x)


def normmax(x):
    return getattr(lib(), "or_normmax_" + _sfx(x))(len(x), x)


def axpy(a, x, y):
    getattr(lib(), "or_axpy_" + _sfx(x))(len(x), a, x, y)
    return y


def scaleadd(a, x, y):
    getattr(lib(), "or_scaleadd_" + _sfx(x))(len(x), a, x, y)
    return y


def scale(a, y):
    getattr(lib(), "or_scale_" + _sfx(y))(len(y), a, y)
    return y


def ewmult(x, y):
    getattr(lib(), "or_ewmult_" + _sfx(x))(len(x), x, y)
    return y


def invdiag(rp, ci, v):
    d = np.empty(len(rp) - 1, v.dtype)
    getattr(lib(), "or_invdiag_" + _sfx(v))(len(rp) - 1, rp, ci, v, d)
    return d


def jacobi_apply(invd, r):
    z = np.empty_like(r)
    getattr(lib(), "or_jacobi_apply_" + _sfx(r))(len(r), invd, r, z)
    return z


def sgs_apply(rp, ci, v, r):
    z = np.empty_like(r)
    getattr(lib(), "or_sgs_apply_" + _sfx(v))(len(r), rp, ci, v, r, z)
    return z


def block_sgs_apply(rp, ci, v, num_blocks, r):
    z = np.empty_like(r)
    getattr(lib(), "or_block_sgs_apply_" + _sfx(v))(len(r), rp, ci, v, int(num_blocks), r, z)
    return z


def mcsgs_apply(rp, ci, v, coloring, r, omega=1.0):
    nc, csz, perm, color = coloring
    z = np.empty_like(r)
    getattr(lib(), "or_mcsgs_apply_" + _sfx(v))(len(r), rp, ci, v, nc, csz,
                                                perm, color, omega, r, z)
    return z


def cast_d2s(x):
    y = np.empty(len(x), np.float32)
    lib().or_cast_d2s(len(x), x, y)
    return y


def cast_s2d(x):
    y = np.empty(len(x), np.float64)
    lib().or_cast_s2d(len(x), x, y)
    return y


# ---- solvers -------------------------------------------------------------------
def _prec_args(rp, ci, v, prec, coloring, num_blocks=1):
    invd = invdiag(rp, ci, v) if prec == PREC_JACOBI else None
    if prec == PREC_BSGS:
        return invd, int(num_blocks), None, None, None
    if prec == PREC_MCSGS:
        if coloring is None:
            coloring = multicoloring(rp, ci)
        nc, csz, perm, color = coloring
        return invd, nc, csz, perm, color
    return invd, 0, None, None, None


class _SolverThreads:
    """For the duration of a call: the solvers' SpMV row loop on `n` threads
    (CPUopenmp_CSR_lMatrix::VectorMult; the same per-row sums, so the same
    results as one thread) and the CG's vector operations on `nb` threads
    (1: sequential, the reference's CPU path; more: a faster checker whose
    dot products sum per-thread blocks, so not the sequential bits)."""

    def __init__(self, n, nb=1):
        self.n, self.nb = max(int(n), 1), max(int(nb), 1)

    def __enter__(self):
        self.prev = (lib().or_get_solver_threads(), lib().or_get_blas_threads())
        lib().or_set_solver_threads(self.n)
        lib().or_set_blas_threads(self.nb)

    def __exit__(self, *exc):
        lib().or_set_solver_threads(self.prev[0])
        lib().or_set_blas_threads(self.prev[1])


def cg(rp, ci, v, b, x0=None, prec=PREC_JACOBI, maxit=10000, atol=0.0,
       rtol=1e-8, dtol=1e6, coloring=None, history=False, num_blocks=1, threads=1,
       blas_threads=1):
    """CG::Solve with the reference semantics; returns (x, info dict).
    num_blocks: the block count of PREC_BSGS; threads: threads of the SpMV
    row loop (the CPUopenmp platform's matrix; same results); blas_threads:
    threads of the vector operations (see _SolverThreads)."""
    with _SolverThreads(threads, blas_threads):
        return _cg(rp, ci, v, b, x0, prec, maxit, atol, rtol, dtol, coloring, history,
                   num_blocks)


def _cg(rp, ci, v, b, x0, prec, maxit, atol, rtol, dtol, coloring, history, num_blocks):
    n = len(rp) - 1
    x = np.zeros(n, v.dtype) if x0 is None else np.array(x0, v.dtype)
    invd, nc, csz, perm, color = _prec_args(rp, ci, v, prec, coloring, num_blocks)
    it, r0, rf = C.c_int(), C.c_double(), C.c_double()
    hist = np.zeros(maxit + 1, np.float64) if history else None
    st = getattr(lib(), "or_cg_" + _sfx(v))(
        n, rp, ci, v, prec, _ptr(invd), nc, _ptr(csz), _ptr(perm), _ptr(color),
        np.ascontiguousarray(b, v.dtype), x, maxit, atol, rtol, dtol,
        C.byref(it), C.byref(r0), C.byref(rf), _ptr(hist))
    info = dict(status=st, iters=it.value, res_init=r0.value, res_final=rf.value)
    if history:
        info["history"] = hist[: it.value + 1].copy()
    return x, info


def async_jacobi(rp, ci, v, b, x0=None, block_rows=256, n_global=1, n_block=1, n_inner=5,
                 mode=0, w=1.0, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6):
    """AsynchronousIterationGPU made deterministic (blocks of rows, outside
    couplings from each launch's input, synchronous inner sweeps); mode 0/1
    solve (normal/damped), 2/3 smooth.  Returns (x, info dict)."""
    n = len(rp) - 1
    x = np.zeros(n, v.dtype) if x0 is None else np.array(x0, v.dtype)
    it, r0, rf = C.c_int(), C.c_double(), C.c_double()
    st = getattr(lib(), "or_async_jacobi_" + _sfx(v))(
        n, rp, ci, v, np.ascontiguousarray(b, v.dtype), x, int(block_rows), int(n_global),
        int(n_block), int(n_inner), int(mode), w, int(maxit), atol, rtol, dtol, C.byref(it),
        C.byref(r0), C.byref(rf))
    return x, dict(status=st, iters=it.value, res_init=r0.value, res_final=rf.value)


def bicgstab(rp, ci, v, b, x0=None, prec=PREC_NONE, maxit=10000, atol=0.0, rtol=1e-8,
             dtol=1e6, coloring=None):
    """BiCGSTAB<LAD>::Solve with the reference semantics (right
    preconditioning); returns (x, info dict)."""
    n = len(rp) - 1
    x = np.zeros(n, v.dtype) if x0 is None else np.array(x0, v.dtype)
    invd, nc, csz, perm, color = _prec_args(rp, ci, v, prec, coloring)
    it, r0, rf = C.c_int(), C.c_double(), C.c_double()
    st = getattr(lib(), "or_bicgstab_" + _sfx(v))(
        n, rp, ci, v, prec, _ptr(invd), nc, _ptr(csz), _ptr(perm), _ptr(color),
        np.ascontiguousarray(b, v.dtype), x, maxit, atol, rtol, dtol, C.byref(it),
        C.byref(r0), C.byref(rf))
    return x, dict(status=st, iters=it.value, res_init=r0.value, res_final=rf.value)


def gmres(rp, ci, v, b, x0=None, prec=PREC_NONE, basis=100, flexible=False,
          maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6, coloring=None):
    n = len(rp) - 1
    x = np.zeros(n, v.dtype) if x0 is None else np.array(x0, v.dtype)
    invd, nc, csz, perm, color = _prec_args(rp, ci, v, prec, coloring)
    it, r0, rf = C.c_int(), C.c_double(), C.c_double()
    st = getattr(lib(), "or_gmres_" + _sfx(v))(
        n, rp, ci, v, prec, _ptr(invd), nc, _ptr(csz), _ptr(perm), _ptr(color),
        basis, int(flexible), np.ascontiguousarray(b, v.dtype), x, maxit, atol,
        rtol, dtol, C.byref(it), C.byref(r0), C.byref(rf))
    return x, dict(status=st, iters=it.value, res_init=r0.value,
                   res_final=rf.value)


def mpir(rp, ci, vd, b, x0=None, max_outer=50, atol=0.0, rtol=1e-10,
         inner_maxit=10000, inner_rtol=1e-4, threads=1, blas_threads=1):
    with _SolverThreads(threads, blas_threads):
        return _mpir(rp, ci, vd, b, x0, max_outer, atol, rtol, inner_maxit, inner_rtol)


def _mpir(rp, ci, vd, b, x0, max_outer, atol, rtol, inner_maxit, inner_rtol):
    n = len(rp) - 1
    x = np.zeros(n, np.float64) if x0 is None else np.array(x0, np.float64)
    vs = vd.astype(np.float32)
    oi, ii, r0, rf = C.c_int(), C.c_int(), C.c_double(), C.c_double()
    st = lib().or_mpir_ds(n, rp, ci, vd, vs, np.ascontiguousarray(b), x,
                          max_outer, atol, rtol, inner_maxit, inner_rtol,
                          C.byref(oi), C.byref(ii), C.byref(r0), C.byref(rf))
    return x, dict(status=st, outer=oi.value, inner_total=ii.value,
                   res_init=r0.value, res_final=rf.value)


# ---- MatrixMarket files ----------------------------------------------------------
def mm_read(path):
    """CPU_COO_lMatrix::ReadFile (lmatrix_coo_cpu.cc:123-243): banner, size
    line, 1-based "row col value" entries; a symmetric file adds the mirror
    of every off-diagonal entry right after it.  Skew-symmetric mirrors get
    the opposite sign (the MatrixMarket definition; the reference keeps only
    the stored triangle).  Returns (nrow, ncol, rows, cols, vals)."""
    with open(path) as f:
        banner = f.readline().split()
        if len(banner) != 5 or banner[0].lower() != "%%matrixmarket":
            raise ValueError("no MatrixMarket banner")
        obj, fmt, field, sym = (t.lower() for t in banner[1:])
        if obj != "matrix" or fmt != "coordinate" or field not in ("real", "integer", "double"):
            raise ValueError("unsupported MatrixMarket type")
        if sym not in ("general", "symmetric", "skew-symmetric"):
            raise ValueError("unsupported symmetry")
        lines = (ln for ln in f if ln.strip() and not ln.lstrip().startswith("%"))
        m, n, nz = (int(t) for t in next(lines).split()[:3])
        if sym != "general" and m != n:
            raise ValueError("a symmetric matrix must be square")
        rows, cols, vals = [], [], []
        for _ in range(nz):
            t = next(lines).split()
            i, j, v = int(t[0]) - 1, int(t[1]) - 1, float(t[2])
            if not (0 <= i < m and 0 <= j < n):
                raise ValueError("entry out of range")
            rows.append(i); cols.append(j); vals.append(v)
            if sym != "general" and i != j:
                rows.append(j); cols.append(i); vals.append(-v if sym == "skew-symmetric" else v)
    return (m, n, np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals, np.float64))


def coo_to_csr_merge(nrow, rows, cols, vals):
    """Sorted CSR with duplicates summed in their order of appearance (the
    product's structure build; coo_2_csr, lmatrix_formats.h:320-456,
    sorts the columns of each row)."""
    order = np.lexsort((np.arange(len(rows)), cols, rows))
    r, c, v = rows[order], cols[order], vals[order]
    rp = np.zeros(nrow + 1, np.int32)
    ci, vv = [], []
    for k in range(len(r)):
        if k > 0 and r[k] == r[k - 1] and c[k] == c[k - 1]:
            vv[-1] = vv[-1] + v[k]
        else:
            ci.append(c[k]); vv.append(v[k]); rp[r[k] + 1] += 1
    return np.cumsum(rp).astype(np.int32), np.array(ci, np.int32), np.array(vv, np.float64)


def fe_assemble_cells(nnode, dofs, E=None, factor=None, cell_mats=None, dtype=np.float64):
    """CSR of a general mesh (cell -> dof table `dofs`, ncell x nloc): the
    host assembly of the reference (pattern, then add_values cell by
    cell)."""
    dofs = np.ascontiguousarray(dofs, np.int32)
    ncell, nloc = dofs.shape
    fn = getattr(lib(), "or_fe_assemble_cells_" + ("d" if dtype == np.float64 else "s"))
    cm = None if cell_mats is None else np.ascontiguousarray(cell_mats, dtype)
    e = None if E is None else np.ascontiguousarray(E, dtype)
    f = None if factor is None else np.ascontiguousarray(factor, dtype)
    rp = np.zeros(nnode + 1, np.int32)
    if fn(nnode, ncell, nloc, dofs, _ptr(cm), _ptr(e), _ptr(f), rp, None, None):
        raise ValueError("dof outside the mesh")
    ci = np.empty(rp[-1], np.int32)
    v = np.empty(rp[-1], dtype)
    fn(nnode, ncell, nloc, dofs, _ptr(cm), _ptr(e), _ptr(f), rp, _ptr(ci), _ptr(v))
    return rp, ci, v


def fe_cell_dofs(order, dim, n):
    """Cell -> dof table of the structured mesh (cells and their local
    nodes lexicographic, x fastest)."""
    p = order * n + 1
    q = order + 1
    c = np.arange(n)
    loc = np.arange(q)
    if dim == 2:
        cy, cx = np.meshgrid(c, c, indexing="ij")
        ly, lx = np.meshgrid(loc, loc, indexing="ij")
        gx = order * cx.reshape(-1, 1) + lx.reshape(1, -1)
        gy = order * cy.reshape(-1, 1) + ly.reshape(1, -1)
        return (gx + p * gy).astype(np.int32)
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    lz, ly, lx = np.meshgrid(loc, loc, loc, indexing="ij")
    gx = order * cx.reshape(-1, 1) + lx.reshape(1, -1)
    gy = order * cy.reshape(-1, 1) + ly.reshape(1, -1)
    gz = order * cz.reshape(-1, 1) + lz.reshape(1, -1)
    return (gx + p * (gy + p * gz)).astype(np.int32)


# ---- geometric multigrid -------------------------------------------------------------
def residual_sub(rp, ci, v, b, x):
    """kernel_gmg_compute_residual_no_offdiag (cuda_gmg_kernel.h:46-64)."""
    r = np.empty(len(rp) - 1, v.dtype)
    getattr(lib(), "or_residual_sub_" + _sfx(v))(len(rp) - 1, rp, ci, v, b, x, r)
    return r


def jacobi_damped(rp, ci, v, invd, w, b, x):
    """One synchronous sweep of kernel_async_inner_damped_iter_no_offdiag
    (cuda_async_iter_kernels.h:116-144)."""
    out = np.empty(len(rp) - 1, v.dtype)
    getattr(lib(), "or_jacobi_damped_" + _sfx(v))(len(rp) - 1, rp, ci, v, invd, w, b, x, out)
    return out


def _lv_residual(L, b, x):
    """Level residual: the GMG kernel's term-by-term form for a CSR level;
    VectorMultAdd(-1, x, 1, r = b) of the stencil for a matrix-free one
    (gmg.h:1110-1119)."""
    if "S" in L:
        return stencil_vmadd(*L["S"], x, y=b, alpha=-1.0, beta=1.0)
    return residual_sub(*L["A"], b, x)


def _lv_sweep(L, w, b, x):
    """Damped Jacobi (cuda_async_iter_kernels.h:116-144) on a CSR level;
    Richardson with the Jacobi preconditioner (richardson.h:136-251,
    x + w (r .* invd)) on a matrix-free one."""
    if "S" in L:
        return x + w * (_lv_residual(L, b, x) * L["invd"])
    return jacobi_damped(*L["A"], L["invd"], w, b, x)


def _lv_first(L, w, b):
    """The sweep from x = 0 without the product with zeros."""
    return w * (b * L["invd"]) if "S" in L else w * L["invd"] * b


def dense_matvec(C, b, parts=8):
    """y = C b as the dense coarse solve computes it: the columns cut into
    `parts` blocks of ceil(n / parts), each block's part of a row summed from
    0 in ascending column order (the CSR row loop over the block), and the
    row's sum p_0 + p_1 + ... in block order (an empty block adds 0)."""
    C = np.ascontiguousarray(C, np.float64)
    b = np.asarray(b, np.float64)
    n = C.shape[0]
    cb = (n + parts - 1) // parts
    y = None
    for k in range(parts):
        j0, j1 = min(k * cb, n), min((k + 1) * cb, n)
        w = j1 - j0
        if w > 0:
            rp = np.arange(0, n * w + 1, w, dtype=np.int32)
            ci = np.tile(np.arange(w, dtype=np.int32), n)
            p = spmv(rp, ci, np.ascontiguousarray(C[:, j0:j1]).ravel(),
                     np.ascontiguousarray(b[j0:j1]))
        else:
            p = np.zeros(n)
        y = p if y is None else y + p
    return y


def mg_cycle(levels, b, pre=1, post=1, coarse=4, omega=0.7, coarse_omega=0.7, l=0):
    """One V-cycle from zero (GMGBase::SolveImplLevel, gmg.h:1070-1187, with
    the GPU kernels of cuda_gmg_kernel.h).  levels[l] is a dict with the
    operator A=(rp, ci, v) or, matrix-free on level 0, S=(order, dim, nc,
    E, factor, fixed) for stencil_vmadd; invd; above the coarsest level R
    and P as (rp, ci, v); on the coarsest level optionally Cinv, a dense
    inverse that replaces its sweeps (x = Cinv b)."""
    L = levels[l]
    last = l == len(levels) - 1
    if last and "Cinv" in L:
        return dense_matvec(L["Cinv"], b)
    w = coarse_omega if last else omega
    x = _lv_first(L, w, b)
    for _ in range((coarse if last else pre) - 1):
        x = _lv_sweep(L, w, b, x)
    if last:
        return x
    r = _lv_residual(L, b, x)
    bc = spmv(*L["R"], r)
    xc = mg_cycle(levels, bc, pre, post, coarse, omega, coarse_omega, l + 1)
    spmv_add(*L["P"], xc, x)
    for _ in range(post):
        x = _lv_sweep(L, w, b, x)
    return x


def cg_prec(A, b, prec, x0=None, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6):
    """The preconditioned branch of or_cg (CG::SolvePrecond, cg.h:165-246)
    with z = prec(r) for any callable preconditioner; A is (rp, ci, v) or a
    callable y = A x."""
    op = A if callable(A) else (lambda x: spmv(*A, x))
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    r = op(x)
    scaleadd(-1.0, np.ascontiguousarray(b, np.float64), r)
    res0 = res = float(np.sqrt(dot(r, r)))
    hist = [res]

    def check(it, res):
        if res <= atol or res < rtol * res0:
            return 0
        if res > dtol * res0:
            return 2
        if it >= maxit:
            return 1
        return -1

    it = 0
    conv = check(0, res)
    z = prec(r)
    p = z.copy()
    beta = dot(r, z)
    while conv == -1:
        it += 1
        Ap = op(p)
        alpha = beta / dot(Ap, p)
        axpy(alpha, p, x)
        axpy(-alpha, Ap, r)
        res = float(np.sqrt(dot(r, r)))
        hist.append(res)
        conv = check(it, res)
        if conv != -1:
            break
        z = prec(r)
        gamma = dot(r, z)
        scaleadd(gamma / beta, z, p)
        beta = gamma
    return x, dict(status=conv, iters=it, res_init=res0, res_final=res, history=hist)


def mg_solve(levels, b, x0=None, maxit=1000, atol=0.0, rtol=1e-8, dtol=1e6, **cycle):
    """Multigrid as the solver (GMGBase::SolveImpl, gmg.h:979-1066) in the
    correction form x = x + cycle(r), r = b - A x term by term; the
    stopping test IterateControl::Check on ||r|| with res0 = ||b - A x0||."""
    x = np.zeros(len(b)) if x0 is None else np.array(x0, np.float64)
    r = _lv_residual(levels[0], b, x)
    res0 = res = float(np.sqrt(dot(r, r)))

    def check(it, res):
        if res <= atol or res < rtol * res0:
            return 0
        if res > dtol * res0:
            return 2
        if it >= maxit:
            return 1
        return -1

    it = 0
    conv = check(0, res)
    while conv == -1:
        it += 1
        x = x + mg_cycle(levels, r, **cycle)
        r = _lv_residual(levels[0], b, x)
        res = float(np.sqrt(dot(r, r)))
        conv = check(it, res)
    return x, dict(status=conv, iters=it, res_init=res0, res_final=res)


def mg_levels(A, R, P, cell_mats, cell_dofs, fixed):
    """Host hierarchy for mg_cycle: level 0 is A = (rp, ci, v); level l > 0
    is assembled cell by cell (fe_assemble_cells) from cell_mats[l-1] over
    the cell -> dof table cell_dofs[l-1], with unit Dirichlet rows at
    fixed[l-1]; R[l], P[l] as (rp, ci, v)."""
    levels = [dict(A=A, invd=invdiag(*A))]
    for l in range(1, len(R) + 1):
        rpc, cic, vc = fe_assemble_cells(len(R[l - 1][0]) - 1, cell_dofs[l - 1],
                                         cell_mats=cell_mats[l - 1])
        zero_rows(fixed[l - 1], rpc, cic, vc)
        levels[l - 1]["R"] = R[l - 1]
        levels[l - 1]["P"] = P[l - 1]
        levels.append(dict(A=(rpc, cic, vc), invd=invdiag(rpc, cic, vc)))
    return levels
