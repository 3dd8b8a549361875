"""Python binding of the gfx950 lmp platform (libhflmp_hip.so) and the
structured FE generator (libhflmp_gen.so).

A thin ctypes layer over the C ABI in include/hflmp_hip.h: device buffers,
the CSR matrix arrays, the BLAS-1 / SpMV / Jacobi entry points and the
device-resident solvers.  There is no CPU fallback: if the native library
is missing or no GPU is visible, constructing a Context raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libhflmp_hip.so")
GEN_LIB = os.path.join(LIB_DIR, "libhflmp_gen.so")
LMP_LIB = os.path.join(LIB_DIR, "libhflmp_lmp.so")

HFLMP_OK = 0
STATUS = {0: "converged", 1: "max iterations", 2: "diverged", 3: "breakdown"}

_hip = None
_gen = None


class HflmpError(RuntimeError):
    pass


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("res_init", C.c_double),
                ("res_final", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MpInfo(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters", C.c_int), ("status", C.c_int),
                ("res_init", C.c_double), ("res_final", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _vp(x):
    return C.c_void_p(x)


def hip_lib():
    """Load libhflmp_hip.so (raises if it was not built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise HflmpError("%s not built: run __graft_entry__.build()" % HIP_LIB)
        L = C.CDLL(HIP_LIB)
        vp, i, sz, d, f = C.c_void_p, C.c_int, C.c_size_t, C.c_double, C.c_float
        L.hflmp_ctx_last_error.restype = C.c_char_p
        L.hflmp_ctx_last_error.argtypes = [vp]
        L.hflmp_ctx_create.argtypes = [i, C.POINTER(vp)]
        for n in ("hflmp_ctx_destroy", "hflmp_ctx_sync"):
            getattr(L, n).argtypes = [vp]
        L.hflmp_device_count.argtypes = [C.POINTER(i)]
        L.hflmp_malloc.argtypes = [vp, sz, C.POINTER(vp)]
        L.hflmp_free.argtypes = [vp, vp]
        for n in ("h2d", "d2h", "d2d"):
            getattr(L, "hflmp_memcpy_" + n).argtypes = [vp, vp, vp, sz]
        L.hflmp_fill_zero.argtypes = [vp, vp, sz]
        L.hflmp_event_create.argtypes = [vp, C.POINTER(vp)]
        L.hflmp_event_record.argtypes = [vp, vp]
        L.hflmp_event_elapsed_ms.argtypes = [vp, vp, vp, C.POINTER(C.c_float)]
        L.hflmp_event_destroy.argtypes = [vp, vp]
        L.hflmp_pcg_workspace_bytes.restype = sz
        L.hflmp_pcg_workspace_bytes.argtypes = [i, i, i]
        for s, ct in (("f64", d), ("f32", f)):
            g = lambda n: getattr(L, "hflmp_%s_%s" % (n, s))
            g("dot").argtypes = [vp, i, vp, vp, C.POINTER(ct)]
            g("dot_dev").argtypes = [vp, i, vp, vp, vp]
            for n in ("nrm2", "norm1", "normmax"):
                g(n).argtypes = [vp, i, vp, C.POINTER(ct)]
            for n in ("argmax_abs", "argmin_abs", "count_nan"):
                g(n).argtypes = [vp, i, vp, C.POINTER(i)]
            for n in ("axpy", "scaleadd"):
                g(n).argtypes = [vp, i, ct, vp, vp]
            g("scal").argtypes = [vp, i, ct, vp]
            for n in ("ewmul", "copy"):
                g(n).argtypes = [vp, i, vp, vp]
            g("rot").argtypes = [vp, i, vp, vp, ct, ct]
            g("jacobi_apply").argtypes = [vp, i, vp, vp, vp]
            for n in ("gather", "scatter", "scatter_add"):
                g(n).argtypes = [vp, i, vp, vp, vp]
            for n in ("csr_spmv", "csr_spmv_add", "csr_spmv_nodiag"):
                g(n).argtypes = [vp, i, i, i, vp, vp, vp, vp, vp]
            g("csr_extract_invdiag").argtypes = [vp, i, vp, vp, vp, vp, C.POINTER(i)]
            g("pcg_jacobi").argtypes = [vp, i, i, vp, vp, vp, vp, vp, vp, i, d, d, d,
                                        vp, sz, C.POINTER(SolveInfo)]
        L.hflmp_async_workspace_bytes.restype = sz
        L.hflmp_async_workspace_bytes.argtypes = [i, i]
        for s in ("f64", "f32"):
            getattr(L, "hflmp_async_jacobi_" + s).argtypes = [
                vp, i, i, vp, vp, vp, vp, vp, vp, i, d, d, d, vp, sz, C.POINTER(SolveInfo)]
        L.hflmp_batched_mpcg_workspace_bytes.restype = sz
        L.hflmp_batched_mpcg_workspace_bytes.argtypes = [vp, i, i]
        L.hflmp_batched_mpcg_jacobi.argtypes = [vp, i, i, i, vp, vp, vp, vp, vp, vp, vp, i, d, d,
                                                i, d, vp, sz, vp, vp, vp, vp, vp, C.POINTER(d)]
        L.hflmp_batched_mpcg_jacobi_plan.argtypes = [vp, vp, i, i, i, vp, vp, vp, vp, vp, vp,
                                                     vp, i, d, d, i, d, vp, sz, vp, vp, vp, vp,
                                                     vp, C.POINTER(d)]
        L.hflmp_spmm_plan_create.argtypes = [vp, i, i, i, vp, vp, i, i, i, vp]
        L.hflmp_spmm_plan_destroy.argtypes = [vp]
        L.hflmp_spmm_plan_destroy.restype = None
        L.hflmp_spmm_plan_info.argtypes = [vp, C.POINTER(i), C.POINTER(i),
                                           C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.hflmp_spmm_plan_geometry.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        for s in ("f64", "f32"):
            getattr(L, "hflmp_plan_spmm_" + s).argtypes = [vp, vp, vp, vp, vp, vp, i]
        for s in ("f64", "f32"):
            getattr(L, "hflmp_csr_spmm_" + s).argtypes = [vp, i, i, i, i, vp, vp, vp, vp, vp]
            getattr(L, "hflmp_batched_pcg_jacobi_" + s).argtypes = [
                vp, i, i, i, vp, vp, vp, vp, vp, vp, i, d, d, d, vp, sz, vp, vp, vp, vp,
                C.POINTER(d)]
            getattr(L, "hflmp_batched_pcg_jacobi_plan_" + s).argtypes = [
                vp, vp, i, i, i, vp, vp, vp, vp, vp, vp, i, d, d, d, vp, sz, vp, vp, vp, vp,
                C.POINTER(d)]
        for s in ("f64", "f32"):
            getattr(L, "hflmp_sell_create_" + s).argtypes = [vp, i, i, i, vp, vp, vp, i, vp]
            getattr(L, "hflmp_sell_set_values_" + s).argtypes = [vp, vp, vp, vp]
            for nm in ("hflmp_sell_spmv_", "hflmp_sell_spmv_add_"):
                getattr(L, nm + s).argtypes = [vp, vp, vp, vp]
            getattr(L, "hflmp_pcg_jacobi_sell_" + s).argtypes = [vp, vp, vp, vp, vp, i, d, d, d,
                                                                  vp, sz, vp]
        L.hflmp_sell_get_info.argtypes = [vp, vp]
        for s in ("f64", "f32"):
            getattr(L, "hflmp_stencil_create_" + s).argtypes = [vp, vp, vp, vp, vp, i, vp, vp]
            getattr(L, "hflmp_stencil_set_factors_" + s).argtypes = [vp, vp, vp]
            getattr(L, "hflmp_stencil_apply_" + s).argtypes = [vp, vp, vp, vp]
            ct = d if s == "f64" else f
            getattr(L, "hflmp_stencil_apply_add_" + s).argtypes = [vp, vp, ct, vp, ct, vp]
            getattr(L, "hflmp_stencil_extract_invdiag_" + s).argtypes = [vp, vp, vp,
                                                                         C.POINTER(i)]
            getattr(L, "hflmp_pcg_jacobi_stencil_" + s).argtypes = [vp, vp, vp, vp, vp, i, d, d,
                                                                     d, vp, sz, vp]
        L.hflmp_stencil_get_info.argtypes = [vp, vp]
        L.hflmp_fe_mesh_create.argtypes = [vp, i, i, i, vp, vp]
        L.hflmp_fe_mesh_nnz.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.hflmp_fe_mesh_pattern.argtypes = [vp, vp, vp, vp]
        for s in ("f64", "f32"):
            getattr(L, "hflmp_fe_assemble_" + s).argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.hflmp_fe_mesh_destroy.argtypes = [vp]
        L.hflmp_fe_mesh_destroy.restype = None
        L.hflmp_stencil_destroy.argtypes = [vp]
        L.hflmp_stencil_destroy.restype = None
        L.hflmp_pcg_stencil_workspace_bytes.restype = sz
        L.hflmp_pcg_stencil_workspace_bytes.argtypes = [C.c_longlong, i]
        L.hflmp_sell_destroy.argtypes = [vp]
        L.hflmp_sell_destroy.restype = None
        L.hflmp_mpcg_sell_workspace_bytes.restype = sz
        L.hflmp_mpcg_sell_workspace_bytes.argtypes = [i]
        L.hflmp_mpcg_jacobi_sell.argtypes = [vp, vp, vp, vp, vp, vp, i, d, d, i, d, i, vp, sz, vp]
        L.hflmp_server_create.argtypes = [vp, vp, vp]
        L.hflmp_server_destroy.argtypes = [vp]
        L.hflmp_server_destroy.restype = None
        L.hflmp_server_footprint.restype = sz
        L.hflmp_server_footprint.argtypes = [i, i, vp, vp, i]
        L.hflmp_server_submit.argtypes = [vp, i, i, vp, vp, vp, vp, vp, d, d, i, i, vp]
        L.hflmp_server_step.argtypes = [vp, vp]
        L.hflmp_server_poll.argtypes = [vp, C.c_longlong, vp]
        L.hflmp_server_release.argtypes = [vp, C.c_longlong]
        L.hflmp_server_stats.argtypes = [vp, vp]
        L.hflmp_batched_pcg_workspace_bytes.restype = sz
        L.hflmp_batched_pcg_workspace_bytes.argtypes = [vp, i, i, i]
        L.hflmp_mpcg_workspace_bytes.restype = sz
        L.hflmp_mpcg_workspace_bytes.argtypes = [i, i]
        L.hflmp_mpcg_jacobi.argtypes = [vp, i, i, vp, vp, vp, vp, vp, vp, vp, i, d, d, i, d,
                                        i, vp, sz, C.POINTER(MpInfo)]
        L.hflmp_mcsgs_create_f64.argtypes = [vp, i, i, vp, vp, vp, i, vp, vp, d,
                                             C.POINTER(vp)]
        L.hflmp_mcsgs_apply_f64.argtypes = [vp, vp, vp, vp]
        L.hflmp_mcsgs_destroy.argtypes = [vp, vp]
        for s_ in ("f64", "f32"):
            getattr(L, "hflmp_sgs_create_" + s_).argtypes = [vp, i, i, vp, vp, vp, vp, vp, i,
                                                              C.POINTER(vp)]
            getattr(L, "hflmp_sgs_apply_" + s_).argtypes = [vp, vp, vp, vp]
        L.hflmp_sgs_get_info.argtypes = [vp, vp]
        L.hflmp_sgs_destroy.argtypes = [vp]
        L.hflmp_sgs_destroy.restype = None
        L.hflmp_cast_f64_to_f32.argtypes = [vp, i, vp, vp]
        L.hflmp_cast_f32_to_f64.argtypes = [vp, i, vp, vp]
        L.hflmp_cast_csr_f64_to_f32.argtypes = [vp, i, vp, vp]
        L.hflmp_csr_diag_ptr.argtypes = [vp, i, vp, vp, vp]
        L.hflmp_csr_zero_rows_f64.argtypes = [vp, i, vp, vp, vp, vp, d]
        L.hflmp_mg_create_f64.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.hflmp_mg_create_stencil_f64.argtypes = [vp, vp, i, vp, vp, vp, vp, vp]
        L.hflmp_mg_destroy.argtypes = [vp]
        L.hflmp_mg_set_coarse_inverse_f64.argtypes = [vp, vp, vp]
        L.hflmp_mg_set_structured_q1.argtypes = [vp, vp, i, i]
        L.hflmp_mg_destroy.restype = None
        L.hflmp_mg_apply_f64.argtypes = [vp, vp, vp, vp]
        L.hflmp_mg_workspace_bytes.argtypes = [vp]
        L.hflmp_mg_workspace_bytes.restype = sz
        for n in ("hflmp_pcg_mg_f64", "hflmp_mg_solve_f64"):
            getattr(L, n).argtypes = [vp, vp, vp, vp, i, d, d, d, vp, sz, vp]
        _hip = L
    return _hip


def gen_lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_LIB):
            raise HflmpError("%s not built: run __graft_entry__.build()" % GEN_LIB)
        L = C.CDLL(GEN_LIB)
        L.hflmp_gen_fe_sizes.argtypes = [C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.hflmp_gen_fe_csr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        L.hflmp_gen_kappa.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64,
                                      C.c_void_p]
        L.hflmp_gen_rhs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_void_p]
        L.hflmp_gen_set_threads.argtypes = [C.c_int]
        L.hflmp_gen_fe_element.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        _gen = L
    return _gen


# ---- structured FE systems (host) ---------------------------------------------
def fe_sizes(order, dim, n):
    N, Z = C.c_int64(), C.c_int64()
    if gen_lib().hflmp_gen_fe_sizes(order, dim, n, C.byref(N), C.byref(Z)):
        raise ValueError("unsupported element order=%d dim=%d n=%d" % (order, dim, n))
    return N.value, Z.value


def fe_kappa(dim, n, sigma, seed):
    k = np.empty(n ** dim, np.float64)
    gen_lib().hflmp_gen_kappa(dim, n, sigma, seed, k.ctypes.data)
    return k


def fe_system(order, dim, n, kappa=None):
    """(row_ptr, col, val) of the Dirichlet FE stiffness matrix."""
    N, Z = fe_sizes(order, dim, n)
    rp = np.empty(N + 1, np.int32)
    ci = np.empty(Z, np.int32)
    v = np.empty(Z, np.float64)
    kp = None
    if kappa is not None:
        kappa = np.ascontiguousarray(kappa, np.float64)
        kp = kappa.ctypes.data
    rc = gen_lib().hflmp_gen_fe_csr(order, dim, n, kp, rp.ctypes.data, ci.ctypes.data,
                                    v.ctypes.data)
    if rc:
        raise HflmpError("hflmp_gen_fe_csr failed (%d)" % rc)
    return rp, ci, v


def fe_element(order, dim, n):
    """Element stiffness (nloc x nloc, h included) entered per cell with
    factor kappa: the stencil operator's local matrix."""
    q = order + 1
    E = np.empty((q ** dim) ** 2, np.float64)
    if gen_lib().hflmp_gen_fe_element(order, dim, n, E.ctypes.data) < 0:
        raise ValueError("unsupported element order=%d dim=%d n=%d" % (order, dim, n))
    return E.reshape(q ** dim, q ** dim)


def fe_fixed(order, dim, n):
    """Indices of the boundary (Dirichlet) nodes, ascending."""
    p = order * n + 1
    g = np.arange(p)
    edge = (g == 0) | (g == p - 1)
    if dim == 2:
        m = edge[None, :] | edge[:, None]
    else:
        m = edge[None, None, :] | edge[None, :, None] | edge[:, None, None]
    return np.nonzero(m.ravel())[0].astype(np.int32)


def fe_cell_dofs(order, dim, n):
    """Cell -> dof table of the structured mesh: cells and their local
    nodes lexicographic (x fastest), int32 array ncell x (order+1)^dim."""
    p = order * n + 1
    loc = np.arange(order + 1)
    c = np.arange(n)
    if dim == 2:
        cy, cx = np.meshgrid(c, c, indexing="ij")
        ly, lx = np.meshgrid(loc, loc, indexing="ij")
        g = [order * cx.reshape(-1, 1) + lx.reshape(1, -1),
             order * cy.reshape(-1, 1) + ly.reshape(1, -1)]
        return (g[0] + p * g[1]).astype(np.int32)
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    lz, ly, lx = np.meshgrid(loc, loc, loc, indexing="ij")
    g = [order * a.reshape(-1, 1) + b.reshape(1, -1) for a, b in ((cx, lx), (cy, ly), (cz, lz))]
    return (g[0] + p * (g[1] + p * g[2])).astype(np.int32)


def fe_rhs(order, dim, n, seed):
    N, _ = fe_sizes(order, dim, n)
    b = np.empty(N, np.float64)
    gen_lib().hflmp_gen_rhs(order, dim, n, seed, b.ctypes.data)
    return b


# ---- device context and buffers ---------------------------------------------------
def device_count():
    n = C.c_int()
    hip_lib().hflmp_device_count(C.byref(n))
    return n.value


class Context:
    """One device, one stream (hflmp_ctx)."""

    def __init__(self, device=0):
        L = hip_lib()
        h = C.c_void_p()
        rc = L.hflmp_ctx_create(device, C.byref(h))
        if rc != HFLMP_OK:
            raise HflmpError("hflmp_ctx_create(%d) failed (%d): no usable GPU" % (device, rc))
        self.h = h
        self.L = L

    def check(self, rc, what=""):
        if rc != HFLMP_OK:
            msg = self.L.hflmp_ctx_last_error(self.h)
            err = HflmpError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))
            err.code = rc
            raise err

    def sync(self):
        self.check(self.L.hflmp_ctx_sync(self.h), "sync")

    def close(self):
        if self.h:
            self.L.hflmp_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def event(self):
        return Event(self)

    def empty(self, n, dtype):
        return DeviceArray(self, n, dtype)

    def zeros(self, n, dtype):
        a = DeviceArray(self, n, dtype)
        if a.nbytes:
            self.check(self.L.hflmp_fill_zero(self.h, a.ptr, a.nbytes), "fill_zero")
        return a

    def upload(self, host):
        host = np.ascontiguousarray(host)
        a = DeviceArray(self, host.size, host.dtype)
        if a.nbytes:
            self.check(self.L.hflmp_memcpy_h2d(self.h, a.ptr, host.ctypes.data, a.nbytes),
                       "memcpy_h2d")
        return a


class Event:
    """A HIP event recorded on the context's stream."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(ctx.L.hflmp_event_create(ctx.h, C.byref(self.h)), "event_create")

    def record(self):
        self.ctx.check(self.ctx.L.hflmp_event_record(self.ctx.h, self.h), "event_record")

    def elapsed_ms(self, start):
        ms = C.c_float()
        self.ctx.check(self.ctx.L.hflmp_event_elapsed_ms(self.ctx.h, start.h, self.h,
                                                         C.byref(ms)), "event_elapsed")
        return ms.value

    def destroy(self):
        if self.h:
            self.ctx.L.hflmp_event_destroy(self.ctx.h, self.h)
            self.h = None


class DeviceArray:
    """A device allocation holding `n` items of `dtype` (freed on close)."""

    def __init__(self, ctx, n, dtype):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.L.hflmp_malloc(ctx.h, self.nbytes, C.byref(p)), "malloc(%d)" % self.nbytes)
        self.ptr = p

    def download(self):
        out = np.empty(self.n, self.dtype)
        if self.nbytes:
            self.ctx.check(self.ctx.L.hflmp_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                                       self.nbytes), "memcpy_d2h")
        return out

    def copy_from_host(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.size == self.n
        self.ctx.check(self.ctx.L.hflmp_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data,
                                                   self.nbytes), "memcpy_h2d")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            self.ctx.L.hflmp_free(self.ctx.h, self.ptr)
        self.ptr = C.c_void_p()


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


class CsrMatrix:
    """Device CSR arrays (int32 row_ptr/col, float64 or float32 values)."""

    def __init__(self, ctx, row_ptr, col, val, ncol=None):
        self.ctx = ctx
        self.nrow = len(row_ptr) - 1
        self.ncol = self.nrow if ncol is None else ncol
        self.nnz = len(col)
        self.rp = ctx.upload(np.asarray(row_ptr, np.int32))
        self.ci = ctx.upload(np.asarray(col, np.int32))
        self.val = ctx.upload(np.asarray(val))
        self.dtype = self.val.dtype

    def astype_f32(self):
        """A float32 copy of the values sharing row_ptr/col."""
        m = CsrMatrix.__new__(CsrMatrix)
        m.ctx, m.nrow, m.ncol, m.nnz = self.ctx, self.nrow, self.ncol, self.nnz
        m.rp, m.ci = self.rp, self.ci
        m.val = self.ctx.empty(self.nnz, np.float32)
        self.ctx.check(self.ctx.L.hflmp_cast_csr_f64_to_f32(self.ctx.h, self.nnz, self.val.ptr,
                                                            m.val.ptr), "cast_csr")
        m.dtype = np.dtype(np.float32)
        return m

    def free(self):
        for a in (self.rp, self.ci, self.val):
            a.free()


# ---- operations ---------------------------------------------------------------------
def _call(ctx, name, *args):
    ctx.check(getattr(ctx.L, name)(ctx.h, *args), name)


def spmv(ctx, A, x, y, mode="assign"):
    name = {"assign": "csr_spmv", "add": "csr_spmv_add", "nodiag": "csr_spmv_nodiag"}[mode]
    _call(ctx, "hflmp_%s_%s" % (name, _sfx(A.dtype)), A.nrow, A.ncol, A.nnz, A.rp.ptr, A.ci.ptr,
          A.val.ptr, x.ptr, y.ptr)


def invdiag(ctx, A):
    d = ctx.empty(A.nrow, A.dtype)
    bad = C.c_int()
    _call(ctx, "hflmp_csr_extract_invdiag_" + _sfx(A.dtype), A.nrow, A.rp.ptr, A.ci.ptr,
          A.val.ptr, d.ptr, C.byref(bad))
    return d


def dot(ctx, x, y):
    ct = C.c_double if x.dtype == np.float64 else C.c_float
    r = ct()
    _call(ctx, "hflmp_dot_" + _sfx(x.dtype), x.n, x.ptr, y.ptr, C.byref(r))
    return r.value


def nrm2(ctx, x):
    ct = C.c_double if x.dtype == np.float64 else C.c_float
    r = ct()
    _call(ctx, "hflmp_nrm2_" + _sfx(x.dtype), x.n, x.ptr, C.byref(r))
    return r.value


def pcg_jacobi(ctx, A, invd, b, x, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6, work=None):
    """Device-resident Jacobi-PCG; x holds x0 on entry and the solution on exit."""
    vb = A.dtype.itemsize
    need = ctx.L.hflmp_pcg_workspace_bytes(A.nrow, A.nnz, vb)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    info = SolveInfo()
    try:
        _call(ctx, "hflmp_pcg_jacobi_" + _sfx(A.dtype), A.nrow, A.nnz, A.rp.ptr, A.ci.ptr,
              A.val.ptr, invd.ptr, b.ptr, x.ptr, maxit, atol, rtol, dtol, work.ptr,
              work.nbytes, C.byref(info))
    finally:
        if own:
            work.free()
    return info.as_dict()


class AsyncParams(C.Structure):
    _fields_ = [("block_rows", C.c_int), ("n_global", C.c_int), ("n_block", C.c_int),
                ("n_inner", C.c_int), ("mode", C.c_int), ("w", C.c_double)]


def async_jacobi(ctx, A, b, x, block_rows=256, n_global=1, n_block=1, n_inner=5, mode=0, w=1.0,
                 maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6):
    """Block-asynchronous Jacobi (hflmp_async_jacobi_*): mode 0/1 solve
    (normal/damped), 2/3 smooth; x holds x0 on entry and the result on exit."""
    need = ctx.L.hflmp_async_workspace_bytes(A.nrow, A.dtype.itemsize)
    work = ctx.empty(need, np.uint8)
    prm = AsyncParams(int(block_rows), int(n_global), int(n_block), int(n_inner), int(mode),
                      float(w))
    info = SolveInfo()
    try:
        _call(ctx, "hflmp_async_jacobi_" + _sfx(A.dtype), A.nrow, A.nnz, A.rp.ptr, A.ci.ptr,
              A.val.ptr, b.ptr, x.ptr, C.byref(prm), int(maxit), float(atol), float(rtol),
              float(dtol), work.ptr, work.nbytes, C.byref(info))
    finally:
        work.free()
    return info.as_dict()


def spmm(ctx, A, X, Y, nrhs):
    """Y = A X for row-major blocks of nrhs <= 64 vectors (device arrays)."""
    _call(ctx, "hflmp_csr_spmm_" + _sfx(A.dtype), A.nrow, A.ncol, A.nnz, nrhs, A.rp.ptr,
          A.ci.ptr, A.val.ptr, X.ptr, Y.ptr)


class SpmmPlan:
    """Row-block layout of a device CSR matrix for Y = A X (hflmp_spmm_plan_*):
    each block (a run of consecutive rows, or a patch of runs on consecutive
    lines of the numbering) has its distinct columns listed once and its X
    rows staged in LDS.  Blocks are sized for products of nrhs vectors of
    value_bytes."""

    def __init__(self, ctx, A, nrhs, value_bytes=None, lds_budget=0):
        self.ctx = ctx
        self.nrow, self.ncol, self.nnz = A.nrow, A.ncol, A.nnz
        self.h = C.c_void_p()
        vb = A.dtype.itemsize if value_bytes is None else value_bytes
        _call(ctx, "hflmp_spmm_plan_create", A.nrow, A.ncol, A.nnz, A.rp.ptr, A.ci.ptr,
              int(nrhs), int(vb), int(lds_budget), C.byref(self.h))

    def info(self):
        r, u = C.c_int(), C.c_int()
        n, b = C.c_longlong(), C.c_longlong()
        rc = self.ctx.L.hflmp_spmm_plan_info(self.h, C.byref(r), C.byref(u), C.byref(n),
                                             C.byref(b))
        if rc:
            raise HflmpError("hflmp_spmm_plan_info failed (%d)" % rc)
        L, a, nl = C.c_int(), C.c_int(), C.c_int()
        rc = self.ctx.L.hflmp_spmm_plan_geometry(self.h, C.byref(L), C.byref(a), C.byref(nl))
        if rc:
            raise HflmpError("hflmp_spmm_plan_geometry failed (%d)" % rc)
        return {"rows_per_block": r.value, "max_cols_per_block": u.value,
                "list_entries": n.value, "device_bytes": b.value,
                "line_rows": L.value, "run_rows": a.value, "runs": nl.value}

    def spmm(self, A, X, Y, nrhs):
        """Y = A X through the plan (A: the matrix it was built from)."""
        _call(self.ctx, "hflmp_plan_spmm_" + _sfx(A.dtype), self.h, A.rp.ptr, A.val.ptr,
              X.ptr, Y.ptr, int(nrhs))

    def free(self):
        if self.h:
            self.ctx.L.hflmp_spmm_plan_destroy(self.h)
            self.h = C.c_void_p()


def batched_mpcg_jacobi(ctx, A, A32, invd32, B, X, nrhs, max_outer=50, atol=0.0, rtol=1e-8,
                        inner_maxit=100000, inner_rtol=1e-4, work=None, plan32=None):
    """FP32-inner / FP64-refined Jacobi-CG on nrhs right-hand sides (MPIR per
    column; A FP64, A32 its FP32 copy sharing the pattern, invd32 its inverse
    diagonal); returns per-column corrections / inner iterations / status /
    residuals and the time.  plan32: a SpmmPlan of A32 (value_bytes 4) for
    the inner solves' products."""
    need = ctx.L.hflmp_batched_mpcg_workspace_bytes(ctx.h, A.nrow, nrhs)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    out = np.zeros(nrhs, np.int32)
    inn = np.zeros(nrhs, np.int32)
    st = np.zeros(nrhs, np.int32)
    r0 = np.zeros(nrhs)
    rf = np.zeros(nrhs)
    sec = C.c_double()
    args = (A.nrow, A.nnz, nrhs, A.rp.ptr, A.ci.ptr, A.val.ptr, A32.val.ptr, invd32.ptr, B.ptr,
            X.ptr, int(max_outer), float(atol), float(rtol), int(inner_maxit),
            float(inner_rtol), work.ptr, work.nbytes, out.ctypes.data, inn.ctypes.data,
            st.ctypes.data, r0.ctypes.data, rf.ctypes.data, C.byref(sec))
    try:
        if plan32 is None:
            _call(ctx, "hflmp_batched_mpcg_jacobi", *args)
        else:
            _call(ctx, "hflmp_batched_mpcg_jacobi_plan", plan32.h, *args)
    finally:
        if own:
            work.free()
    return dict(outer=out, inner_iters=inn, status=st, res_init=r0, res_final=rf,
                seconds=sec.value)


def batched_pcg_jacobi(ctx, A, invd, B, X, nrhs, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6,
                       work=None, plan=None):
    """Lockstep Jacobi-PCG on nrhs right-hand sides (row-major device blocks
    B, X); returns per-column iterations/status/residuals and the time.
    plan: a SpmmPlan of A for nrhs vectors, through which the iteration's
    products go (hflmp_batched_pcg_jacobi_plan_*)."""
    need = ctx.L.hflmp_batched_pcg_workspace_bytes(ctx.h, A.nrow, nrhs, A.dtype.itemsize)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    it = np.zeros(nrhs, np.int32)
    st = np.zeros(nrhs, np.int32)
    r0 = np.zeros(nrhs)
    rf = np.zeros(nrhs)
    sec = C.c_double()
    args = (A.nrow, A.nnz, nrhs, A.rp.ptr, A.ci.ptr, A.val.ptr, invd.ptr, B.ptr, X.ptr, maxit,
            atol, rtol, dtol, work.ptr, work.nbytes, it.ctypes.data, st.ctypes.data,
            r0.ctypes.data, rf.ctypes.data, C.byref(sec))
    try:
        if plan is None:
            _call(ctx, "hflmp_batched_pcg_jacobi_" + _sfx(A.dtype), *args)
        else:
            _call(ctx, "hflmp_batched_pcg_jacobi_plan_" + _sfx(A.dtype), plan.h, *args)
    finally:
        if own:
            work.free()
    return dict(iters=it, status=st, res_init=r0, res_final=rf, seconds=sec.value)


MP_INT32_COLS = 1
MP_RELIABLE = 2


def mpcg_jacobi(ctx, A, A32, invd32, b, x, max_outer=50, atol=0.0, rtol=1e-10,
                inner_maxit=20000, inner_rtol=1e-5, flags=0, work=None):
    """FP32-inner / FP64-refined Jacobi-CG on the device (hflmp_mpcg_jacobi);
    A is the FP64 matrix, A32 its FP32 copy (A.astype_f32()), invd32 the FP32
    inverse diagonal; x holds x0 and receives the solution.  flags
    MP_INT32_COLS keeps int32 columns in the FP32 operator."""
    need = ctx.L.hflmp_mpcg_workspace_bytes(A.nrow, A.nnz)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    info = MpInfo()
    try:
        _call(ctx, "hflmp_mpcg_jacobi", A.nrow, A.nnz, A.rp.ptr, A.ci.ptr, A.val.ptr,
              A32.val.ptr, invd32.ptr, b.ptr, x.ptr, max_outer, atol, rtol, inner_maxit,
              inner_rtol, flags, work.ptr, work.nbytes, C.byref(info))
    finally:
        if own:
            work.free()
    return info.as_dict()


class SellInfo(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("nnz", C.c_int), ("chunk", C.c_int),
                ("col_bytes", C.c_int), ("value_bytes", C.c_int), ("entries", C.c_longlong),
                ("device_bytes", C.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SELL_COL16 = 1
SELL_ANY_PADDING = 2
EINVAL = -1
EFORMAT = -6


class SellMatrix:
    """SELL-64 device layout of a device CSR matrix (hflmp_sell_*), built on
    the device; flags SELL_COL16 asks for int16 slice-offset columns."""

    def __init__(self, ctx, A, flags=0):
        self.ctx = ctx
        self.nrow, self.ncol, self.nnz = A.nrow, A.ncol, A.nnz
        self.dtype = A.dtype
        self.h = C.c_void_p()
        _call(ctx, "hflmp_sell_create_" + _sfx(A.dtype), A.nrow, A.ncol, A.nnz, A.rp.ptr,
              A.ci.ptr, A.val.ptr, int(flags), C.byref(self.h))

    def info(self):
        inf = SellInfo()
        rc = self.ctx.L.hflmp_sell_get_info(self.h, C.byref(inf))
        if rc:
            raise HflmpError("hflmp_sell_get_info failed (%d)" % rc)
        return inf.as_dict()

    def spmv(self, x, y, add=False):
        name = "hflmp_sell_spmv_add_" if add else "hflmp_sell_spmv_"
        _call(self.ctx, name + _sfx(self.dtype), self.h, x.ptr, y.ptr)

    def set_values(self, A):
        _call(self.ctx, "hflmp_sell_set_values_" + _sfx(self.dtype), self.h, A.rp.ptr,
              A.val.ptr)

    def free(self):
        if self.h:
            self.ctx.L.hflmp_sell_destroy(self.h)
            self.h = C.c_void_p()


def pcg_jacobi_sell(ctx, S, invd, b, x, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6,
                    work=None):
    """Device-resident Jacobi-PCG over a SellMatrix."""
    need = ctx.L.hflmp_pcg_workspace_bytes(S.nrow, S.nnz, S.dtype.itemsize)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    info = SolveInfo()
    try:
        _call(ctx, "hflmp_pcg_jacobi_sell_" + _sfx(S.dtype), S.h, invd.ptr, b.ptr, x.ptr, maxit,
              atol, rtol, dtol, work.ptr, work.nbytes, C.byref(info))
    finally:
        if own:
            work.free()
    return info.as_dict()


class FeMesh:
    """Device assembly of a general mesh (hflmp_fe_*): cell_dofs is a host
    ncell x nloc array of node indices."""

    def __init__(self, ctx, nnode, cell_dofs):
        self.ctx = ctx
        dofs = np.ascontiguousarray(cell_dofs, np.int32)
        self.nnode = int(nnode)
        self.ncell, self.nloc = dofs.shape
        d = ctx.upload(dofs)
        self.h = C.c_void_p()
        try:
            _call(ctx, "hflmp_fe_mesh_create", self.nnode, self.ncell, self.nloc, d.ptr,
                  C.byref(self.h))
        finally:
            d.free()
        z = C.c_longlong()
        ctx.check(ctx.L.hflmp_fe_mesh_nnz(self.h, C.byref(z)), "fe_mesh_nnz")
        self.nnz = z.value

    def pattern(self):
        rp = self.ctx.empty(self.nnode + 1, np.int32)
        ci = self.ctx.empty(max(self.nnz, 1), np.int32)
        _call(self.ctx, "hflmp_fe_mesh_pattern", self.h, rp.ptr, ci.ptr)
        return rp, ci

    def assemble(self, rp, ci, E=None, factor=None, cell_mats=None, dtype=np.float64):
        dt = np.dtype(dtype)
        up = [None if a is None else self.ctx.upload(np.ascontiguousarray(a, dt))
              for a in (cell_mats, E, factor)]
        val = self.ctx.empty(max(self.nnz, 1), dt)
        try:
            _call(self.ctx, "hflmp_fe_assemble_" + _sfx(dt), self.h,
                  *[None if a is None else a.ptr for a in up], rp.ptr, ci.ptr, val.ptr)
            self.ctx.sync()
        finally:
            for a in up:
                if a is not None:
                    a.free()
        return val

    def free(self):
        if self.h:
            self.ctx.L.hflmp_fe_mesh_destroy(self.h)
            self.h = C.c_void_p()


def csr_from_device(ctx, nrow, ncol, nnz, rp, ci, val):
    """A CsrMatrix over device arrays that it takes ownership of."""
    m = CsrMatrix.__new__(CsrMatrix)
    m.ctx, m.nrow, m.ncol, m.nnz = ctx, nrow, ncol, int(nnz)
    m.rp, m.ci, m.val = rp, ci, val
    m.dtype = val.dtype
    return m


def assemble_q1(ctx, dim, n, cell_mats):
    """Device assembly of the structured Q1 mesh of n^dim cells from one
    matrix per cell (ncell x nloc x nloc, cells and local nodes
    lexicographic), with unit Dirichlet rows on the boundary: the operator
    of a multigrid level."""
    nnode = (n + 1) ** dim
    mesh = FeMesh(ctx, nnode, fe_cell_dofs(1, dim, n))
    rp = ci = val = fx = None
    try:
        rp, ci = mesh.pattern()
        val = mesh.assemble(rp, ci, cell_mats=np.ascontiguousarray(cell_mats, np.float64))
        fixed = fe_fixed(1, dim, n)
        fx = ctx.upload(fixed)
        _call(ctx, "hflmp_csr_zero_rows_f64", len(fixed), fx.ptr, rp.ptr, ci.ptr, val.ptr, 1.0)
        ctx.sync()
    except Exception:
        for a in (rp, ci, val):
            if a is not None:
                a.free()
        raise
    finally:
        mesh.free()
        if fx is not None:
            fx.free()
    return csr_from_device(ctx, nnode, nnode, mesh.nnz, rp, ci, val)


class MgParams(C.Structure):
    _fields_ = [("pre_smooth", C.c_int), ("post_smooth", C.c_int), ("coarse_sweeps", C.c_int),
                ("omega", C.c_double), ("coarse_omega", C.c_double)]


class CsrDesc(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("nnz", C.c_int),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]

    @classmethod
    def of(cls, A):
        return cls(A.nrow, A.ncol, A.nnz, A.rp.ptr, A.ci.ptr, A.val.ptr)


class Multigrid:
    """Device V-cycle over the levels A[0] (finest) .. A[L-1] with
    restrictions R[l] and prolongations P[l] (device CsrMatrix objects,
    float64; copied at creation).  With `stencil` (an FP64
    StencilOperator, kept referenced) the finest operator is matrix-free
    and A[0] may be None.  `owned` matrices are freed with it."""

    def __init__(self, ctx, A, R, P, pre_smooth=1, post_smooth=1, coarse_sweeps=4, omega=0.7,
                 coarse_omega=0.7, owned=(), stencil=None):
        self.ctx = ctx
        self.nlevels = len(A)
        self.stencil = stencil
        self.n = stencil.nrow if stencil is not None else A[0].nrow
        self.params = MgParams(pre_smooth, post_smooth, coarse_sweeps, omega, coarse_omega)
        descs = [CsrDesc(self.n, self.n, 0, None, None, None) if (l == 0 and m is None)
                 else CsrDesc.of(m) for l, m in enumerate(A)]
        a = (CsrDesc * len(A))(*descs)
        r = (CsrDesc * max(len(R), 1))(*[CsrDesc.of(m) for m in R])
        p = (CsrDesc * max(len(P), 1))(*[CsrDesc.of(m) for m in P])
        self.h = C.c_void_p()
        self.owned = list(owned)
        if stencil is not None:
            _call(ctx, "hflmp_mg_create_stencil_f64", stencil.h, self.nlevels, a, r, p,
                  C.byref(self.params), C.byref(self.h))
        else:
            _call(ctx, "hflmp_mg_create_f64", self.nlevels, a, r, p, C.byref(self.params),
                  C.byref(self.h))
        self.level_rows = [d.nrow for d in descs]

    def set_structured(self, dim, ncells):
        """Transfers of the structured Q1 hierarchy without reading R and P
        (dim 0: through R and P again)."""
        _call(self.ctx, "hflmp_mg_set_structured_q1", self.h, dim, ncells)

    def set_coarse_inverse(self, C):
        """Solve the coarsest level with the dense inverse C (host array,
        n x n) instead of sweeps; None returns to the sweeps."""
        if C is None:
            _call(self.ctx, "hflmp_mg_set_coarse_inverse_f64", self.h, None)
            return
        d = self.ctx.upload(np.ascontiguousarray(C, np.float64))
        try:
            _call(self.ctx, "hflmp_mg_set_coarse_inverse_f64", self.h, d.ptr)
            self.ctx.sync()
        finally:
            d.free()

    def apply(self, r, z):
        """z = M r (one V-cycle from zero); r and z device arrays."""
        _call(self.ctx, "hflmp_mg_apply_f64", self.h, r.ptr, z.ptr)

    def workspace_bytes(self):
        return self.ctx.L.hflmp_mg_workspace_bytes(self.h)

    def _run(self, name, b, x, maxit, atol, rtol, dtol, work):
        own = work is None
        if own:
            work = self.ctx.empty(self.workspace_bytes(), np.uint8)
        info = SolveInfo()
        try:
            _call(self.ctx, name, self.h, b.ptr, x.ptr, maxit, atol, rtol, dtol, work.ptr,
                  work.nbytes, C.byref(info))
        finally:
            if own:
                work.free()
        return info.as_dict()

    def pcg(self, b, x, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6, work=None):
        """CG preconditioned by one V-cycle per iteration, device-resident."""
        return self._run("hflmp_pcg_mg_f64", b, x, maxit, atol, rtol, dtol, work)

    def solve(self, b, x, maxit=1000, atol=0.0, rtol=1e-8, dtol=1e6, work=None):
        """Multigrid iteration x = x + M (b - A x) until the residual test."""
        return self._run("hflmp_mg_solve_f64", b, x, maxit, atol, rtol, dtol, work)

    def free(self):
        if self.h:
            self.ctx.L.hflmp_mg_destroy(self.h)
            self.h = C.c_void_p()
        for m in self.owned:
            m.free()
        self.owned = []


class StencilDesc(C.Structure):
    _fields_ = [("order", C.c_int), ("dim", C.c_int), ("ncell", C.c_int * 3),
                ("nmat", C.c_int)]


class StencilInfo(C.Structure):
    _fields_ = [("order", C.c_int), ("dim", C.c_int), ("nloc", C.c_int), ("nmat", C.c_int),
                ("ncell", C.c_int * 3), ("nnode", C.c_longlong), ("ncell_total", C.c_longlong),
                ("nfixed", C.c_int), ("value_bytes", C.c_int), ("device_bytes", C.c_size_t)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["ncell"] = list(self.ncell)
        return d


class StencilOperator:
    """Matrix-free stencil operator on a structured Q1/Q2 mesh
    (hflmp_stencil_*): local_matrices (host, nmat x nloc x nloc), cell
    factors and the optional per-cell matrix index and fixed dof list are
    host arrays uploaded here."""

    def __init__(self, ctx, order, dim, ncell, local_matrices, cell_factor, fixed=None,
                 cell_lm=None, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        E = np.ascontiguousarray(local_matrices, self.dtype)
        q = order + 1
        nloc = q ** dim
        nmat = E.size // (nloc * nloc)
        nc = list(ncell) + [1] * (3 - len(ncell))
        desc = StencilDesc(order, dim, (C.c_int * 3)(*nc), nmat)
        f = ctx.upload(np.ascontiguousarray(cell_factor, self.dtype))
        fx = None if fixed is None or len(fixed) == 0 else \
            ctx.upload(np.ascontiguousarray(fixed, np.int32))
        lm = None if cell_lm is None else ctx.upload(np.ascontiguousarray(cell_lm, np.uint8))
        self.h = C.c_void_p()
        try:
            _call(ctx, "hflmp_stencil_create_" + _sfx(self.dtype), C.byref(desc),
                  E.ctypes.data, None if lm is None else lm.ptr, f.ptr,
                  0 if fx is None else len(fixed), None if fx is None else fx.ptr,
                  C.byref(self.h))
        finally:
            for a in (f, fx, lm):
                if a is not None:
                    a.free()
        inf = self.info()
        self.nrow = inf["nnode"]

    def info(self):
        inf = StencilInfo()
        rc = self.ctx.L.hflmp_stencil_get_info(self.h, C.byref(inf))
        if rc:
            raise HflmpError("hflmp_stencil_get_info failed (%d)" % rc)
        return inf.as_dict()

    def apply(self, x, y):
        _call(self.ctx, "hflmp_stencil_apply_" + _sfx(self.dtype), self.h, x.ptr, y.ptr)

    def apply_add(self, alpha, x, beta, y):
        _call(self.ctx, "hflmp_stencil_apply_add_" + _sfx(self.dtype), self.h, alpha, x.ptr,
              beta, y.ptr)

    def invdiag(self):
        d = self.ctx.empty(self.nrow, self.dtype)
        bad = C.c_int()
        _call(self.ctx, "hflmp_stencil_extract_invdiag_" + _sfx(self.dtype), self.h, d.ptr,
              C.byref(bad))
        return d

    def set_factors(self, cell_factor):
        f = self.ctx.upload(np.ascontiguousarray(cell_factor, self.dtype))
        try:
            _call(self.ctx, "hflmp_stencil_set_factors_" + _sfx(self.dtype), self.h, f.ptr)
            self.ctx.sync()
        finally:
            f.free()

    def free(self):
        if self.h:
            self.ctx.L.hflmp_stencil_destroy(self.h)
            self.h = C.c_void_p()


def pcg_jacobi_stencil(ctx, S, invd, b, x, maxit=10000, atol=0.0, rtol=1e-8, dtol=1e6,
                       work=None):
    """Device-resident Jacobi-PCG over a StencilOperator."""
    need = ctx.L.hflmp_pcg_stencil_workspace_bytes(S.nrow, S.dtype.itemsize)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    info = SolveInfo()
    try:
        _call(ctx, "hflmp_pcg_jacobi_stencil_" + _sfx(S.dtype), S.h, invd.ptr, b.ptr, x.ptr,
              maxit, atol, rtol, dtol, work.ptr, work.nbytes, C.byref(info))
    finally:
        if own:
            work.free()
    return info.as_dict()


def mpcg_jacobi_sell(ctx, S64, S32, invd32, b, x, max_outer=50, atol=0.0, rtol=1e-10,
                     inner_maxit=20000, inner_rtol=1e-5, flags=0, work=None):
    """FP32-inner / FP64-refined Jacobi-CG with both operators in SELL-64
    layout (S64: FP64 values, S32: FP32 values); flags MP_RELIABLE keeps the
    inner search direction across corrections."""
    need = ctx.L.hflmp_mpcg_sell_workspace_bytes(S64.nrow)
    own = work is None
    if own:
        work = ctx.empty(need, np.uint8)
    info = MpInfo()
    try:
        _call(ctx, "hflmp_mpcg_jacobi_sell", S64.h, S32.h, invd32.ptr, b.ptr, x.ptr, max_outer,
              atol, rtol, inner_maxit, inner_rtol, int(flags), work.ptr, work.nbytes,
              C.byref(info))
    finally:
        if own:
            work.free()
    return info.as_dict()


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("chunk_iters", C.c_int),
                ("use_graphs", C.c_int), ("mixed_min_nnz", C.c_int), ("inner_rtol", C.c_double),
                ("mixed_max_nnz", C.c_int), ("solo_min_nnz", C.c_int)]


class ServerResult(C.Structure):
    _fields_ = [("state", C.c_int), ("iters", C.c_int), ("status", C.c_int),
                ("res_init", C.c_double), ("res_final", C.c_double),
                ("zero_diagonals", C.c_int), ("footprint", C.c_size_t),
                ("queued_seconds", C.c_double), ("latency_seconds", C.c_double),
                ("precision", C.c_int), ("corrections", C.c_int), ("col_bytes", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int), ("active", C.c_int), ("active_units", C.c_int),
                ("arena_free", C.c_size_t), ("arena_largest_free", C.c_size_t),
                ("steps", C.c_longlong), ("graph_captures", C.c_longlong),
                ("peak_active", C.c_int), ("peak_arena_in_use", C.c_size_t),
                ("admit_seconds", C.c_double), ("capture_seconds", C.c_double),
                ("wait_seconds", C.c_double), ("solo_runs", C.c_longlong),
                ("join_seconds", C.c_double), ("retire_seconds", C.c_double),
                ("chunk_device_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def server_footprint(rp, ci, precision=64):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    return hip_lib().hflmp_server_footprint(len(rp) - 1, len(ci), rp.ctypes.data,
                                            ci.ctypes.data, int(precision))


class Server:
    """Device-resident solve server (hflmp_server_*): submit host CSR
    systems, step, collect x.  Keeps each request's host arrays alive until
    its result is taken."""

    def __init__(self, ctx, hbm_budget, max_active=64, chunk_iters=16, use_graphs=True,
                 mixed_min_nnz=0, inner_rtol=1e-2, mixed_max_nnz=0, solo_min_nnz=0):
        self.ctx = ctx
        cfg = ServerConfig(int(hbm_budget), int(max_active), int(chunk_iters),
                           1 if use_graphs else 0, int(mixed_min_nnz), float(inner_rtol),
                           int(mixed_max_nnz), int(solo_min_nnz))
        self.h = C.c_void_p()
        ctx.check(ctx.L.hflmp_server_create(ctx.h, C.byref(cfg), C.byref(self.h)),
                  "server_create")
        self.keep = {}

    def _check(self, rc, what):
        if rc:
            self.ctx.check(rc, what)

    def submit(self, rp, ci, v, b, rtol=1e-8, atol=0.0, maxit=100000, precision=64):
        """precision 64 (FP64 CG), 32 (FP32-inner / FP64-refined) or 0 (by
        size: mixed for mixed_min_nnz <= nnz < mixed_max_nnz)."""
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        v = np.ascontiguousarray(v, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(len(rp) - 1)
        t = C.c_longlong()
        self._check(self.ctx.L.hflmp_server_submit(
            self.h, len(rp) - 1, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
            b.ctypes.data, x.ctypes.data, float(atol), float(rtol), int(maxit), int(precision),
            C.byref(t)),
            "server_submit")
        self.keep[t.value] = (rp, ci, v, b, x)
        return t.value

    def step(self):
        n = C.c_int()
        self._check(self.ctx.L.hflmp_server_step(self.h, C.byref(n)), "server_step")
        return n.value

    def poll(self, ticket):
        r = ServerResult()
        self._check(self.ctx.L.hflmp_server_poll(self.h, int(ticket), C.byref(r)), "server_poll")
        return r.as_dict()

    def take(self, ticket):
        """x and the result record of a done ticket (then forgotten)."""
        info = self.poll(ticket)
        if info["state"] != 2:
            raise HflmpError("ticket %d is not done" % ticket)
        x = self.keep.pop(ticket)[4]
        self._check(self.ctx.L.hflmp_server_release(self.h, int(ticket)), "server_release")
        return x, info

    def stats(self):
        s = ServerStats()
        self._check(self.ctx.L.hflmp_server_stats(self.h, C.byref(s)), "server_stats")
        return s.as_dict()

    def close(self):
        if self.h:
            self.ctx.L.hflmp_server_destroy(self.h)
            self.h = C.c_void_p()


class McSgs:
    """Multicoloured symmetric Gauss-Seidel on a device CSR matrix
    (hflmp_mcsgs_*); `coloring` = (ncolors, color_sizes, perm) on the host."""

    def __init__(self, ctx, A, coloring, omega=1.0):
        nc, csz, perm = coloring[:3]
        csz = np.ascontiguousarray(csz, np.int32)
        perm = np.ascontiguousarray(perm, np.int32)
        self.ctx = ctx
        self.h = C.c_void_p()
        _call(ctx, "hflmp_mcsgs_create_f64", A.nrow, A.nnz, A.rp.ptr, A.ci.ptr, A.val.ptr,
              int(nc), csz.ctypes.data, perm.ctypes.data, float(omega), C.byref(self.h))

    def apply(self, r, z):
        _call(self.ctx, "hflmp_mcsgs_apply_f64", self.h, r.ptr, z.ptr)

    def destroy(self):
        if self.h:
            self.ctx.L.hflmp_mcsgs_destroy(self.ctx.h, self.h)
            self.h = None


class SgsInfo(C.Structure):
    _fields_ = [("nrow", C.c_int), ("num_blocks", C.c_int), ("levels_forward", C.c_int),
                ("levels_backward", C.c_int), ("launches", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Sgs:
    """Symmetric Gauss-Seidel in the sequential row order (num_blocks = 1)
    or its block form (hflmp_sgs_*) on a device CSR matrix.  `host_pattern`
    (row_ptr, col) spares the download of the pattern."""

    def __init__(self, ctx, A, num_blocks=1, host_pattern=None):
        self.ctx, self.A = ctx, A
        self.h = C.c_void_p()
        hrp = hci = None
        if host_pattern is not None:
            self._hp = (np.ascontiguousarray(host_pattern[0], np.int32),
                        np.ascontiguousarray(host_pattern[1], np.int32))
            hrp, hci = self._hp[0].ctypes.data, self._hp[1].ctypes.data
        _call(ctx, "hflmp_sgs_create_" + _sfx(A.dtype), A.nrow, A.nnz, A.rp.ptr, A.ci.ptr,
              A.val.ptr, hrp, hci, int(num_blocks), C.byref(self.h))

    def apply(self, r, z):
        _call(self.ctx, "hflmp_sgs_apply_" + _sfx(self.A.dtype), self.h, r.ptr, z.ptr)

    def info(self):
        i = SgsInfo()
        self.ctx.check(self.ctx.L.hflmp_sgs_get_info(self.h, C.byref(i)), "sgs_get_info")
        return i.as_dict()

    def destroy(self):
        if self.h:
            self.ctx.L.hflmp_sgs_destroy(self.h)
            self.h = None


# ---- C++ lmp layer (libhflmp_lmp.so) ---------------------------------------------------
LMP_CG, LMP_GMRES, LMP_FGMRES, LMP_MPIR, LMP_BICGSTAB = 0, 1, 2, 3, 4
LMP_PREC_NONE, LMP_PREC_JACOBI, LMP_PREC_MCSGS, LMP_PREC_GMG = 0, 1, 2, 3
LMP_PREC_SGS, LMP_PREC_BSGS = 4, 5


class LmpOptions(C.Structure):
    _fields_ = [("solver", C.c_int), ("precond", C.c_int), ("precision", C.c_int),
                ("maxit", C.c_int), ("atol", C.c_double), ("rtol", C.c_double),
                ("dtol", C.c_double), ("basis", C.c_int), ("omega", C.c_double),
                ("mpir_max_outer", C.c_int), ("mpir_atol", C.c_double),
                ("mpir_rtol", C.c_double), ("solver_inner", C.c_int), ("sgs_blocks", C.c_int)]


class LmpResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("outer", C.c_int),
                ("res_init", C.c_double), ("res_final", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lmp = None


def lmp_lib():
    global _lmp
    if _lmp is None:
        if not os.path.exists(LMP_LIB):
            raise HflmpError("%s not built: run __graft_entry__.build()" % LMP_LIB)
        L = C.CDLL(LMP_LIB)
        L.hflmp_lmp_default_options.argtypes = [C.POINTER(LmpOptions)]
        L.hflmp_lmp_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(LmpOptions), C.POINTER(LmpResult)]
        L.hflmp_lmp_vector_mult.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p]
        L.hflmp_lmp_solve_gmg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(MgParams), C.c_void_p,
                                          C.POINTER(LmpOptions), C.POINTER(LmpResult)]
        L.hflmp_lmp_last_error.restype = C.c_char_p
        ip = C.POINTER(C.c_int)
        L.hflmp_lmp_mm_read_coo.argtypes = [C.c_char_p, ip, ip, ip, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int]
        L.hflmp_lmp_mm_write_csr.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p]
        L.hflmp_lmp_file_spmv.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_char_p,
                                          C.c_char_p, ip, ip]
        _lmp = L
    return _lmp


def lmp_options(**kw):
    o = LmpOptions()
    lmp_lib().hflmp_lmp_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def lmp_solve(rp, ci, v, b, x0=None, device=0, **kw):
    """Solve with the C++ solver templates over CoupledMatrix/CoupledVector on
    the HIP platform; returns (x, result dict)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(rp) - 1) if x0 is None else np.array(x0, np.float64)
    o = lmp_options(**kw)
    r = LmpResult()
    rc = lmp_lib().hflmp_lmp_solve(device, len(rp) - 1, len(ci), rp.ctypes.data,
                                   ci.ctypes.data, v.ctypes.data, b.ctypes.data,
                                   x.ctypes.data, C.byref(o), C.byref(r))
    if rc:
        raise HflmpError("hflmp_lmp_solve: " + lmp_lib().hflmp_lmp_last_error().decode())
    return x, r.as_dict()


def _host_csr(m):
    """(CsrDesc over host arrays, arrays kept alive) of a scipy matrix or a
    (row_ptr, col, val, ncol) tuple."""
    if isinstance(m, tuple):
        rp, ci, v, ncol = m
    else:
        m = m.tocsr()
        rp, ci, v, ncol = m.indptr, m.indices, m.data, m.shape[1]
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    return CsrDesc(len(rp) - 1, int(ncol), int(rp[-1]), rp.ctypes.data, ci.ctypes.data,
                   v.ctypes.data), (rp, ci, v)


def lmp_solve_gmg(rp, ci, v, b, coarse, R, P, x0=None, coarse_inverse=None, pre_smooth=1,
                  post_smooth=1, coarse_sweeps=4, omega=0.7, coarse_omega=0.7, device=0, **kw):
    """Solve with a C++ solver template (kw solver=, default CG)
    preconditioned by PreconditionerGMG: coarse[l], R[l], P[l] are host
    matrices (scipy, or (row_ptr, col, val, ncol) tuples) of level l + 1 and
    of the transfers between levels l and l + 1; returns (x, result dict)."""
    if not (len(coarse) == len(R) == len(P)):
        raise ValueError("lmp_solve_gmg: need one R and one P per coarse level")
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(rp) - 1) if x0 is None else np.array(x0, np.float64)
    keep = []

    def descs(ms):
        arr = (CsrDesc * max(1, len(ms)))()
        for i, m in enumerate(ms):
            arr[i], k = _host_csr(m)
            keep.append(k)
        return arr

    A_, R_, P_ = descs(coarse), descs(R), descs(P)
    cinv = None
    if coarse_inverse is not None:
        cinv = np.ascontiguousarray(coarse_inverse, np.float64)
        if not coarse:
            raise ValueError("lmp_solve_gmg: a dense coarse solve needs two levels or more")
        n = coarse[-1].shape[0] if not isinstance(coarse[-1], tuple) else len(coarse[-1][0]) - 1
        if cinv.shape != (n, n):
            raise ValueError("lmp_solve_gmg: coarse_inverse must be %d x %d" % (n, n))
    prm = MgParams(pre_smooth, post_smooth, coarse_sweeps, omega, coarse_omega)
    kw.setdefault("solver", LMP_CG)
    o = lmp_options(**kw)
    r = LmpResult()
    _lmp_check(lmp_lib().hflmp_lmp_solve_gmg(
        device, len(rp) - 1, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
        b.ctypes.data, x.ctypes.data, len(coarse) + 1, C.addressof(A_), C.addressof(R_),
        C.addressof(P_), C.byref(prm), None if cinv is None else cinv.ctypes.data,
        C.byref(o), C.byref(r)), "hflmp_lmp_solve_gmg")
    return x, r.as_dict()


def lmp_vector_mult(rp, ci, v, x, device=0):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(len(rp) - 1)
    rc = lmp_lib().hflmp_lmp_vector_mult(device, len(rp) - 1, len(ci), rp.ctypes.data,
                                         ci.ctypes.data, v.ctypes.data, x.ctypes.data,
                                         y.ctypes.data)
    if rc:
        raise HflmpError("hflmp_lmp_vector_mult: " + lmp_lib().hflmp_lmp_last_error().decode())
    return y


def _lmp_check(rc, what):
    if rc:
        raise HflmpError(what + ": " + lmp_lib().hflmp_lmp_last_error().decode())


def mm_read_coo(path):
    """MatrixMarket coordinate file as the lmp reader sees it: (nrow, ncol,
    rows, cols, vals), 0-based, in file order, mirrors of a symmetric file
    after their entry."""
    L = lmp_lib()
    m, n, k = C.c_int(), C.c_int(), C.c_int()
    p = os.fsencode(path)
    _lmp_check(L.hflmp_lmp_mm_read_coo(p, C.byref(m), C.byref(n), C.byref(k), None, None, None,
                                       0), "mm_read_coo")
    r = np.empty(k.value, np.int32)
    c = np.empty(k.value, np.int32)
    v = np.empty(k.value, np.float64)
    _lmp_check(L.hflmp_lmp_mm_read_coo(p, C.byref(m), C.byref(n), C.byref(k), r.ctypes.data,
                                       c.ctypes.data, v.ctypes.data, k.value), "mm_read_coo")
    return m.value, n.value, r, c, v


def mm_write_csr(path, rp, ci, v, ncol=None):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    nrow = len(rp) - 1
    _lmp_check(lmp_lib().hflmp_lmp_mm_write_csr(os.fsencode(path), nrow,
                                                nrow if ncol is None else ncol, rp.ctypes.data,
                                                ci.ctypes.data, v.ctypes.data), "mm_write_csr")


def lmp_file_spmv(matrix_file, x_file, y_file, matrix_out=None, device=0):
    """y = A x on the HIP platform with A, x read from files (ReadFile) and y
    written (WriteFile); returns (nrow, nnz)."""
    nrow, nnz = C.c_int(), C.c_int()
    _lmp_check(lmp_lib().hflmp_lmp_file_spmv(
        device, os.fsencode(matrix_file), os.fsencode(x_file), os.fsencode(y_file),
        None if matrix_out is None else os.fsencode(matrix_out), C.byref(nrow), C.byref(nnz)),
        "lmp_file_spmv")
    return nrow.value, nnz.value
