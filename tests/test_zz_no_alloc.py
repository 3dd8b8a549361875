"""The allocation-free Jacobi-PCG (hflmp_pcg_jacobi_sell_noalloc_f64, no
reference counterpart) against hflmp_pcg_jacobi_sell_f64 with a workspace:
the same kernels and grids, so x, the iteration count and the residuals are
bit-identical, on two calls in a row, the second on a smaller system (the
static workspace keeps the first call's contents).  Sizes beyond its
262144 rows are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from ba_amd import hflmp

pytestmark = pytest.mark.gpu


def _noalloc(ctx, S, invd, b, x, rtol=1e-8, maxit=20000):
    f = ctx.L.hflmp_pcg_jacobi_sell_noalloc_f64
    f.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(hflmp.SolveInfo)]
    info = hflmp.SolveInfo()
    rc = f(ctx.h, S.h, invd.ptr, b.ptr, x.ptr, maxit, 0.0, rtol, 1e6, C.byref(info))
    return rc, info.as_dict()


def _case(ctx, dim, n, seed, flags):
    rp, ci, v = hflmp.fe_system(1, dim, n, hflmp.fe_kappa(dim, n, 1.0, seed))
    b = hflmp.fe_rhs(1, dim, n, seed + 1)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    S = hflmp.SellMatrix(ctx, A, flags)
    invd = hflmp.invdiag(ctx, A)
    db = ctx.upload(b)
    x_ref = ctx.zeros(A.nrow, np.float64)
    ref = hflmp.pcg_jacobi_sell(ctx, S, invd, db, x_ref, rtol=1e-8, maxit=20000)
    x = ctx.zeros(A.nrow, np.float64)
    rc, info = _noalloc(ctx, S, invd, db, x)
    assert rc == 0, ctx.L.hflmp_ctx_last_error(ctx.h)
    assert info["status"] == ref["status"] == 0
    assert info["iters"] == ref["iters"]
    assert info["res_init"] == ref["res_init"] and info["res_final"] == ref["res_final"]
    assert np.array_equal(x.download(), x_ref.download())
    return A.nrow


def test_noalloc_pcg_matches_workspace_pcg(ctx):
    # first a 3D system with int16 columns, then a smaller 2D one with int32
    # columns: the second call runs over what the first left in the array
    n1 = _case(ctx, 3, 40, 5, hflmp.SELL_COL16)
    n2 = _case(ctx, 2, 60, 6, 0)
    assert n1 > n2


def test_noalloc_pcg_refuses_oversize(ctx):
    # 2D Q1 512^2: 263169 rows > 262144; refused before any launch
    rp, ci, v = hflmp.fe_system(1, 2, 512)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    S = hflmp.SellMatrix(ctx, A, 0)
    invd = hflmp.invdiag(ctx, A)
    db, x = ctx.upload(np.ones(A.nrow)), ctx.zeros(A.nrow, np.float64)
    rc, _ = _noalloc(ctx, S, invd, db, x)
    assert rc == -1  # HFLMP_EINVAL
    assert b"262144" in ctx.L.hflmp_ctx_last_error(ctx.h)
    assert not x.download().any()
