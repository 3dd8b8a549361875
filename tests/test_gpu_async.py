"""Block-asynchronous Jacobi (AsynchronousIterationGPU made deterministic)
on the device against the oracle's restatement (pinned in
tests/test_oracle.py by the one-block / one-sweep / one-row-block closed
forms and a dense two-block iteration): the iterates of the smoothing modes
are bit-identical for blocks held in LDS (up to 2048 rows), for larger
blocks swept through global memory and for one block of all rows; the solve
modes stop within one round of the oracle's inner-iteration count with the
solution within the residual tolerance."""
import numpy as np
import pytest

from ba_amd import hflmp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _system(order, dim, n, seed):
    rp, ci, v = hflmp.fe_system(order, dim, n, hflmp.fe_kappa(dim, n, 1.0, seed))
    b = hflmp.fe_rhs(order, dim, n, seed + 1)
    return rp, ci, v, b


def _run(ctx, rp, ci, v, b, x0, **kw):
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    db, dx = ctx.upload(b.astype(v.dtype)), ctx.upload(x0.astype(v.dtype))
    info = hflmp.async_jacobi(ctx, A, db, dx, **kw)
    return dx.download(), info


@pytest.mark.parametrize("brows", [1, 37, 256, 2048, 3000, 10 ** 9])
@pytest.mark.parametrize("mode,w", [(2, 1.0), (3, 0.8)])
def test_async_smooth_bit_exact(ctx, brows, mode, w):
    # 2D Q1 80^2: 6561 rows -- blocks in LDS (<= 2048 rows), one larger
    # block size through global sweeps, and one block of all rows
    rp, ci, v, b = _system(1, 2, 80, 3)
    x0 = np.random.default_rng(brows % 1000).uniform(-1, 1, len(rp) - 1)
    kw = dict(block_rows=brows, n_global=2, n_block=2, n_inner=3, mode=mode, w=w)
    x, info = _run(ctx, rp, ci, v, b, x0, **kw)
    xo, io = O.async_jacobi(rp, ci, v, b, x0, **kw)
    assert info["iters"] == io["iters"] == 12
    assert np.array_equal(x, xo)


def test_async_smooth_3d_fp32(ctx):
    rp, ci, v, b = _system(1, 3, 20, 5)
    v32, b32 = v.astype(np.float32), b.astype(np.float32)
    x0 = np.random.default_rng(2).uniform(-1, 1, len(rp) - 1).astype(np.float32)
    kw = dict(block_rows=512, n_inner=4, mode=3, w=0.75)
    x, _ = _run(ctx, rp, ci, v32, b32, x0, **kw)
    xo, _ = O.async_jacobi(rp, ci, v32, b32, x0, **kw)
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("mode,w,brows", [(0, 1.0, 256), (1, 0.9, 1024), (1, 0.9, 5000)])
def test_async_solve_vs_oracle(ctx, mode, w, brows):
    rp, ci, v, b = _system(1, 2, 24, 9)
    n_inner = 3
    kw = dict(block_rows=brows, n_inner=n_inner, mode=mode, w=w, rtol=1e-6, maxit=50000)
    x, info = _run(ctx, rp, ci, v, b, np.zeros(len(rp) - 1), **kw)
    xo, io = O.async_jacobi(rp, ci, v, b, None, **kw)
    assert info["status"] == 0 and io["status"] == 0
    # the residual's squares are summed in another order, so the stopping
    # test may flip one round (n_inner sweeps) earlier or later
    assert abs(info["iters"] - io["iters"]) <= n_inner, (info["iters"], io["iters"])
    assert info["res_init"] == pytest.approx(io["res_init"], rel=1e-13)
    r = b - O.spmv(rp, ci, v, x)
    assert np.linalg.norm(r) < 1e-6 * np.linalg.norm(b) * 1.01
    if info["iters"] == io["iters"]:
        assert np.array_equal(x, xo)  # the iterates themselves are bit-identical
    else:
        assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo)


def test_async_refuses_zero_diagonal(ctx):
    rp, ci, v, b = _system(1, 2, 6, 1)
    v = v.copy()
    k = int(rp[10]) + list(ci[rp[10]:rp[11]]).index(10)
    v[k] = 0.0
    with pytest.raises(hflmp.HflmpError, match="diagonal"):
        _run(ctx, rp, ci, v, b, np.zeros(len(rp) - 1), mode=0)
