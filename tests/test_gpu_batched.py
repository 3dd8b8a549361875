"""Several right-hand sides at once against the CPU oracle run one column at
a time: the SpMM is bit-identical to the single-vector SpMV of each column,
and every column of the lockstep Jacobi-PCG matches the oracle's CG on that
right-hand side (iterations within 2, solution within 1e-8)."""
import numpy as np
import pytest

from ba_amd import hflmp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("k", [64, 7, 1])
def test_spmm_columns_bit_exact(ctx, k, dtype):
    rp, ci, v = hflmp.fe_system(2, 2, 40, hflmp.fe_kappa(2, 40, 1.0, 3))
    v = v.astype(dtype)
    N = len(rp) - 1
    X = np.random.default_rng(k).uniform(-1, 1, (N, k)).astype(dtype)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    dX, dY = ctx.upload(X.ravel()), ctx.empty(N * k, dtype)
    hflmp.spmm(ctx, A, dX, dY, k)
    Y = dY.download().reshape(N, k)
    for c in range(k):
        assert np.array_equal(Y[:, c], O.spmv(rp, ci, v, np.ascontiguousarray(X[:, c])))


def test_spmm_long_rows(ctx):
    # rows longer than a wave (64) take several broadcast rounds
    from tests.helpers import random_csr
    rng = np.random.default_rng(0)
    lengths = rng.integers(0, 200, 300)
    lengths[5] = 0
    rp, ci, v = random_csr(300, 500, lengths, seed=1)
    k = 33
    X = rng.uniform(-1, 1, (500, k))
    A = hflmp.CsrMatrix(ctx, rp, ci, v, ncol=500)
    dX, dY = ctx.upload(X.ravel()), ctx.empty(300 * k, np.float64)
    hflmp.spmm(ctx, A, dX, dY, k)
    Y = dY.download().reshape(300, k)
    for c in (0, 13, 32):
        assert np.array_equal(Y[:, c], O.spmv(rp, ci, v, np.ascontiguousarray(X[:, c])))


@pytest.mark.parametrize("k,planned", [(64, False), (5, False), (64, True), (5, True)])
def test_batched_pcg_vs_per_column_oracle(ctx, k, planned):
    # planned: the iteration's products through the block layout (patches
    # of the 97-node lines of this Q2 mesh)
    order, dim, n = 2, 2, 48
    rp, ci, v = hflmp.fe_system(order, dim, n)
    N = len(rp) - 1
    B = np.stack([hflmp.fe_rhs(order, dim, n, 100 + c) for c in range(k)], axis=1)
    B[:, 3] = 0.0  # a zero right-hand side converges at iteration 0
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    invd = hflmp.invdiag(ctx, A)
    dB, dX = ctx.upload(B.ravel()), ctx.zeros(N * k, np.float64)
    P = hflmp.SpmmPlan(ctx, A, k) if planned else None
    if planned:
        assert P.info()["line_rows"] == 97 and P.info()["runs"] > 1
    info = hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX, k, rtol=1e-8, maxit=5000, plan=P)
    if P is not None:
        P.free()
    X = dX.download().reshape(N, k)
    for c in range(k):
        b = np.ascontiguousarray(B[:, c])
        xo, io = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-8, maxit=5000)
        assert info["status"][c] == 0
        assert abs(info["iters"][c] - io["iters"]) <= 2, (c, info["iters"][c], io["iters"])
        nrm = np.linalg.norm(xo)
        assert np.linalg.norm(X[:, c] - xo) <= 1e-8 * max(nrm, 1e-300) or nrm == 0
    assert info["iters"][3] == 0


@pytest.mark.parametrize("planned", [False, True])
def test_batched_pcg_fp32(ctx, planned):
    rp, ci, v = hflmp.fe_system(1, 2, 40)
    N = len(rp) - 1
    k = 16
    B = np.stack([hflmp.fe_rhs(1, 2, 40, c) for c in range(k)], axis=1).astype(np.float32)
    v32 = v.astype(np.float32)
    A = hflmp.CsrMatrix(ctx, rp, ci, v32)
    invd = hflmp.invdiag(ctx, A)
    dB, dX = ctx.upload(B.ravel()), ctx.zeros(N * k, np.float32)
    P = hflmp.SpmmPlan(ctx, A, k) if planned else None
    info = hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX, k, rtol=1e-5, maxit=5000, plan=P)
    if P is not None:
        P.free()
    X = dX.download().reshape(N, k)
    for c in (0, 7, 15):
        xo, io = O.cg(rp, ci, v32, np.ascontiguousarray(B[:, c]), prec=O.PREC_JACOBI, rtol=1e-5)
        assert abs(info["iters"][c] - io["iters"]) <= 2
        assert np.linalg.norm(X[:, c] - xo) <= 1e-4 * np.linalg.norm(xo)


@pytest.mark.parametrize("lengths", [(32, 33, 64, 65, 0, 1, 31)])
def test_spmm_batch_boundaries(ctx, lengths):
    # rows at the edges of the gather batches (32 entries at once, a second
    # batch up to 64, the generic loop past 64) and empty rows
    from tests.helpers import random_csr
    L = np.array(lengths * 20)
    rp, ci, v = random_csr(len(L), 300, L, seed=4)
    k = 64
    X = np.random.default_rng(5).uniform(-1, 1, (300, k))
    A = hflmp.CsrMatrix(ctx, rp, ci, v, ncol=300)
    dX, dY = ctx.upload(X.ravel()), ctx.empty(len(L) * k, np.float64)
    hflmp.spmm(ctx, A, dX, dY, k)
    Y = dY.download().reshape(len(L), k)
    for c in (0, 31, 63):
        assert np.array_equal(Y[:, c], O.spmv(rp, ci, v, np.ascontiguousarray(X[:, c])))


@pytest.mark.parametrize("k,planned", [(16, False), (5, False), (16, True)])
def test_batched_mixed_vs_mpir_oracle(ctx, k, planned):
    """FP32-inner / FP64-refined blocked CG against the oracle's MPIR
    (or_mpir_ds: FP64 residual, FP32 Jacobi-PCG correction) column by column:
    corrections within 1, the true FP64 residual below 1e-8, and x within
    1e-6 of the oracle's and of the FP64 CPU CG solve (planned: the inner
    products through the FP32 block layout)."""
    rp, ci, v = hflmp.fe_system(2, 2, 24, hflmp.fe_kappa(2, 24, 1.0, 41))
    N = len(rp) - 1
    B = np.stack([hflmp.fe_rhs(2, 2, 24, 100 + c) for c in range(k)], axis=1)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    A32 = A.astype_f32()
    invd32 = hflmp.invdiag(ctx, A32)
    dB, dX = ctx.upload(np.ascontiguousarray(B).ravel()), ctx.zeros(N * k, np.float64)
    P = hflmp.SpmmPlan(ctx, A32, k, 4) if planned else None
    info = hflmp.batched_mpcg_jacobi(ctx, A, A32, invd32, dB, dX, k, rtol=1e-8,
                                     inner_rtol=1e-4, plan32=P)
    if P is not None:
        P.free()
    X = dX.download().reshape(N, k)
    for c in range(k):
        b = np.ascontiguousarray(B[:, c])
        xo, oo = O.mpir(rp, ci, v, b, max_outer=50, rtol=1e-8, inner_rtol=1e-4,
                        inner_maxit=100000)
        assert info["status"][c] == 0
        assert abs(int(info["outer"][c]) - oo["outer"]) <= 1, (info["outer"][c], oo)
        xg = np.ascontiguousarray(X[:, c])
        r = b - O.spmv(rp, ci, v, xg)
        assert np.linalg.norm(r) < 1e-8 * np.linalg.norm(b)
        assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
        xc, _ = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-10, maxit=100000)
        assert np.linalg.norm(xg - xc) <= 1e-6 * np.linalg.norm(xc)


def test_batched_pcg_plan_matches_csr_form(ctx):
    # the same solve with the products through the block layout: equal
    # iteration counts and x to rounding (only the order of the p.Ap partial
    # sums differs); a plan built for another matrix is refused
    rp, ci, v = hflmp.fe_system(2, 2, 40)
    N, k = len(rp) - 1, 64
    B = np.stack([hflmp.fe_rhs(2, 2, 40, 200 + c) for c in range(k)], axis=1)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    invd = hflmp.invdiag(ctx, A)
    dB = ctx.upload(B.ravel())
    dX1, dX2 = ctx.zeros(N * k, np.float64), ctx.zeros(N * k, np.float64)
    i1 = hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX1, k, rtol=1e-8, maxit=20000)
    P = hflmp.SpmmPlan(ctx, A, k)
    i2 = hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX2, k, rtol=1e-8, maxit=20000, plan=P)
    assert (i1["status"] == 0).all() and (i2["status"] == 0).all()
    assert np.abs(i1["iters"] - i2["iters"]).max() <= 1
    X1, X2 = dX1.download().reshape(N, k), dX2.download().reshape(N, k)
    assert np.linalg.norm(X1 - X2) <= 1e-9 * np.linalg.norm(X1)
    rp2, ci2, v2 = hflmp.fe_system(2, 2, 30)
    A2 = hflmp.CsrMatrix(ctx, rp2, ci2, v2)
    P2 = hflmp.SpmmPlan(ctx, A2, k)
    with pytest.raises(hflmp.HflmpError):
        hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX2, k, rtol=1e-8, maxit=5000, plan=P2)
    P.free(); P2.free(); A2.free(); A.free()
