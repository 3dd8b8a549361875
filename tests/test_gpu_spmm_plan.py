"""The row-block SpMM layout (hflmp_spmm_plan_*): every column of Y = A X
through the plan is bit-identical to the oracle's CSR SpMV of that column
(the order of CPUsimple_CSR_lMatrix::VectorMult), for FE patterns, ragged
and empty rows and every block size the builder can pick; matrices whose
blocks span too many columns are refused, as are columns out of range."""
import ctypes as C

import numpy as np
import pytest

from ba_amd import hflmp
from oracle import oracle as O
from tests.helpers import random_csr

pytestmark = pytest.mark.gpu


def _check(ctx, rp, ci, v, ncol, k, dtype, budget=0):
    N = len(rp) - 1
    v = v.astype(dtype)
    X = np.random.default_rng(k + N).uniform(-1, 1, (ncol, k)).astype(dtype)
    A = hflmp.CsrMatrix(ctx, rp, ci, v, ncol=ncol)
    P = hflmp.SpmmPlan(ctx, A, k, np.dtype(dtype).itemsize, budget)
    dX, dY = ctx.upload(X.ravel()), ctx.empty(N * k, dtype)
    try:
        P.spmm(A, dX, dY, k)
        Y = dY.download().reshape(N, k)
        for c in range(k):
            yo = O.spmv(rp, ci, v, np.ascontiguousarray(X[:, c]))
            assert np.array_equal(Y[:, c], yo), (c, np.max(np.abs(Y[:, c] - yo)))
        # the plan's product is the CSR SpMM's, bit for bit
        dZ = ctx.empty(N * k, dtype)
        hflmp.spmm(ctx, A, dX, dZ, k)
        assert np.array_equal(dZ.download().reshape(N, k), Y)
        dZ.free()
        return P.info()
    finally:
        P.free()
        dX.free()
        dY.free()
        A.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("k", [64, 33, 7, 1])
def test_plan_fe_q2_columns_bit_exact(ctx, k, dtype):
    rp, ci, v = hflmp.fe_system(2, 2, 40, hflmp.fe_kappa(2, 40, 1.0, 3))
    info = _check(ctx, rp, ci, v, len(rp) - 1, k, dtype)
    assert 4 <= info["rows_per_block"] <= 64
    assert info["rows_per_block"] == info["run_rows"] * info["runs"]
    assert info["max_cols_per_block"] <= 128


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plan_fe_q1_3d(ctx, dtype):
    rp, ci, v = hflmp.fe_system(1, 3, 20)
    info = _check(ctx, rp, ci, v, len(rp) - 1, 64, dtype)
    if dtype == np.float32:
        # up to 128 columns per block: patches of a plane of the numbering,
        # whose lines of 21 nodes are found
        assert info["line_rows"] == 21 and info["runs"] > 1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plan_q2_patches(ctx, dtype):
    # 2D Q2 on 40^2 cells: lines of 81 nodes (neither 81 nor the 81 lines
    # divide into whole patches), so edge patches are clipped at the line
    # end and the last line group is partial
    rp, ci, v = hflmp.fe_system(2, 2, 40, hflmp.fe_kappa(2, 40, 1.0, 5))
    info = _check(ctx, rp, ci, v, len(rp) - 1, 64, dtype)
    assert info["line_rows"] == 81 and info["runs"] > 1
    assert info["rows_per_block"] == info["run_rows"] * info["runs"]
    # fewer staged X rows than runs of consecutive rows would need
    assert info["list_entries"] < 4 * (len(rp) - 1)


def _shifted(rp, ci, v, m):
    """m rows holding only their diagonal, then the matrix (indices + m):
    the lines of the numbering no longer start at a multiple of L."""
    rp2 = np.concatenate([np.arange(m, dtype=np.int32), rp + m]).astype(np.int32)
    ci2 = np.concatenate([np.arange(m, dtype=np.int32), ci + m]).astype(np.int32)
    v2 = np.concatenate([np.full(m, 2.0), v])
    return rp2, ci2, v2


@pytest.mark.parametrize("m", [3, 40])
def test_plan_patches_shifted_lines(ctx, m):
    rp, ci, v = _shifted(*hflmp.fe_system(2, 2, 24), m)
    info = _check(ctx, rp, ci, v, len(rp) - 1, 48, np.float64)
    assert info["line_rows"] == 49


@pytest.mark.parametrize("budget", [8192, 16384, 160 * 1024])
def test_plan_block_sizes(ctx, budget):
    # a smaller LDS budget forces smaller blocks; the products do not change
    rp, ci, v = hflmp.fe_system(1, 2, 64)
    info = _check(ctx, rp, ci, v, len(rp) - 1, 16, np.float64, budget=budget)
    assert info["max_cols_per_block"] * 16 * 8 <= budget


def _banded_csr(nrow, ncol, lengths, half, seed):
    """CSR whose row i holds lengths[i] distinct sorted columns drawn from
    [i - half, i + half] (clipped to [0, ncol))."""
    rng = np.random.default_rng(seed)
    rp = np.zeros(nrow + 1, np.int32)
    cols = []
    for i in range(nrow):
        lo, hi = max(0, i - half), min(ncol, i + half + 1)
        L = min(int(lengths[i]), hi - lo)
        c = np.sort(rng.choice(np.arange(lo, hi), size=L, replace=False)) if L else []
        cols.append(np.asarray(c, np.int32))
        rp[i + 1] = rp[i] + L
    ci = np.concatenate(cols).astype(np.int32)
    return rp, ci, rng.uniform(-1, 1, len(ci))


def test_plan_ragged_and_empty_rows(ctx):
    rng = np.random.default_rng(5)
    nrow, ncol = 333, 900
    lengths = rng.integers(0, 21, nrow)
    lengths[[0, 7, 8, 9, 332]] = 0
    rp, ci, v = _banded_csr(nrow, ncol, lengths, 25, seed=2)
    _check(ctx, rp, ci, v, ncol, 5, np.float64)
    # FP32 blocks take up to 128 columns: a row longer than a wave
    lengths = rng.integers(0, 41, nrow)
    lengths[100] = 100
    rp, ci, v = _banded_csr(nrow, ncol, lengths, 60, seed=3)
    _check(ctx, rp, ci, v, ncol, 64, np.float32, budget=160 * 1024)


def test_plan_refuses_scattered_columns(ctx):
    # random columns over the whole range: four rows already exceed the
    # columns one block may stage
    lengths = np.full(200, 40)
    rp, ci, v = random_csr(200, 900, lengths, seed=4)
    A = hflmp.CsrMatrix(ctx, rp, ci, v, ncol=900)
    with pytest.raises(hflmp.HflmpError) as e:
        hflmp.SpmmPlan(ctx, A, 8, 8)
    assert "(-6)" in str(e.value)  # HFLMP_EFORMAT
    A.free()


def test_plan_rectangular_leading_empty_block(ctx):
    # the first block has no entries at all; columns beyond nrow
    nrow, ncol = 70, 5000
    lengths = np.zeros(nrow, np.int64)
    lengths[20:] = 3
    rp, ci, v = random_csr(nrow, ncol, lengths, seed=3)
    ci[-1] = ncol - 1
    _check(ctx, rp, ci, v, ncol, 8, np.float64)


def test_plan_refuses_wide_blocks(ctx):
    # every row spans the whole column range: no block fits the budget
    nrow, ncol = 64, 70000
    rp = np.arange(0, 2 * nrow + 1, 2, dtype=np.int32)
    ci = np.empty(2 * nrow, np.int32)
    ci[0::2] = 0
    ci[1::2] = ncol - 1
    v = np.ones(2 * nrow)
    A = hflmp.CsrMatrix(ctx, rp, ci, v, ncol=ncol)
    with pytest.raises(hflmp.HflmpError) as e:
        hflmp.SpmmPlan(ctx, A, 64, 8)
    assert "(-6)" in str(e.value)  # HFLMP_EFORMAT
    A.free()


def test_plan_refuses_bad_column(ctx):
    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 1, 1], np.int32)
    v = np.ones(3)
    A = hflmp.CsrMatrix(ctx, rp, ci, v, ncol=2)
    # a column past ncol, written into the device copy behind the matrix's back
    bad = ctx.upload(np.array([0, 1, 2], np.int32))
    h = C.c_void_p()
    rc = ctx.L.hflmp_spmm_plan_create(ctx.h, 2, 2, 3, A.rp.ptr, bad.ptr, 8, 8, 0, C.byref(h))
    assert rc == -1  # HFLMP_EINVAL
    assert not h.value
    bad.free()
    A.free()
