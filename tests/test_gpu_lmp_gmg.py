"""PreconditionerGMG in the C++ lmp layer: the unchanged CG<LAD> template
over CoupledMatrix / CoupledVector, preconditioned by one device V-cycle
per iteration, against the oracle's CG with the oracle's V-cycle and
against the device-resident hflmp_pcg_mg_f64 (within 2 iterations and
1e-8); FGMRES with it by the true residual; the refusals."""
import numpy as np
import pytest
import scipy.sparse as sp

from ba_amd import hflmp, mg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PRM = dict(pre_smooth=1, post_smooth=1, coarse_sweeps=8, omega=0.8, coarse_omega=0.8)
OKW = dict(pre=1, post=1, coarse=8, omega=0.8, coarse_omega=0.8)


def _hierarchy(dim, n, seed=7):
    kap = hflmp.fe_kappa(dim, n, 1.0, seed)
    E = hflmp.fe_element(1, dim, n)
    A = hflmp.fe_system(1, dim, n, kap)
    sizes = mg.level_sizes(n)
    mats = mg.coarse_cell_mats(dim, n, E, kap, sizes)
    R = [mg.restriction(dim, nc) for nc in sizes[1:]]
    P = [mg.prolongation(dim, nc) for nc in sizes[1:]]
    levels = O.mg_levels(A, R, P, mats, [hflmp.fe_cell_dofs(1, dim, nc) for nc in sizes[1:]],
                         [hflmp.fe_fixed(1, dim, nc) for nc in sizes[1:]])
    rows = [(s + 1) ** dim for s in sizes]
    coarse = [lv["A"] + (rows[l],) for l, lv in enumerate(levels) if l > 0]
    Rt = [R[l] + (rows[l],) for l in range(len(R))]
    Pt = [P[l] + (rows[l + 1],) for l in range(len(P))]
    return kap, E, A, levels, coarse, Rt, Pt


@pytest.mark.parametrize("case", [(2, 64), (3, 16), (3, 32)])
def test_lmp_cg_gmg_vs_oracle_and_device_pcg(ctx, case):
    dim, n = case
    print("lmp cg gmg", case, flush=True)
    kap, E, A, levels, coarse, R, P = _hierarchy(dim, n)
    b = hflmp.fe_rhs(1, dim, n, 5)
    x, info = hflmp.lmp_solve_gmg(*A, b, coarse, R, P, rtol=1e-8, maxit=500, **PRM)
    xo, io = O.cg_prec(A, b, lambda r: O.mg_cycle(levels, r, **OKW), rtol=1e-8)
    print("iterations: lmp %d, oracle %d" % (info["iters"], io["iters"]))
    assert info["status"] == 0 and io["status"] == 0
    assert abs(info["iters"] - io["iters"]) <= 2, (info, io)
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    # the device-resident PCG over the same hierarchy
    dA = hflmp.CsrMatrix(ctx, *A)
    M = mg.build_structured(ctx, dim, n, E, kap, dA, structured_transfers=False, **PRM)
    try:
        db, dx = ctx.upload(b), ctx.zeros(len(b), np.float64)
        dinfo = M.pcg(db, dx, rtol=1e-8)
        assert abs(info["iters"] - dinfo["iters"]) <= 2, (info, dinfo)
        assert np.linalg.norm(x - dx.download()) <= 1e-8 * np.linalg.norm(xo)
        db.free(), dx.free()
    finally:
        M.free()
        dA.free()


def test_lmp_cg_gmg_dense_coarse(ctx):
    dim, n = 3, 16
    kap, E, A, levels, coarse, R, P = _hierarchy(dim, n)
    rpc, cic, vc, nc = coarse[-1]
    Cinv = np.linalg.inv(sp.csr_matrix((vc, cic, rpc), shape=(nc, nc)).toarray())
    levels[-1]["Cinv"] = Cinv
    b = hflmp.fe_rhs(1, dim, n, 3)
    x, info = hflmp.lmp_solve_gmg(*A, b, coarse, R, P, coarse_inverse=Cinv, rtol=1e-8,
                                  maxit=500, **PRM)
    xo, io = O.cg_prec(A, b, lambda r: O.mg_cycle(levels, r, **OKW), rtol=1e-8)
    assert info["status"] == 0 and io["status"] == 0
    assert abs(info["iters"] - io["iters"]) <= 2, (info, io)
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_lmp_fgmres_gmg_true_residual():
    dim, n = 3, 16
    kap, E, A, levels, coarse, R, P = _hierarchy(dim, n, seed=3)
    b = hflmp.fe_rhs(1, dim, n, 4)
    x, info = hflmp.lmp_solve_gmg(*A, b, coarse, R, P, solver=hflmp.LMP_FGMRES, basis=30,
                                  rtol=1e-9, maxit=500, **PRM)
    _, jac = hflmp.lmp_solve(*A, b, solver=hflmp.LMP_FGMRES, precond=hflmp.LMP_PREC_JACOBI,
                             basis=30, rtol=1e-9, maxit=5000)
    assert info["status"] == 0
    assert np.linalg.norm(b - O.spmv(*A, x)) <= 2e-9 * info["res_init"]
    assert info["iters"] * 3 < jac["iters"], (info, jac)


def test_lmp_gmg_refusals():
    dim, n = 2, 16
    kap, E, A, levels, coarse, R, P = _hierarchy(dim, n)
    b = hflmp.fe_rhs(1, dim, n, 1)
    with pytest.raises(hflmp.HflmpError, match="hierarchy"):
        hflmp.lmp_solve(*A, b, precond=hflmp.LMP_PREC_GMG)
    with pytest.raises(ValueError):
        hflmp.lmp_solve_gmg(*A, b, coarse, R[:-1], P)
    rp, ci, v, nr = coarse[0]
    bad = (rp, ci, v, 0)     # no columns
    with pytest.raises(hflmp.HflmpError, match="coarse operator"):
        hflmp.lmp_solve_gmg(*A, b, [bad] + coarse[1:], R, P)
    with pytest.raises(hflmp.HflmpError, match="bad arguments"):
        hflmp.lmp_solve_gmg(*A, b, coarse, R, P, solver=hflmp.LMP_MPIR)
    # a column past the level's width is refused on the host
    badc = ci.copy()
    badc[5] = nr
    with pytest.raises(hflmp.HflmpError, match="out of range"):
        hflmp.lmp_solve_gmg(*A, b, [(rp, badc, v, nr)] + coarse[1:], R, P)
    with pytest.raises(ValueError):
        hflmp.lmp_solve_gmg(*A, b, [], [], [], coarse_inverse=np.eye(len(b)))
