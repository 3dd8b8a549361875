"""Geometric multigrid on the CPU: the hierarchy builder (ba_amd/mg.py)
against independent constructions, and the oracle's V-cycle, residual and
damped Jacobi sweep against what the mathematics fixes."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ba_amd import hflmp, mg
from oracle import oracle as O


def _sp(rp, ci, v, ncol=None):
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n if ncol is None else ncol))


def _coords(dim, n):
    """Node coordinates (x fastest) of the (n+1)^dim grid on [0,1]^dim."""
    g = np.arange(n + 1) / n
    if dim == 2:
        y, x = np.meshgrid(g, g, indexing="ij")
        return [x.ravel(), y.ravel()]
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return [x.ravel(), y.ravel(), z.ravel()]


def _hierarchy(dim, n, sigma=1.0, seed=7, nlevels=None):
    kap = O.fe_kappa(dim, n, sigma, seed) if sigma is not None else np.ones(n ** dim)
    E = O.fe_element(1, dim, n)
    A = O.fe_system(1, dim, n, kap)
    sizes = mg.level_sizes(n, nlevels)
    mats = mg.coarse_cell_mats(dim, n, E, kap, sizes)
    R = [mg.restriction(dim, nc) for nc in sizes[1:]]
    P = [mg.prolongation(dim, nc) for nc in sizes[1:]]
    levels = O.mg_levels(A, R, P, mats, [hflmp.fe_cell_dofs(1, dim, nc) for nc in sizes[1:]],
                         [hflmp.fe_fixed(1, dim, nc) for nc in sizes[1:]])
    return kap, E, A, sizes, mats, levels


@pytest.mark.parametrize("dim", [2, 3])
def test_interpolation_reproduces_multilinear(dim):
    nc = 4
    Pf = mg.prolongation_matrix(dim, nc, interior=False)
    f = lambda X: np.prod([(1.0 + (d + 2) * X[d]) for d in range(dim)], axis=0)
    assert np.allclose(Pf @ f(_coords(dim, nc)), f(_coords(dim, 2 * nc)), rtol=0, atol=1e-14)
    # rows sum to one (constants reproduced)
    assert np.allclose(np.asarray(Pf.sum(axis=1)).ravel(), 1.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_prolongation_drops_dirichlet(dim):
    nc = 4
    Pf = mg.prolongation_matrix(dim, nc, interior=False).toarray()
    rp, ci, v = mg.prolongation(dim, nc)
    P = _sp(rp, ci, v, (nc + 1) ** dim).toarray()
    bf = np.zeros((2 * nc + 1) ** dim, bool)
    bf[hflmp.fe_fixed(1, dim, 2 * nc)] = True
    bc = np.zeros((nc + 1) ** dim, bool)
    bc[hflmp.fe_fixed(1, dim, nc)] = True
    expect = Pf.copy()
    expect[bf, :] = 0
    expect[:, bc] = 0
    assert np.array_equal(P, expect)
    assert np.all(np.diff(rp)[bf] == 0)
    rr, rc, rv = mg.restriction(dim, nc)
    assert np.array_equal(_sp(rr, rc, rv, (2 * nc + 1) ** dim).toarray(), P.T)
    for a, b in ((rp, ci), (rr, rc)):   # rows sorted
        assert all(np.all(np.diff(b[a[i]:a[i + 1]]) > 0) for i in range(len(a) - 1))


@pytest.mark.parametrize("dim", [2, 3])
def test_child_interp_partition_of_unity(dim):
    Q = mg.child_interp(dim)
    assert np.allclose(Q.sum(axis=2), 1.0)
    # a child's corner that is a coarse corner takes exactly that node
    nloc = 2 ** dim
    for k in range(nloc):
        assert Q[k, k, k] == 1.0


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 8)])
def test_galerkin_coarse_operator(dim, n):
    """The cell-by-cell coarse operators equal P^T A P of the level above
    on the interior nodes (scipy's sparse product), two levels deep."""
    _, _, A, sizes, _, levels = _hierarchy(dim, n, nlevels=3)
    Af = _sp(*A)
    for l in range(2):
        nc = sizes[l + 1]
        Ac = _sp(*levels[l + 1]["A"])
        Pm = _sp(*levels[l]["P"], (nc + 1) ** dim)
        G = (Pm.T @ Af @ Pm).toarray()
        inner = np.ones((nc + 1) ** dim, bool)
        inner[hflmp.fe_fixed(1, dim, nc)] = False
        D = Ac.toarray()[np.ix_(inner, inner)]
        assert np.allclose(D, G[np.ix_(inner, inner)], rtol=0, atol=1e-12 * np.abs(D).max())
        Af = Ac


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 8)])
def test_constant_coefficient_coarse_equals_rediscretisation(dim, n):
    """With kappa = 1 the Galerkin operator of nested Q1 spaces is the Q1
    stiffness matrix of the coarse mesh: the generator's own system."""
    _, _, _, sizes, _, levels = _hierarchy(dim, n, sigma=None, nlevels=3)
    for l in (1, 2):
        rp, ci, v = levels[l]["A"]
        rg, cg, vg = O.fe_system(1, dim, sizes[l])
        assert np.array_equal(rp, rg) and np.array_equal(ci, cg)
        assert np.allclose(v, vg, rtol=0, atol=1e-13 * np.abs(vg).max())


def test_level_sizes():
    assert mg.level_sizes(128) == [128, 64, 32, 16, 8, 4, 2]
    assert mg.level_sizes(48, min_cells=4) == [48, 24, 12, 6]
    assert mg.level_sizes(12) == [12, 6, 3]
    assert mg.level_sizes(64, nlevels=3) == [64, 32, 16]
    with pytest.raises(ValueError):
        mg.level_sizes(12, nlevels=4)
    with pytest.raises(ValueError):
        mg.coarse_cells(2, 5, E=np.eye(4), kappa=np.ones(25))


def test_residual_and_damped_sweep():
    rp, ci, v = O.fe_system(1, 3, 6, O.fe_kappa(3, 6, 1.0, 2))
    A = _sp(rp, ci, v)
    rng = np.random.default_rng(1)
    x, b = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    r = O.residual_sub(rp, ci, v, b, x)
    assert np.allclose(r, b - A @ x, rtol=0, atol=1e-13)
    invd = O.invdiag(rp, ci, v)
    w = 0.7
    y = O.jacobi_damped(rp, ci, v, invd, w, b, x)
    assert np.allclose(y, x + w * invd * (b - A @ x), rtol=0, atol=1e-13)
    # the exact solution is a fixed point of the sweep
    xs = spla.spsolve(A.tocsc(), b)
    assert np.allclose(O.jacobi_damped(rp, ci, v, invd, w, b, xs), xs, rtol=0, atol=1e-12)
    # x = 0: the sweep is (w invd) b, the cycle's first step
    z = np.zeros_like(b)
    assert np.allclose(O.jacobi_damped(rp, ci, v, invd, w, b, z), w * invd * b, rtol=0,
                       atol=0)


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 8)])
def test_vcycle_symmetric_positive(dim, n):
    _, _, A, _, _, levels = _hierarchy(dim, n)
    N = len(A[0]) - 1
    inner = np.ones(N, bool)
    inner[hflmp.fe_fixed(1, dim, n)] = False
    rng = np.random.default_rng(3)
    u, w = rng.standard_normal(N) * inner, rng.standard_normal(N) * inner
    Mu, Mw = O.mg_cycle(levels, u), O.mg_cycle(levels, w)
    assert np.all(Mu[~inner] == 0) and np.all(Mw[~inner] == 0)
    assert abs(u @ Mw - w @ Mu) <= 1e-12 * abs(u @ Mw)
    assert u @ Mu > 0 and w @ Mw > 0
    # linear in its argument
    assert np.allclose(O.mg_cycle(levels, 2.0 * u - w), 2.0 * Mu - Mw, rtol=0, atol=1e-12)


def test_vcycle_single_level_is_jacobi():
    rp, ci, v = O.fe_system(1, 2, 8)
    levels = [dict(A=(rp, ci, v), invd=O.invdiag(rp, ci, v))]
    b = O.fe_rhs(1, 2, 8, 4)
    x = O.mg_cycle(levels, b, coarse=3, coarse_omega=0.6)
    y = 0.6 * levels[0]["invd"] * b
    for _ in range(2):
        y = O.jacobi_damped(rp, ci, v, levels[0]["invd"], 0.6, b, y)
    assert np.array_equal(x, y)


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 16)])
def test_mg_pcg_and_solver_converge(dim, n):
    kap, _, A, _, _, levels = _hierarchy(dim, n)
    b = O.fe_rhs(1, dim, n, 5)
    xs = spla.spsolve(_sp(*A).tocsc(), b)
    _, jac = O.cg(*A, b)
    x, info = O.cg_prec(A, b, lambda r: O.mg_cycle(levels, r, 1, 1, 8, 0.8, 0.8))
    assert info["status"] == 0 and info["iters"] * 3 < jac["iters"]
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)
    x2, info2 = O.mg_solve(levels, b, rtol=1e-8, pre=1, post=1, coarse=8, omega=0.8,
                           coarse_omega=0.8)
    assert info2["status"] == 0 and info2["iters"] < 60
    assert np.linalg.norm(x2 - xs) <= 1e-6 * np.linalg.norm(xs)


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 8)])
def test_matrix_free_level_agrees_with_assembled(dim, n):
    """The finest level given as the stencil (Richardson-Jacobi smoothing,
    VectorMultAdd residual) is the same operator as the assembled level with
    the damped Jacobi step: the cycles agree to rounding, and stay
    symmetric."""
    kap, E, A, _, _, levels = _hierarchy(dim, n)
    fixed = hflmp.fe_fixed(1, dim, n)
    mf = [dict(levels[0])] + levels[1:]
    del mf[0]["A"]
    mf[0]["S"] = (1, dim, [n] * dim, E, kap, fixed)
    N = len(A[0]) - 1
    inner = np.ones(N, bool)
    inner[fixed] = False
    rng = np.random.default_rng(5)
    u, w = rng.standard_normal(N) * inner, rng.standard_normal(N) * inner
    for kw in (dict(), dict(pre=2, post=2, coarse=3, omega=0.6, coarse_omega=0.9)):
        a, m = O.mg_cycle(levels, u, **kw), O.mg_cycle(mf, u, **kw)
        assert np.allclose(m, a, rtol=0, atol=1e-12 * np.abs(a).max())
    Mu, Mw = O.mg_cycle(mf, u), O.mg_cycle(mf, w)
    assert abs(u @ Mw - w @ Mu) <= 1e-12 * abs(u @ Mw)
    # the stencil's residual is b - A x
    r = O._lv_residual(mf[0], w, u)
    assert np.allclose(r, w - O.spmv(*A, u), rtol=0, atol=1e-12)


def test_dense_coarse_solve():
    rp, ci, v = O.fe_system(1, 2, 6, O.fe_kappa(2, 6, 1.0, 4))
    N = len(rp) - 1
    Cinv = np.linalg.inv(_sp(rp, ci, v).toarray())
    b = O.fe_rhs(1, 2, 6, 9)
    y = O.dense_matvec(Cinv, b)
    assert np.allclose(y, Cinv @ b, rtol=0, atol=1e-14 * np.abs(Cinv @ b).max())
    # an exact coarsest solve on a single level is the exact preconditioner
    lv = [dict(A=(rp, ci, v), invd=O.invdiag(rp, ci, v), Cinv=Cinv)]
    x, info = O.cg_prec((rp, ci, v), b, lambda r: O.mg_cycle(lv, r), rtol=1e-10)
    assert info["iters"] <= 2 and info["status"] == 0
    assert np.allclose(_sp(rp, ci, v) @ x, b, rtol=0, atol=1e-10 * np.abs(b).max())
    assert N == Cinv.shape[0]
