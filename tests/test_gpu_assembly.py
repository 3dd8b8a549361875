"""Finite-element assembly on the device against the oracle's host
assembly (pattern of the cell couplings, then add_values cell by cell):
pattern and values bit-identical for the structured Q1/Q2 meshes (and,
with the Dirichlet rows applied, equal to the generator's matrix), for
random unstructured meshes with per-cell matrices, FP64 and FP32, and the
refusals."""
import numpy as np
import pytest

from ba_amd import hflmp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _device_csr(ctx, mesh, **kw):
    rp, ci = mesh.pattern()
    val = mesh.assemble(rp, ci, **kw)
    out = rp.download(), ci.download()[:mesh.nnz], val.download()[:mesh.nnz]
    rp.free(); ci.free(); val.free()
    return out


@pytest.mark.parametrize("case", [(1, 2, 1), (1, 2, 23), (1, 3, 9), (2, 2, 11), (1, 3, 40)])
def test_structured_equals_generator(ctx, case):
    order, dim, n = case
    kap = O.fe_kappa(dim, n, 1.0, 5)
    E = O.fe_element(order, dim, n)
    dofs = O.fe_cell_dofs(order, dim, n)
    N = (order * n + 1) ** dim
    mesh = hflmp.FeMesh(ctx, N, dofs)
    try:
        rp, ci, v = _device_csr(ctx, mesh, E=E, factor=kap)
    finally:
        mesh.free()
    ro, co, vo = O.fe_assemble_cells(N, dofs, E=E, factor=kap)
    assert np.array_equal(rp, ro) and np.array_equal(ci, co) and np.array_equal(v, vo)
    fixed = np.nonzero(O.fe_boundary_mask(order, dim, n))[0].astype(np.int32)
    O.zero_rows(fixed, rp, ci, v)
    rq, cq, vq = hflmp.fe_system(order, dim, n, kap)
    assert np.array_equal(rp, rq) and np.array_equal(ci, cq) and np.array_equal(v, vq)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nnode,ncell,nloc", [(50, 80, 4), (400, 300, 10), (1000, 2000, 8),
                                               (7, 1, 3)])
def test_unstructured_cell_matrices(ctx, dtype, nnode, ncell, nloc):
    rng = np.random.default_rng(nnode + ncell)
    # cells of distinct random nodes (tetrahedra-like), random dense matrices
    dofs = np.stack([rng.choice(nnode, nloc, replace=False) for _ in range(ncell)]).astype(np.int32)
    mats = rng.uniform(-1, 1, (ncell, nloc, nloc)).astype(dtype)
    mesh = hflmp.FeMesh(ctx, nnode, dofs)
    try:
        rp, ci, v = _device_csr(ctx, mesh, cell_mats=mats, dtype=dtype)
        ro, co, vo = O.fe_assemble_cells(nnode, dofs, cell_mats=mats, dtype=dtype)
        assert mesh.nnz == len(co)
        assert np.array_equal(rp, ro) and np.array_equal(ci, co) and np.array_equal(v, vo)
        # one shared matrix with per-cell factors
        E = rng.uniform(-1, 1, (nloc, nloc)).astype(dtype)
        f = rng.uniform(0.5, 2, ncell).astype(dtype)
        rp, ci, v = _device_csr(ctx, mesh, E=E, factor=f, dtype=dtype)
        ro, co, vo = O.fe_assemble_cells(nnode, dofs, E=E, factor=f, dtype=dtype)
        assert np.array_equal(v, vo)
    finally:
        mesh.free()


def test_isolated_nodes_and_refusals(ctx):
    # node 3 belongs to no cell: an empty row
    dofs = np.array([[0, 1, 2], [2, 4, 1]], np.int32)
    mesh = hflmp.FeMesh(ctx, 5, dofs)
    try:
        rp, ci, v = _device_csr(ctx, mesh, E=np.ones((3, 3)))
        ro, co, vo = O.fe_assemble_cells(5, dofs, E=np.ones((3, 3)))
        assert rp[4] == rp[3] and np.array_equal(rp, ro) and np.array_equal(v, vo)
    finally:
        mesh.free()
    with pytest.raises(hflmp.HflmpError):
        hflmp.FeMesh(ctx, 4, np.array([[0, 4]], np.int32))       # dof outside the mesh
    with pytest.raises(hflmp.HflmpError):
        hflmp.FeMesh(ctx, 40, np.arange(33, dtype=np.int32).reshape(1, 33))   # nloc > 32
