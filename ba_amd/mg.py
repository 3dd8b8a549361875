"""Geometric multigrid hierarchies of structured Q1 meshes for the device
V-cycle (hflmp_mg_*, include/hflmp_hip.h).

The reference's GMG (linear_solver/gmg.h) takes its level operators,
restrictions and prolongations from a mesh hierarchy of the application.
Here the hierarchy of a structured Q1 mesh of n^dim cells (n divisible by
2 on every level) is built directly:

* P (fine x coarse) interpolates a coarse Q1 function at the fine nodes
  (weights 1, 1/2, 1/4, 1/8), with the rows of fine Dirichlet nodes and the
  columns of coarse Dirichlet nodes left empty, so a correction never
  touches a boundary value;
* R = P^T;
* the coarse operator is the Galerkin product P^T A P, assembled cell by
  cell: a coarse cell's matrix is sum_k Q_k^T M_k Q_k over its 2^dim
  children k with matrices M_k, Q_k interpolating the coarse cell's nodes
  at child k's nodes.  On the finest level M_k = kappa_k E, the cell
  matrices the fine operator was assembled from.  The coarse CSR matrix is
  then assembled on the device (hflmp_fe_assemble_f64 with cell matrices)
  and its Dirichlet rows replaced by unit rows, as on the finest level.

Host code only builds these inputs; every product of the cycle runs in
libhflmp_hip.so.
"""
import numpy as np
import scipy.sparse as sp

from . import hflmp


def interp_1d(nc):
    """(2 nc + 1) x (nc + 1) linear interpolation: fine node 2i is coarse
    node i, fine node 2i+1 the mean of coarse nodes i and i+1."""
    nf = 2 * nc + 1
    rows, cols, vals = [], [], []
    for f in range(nf):
        if f % 2 == 0:
            rows.append(f), cols.append(f // 2), vals.append(1.0)
        else:
            rows += [f, f]
            cols += [f // 2, f // 2 + 1]
            vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc + 1))


def _interior(dim, n):
    p = n + 1
    g = np.arange(p)
    inner = (g > 0) & (g < p - 1)
    m = inner
    for _ in range(dim - 1):
        m = np.multiply.outer(inner, m)
    return m.ravel()


def _csr_arrays(M):
    M = sp.csr_matrix(M)
    M.eliminate_zeros()
    M.sort_indices()
    return (M.indptr.astype(np.int32), M.indices.astype(np.int32),
            M.data.astype(np.float64))


def prolongation_matrix(dim, nc, interior=True):
    """scipy CSR of the Q1 interpolation from (nc+1)^dim coarse nodes to
    (2nc+1)^dim fine nodes, nodes lexicographic with x fastest; with
    `interior`, Dirichlet rows and columns are emptied."""
    P1 = interp_1d(nc)
    P = P1
    for _ in range(dim - 1):
        P = sp.kron(P1, P, format="csr")
    if interior:
        P = sp.diags(_interior(dim, 2 * nc).astype(np.float64)) @ P @ sp.diags(
            _interior(dim, nc).astype(np.float64))
    return sp.csr_matrix(P)


def prolongation(dim, nc):
    """(row_ptr, col, val) of P for a coarse mesh of nc^dim cells."""
    return _csr_arrays(prolongation_matrix(dim, nc))


def restriction(dim, nc):
    """(row_ptr, col, val) of R = P^T, rows sorted."""
    return _csr_arrays(prolongation_matrix(dim, nc).T)


def child_interp(dim):
    """Q[k] (nloc x nloc): the coarse cell's Q1 basis at the local nodes of
    child k; child k = ox + 2 oy (+ 4 oz) and local nodes lexicographic."""
    nloc = 2 ** dim
    Q = np.zeros((nloc, nloc, nloc))
    bits = lambda v: [(v >> d) & 1 for d in range(dim)]
    for k in range(nloc):
        o = bits(k)
        for a in range(nloc):
            t = [(o[d] + bits(a)[d]) / 2.0 for d in range(dim)]
            for b in range(nloc):
                w = 1.0
                for d in range(dim):
                    w *= t[d] if bits(b)[d] else 1.0 - t[d]
                Q[k, a, b] = w
    return Q


def _children(arr, dim, nf):
    """(nc^dim, 2^dim, ...) view of per-cell data of an nf^dim mesh,
    grouped by coarse cell (lexicographic) and child k."""
    nc = nf // 2
    tail = arr.shape[1:]
    if dim == 2:
        a = arr.reshape((nc, 2, nc, 2) + tail)          # [cy][oy][cx][ox]
        a = a.transpose((0, 2, 1, 3) + tuple(range(4, 4 + len(tail))))
    else:
        a = arr.reshape((nc, 2, nc, 2, nc, 2) + tail)   # [cz][oz][cy][oy][cx][ox]
        a = a.transpose((0, 2, 4, 1, 3, 5) + tuple(range(6, 6 + len(tail))))
    return a.reshape((nc ** dim, 2 ** dim) + tail)


def coarse_cells(dim, nf, cell_mats=None, E=None, kappa=None):
    """Cell matrices of the (nf/2)^dim mesh from those of the nf^dim mesh:
    either cell_mats (ncell x nloc x nloc) or kappa_c E."""
    if nf % 2:
        raise ValueError("coarse_cells: %d cells per direction is odd" % nf)
    Q = child_interp(dim)
    if cell_mats is None:
        E = np.asarray(E, np.float64)
        Eh = np.einsum("kia,ij,kjb->kab", Q, E, Q)
        kc = _children(np.asarray(kappa, np.float64).reshape(-1), dim, nf)
        return np.einsum("ck,kab->cab", kc, Eh)
    Mc = _children(np.asarray(cell_mats, np.float64), dim, nf)
    return np.einsum("kia,ckij,kjb->cab", Q, Mc, Q)


def level_sizes(n, nlevels=None, min_cells=2):
    """Cells per direction of each level: n, n/2, ... while even and at
    least min_cells (or exactly nlevels levels)."""
    sizes = [n]
    while (nlevels is None or len(sizes) < nlevels) and sizes[-1] % 2 == 0 \
            and sizes[-1] // 2 >= min_cells:
        sizes.append(sizes[-1] // 2)
    if nlevels is not None and len(sizes) != nlevels:
        raise ValueError("level_sizes: %d cells cannot be halved into %d levels"
                         % (n, nlevels))
    return sizes


def coarse_cell_mats(dim, n, E, kappa, sizes):
    """Cell matrices of levels 1.. (host arrays)."""
    out = []
    M = None
    nf = n
    for nc in sizes[1:]:
        M = coarse_cells(dim, nf, E=E, kappa=kappa) if M is None else \
            coarse_cells(dim, nf, cell_mats=M)
        out.append(M)
        nf = nc
    return out


def dense_inverse(A):
    """Dense inverse of a (small) device CsrMatrix, computed on the host."""
    rp, ci, v = A.rp.download(), A.ci.download()[:A.nnz], A.val.download()[:A.nnz]
    D = sp.csr_matrix((v, ci, rp), shape=(A.nrow, A.ncol)).toarray()
    return np.linalg.inv(D)


def build_structured(ctx, dim, n, E, kappa, A, nlevels=None, min_cells=2, stencil=None,
                     coarse="sweeps", structured_transfers=True, **params):
    """Multigrid over the structured Q1 mesh of n^dim cells whose finest
    operator A (a device CsrMatrix) was assembled from kappa_c E with unit
    Dirichlet rows on the boundary, or is given matrix-free as `stencil`
    (a StencilOperator with local matrix E and cell factors kappa; A is
    then None).  coarse="dense" solves the coarsest level with its dense
    inverse instead of sweeps; structured_transfers restricts and
    prolongates without reading R and P (same results).  Returns an
    hflmp.Multigrid that owns the coarse operators, R and P."""
    if coarse not in ("sweeps", "dense"):
        raise ValueError("build_structured: coarse must be 'sweeps' or 'dense'")
    sizes = level_sizes(n, nlevels, min_cells)
    if coarse == "dense" and len(sizes) < 2:
        raise ValueError("build_structured: a dense coarse solve needs two levels or more")
    mats = coarse_cell_mats(dim, n, E, kappa, sizes)
    As, Rs, Ps = [A], [], []
    owned = []
    try:
        for l, nc in enumerate(sizes[1:]):
            rp, ci, v = restriction(dim, nc)
            Rs.append(hflmp.CsrMatrix(ctx, rp, ci, v, ncol=(sizes[l] + 1) ** dim))
            rp, ci, v = prolongation(dim, nc)
            Ps.append(hflmp.CsrMatrix(ctx, rp, ci, v, ncol=(nc + 1) ** dim))
            owned += [Rs[-1], Ps[-1]]
            As.append(hflmp.assemble_q1(ctx, dim, nc, mats[l]))
            owned.append(As[-1])
        M = hflmp.Multigrid(ctx, As, Rs, Ps, owned=owned, stencil=stencil, **params)
        try:
            if coarse == "dense":
                M.set_coarse_inverse(dense_inverse(As[-1]))
            if structured_transfers:
                M.set_structured(dim, n)
        except Exception:
            M.free()
            raise
        return M
    except Exception:
        for m in owned:
            m.free()
        raise
