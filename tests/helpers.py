"""Shared test inputs and tolerance checks."""
import numpy as np


def random_csr(nrow, ncol, lengths, seed, dtype=np.float64):
    """CSR with the given row lengths, distinct sorted columns, values in
    [-1, 1].  Rows may be empty; lengths must be <= ncol."""
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, np.int64)
    rp = np.zeros(nrow + 1, np.int32)
    rp[1:] = np.cumsum(lengths)
    ci = np.empty(int(rp[-1]), np.int32)
    for i in range(nrow):
        L = int(lengths[i])
        if L:
            cols = np.sort(rng.choice(ncol, size=L, replace=False))
            ci[rp[i]:rp[i + 1]] = cols
    v = rng.uniform(-1, 1, size=len(ci)).astype(dtype)
    return rp, ci, v


def spmv_abs(rp, ci, v, x):
    """(|A| |x|)_i, the scale of the per-entry SpMV tolerance."""
    import scipy.sparse as sp
    n = len(rp) - 1
    A = sp.csr_matrix((np.abs(v).astype(np.float64), ci, rp), shape=(n, len(x)))
    return A @ np.abs(x.astype(np.float64))


def assert_spmv_close(y, yref, scale, rel=1e-13):
    """|y - yref|_i <= rel * (|A||x|)_i for every row."""
    err = np.abs(y.astype(np.float64) - yref.astype(np.float64))
    bad = err > rel * scale + 1e-300
    assert not bad.any(), "rows %s: err %s scale %s" % (
        np.nonzero(bad)[0][:5], err[bad][:5], scale[bad][:5])


def sine_mode(order, dim, n, modes):
    """Product of sines on the lexicographic node grid; vanishes on the
    boundary."""
    p = order * n + 1
    t = np.arange(p) / (p - 1)
    f = [np.sin(np.pi * modes[d] * t) for d in range(dim)]
    if dim == 2:
        return np.einsum("y,x->yx", f[1], f[0]).ravel()
    return np.einsum("z,y,x->zyx", f[2], f[1], f[0]).ravel()


def q1_sine_eigenvalue(dim, n, modes):
    """Eigenvalue of the assembled Q1 stiffness matrix (interior rows) for a
    sine mode: A = sum_d K_d (x) M_others with 1D stencils
    K = (1/h)[-1, 2, -1], M = (h/6)[1, 4, 1]."""
    h = 1.0 / n
    th = [np.pi * m / n for m in modes]
    k = [(2 - 2 * np.cos(a)) / h for a in th]
    m = [h * (4 + 2 * np.cos(a)) / 6 for a in th]
    lam = 0.0
    for d in range(dim):
        term = k[d]
        for e in range(dim):
            if e != d:
                term *= m[e]
        lam += term
    return lam
