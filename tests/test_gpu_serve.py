"""The solve server: every FP64 request against the oracle's Jacobi-CG
(iterations within 2, solution within 1e-8); every FP32-inner / FP64-refined
request against its true FP64 residual and the oracle's tight solve; each
result bit-identical alone and in a mix (with and without graph replay, in
both precisions); admission by HBM footprint (requests wait, then run, with
the same results); and its edge cases."""
import numpy as np
import pytest

from ba_amd import hflmp
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MB = 1 << 20


def request(dim, n, seed):
    kap = hflmp.fe_kappa(dim, n, 1.0, seed)
    rp, ci, v = hflmp.fe_system(1, dim, n, kap)
    b = hflmp.fe_rhs(1, dim, n, seed + 1)
    return rp, ci, v, b


MIX = [(2, 8, 11), (2, 40, 12), (3, 6, 13), (2, 17, 14), (3, 11, 15), (2, 64, 16), (2, 3, 17),
       (3, 16, 18)]


def run_all(server, reqs, submit_every=0, rtol=1e-8, precision=64):
    """Submit the requests (all at once, or one per `submit_every` steps
    while stepping) and step until all are done."""
    tickets, pending = [], list(reqs)
    steps = 0
    while pending or any(server.poll(t)["state"] != 2 for t in tickets):
        if pending and (submit_every == 0 or steps % submit_every == 0):
            batch = pending if submit_every == 0 else pending[:1]
            for r in batch:
                tickets.append(server.submit(*r, rtol=rtol, precision=precision))
            pending = [] if submit_every == 0 else pending[1:]
        server.step()
        steps += 1
        assert steps < 100000
    return [server.take(t) for t in tickets]


def test_serve_vs_oracle(ctx):
    reqs = [request(*m) for m in MIX]
    sv = hflmp.Server(ctx, 256 * MB, max_active=8, chunk_iters=8)
    try:
        out = run_all(sv, reqs)
    finally:
        sv.close()
    for (rp, ci, v, b), (x, info) in zip(reqs, out):
        xo, io = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-8, maxit=100000)
        assert info["status"] == 0
        assert abs(info["iters"] - io["iters"]) <= 2, (info["iters"], io["iters"])
        assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
        assert info["latency_seconds"] >= info["queued_seconds"] >= 0


@pytest.mark.parametrize("graphs", [True, False])
def test_serve_alone_equals_mix(ctx, graphs):
    reqs = [request(*m) for m in MIX]
    probe = request(3, 12, 99)
    sv = hflmp.Server(ctx, 256 * MB, max_active=1, chunk_iters=5, use_graphs=graphs)
    try:
        (x_alone, i_alone), = run_all(sv, [probe])
    finally:
        sv.close()
    # the probe arrives in the middle of a staggered mix, with another chunk
    # length and more slots
    mixed = reqs[:4] + [probe] + reqs[4:]
    sv = hflmp.Server(ctx, 256 * MB, max_active=6, chunk_iters=7, use_graphs=not graphs)
    try:
        out = run_all(sv, mixed, submit_every=3)
        st = sv.stats()
    finally:
        sv.close()
    x_mix, i_mix = out[4]
    assert i_mix["iters"] == i_alone["iters"]
    assert np.array_equal(x_mix, x_alone)
    assert st["active"] == 0 and st["queued"] == 0


def csr_share(rp, ci):
    """Arena bytes of a system's CSR arrays (row_ptr, col, FP64 values, each
    rounded to 256 B), returned to the arena once it is admitted."""
    n, nnz = len(rp) - 1, len(ci)
    return ((4 * (n + 1) + 255) // 256 + (4 * nnz + 255) // 256 + (8 * nnz + 255) // 256) * 256


def test_serve_admission_by_footprint(ctx):
    reqs = [request(2, 100, 40 + k) for k in range(6)]
    foot = hflmp.server_footprint(reqs[0][0], reqs[0][1])
    held = foot - csr_share(reqs[0][0], reqs[0][1])
    # room for two systems at a time: one admitted (its CSR share returned)
    # and a second one's whole footprint, but not a third
    budget = max(held + foot + held // 2, MB)
    sv = hflmp.Server(ctx, budget, max_active=16, chunk_iters=16)
    try:
        tickets = [sv.submit(*r) for r in reqs]
        sv.step()
        st = sv.stats()
        assert st["active"] == 2 and st["queued"] == 4
        while any(sv.poll(t)["state"] != 2 for t in tickets):
            sv.step()
            assert sv.stats()["active"] <= 2
        out = [sv.take(t) for t in tickets]
    finally:
        sv.close()
    ref = hflmp.Server(ctx, 64 * budget, max_active=16, chunk_iters=16)
    try:
        out2 = run_all(ref, reqs)
    finally:
        ref.close()
    for (x, i), (x2, i2) in zip(out, out2):
        assert np.array_equal(x, x2) and i["iters"] == i2["iters"]
        assert i["footprint"] == foot
    # a system larger than the whole budget is refused at submission
    big = request(2, 300, 7)
    small = hflmp.Server(ctx, 4 * MB, max_active=4, chunk_iters=4)
    try:
        with pytest.raises(hflmp.HflmpError):
            small.submit(*big)
    finally:
        small.close()


def test_serve_edge_cases(ctx):
    rp, ci, v, b = request(2, 10, 3)
    zero = np.zeros_like(b)
    sv = hflmp.Server(ctx, 64 * MB, max_active=4, chunk_iters=4)
    try:
        (x0, i0), (x1, i1) = run_all(sv, [(rp, ci, v, zero), (rp, ci, v, b)])
        assert i0["iters"] == 0 and i0["status"] == 0 and not x0.any()
        assert i1["status"] == 0 and i1["iters"] > 0
        # an iteration cap: status "max iterations"
        t = sv.submit(rp, ci, v, b, rtol=1e-14, maxit=3)
        while sv.poll(t)["state"] != 2:
            sv.step()
        x, info = sv.take(t)
        assert info["status"] == 1 and info["iters"] == 3
        with pytest.raises(hflmp.HflmpError):
            sv.poll(t)  # released
        # a row longer than 255 entries is refused
        n = 300
        rp2 = np.array([0, n] + list(range(n + 1, 2 * n)), np.int32)
        ci2 = np.concatenate([np.arange(n), np.arange(1, n)]).astype(np.int32)
        with pytest.raises(hflmp.HflmpError):
            sv.submit(rp2, ci2, np.ones(len(ci2)), np.ones(n))
    finally:
        sv.close()


def _wide_band_system(m=200, far=35000, seed=5):
    """SPD 2D Laplacian on m^2 nodes plus symmetric couplings between nodes
    `far` apart, so some columns are not within int16 of their slice."""
    import scipy.sparse as sp
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    L = sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m))
    N = m * m
    idx = np.arange(0, N - far, 97)
    Cf = sp.csr_matrix((np.full(len(idx), -0.1), (idx, idx + far)), shape=(N, N))
    extra = np.asarray(abs(Cf).sum(axis=1)).ravel() + np.asarray(abs(Cf).sum(axis=0)).ravel()
    A = sp.csr_matrix(L + Cf + Cf.T + sp.diags(extra))
    A.sort_indices()
    b = np.random.default_rng(seed).uniform(-1, 1, N)
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64),
            b)


def test_serve_fp64_column_widths(ctx):
    """The FP64 operator takes int16 columns where every column fits and
    int32 otherwise; either way the result matches the single device solve
    (same iterations; x to rounding, as the two reduce their dot products
    in different orders)."""
    narrow = request(2, 40, 12)
    wide = _wide_band_system()
    rp, ci = wide[0], wide[1]
    base = (np.repeat(np.arange(len(rp) - 1), np.diff(rp)) // 64) * 64
    assert np.abs(ci.astype(np.int64) - base).max() > 32767   # the band does not fit
    sv = hflmp.Server(ctx, 256 * MB, max_active=4, chunk_iters=6)
    try:
        out = run_all(sv, [narrow, wide, narrow])
    finally:
        sv.close()
    assert [i["col_bytes"] for _, i in out] == [2, 4, 2]
    for (rp, ci, v, b), (x, info) in zip([narrow, wide], out[:2]):
        A = hflmp.CsrMatrix(ctx, rp, ci, v)
        S = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
        invd = hflmp.invdiag(ctx, A)
        db, dx = ctx.upload(b), ctx.zeros(A.nrow, np.float64)
        si = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, rtol=1e-8, maxit=100000)
        assert S.info()["col_bytes"] == info["col_bytes"]
        assert info["status"] == 0 and si["status"] == 0
        assert abs(info["iters"] - si["iters"]) <= 1
        xs = dx.download()
        assert np.linalg.norm(x - xs) <= 1e-12 * np.linalg.norm(xs)
        S.free(); A.free(); invd.free(); db.free(); dx.free()
    xo, io = O.cg(*wide, prec=O.PREC_JACOBI, rtol=1e-8, maxit=100000)
    assert abs(out[1][1]["iters"] - io["iters"]) <= 2
    assert np.linalg.norm(out[1][0] - xo) <= 1e-8 * np.linalg.norm(xo)


def test_serve_mixed_column_widths(ctx):
    """A mixed system's FP32 operator takes int32 columns where its band
    does not fit int16 (decided on the device, as for the FP64 operator).
    In a mix with int16 systems, submitted one per step so that the
    admission thread's work and the joins interleave with running chunks,
    its result is bit-identical to the same system solved alone, and every
    system meets the tolerance on the true residual.  The chunks' device
    time is reported, within the run's wall time."""
    import time
    narrow = request(2, 40, 12)
    wide = _wide_band_system()
    sv = hflmp.Server(ctx, 256 * MB, max_active=1, chunk_iters=6)
    try:
        (x_alone, i_alone), = run_all(sv, [wide], precision=32)
    finally:
        sv.close()
    assert i_alone["col_bytes"] == 4
    reqs = [narrow, wide, narrow, request(2, 33, 4)]
    sv = hflmp.Server(ctx, 256 * MB, max_active=4, chunk_iters=6)
    t0 = time.perf_counter()
    try:
        out = run_all(sv, reqs, submit_every=1, precision=32)
        st = sv.stats()
    finally:
        sv.close()
    wall = time.perf_counter() - t0
    assert [i["col_bytes"] for _, i in out] == [2, 4, 2, 2]
    x_mix, i_mix = out[1]
    assert i_mix["iters"] == i_alone["iters"] and np.array_equal(x_mix, x_alone)
    for (rp, ci, v, b), (x, info) in zip(reqs, out):
        assert info["status"] == 0 and info["precision"] == 32
        assert np.linalg.norm(b - O.spmv(rp, ci, v, x)) < 1e-8 * np.linalg.norm(b)
    assert 0 < st["chunk_device_seconds"] <= wall


def test_serve_mixed_vs_oracle(ctx):
    reqs = [request(*m) for m in MIX]
    sv = hflmp.Server(ctx, 256 * MB, max_active=8, chunk_iters=8)
    try:
        out = run_all(sv, reqs, precision=32)
    finally:
        sv.close()
    for (rp, ci, v, b), (x, info) in zip(reqs, out):
        assert info["status"] == 0 and info["precision"] == 32
        bn = np.linalg.norm(b)
        # the stopping test is on the FP64 true residual
        assert np.linalg.norm(b - O.spmv(rp, ci, v, x)) < 1e-8 * bn
        # as accurate as FP64 CG at the same tolerance, against a tight solve
        # (plus 1e-8 relative: on the smallest systems FP64 CG is exact)
        xt, _ = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-13, maxit=200000)
        x8, _ = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-8, maxit=200000)
        tn = np.linalg.norm(xt)
        assert np.linalg.norm(x - xt) <= 10 * np.linalg.norm(x8 - xt) + 1e-8 * tn
        assert info["corrections"] >= 1


@pytest.mark.parametrize("graphs", [True, False])
def test_serve_mixed_alone_equals_mix(ctx, graphs):
    reqs = [request(*m) for m in MIX]
    probe = request(2, 50, 77)
    sv = hflmp.Server(ctx, 256 * MB, max_active=1, chunk_iters=6, use_graphs=graphs)
    try:
        (x_alone, i_alone), = run_all(sv, [probe], precision=32)
    finally:
        sv.close()
    sv = hflmp.Server(ctx, 256 * MB, max_active=5, chunk_iters=9, use_graphs=not graphs,
                      mixed_min_nnz=5000)
    try:
        # precision 0: the mix is FP64 below 5000 nonzeros and mixed above
        mixed = reqs[:3] + [probe] + reqs[3:]
        tickets, out = [], []
        for k, r in enumerate(mixed):
            tickets.append(sv.submit(*r, precision=32 if k == 3 else 0))
            sv.step()
        while any(sv.poll(t)["state"] != 2 for t in tickets):
            sv.step()
        out = [sv.take(t) for t in tickets]
    finally:
        sv.close()
    x_mix, i_mix = out[3]
    assert i_mix["iters"] == i_alone["iters"]
    assert i_mix["corrections"] == i_alone["corrections"]
    assert np.array_equal(x_mix, x_alone)
    for (rp, ci, v, b), (x, info) in zip(mixed, out):
        assert info["status"] == 0
        assert info["precision"] == (32 if len(ci) >= 5000 else 64)


def test_serve_precision_window(ctx):
    # precision 0 picks the mixed mode for 5000 <= nnz < 20000 only
    reqs = [request(*m) for m in MIX]
    sv = hflmp.Server(ctx, 256 * MB, max_active=8, chunk_iters=8, mixed_min_nnz=5000,
                      mixed_max_nnz=20000)
    try:
        tickets = [sv.submit(*r, precision=0) for r in reqs]
        while any(sv.poll(t)["state"] != 2 for t in tickets):
            sv.step()
        out = [sv.take(t) for t in tickets]
    finally:
        sv.close()
    modes = [info["precision"] for _, info in out]
    assert modes == [32 if 5000 <= len(r[1]) < 20000 else 64 for r in reqs]
    assert 32 in modes and 64 in modes
    for (rp, ci, v, b), (x, info) in zip(reqs, out):
        assert info["status"] == 0
        # FP64 stops on the recursive residual, mixed on the true one
        assert np.linalg.norm(b - O.spmv(rp, ci, v, x)) < 1.1e-8 * np.linalg.norm(b)


def test_serve_tune_small_sample(ctx):
    """The autotuner on a small mixed sample returns a usable configuration
    (window bounds on class boundaries, a candidate chunk size) and a
    report of every timing it based its choices on."""
    from ba_amd import serve_tune as T
    sample = []
    for dim, n, seed in ((2, 24, 1), (2, 40, 2), (3, 10, 3), (3, 14, 4)):
        kap = hflmp.fe_kappa(dim, n, 1.0, seed)
        rp, ci, v = hflmp.fe_system(1, dim, n, kap)
        sample.append((rp, ci, v, hflmp.fe_rhs(1, dim, n, seed + 10)))
    base = dict(hbm_budget=1 << 30, max_active=16, chunk_iters=16, use_graphs=True,
                mixed_min_nnz=0, inner_rtol=1e-2, mixed_max_nnz=0)
    cfg, rep = T.tune(ctx, sample, base, rtol=1e-8, chunks=(8, 16))
    assert cfg["chunk_iters"] in (8, 16)
    assert (cfg["mixed_min_nnz"], cfg["mixed_max_nnz"]) in T.WINDOWS
    assert set(rep["chunks"]) == {8, 16}
    assert len(rep["windows"]) == len(T.WINDOWS)
    for w in rep["windows"].values():
        assert w["seconds"] > 0 and w["converged"]
    assert cfg["inner_rtol"] in (1e-2, 1e-1) and rep["inner_rtol_chosen"] == cfg["inner_rtol"]
    if cfg["mixed_min_nnz"] > 0:
        assert set(rep["inner_rtol"]) == {"0.01", "0.1"}


@pytest.mark.parametrize("precision,solo", [(64, 0), (32, 0), (64, 1)])
@pytest.mark.parametrize("bad_col", [-1, "nrow"])
def test_serve_refuses_out_of_range_column(ctx, precision, solo, bad_col):
    """A request whose column index lies outside 0..nrow-1 is refused on the
    device at admission (status 4, x = 0) before any product gathers with
    it; a healthy request in the same server is solved as usual (grouped,
    or alone with solo_min_nnz = 1)."""
    rp, ci, v, b = request(2, 9, 31)
    ci_bad = ci.copy()
    k = int(rp[40]) + 1  # an entry of an interior row
    ci_bad[k] = -1 if bad_col == -1 else len(rp) - 1
    good = request(2, 12, 32)
    sv = hflmp.Server(ctx, 64 * MB, max_active=4, chunk_iters=8, solo_min_nnz=solo)
    try:
        out = run_all(sv, [(rp, ci_bad, v, b), good], precision=precision)
    finally:
        sv.close()
    (x_bad, i_bad), (x_good, i_good) = out
    assert i_bad["status"] == 4
    assert not x_bad.any()
    assert i_good["status"] == 0
    grp, gci, gv, gb = good
    # FP64 stops on the recursive residual, mixed on the true one
    assert np.linalg.norm(gb - O.spmv(grp, gci, gv, x_good)) < 1.1e-8 * np.linalg.norm(gb)


def test_serve_csr_share_returns_to_the_arena(ctx):
    # Two copies of one system under a budget that holds one whole block
    # plus all of a second but its CSR share: the first system's CSR arrays
    # go back to the arena once its SELL-64 layout is built, so the second
    # is admitted in the same step and both run at once, with the results
    # they have alone.
    rp, ci, v, b = request(2, 96, 31)
    foot = hflmp.server_footprint(rp, ci)
    csr = csr_share(rp, ci)
    budget = 2 * foot - csr // 2
    assert budget < 2 * foot and budget >= 2 * foot - csr
    sv = hflmp.Server(ctx, budget, max_active=4, chunk_iters=8)
    try:
        out = run_all(sv, [(rp, ci, v, b), (rp, ci, v, b)])
        st = sv.stats()
    finally:
        sv.close()
    assert st["peak_active"] == 2
    assert st["peak_arena_in_use"] <= 2 * foot - csr
    (x1, i1), (x2, i2) = out
    assert i1["status"] == 0 and i2["status"] == 0
    assert np.array_equal(x1, x2) and i1["iters"] == i2["iters"]
    sv1 = hflmp.Server(ctx, budget, max_active=4, chunk_iters=8)
    try:
        (xa, ia), = run_all(sv1, [(rp, ci, v, b)])
    finally:
        sv1.close()
    assert np.array_equal(xa, x1) and ia["iters"] == i1["iters"]


def _single_sell_solve(ctx, rp, ci, v, b, rtol=1e-8, maxit=100000):
    """The single-system device solve over SELL-64 (int16 columns where the
    band fits), with the server's divergence bound (none)."""
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    S = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
    invd = hflmp.invdiag(ctx, A)
    db, dx = ctx.upload(b), ctx.zeros(A.nrow, np.float64)
    try:
        info = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, rtol=rtol, maxit=maxit, dtol=1e300)
        return dx.download(), info, S.info()["col_bytes"]
    finally:
        S.free(); A.free(); invd.free(); db.free(); dx.free()


@pytest.mark.parametrize("graphs", [True, False])
def test_serve_solo_equals_single_solver(ctx, graphs):
    """FP64 systems with at least solo_min_nnz nonzeros run alone, on the
    server's second stream, with the single solver's kernels and grids:
    iterations, residuals and x are those of hflmp_pcg_jacobi_sell_f64 bit
    for bit (int16 and int32 columns), while the smaller systems of the same
    staggered mix run grouped and still match the oracle."""
    big = [request(3, 24, 61), request(2, 150, 62), _wide_band_system()]
    small = [request(*m) for m in MIX[:4]]
    thr = min(len(r[1]) for r in big)
    assert max(len(r[1]) for r in small) < thr
    reqs = small[:2] + big[:2] + small[2:] + big[2:]
    sv = hflmp.Server(ctx, 512 * MB, max_active=8, chunk_iters=8, use_graphs=graphs,
                      solo_min_nnz=thr)
    try:
        out = run_all(sv, reqs, submit_every=2)
        st = sv.stats()
    finally:
        sv.close()
    assert st["solo_runs"] == 3
    assert st["active"] == 0 and st["queued"] == 0
    for (rp, ci, v, b), (x, info) in zip(reqs, out):
        assert info["status"] == 0
        if len(ci) >= thr:
            xs, si, cb = _single_sell_solve(ctx, rp, ci, v, b)
            assert info["col_bytes"] == cb
            assert info["iters"] == si["iters"]
            assert info["res_final"] == si["res_final"] and info["res_init"] == si["res_init"]
            assert np.array_equal(x, xs)
        else:
            xo, io = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-8, maxit=100000)
            assert abs(info["iters"] - io["iters"]) <= 2, (info["iters"], io["iters"])
            assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    assert [i["col_bytes"] for _, i in out if i["col_bytes"] == 4] == [4]


def test_serve_solo_alone_equals_mix(ctx):
    """A solo system's result does not depend on what else the server runs:
    alone, beside grouped systems, or beside another solo system, with
    another chunk length."""
    probe = request(2, 150, 62)
    sv = hflmp.Server(ctx, 256 * MB, max_active=1, chunk_iters=5, solo_min_nnz=100000)
    try:
        (x_alone, i_alone), = run_all(sv, [probe])
    finally:
        sv.close()
    reqs = [request(*m) for m in MIX[:3]] + [probe, request(3, 24, 61)] + \
        [request(*m) for m in MIX[3:6]]
    sv = hflmp.Server(ctx, 512 * MB, max_active=6, chunk_iters=7, solo_min_nnz=100000)
    try:
        out = run_all(sv, reqs, submit_every=3)
        st = sv.stats()
    finally:
        sv.close()
    assert st["solo_runs"] == 2
    x_mix, i_mix = out[3]
    assert i_mix["iters"] == i_alone["iters"]
    assert np.array_equal(x_mix, x_alone)
