#!/usr/bin/env python3
"""Benchmark of the gfx950 lmp platform.

Headline (the timed steps): FP64 SpMV y = A x of the 3D Q1 Poisson matrix
(128^3 hex cells, 129^3 = 2,146,689 DoFs, 57,066,625 nonzeros) with A, x, y
resident in HBM.  A is assembled in CSR and executed through its SELL-64
device layout with int32 columns (built once on the device from the CSR
arrays; products bit-identical to the sequential CSR loop).  `value` is
GB/s of CSR algorithmic bytes: 12 nnz + 20 N + 4 per SpMV (int32 col + FP64
val per nonzero, int32 row_ptr, x read once, y written once).

Reported beside it, each measured in the same run:
  * the SpMV of the CSR-array row-tile kernel and of the compressed SELL-64
    layout with int16 column offsets (the layout HIP_CSR_lMatrix builds
    when every column fits; scored with the same CSR bytes, reported
    separately as a compressed format);
  * parity of both layouts against the oracle (bit-exact);
  * Jacobi-CG to 1e-8 on the same system (CSR and SELL operators);
  * mixed precision: log-normal diffusivity 128^3, FP32-inner / FP64-refined
    CG vs plain FP64 CG, both to a relative residual of 1e-10;
  * batched: 2D Q2 512^2 with 64 right-hand sides to 1e-8 vs one-at-a-time
    CG, FP64 and FP32;
  * serving: a stream of 2D/3D Q1 solves of uneven sizes (FP64 Jacobi-CG to
    1e-8) through the solve server vs one at a time; solves/s and p50/p99
    latency under Poisson arrivals;
  * the oracle's CPU SpMV on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

KAPPA_SEED, RHS_SEED = 20241203, 20241204  # mixed-precision system
BATCH_SEED0 = 1000  # right-hand sides of the batched system: seeds 1000..1063


_T_START = time.perf_counter()


def progress(msg):
    """One line on stderr when a leg or part of one ends (the JSON result
    alone goes to stdout)."""
    print("[bench %7.1f s] %s" % (time.perf_counter() - _T_START, msg), file=sys.stderr,
          flush=True)


def spmv_bytes(N, nnz, vb=8):
    return (4 + vb) * nnz + 4 * (N + 1) + 2 * vb * N


def cg_iter_bytes(N, nnz, vb=8):
    # SpMV + 13 N-vector streams of the fused iteration
    return (4 + vb) * nnz + 4 * (N + 1) + 13 * vb * N


def spmm_bytes(N, nnz, k, vb=8):
    return (4 + vb) * nnz + 4 * (N + 1) + 2 * k * vb * N


def host_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    if not name:
        import platform
        name = platform.processor() or "unknown"
    return {"model": name, "logical_cpus": os.cpu_count(), "threads_used": host_threads()}


def cpu_cg(rp, ci, v, b, rtol, gpu_x=None, gpu_iters=None, threads=None):
    """Jacobi-CG on the host along the reference's CPU path: CG::SolvePrecond
    (linear_solver/cg.h:165-246) over the CPUopenmp platform, i.e. the SpMV
    row loop on every host thread (CPUopenmp_CSR_lMatrix::VectorMult,
    openmp/CPUopenmp_csr_lmatrix.cc:145-165) and the vector operations on one
    thread (CPUopenmp_lvector.cc's omp simd loops).  Timed once, from x = 0;
    with the GPU's solution, the iteration difference and the relative
    error of the GPU x against this solve."""
    from oracle import oracle as O
    th = threads or host_threads()
    t0 = time.perf_counter()
    x, info = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=rtol, maxit=100000, threads=th)
    el = time.perf_counter() - t0
    out = {"seconds_to_tol": round(el, 4), "iters": info["iters"],
           "converged": info["status"] == 0, "spmv_threads": th, "vector_threads": 1,
           "ms_per_iter": round(el * 1e3 / max(info["iters"], 1), 3)}
    if gpu_x is not None:
        out["gpu_iters"] = gpu_iters
        out["iter_diff_gpu_vs_cpu"] = abs(gpu_iters - info["iters"])
        out["rel_error_gpu_x_vs_cpu_x"] = float(np.linalg.norm(gpu_x - x) /
                                               max(np.linalg.norm(x), 1e-300))
    return out, x


def cpu_solve_for_parity(rp, ci, v, b, rtol):
    """A CPU FP64 Jacobi-CG solve used only as the parity reference of a
    GPU result (not timed): SpMV and vector operations on every host thread
    (the dot products sum per-thread blocks)."""
    from oracle import oracle as O
    th = host_threads()
    x, info = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=rtol, maxit=100000, threads=th,
                   blas_threads=th)
    return x, info


def cpu_baseline(rp, ci, v, x, budget_s):
    """Oracle SpMV (reference row-order semantics) on the host cores:
    OpenMP over rows with every available thread, repeated for about
    `budget_s` seconds, plus a few single-thread NAIVE passes."""
    from oracle import oracle as O
    N, nnz = len(rp) - 1, len(ci)
    nb = spmv_bytes(N, nnz)
    y = np.empty(N)
    th = host_threads()
    O.spmv_omp(rp, ci, v, x, y, th)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        O.spmv_omp(rp, ci, v, x, y, th)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    omp_gbs = nb * reps / el / 1e9
    nreps, t0 = 0, time.perf_counter()
    while True:
        O.spmv(rp, ci, v, x)
        nreps += 1
        el1 = time.perf_counter() - t0
        if el1 >= budget_s / 3:
            break
    naive_gbs = nb * nreps / el1 / 1e9
    return ({"value": round(omp_gbs, 3), "unit": "GB/s", "cores": th, "kind": "port",
             "sample": "%d OpenMP SpMVs (%.1f s) of the same 128^3 Q1 FP64 matrix, "
                       "CPUopenmp_CSR_lMatrix::VectorMult semantics" % (reps, el)},
            {"value": round(naive_gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
             "sample": "%d single-thread SpMVs (%.1f s), CPUsimple_CSR_lMatrix::VectorMult "
                       "semantics" % (nreps, el1)})


def measured_traffic(n, kernel):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (scaled by a 1 GiB copy), recorded in profiles/; None if not measured
    for this size and kernel."""
    path = os.path.join(ROOT, "profiles", "spmv_128cube_traffic.json")
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(kernel)
    return round(rec["traffic_bytes"]) if rec else None


def copy_calibration(hflmp, reps):
    """y = x over 2^27 doubles: 1 GiB read + 1 GiB written per launch."""
    n = 1 << 27
    ctx = hflmp.Context(0)
    x, y = ctx.zeros(n, np.float64), ctx.empty(n, np.float64)
    for _ in range(2):
        ctx.check(ctx.L.hflmp_copy_f64(ctx.h, n, x.ptr, y.ptr), "copy")
    e0, e1 = ctx.event(), ctx.event()
    e0.record()
    for _ in range(reps):
        ctx.check(ctx.L.hflmp_copy_f64(ctx.h, n, x.ptr, y.ptr), "copy")
    e1.record()
    ms = e1.elapsed_ms(e0) / reps
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n,
                      "ms_per_copy": ms, "GBps_rw": 16 * n / ms / 1e6}), flush=True)
    x.free(); y.free(); ctx.close()


def time_launches(ctx, fn, reps, warmup=5):
    """Average duration of `fn` over `reps` back-to-back launches, by HIP
    events on the context's stream (the stream the kernels run on)."""
    for _ in range(warmup):
        fn()
    e0, e1 = ctx.event(), ctx.event()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    ctx.sync()
    return e1.elapsed_ms(e0) / 1e3 / reps


def cg_record(info, N, nnz, vb=8):
    it = max(info["iters"], 1)
    from ba_amd import hflmp
    return {"iters": info["iters"], "status": hflmp.STATUS.get(info["status"], "?"),
            "seconds_to_tol": round(info["seconds"], 5),
            "ms_per_iter": round(info["seconds"] * 1e3 / it, 5),
            "GBps_per_iter": round(cg_iter_bytes(N, nnz, vb) * it / info["seconds"] / 1e9, 2),
            "res_final_rel": info["res_final"] / info["res_init"]}


def bench_cg(hflmp, ctx, A, S, Sw, b):
    """Jacobi-CG to 1e-8 through SELL-64 with int16 columns (S, the layout
    HIP_CSR_lMatrix builds), SELL-64 with int32 columns (Sw) and the CSR
    tiles (second solve of each timed)."""
    N, nnz = A.nrow, A.nnz
    invd = hflmp.invdiag(ctx, A)
    db = ctx.upload(b)
    res = {"precond": "jacobi", "rtol": 1e-8}
    xs = None
    for name, run in (("sell", lambda dx: hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx,
                                                                maxit=20000, rtol=1e-8)),
                      ("sell_int32", lambda dx: hflmp.pcg_jacobi_sell(ctx, Sw, invd, db, dx,
                                                                      maxit=20000, rtol=1e-8)),
                      ("csr", lambda dx: hflmp.pcg_jacobi(ctx, A, invd, db, dx, maxit=20000,
                                                          rtol=1e-8))):
        for rep in range(2):
            dx = ctx.zeros(N, np.float64)
            info = run(dx)
            if name == "sell" and rep == 1:
                xs = dx.download()
            dx.free()
        res[name] = cg_record(info, N, nnz)
    res["sell"]["layout"] = "SELL-64, int16 columns" if S.info()["col_bytes"] == 2 \
        else "SELL-64, int32 columns"
    db.free(); invd.free()
    return res, xs


def bench_c1(hflmp, n=256, seed=20241201):
    """C1: 2D Q1 Poisson 256^2 (66k DoFs, 9-point pattern), seeded uniform
    RHS, Jacobi-CG to 1e-8: the reference's CPU path (threaded SpMV,
    one-thread vectors) and the sequential CPUsimple path timed on the host,
    the device-resident solve on the GPU (second of two solves timed), with
    the iteration difference and the x error of the GPU against the CPU."""
    from oracle import oracle as O
    rp, ci, v = hflmp.fe_system(1, 2, n)
    b = hflmp.fe_rhs(1, 2, n, seed)
    N = len(rp) - 1
    ctx = hflmp.Context(0)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    S = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
    invd = hflmp.invdiag(ctx, A)
    db = ctx.upload(b)
    for rep in range(2):
        dx = ctx.zeros(N, np.float64)
        gi = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, maxit=20000, rtol=1e-8)
        xg = dx.download()
        dx.free()
    S.free(); invd.free(); db.free(); A.free()
    ctx.close()
    cpu, _ = cpu_cg(rp, ci, v, b, 1e-8, xg, gi["iters"])
    t0 = time.perf_counter()
    _, i1 = O.cg(rp, ci, v, b, prec=O.PREC_JACOBI, rtol=1e-8, maxit=20000)
    t1 = time.perf_counter() - t0
    return {"workload": "2D Q1 Poisson %d^2 cells (%d DoFs, %d nonzeros), FP64, Jacobi-CG "
                        "to 1e-8, RHS seed %d" % (n, N, len(ci), seed),
            "gpu": {"iters": gi["iters"], "seconds_to_tol": round(gi["seconds"], 5),
                    "operator": "SELL-64, int16 columns"},
            "cpu_reference_path": cpu,
            "cpu_single_thread": {"seconds_to_tol": round(t1, 4), "iters": i1["iters"],
                                  "what": "CPUsimple platform: every operation on one thread"},
            "speedup_gpu_vs_cpu_reference_path": round(cpu["seconds_to_tol"] / gi["seconds"], 2)}


def stencil_bytes(N, ncell, vb=8):
    """Matrix-free product: x read, y written, one factor per cell, a bit
    per node for the fixed dofs."""
    return 2 * vb * N + vb * ncell + 4 * ((N + 31) // 32)


def bench_stencil(hflmp, ctx, n, x, b, steps, sell_cg=None, assembled=None):
    """Matrix-free StencilOperator on the same 3D Q1 128^3 system (constant
    coefficient): product time and roofline, bit-exactness against the
    oracle's cell loop, agreement with the assembled product, and Jacobi-CG
    to 1e-8 through it."""
    from oracle import oracle as O
    E = hflmp.fe_element(1, 3, n)
    fixed = hflmp.fe_fixed(1, 3, n)
    ones = np.ones(n ** 3)
    St = hflmp.StencilOperator(ctx, 1, 3, [n] * 3, E, ones, fixed)
    N = St.nrow
    dx = ctx.upload(x)
    dy = ctx.empty(N, np.float64)
    t = time_launches(ctx, lambda: St.apply(dx, dy), steps)
    nb = stencil_bytes(N, n ** 3)
    out = {"workload": "3D Q1 Poisson %d^3, FP64, matrix-free (element matrix x cell factor)" % n,
           "apply_us": round(t * 1e6, 2), "bytes_per_apply": nb,
           "roofline": {"bound": "hbm", "achieved": round(nb / t / 1e9, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(nb / t / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(n, "stencil_f64")},
           "device_bytes": St.info()["device_bytes"]}
    St.apply(dx, dy)
    y = dy.download()
    yref = O.stencil_vmadd(1, 3, [n] * 3, E, ones, fixed, x)
    out["parity"] = {"bit_exact_vs_oracle_cell_loop": bool(np.array_equal(y, yref))}
    if assembled is not None:
        S, rp, ci, v = assembled
        S.spmv(dx, dy)
        ya = dy.download()
        import scipy.sparse as sp
        scale = sp.csr_matrix((np.abs(v), ci, rp), shape=(N, N)) @ np.abs(x)
        out["parity"]["max_diff_vs_assembled_over_absA_absx"] = float(
            np.max(np.abs(y - ya) / np.maximum(scale, 1e-300)))
    invd = St.invdiag()
    db = ctx.upload(b)
    for rep in range(2):
        dxs = ctx.zeros(N, np.float64)
        info = hflmp.pcg_jacobi_stencil(ctx, St, invd, db, dxs, maxit=20000, rtol=1e-8)
        dxs.free()
    it = max(info["iters"], 1)
    out["cg"] = {"rtol": 1e-8, "iters": info["iters"],
                 "status": hflmp.STATUS.get(info["status"], "?"),
                 "seconds_to_tol": round(info["seconds"], 5),
                 "ms_per_iter": round(info["seconds"] * 1e3 / it, 5)}
    if sell_cg:
        out["cg"]["speedup_vs_sell_cg"] = round(sell_cg["seconds_to_tol"] / info["seconds"], 3)
        out["cg"]["iters_sell_cg"] = sell_cg["iters"]
    St.free(); dx.free(); dy.free(); invd.free(); db.free()
    return out


def bench_mg(hflmp, ctx, n, steps, sell_cg=None):
    """Geometric multigrid (V-cycle, damped Jacobi 1+1, Galerkin coarse
    operators down to 8^3 cells solved by their dense inverse) as the CG
    preconditioner at n^3: time of
    one V-cycle, and CG to 1e-8 against Jacobi-CG on the constant-
    coefficient system and on the log-normal (sigma 1) one.  Second solve
    timed; the hierarchy is built once per system (setup reported)."""
    from ba_amd import mg
    E = hflmp.fe_element(1, 3, n)
    b = hflmp.fe_rhs(1, 3, n, 20241202)
    out = {"workload": "3D Q1 %d^3, FP64, CG preconditioned by one V-cycle, rtol 1e-8" % n,
           "cycle": "V(1,1), damped Jacobi w 0.8, dense inverse on the coarsest level",
           "levels": mg.level_sizes(n, min_cells=8)}
    db = ctx.upload(b)
    for name, kap in (("poisson", np.ones(n ** 3)),
                      ("lognormal_sigma1", hflmp.fe_kappa(3, n, 1.0, KAPPA_SEED))):
        rp, ci, v = hflmp.fe_system(1, 3, n, None if name == "poisson" else kap)
        A = hflmp.CsrMatrix(ctx, rp, ci, v)
        t0 = time.perf_counter()
        M = mg.build_structured(ctx, 3, n, E, kap, A, min_cells=8, coarse="dense",
                                pre_smooth=1, post_smooth=1, omega=0.8)
        ctx.sync()
        setup = time.perf_counter() - t0
        rec = {"setup_s": round(setup, 3)}
        dr = ctx.upload(b)
        dz = ctx.empty(A.nrow, np.float64)
        rec["vcycle_us"] = round(time_launches(ctx, lambda: M.apply(dr, dz), steps) * 1e6, 2)
        for rep in range(2):
            dx = ctx.zeros(A.nrow, np.float64)
            info = M.pcg(db, dx, maxit=2000, rtol=1e-8)
            if rep == 1:
                x = dx.download()
            dx.free()
        rec.update({"iters": info["iters"], "status": hflmp.STATUS.get(info["status"], "?"),
                    "seconds_to_tol": round(info["seconds"], 5),
                    "ms_per_iter": round(info["seconds"] * 1e3 / max(info["iters"], 1), 4),
                    "true_res_rel": None})
        dxr = ctx.upload(x)
        hflmp.spmv(ctx, A, dxr, dz)
        rec["true_res_rel"] = float(np.linalg.norm(b - dz.download()) / np.linalg.norm(b))
        dxr.free()
        jac = sell_cg if name == "poisson" else None
        if jac is None:
            S = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
            invd = hflmp.invdiag(ctx, A)
            for rep in range(2):
                dx = ctx.zeros(A.nrow, np.float64)
                ji = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, maxit=20000, rtol=1e-8)
                dx.free()
            jac = {"iters": ji["iters"], "seconds_to_tol": ji["seconds"]}
            S.free(); invd.free()
        rec["jacobi_sell_cg"] = {"iters": jac["iters"],
                                 "seconds_to_tol": round(jac["seconds_to_tol"], 5)}
        rec["speedup_vs_jacobi_sell_cg"] = round(jac["seconds_to_tol"] / info["seconds"], 3)
        M.free(); A.free()
        # the same with the finest level matrix-free (StencilOperator)
        St = hflmp.StencilOperator(ctx, 1, 3, [n] * 3, E, kap, hflmp.fe_fixed(1, 3, n))
        t0 = time.perf_counter()
        M = mg.build_structured(ctx, 3, n, E, kap, None, stencil=St, min_cells=8,
                                coarse="dense", pre_smooth=1, post_smooth=1, omega=0.8)
        ctx.sync()
        mf = {"setup_s": round(time.perf_counter() - t0, 3),
              "vcycle_us": round(time_launches(ctx, lambda: M.apply(dr, dz), steps) * 1e6, 2)}
        for rep in range(2):
            dx = ctx.zeros(St.nrow, np.float64)
            info = M.pcg(db, dx, maxit=2000, rtol=1e-8)
            dx.free()
        invd = St.invdiag()
        for rep in range(2):
            dx = ctx.zeros(St.nrow, np.float64)
            ji = hflmp.pcg_jacobi_stencil(ctx, St, invd, db, dx, maxit=20000, rtol=1e-8)
            dx.free()
        mf.update({"iters": info["iters"], "status": hflmp.STATUS.get(info["status"], "?"),
                   "seconds_to_tol": round(info["seconds"], 5),
                   "jacobi_stencil_cg": {"iters": ji["iters"],
                                         "seconds_to_tol": round(ji["seconds"], 5)},
                   "speedup_vs_jacobi_sell_cg": round(jac["seconds_to_tol"] / info["seconds"], 3),
                   "speedup_vs_jacobi_stencil_cg": round(ji["seconds"] / info["seconds"], 3)})
        rec["matrix_free_fine_level"] = mf
        out[name] = rec
        M.free(); St.free(); invd.free(); dr.free(); dz.free()
    db.free()
    return out


def bench_assembly(hflmp, n, kappa_seed=KAPPA_SEED):
    """Device assembly of the 3D Q1 n^3 system from its cell -> dof table
    (pattern, then values cell by cell), with the Dirichlet rows applied,
    against the host generator's CSR of the same system (bit-identical)."""
    kap = hflmp.fe_kappa(3, n, 1.0, kappa_seed)
    t0 = time.perf_counter()
    rp_h, ci_h, v_h = hflmp.fe_system(1, 3, n, kap)
    t_host = time.perf_counter() - t0
    dofs = hflmp.fe_cell_dofs(1, 3, n)
    E = hflmp.fe_element(1, 3, n)
    N = (n + 1) ** 3
    fixed = hflmp.fe_fixed(1, 3, n)
    ctx = hflmp.Context(0)
    dd, dE, df, dfix = ctx.upload(dofs), ctx.upload(E), ctx.upload(kap), ctx.upload(fixed)
    times = []
    for rep in range(3):
        ctx.sync()
        t0 = time.perf_counter()
        h = C.c_void_p()
        ctx.check(ctx.L.hflmp_fe_mesh_create(ctx.h, N, dofs.shape[0], dofs.shape[1], dd.ptr,
                                             C.byref(h)), "fe_mesh_create")
        z = C.c_longlong()
        ctx.check(ctx.L.hflmp_fe_mesh_nnz(h, C.byref(z)), "fe_mesh_nnz")
        rp, ci, val = ctx.empty(N + 1, np.int32), ctx.empty(z.value, np.int32), \
            ctx.empty(z.value, np.float64)
        ctx.check(ctx.L.hflmp_fe_mesh_pattern(ctx.h, h, rp.ptr, ci.ptr), "fe_mesh_pattern")
        ctx.check(ctx.L.hflmp_fe_assemble_f64(ctx.h, h, None, dE.ptr, df.ptr, rp.ptr, ci.ptr,
                                              val.ptr), "fe_assemble")
        ctx.check(ctx.L.hflmp_csr_zero_rows_f64(ctx.h, len(fixed), dfix.ptr, rp.ptr, ci.ptr,
                                                val.ptr, 1.0), "zero_rows")
        ctx.sync()
        times.append(time.perf_counter() - t0)
        ctx.L.hflmp_fe_mesh_destroy(h)
        if rep < 2:
            rp.free(); ci.free(); val.free()
    same = (np.array_equal(rp.download(), rp_h) and np.array_equal(ci.download(), ci_h)
            and np.array_equal(val.download(), v_h))
    for a in (rp, ci, val, dd, dE, df, dfix):
        a.free()
    ctx.close()
    return {"workload": "3D Q1 %d^3 log-normal diffusivity: cell -> dof table to CSR with "
                        "Dirichlet rows" % n, "nrow": N, "nnz": int(len(ci_h)),
            "device_seconds": round(min(times), 4),
            "host_generator_seconds": round(t_host, 4),
            "host_generator_threads": host_threads(),
            "bit_identical_to_host_csr": bool(same)}


def bench_mixed(hflmp, n):
    """C3: log-normal diffusivity (sigma = 1), 128^3 Q1, relative residual
    1e-10: plain FP64 CG vs FP32-inner / FP64-refined CG, both through the
    SELL-64 operators with int16 columns; FP64 CG with int32 columns is
    timed beside them."""
    from oracle import oracle as O
    kap = hflmp.fe_kappa(3, n, 1.0, KAPPA_SEED)
    rp, ci, v = hflmp.fe_system(1, 3, n, kap)
    b = hflmp.fe_rhs(1, 3, n, RHS_SEED)
    N, nnz = len(rp) - 1, len(ci)
    ctx = hflmp.Context(0)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    A32 = A.astype_f32()
    invd, invd32 = hflmp.invdiag(ctx, A), hflmp.invdiag(ctx, A32)
    S64 = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
    S64w = hflmp.SellMatrix(ctx, A, 0)
    S32 = hflmp.SellMatrix(ctx, A32, hflmp.SELL_COL16)
    db = ctx.upload(b)
    out = {"workload": "3D Q1 log-normal diffusivity (sigma 1) %d^3, Jacobi, rtol 1e-10" % n,
           "seeds": {"kappa": KAPPA_SEED, "rhs": RHS_SEED},
           "timing": "FP64 (int16 and int32 columns), reliable and restart solves in turn, "
                     "3 rounds after a warm-up round, fastest round of each",
           "fp64_operator": "SELL-64, %s columns" % ("int16" if S64.info()["col_bytes"] == 2
                                                      else "int32"),
           "fp32_operator": "SELL-64, %s columns" % ("int16" if S32.info()["col_bytes"] == 2
                                                      else "int32")}
    # The three solves are timed in turn, three rounds after a warm-up
    # round, and each keeps its fastest round: run-to-run drift of the
    # device then touches all three alike.
    def run_fp64(S=S64):
        dx = ctx.zeros(N, np.float64)
        info = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, maxit=50000, rtol=1e-10)
        x = dx.download()
        dx.free()
        return info, x

    # the residual is replaced when the inner solve has reduced it by
    # inner_rtol: 1e-4 for the restart variant; reliable updates replace it
    # earlier (1e-2, the server's default), before the FP32 recursion drifts
    IRT_RESTART, IRT_RELIABLE = 1e-4, 1e-2

    def run_mixed(flags):
        dxm = ctx.zeros(N, np.float64)
        info = hflmp.mpcg_jacobi_sell(ctx, S64, S32, invd32, db, dxm, rtol=1e-10,
                                      inner_rtol=IRT_RELIABLE if flags & hflmp.MP_RELIABLE
                                      else IRT_RESTART, flags=flags)
        xm = dxm.download()
        dxm.free()
        return info, xm

    best = {}
    for rnd in range(4):
        for key, fn in (("fp64", run_fp64), ("fp64_int32", lambda: run_fp64(S64w)),
                        ("reliable", lambda: run_mixed(hflmp.MP_RELIABLE)),
                        ("restart", lambda: run_mixed(0))):
            info, xv = fn()
            if rnd > 0 and (key not in best or info["seconds"] < best[key][0]["seconds"]):
                best[key] = (info, xv)
    i64, x64 = best["fp64"]
    i64w, x64w = best["fp64_int32"]
    ir, xr = best["reliable"]
    im, xm = best["restart"]
    bn = float(np.linalg.norm(b))
    out["fp64_cg"] = {"iters": i64["iters"], "seconds": round(i64["seconds"], 5),
                      "status": hflmp.STATUS.get(i64["status"], "?"),
                      "true_rel_residual": float(np.linalg.norm(b - O.spmv(rp, ci, v, x64)))
                      / bn}
    out["mixed"] = {"outer": im["outer_iters"], "inner_iters": im["inner_iters"],
                    "inner_rtol": IRT_RESTART, "seconds": round(im["seconds"], 5),
                    "status": hflmp.STATUS.get(im["status"], "?"),
                    "true_rel_residual": float(np.linalg.norm(b - O.spmv(rp, ci, v, xm))) / bn,
                    "rel_error_vs_fp64_gpu": float(np.linalg.norm(xm - x64)
                                                   / np.linalg.norm(x64))}
    out["speedup_mixed_vs_fp64"] = round(i64["seconds"] / im["seconds"], 3)
    out["fp64_cg_int32_columns"] = {
        "what": "the same FP64 CG through SELL-64 with int32 columns (12 B per nonzero)",
        "iters": i64w["iters"], "seconds": round(i64w["seconds"], 5),
        "bit_identical_x": bool(np.array_equal(x64w, x64)),
        "speedup_mixed_vs_it": round(i64w["seconds"] / im["seconds"], 3),
        "speedup_mixed_reliable_vs_it": round(i64w["seconds"] / ir["seconds"], 3)}
    out["mixed_reliable"] = {
        "what": "same, with reliable updates: the inner CG keeps its search direction across "
                "corrections instead of restarting",
        "outer": ir["outer_iters"], "inner_iters": ir["inner_iters"],
        "inner_rtol": IRT_RELIABLE, "seconds": round(ir["seconds"], 5), "status": hflmp.STATUS.get(ir["status"], "?"),
        "true_rel_residual": float(np.linalg.norm(b - O.spmv(rp, ci, v, xr))) / bn,
        "rel_error_vs_fp64_gpu": float(np.linalg.norm(xr - x64) / np.linalg.norm(x64)),
        "speedup_vs_fp64": round(i64["seconds"] / ir["seconds"], 3)}
    # for comparison, not a mixed-precision result: FP64 CG with a multigrid
    # V-cycle (matrix-free finest level) on the same system and tolerance
    from ba_amd import mg
    E = hflmp.fe_element(1, 3, n)
    St = hflmp.StencilOperator(ctx, 1, 3, [n] * 3, E, kap, hflmp.fe_fixed(1, 3, n))
    M = mg.build_structured(ctx, 3, n, E, kap, None, stencil=St, min_cells=8, coarse="dense",
                            pre_smooth=1, post_smooth=1, omega=0.8)
    for rep in range(2):
        dxg = ctx.zeros(N, np.float64)
        ig = M.pcg(db, dxg, maxit=2000, rtol=1e-10)
        if rep == 0:
            dxg.free()
    xg = dxg.download()
    dxg.free(); M.free(); St.free()
    out["fp64_mg_pcg"] = {
        "what": "not mixed precision: FP64 CG preconditioned by one V-cycle (matrix-free "
                "finest level, 5 levels, dense coarse solve), same system and tolerance",
        "iters": ig["iters"], "seconds": round(ig["seconds"], 5),
        "status": hflmp.STATUS.get(ig["status"], "?"),
        "true_rel_residual": float(np.linalg.norm(b - O.spmv(rp, ci, v, xg))) / bn,
        "rel_error_vs_fp64_gpu": float(np.linalg.norm(xg - x64) / np.linalg.norm(x64)),
        "speedup_vs_fp64_jacobi_cg": round(i64["seconds"] / ig["seconds"], 3)}
    for a in (db, invd, invd32):
        a.free()
    S64.free(); S64w.free(); S32.free(); A32.val.free(); A.free()
    ctx.close()
    # every result against the FP64 CPU solve of the same system
    t0 = time.perf_counter()
    xc, ic = cpu_solve_for_parity(rp, ci, v, b, 1e-10)
    nc = float(np.linalg.norm(xc))
    out["fp64_cpu_solve"] = {
        "what": "FP64 Jacobi-CG to 1e-10 on the host (all threads), the reference solution "
                "for the errors below", "iters": ic["iters"],
        "seconds": round(time.perf_counter() - t0, 2),
        "true_rel_residual": float(np.linalg.norm(b - O.spmv(rp, ci, v, xc))) / bn}
    out["rel_error_vs_fp64_cpu"] = {
        "fp64_gpu_cg": float(np.linalg.norm(x64 - xc)) / nc,
        "mixed_restart": float(np.linalg.norm(xm - xc)) / nc,
        "mixed_reliable": float(np.linalg.norm(xr - xc)) / nc,
        "fp64_mg_pcg": float(np.linalg.norm(xg - xc)) / nc}
    return out


def bench_batched(hflmp, k=64, n=512, ncheck_cpu=2, spmm_only=False):
    """C4: 2D Q2 512^2, 64 right-hand sides at once (lockstep Jacobi-PCG)
    vs the same right-hand sides solved one at a time: FP64 to 1e-8, plain
    FP32 to 1e-5, and FP32-inner / FP64-refined (MPIR per column) to 1e-8.
    Every column's solution is compared with the CPU FP64 solve for
    `ncheck_cpu` sampled columns."""
    rp, ci, v = hflmp.fe_system(2, 2, n)
    N, nnz = len(rp) - 1, len(ci)
    B = np.stack([hflmp.fe_rhs(2, 2, n, BATCH_SEED0 + c) for c in range(k)], axis=1)
    if spmm_only:
        return bench_spmm_products(hflmp, rp, ci, v, B, k)
    cols_cpu = [int(c) for c in np.linspace(0, k - 1, ncheck_cpu).round()]
    xcpu = {c: cpu_solve_for_parity(rp, ci, v, np.ascontiguousarray(B[:, c]), 1e-8)[0]
            for c in cols_cpu}

    def cpu_err(X):
        return max(float(np.linalg.norm(X[:, c] - xcpu[c]) / np.linalg.norm(xcpu[c]))
                   for c in cols_cpu)

    ctx = hflmp.Context(0)
    out = {"workload": "2D Q2 Poisson %d^2, %d right-hand sides, Jacobi" % (n, k),
           "nrow": N, "nnz": nnz,
           "cpu_parity_sample": "columns %s solved on the host in FP64 to 1e-8" % cols_cpu}
    for dt, name in ((np.float64, "fp64"), (np.float32, "fp32")):
        rtol = 1e-8 if dt == np.float64 else 1e-5
        A = hflmp.CsrMatrix(ctx, rp, ci, v.astype(dt))
        invd = hflmp.invdiag(ctx, A)
        Bd = B.astype(dt)
        dB = ctx.upload(Bd.ravel())
        dX = ctx.zeros(N * k, dt)
        vb = np.dtype(dt).itemsize
        t_mm = time_launches(ctx, lambda: hflmp.spmm(ctx, A, dB, dX, k), 20)
        # the same product through the row-block layout (X rows staged in LDS)
        Y0 = dX.download()
        ctx.sync()
        tp0 = time.perf_counter()
        P = hflmp.SpmmPlan(ctx, A, k)
        t_build = time.perf_counter() - tp0
        t_plan = time_launches(ctx, lambda: P.spmm(A, dB, dX, k), 20)
        plan_rep = {"us": round(t_plan * 1e6, 2),
                    "GBps": round(spmm_bytes(N, nnz, k, vb) / t_plan / 1e9, 2),
                    "frac_hbm": round(spmm_bytes(N, nnz, k, vb) / t_plan / 1e9 / HBM_PEAK_GBS,
                                      4),
                    "bit_identical_to_csr_spmm": bool(np.array_equal(dX.download(), Y0)),
                    "build_s": round(t_build, 3), **P.info()}
        # the lockstep solve with its products through the block layout
        # (the plan built once, outside the timed solve), then through the
        # CSR walk for comparison
        ctx.check(ctx.L.hflmp_fill_zero(ctx.h, dX.ptr, dX.nbytes), "fill")
        info = hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX, k, rtol=rtol, maxit=50000,
                                        plan=P)
        X = dX.download().reshape(N, k).astype(np.float64)
        ctx.check(ctx.L.hflmp_fill_zero(ctx.h, dX.ptr, dX.nbytes), "fill")
        info_csr = hflmp.batched_pcg_jacobi(ctx, A, invd, dB, dX, k, rtol=rtol, maxit=50000)
        X_csr = dX.download().reshape(N, k).astype(np.float64)
        P.free()
        # one at a time: the same columns through the single-vector solver
        S = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
        singles, t_single = [], 0.0
        ncheck = 8
        for c in range(ncheck):
            db = ctx.upload(np.ascontiguousarray(Bd[:, c]))
            dx = ctx.zeros(N, dt)
            si = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, rtol=rtol, maxit=50000)
            singles.append(si["iters"])
            t_single += si["seconds"]
            db.free(); dx.free()
        t_single_all = t_single / ncheck * k
        out[name] = {"rtol": rtol, "spmm_us": round(t_mm * 1e6, 2),
                     "spmm_GBps": round(spmm_bytes(N, nnz, k, vb) / t_mm / 1e9, 2),
                     "spmm_frac_hbm": round(spmm_bytes(N, nnz, k, vb) / t_mm / 1e9
                                            / HBM_PEAK_GBS, 4),
                     "spmm_row_block_layout": plan_rep,
                     "iters_max": int(info["iters"].max()), "iters_min": int(info["iters"].min()),
                     "all_converged": bool((info["status"] == 0).all()),
                     "seconds_batched": round(info["seconds"], 4),
                     "solves_per_s_batched": round(k / info["seconds"], 3),
                     "products": "block layout (spmm_row_block_layout)",
                     "seconds_batched_csr_products": round(info_csr["seconds"], 4),
                     "csr_products_iters_max": int(info_csr["iters"].max()),
                     "max_rel_diff_vs_csr_products": float(
                         np.linalg.norm(X - X_csr) / np.linalg.norm(X_csr)),
                     "seconds_one_at_a_time_est": round(t_single_all, 4),
                     "one_at_a_time_sample": "%d of the %d columns solved singly "
                                             "(SELL operator), scaled to %d" % (ncheck, k, k),
                     "single_iters": singles,
                     "speedup_batched": round(t_single_all / info["seconds"], 3),
                     "max_rel_error_vs_fp64_cpu": cpu_err(X)}
        S.free(); dB.free(); dX.free(); invd.free(); A.free()
    # FP32-inner / FP64-refined: FP64 residuals and outer test, FP32 lockstep
    # inner CG, to 1e-8 on every column
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    A32 = A.astype_f32()
    invd32 = hflmp.invdiag(ctx, A32)
    dB = ctx.upload(B.ravel())
    dX = ctx.zeros(N * k, np.float64)
    P32 = hflmp.SpmmPlan(ctx, A32, k, 4)
    for rep in range(2):
        ctx.check(ctx.L.hflmp_fill_zero(ctx.h, dX.ptr, dX.nbytes), "fill")
        mi = hflmp.batched_mpcg_jacobi(ctx, A, A32, invd32, dB, dX, k, rtol=1e-8,
                                       inner_rtol=1e-4, plan32=P32)
    P32.free()
    X = dX.download().reshape(N, k)
    bn = np.linalg.norm(B, axis=0)
    from oracle import oracle as O
    tr = max(float(np.linalg.norm(B[:, c] - O.spmv(rp, ci, v, np.ascontiguousarray(X[:, c])))
                   / bn[c]) for c in cols_cpu)
    out["fp32_inner_fp64_refined"] = {
        "rtol": 1e-8, "inner_rtol": 1e-4,
        "outer_max": int(mi["outer"].max()), "inner_iters_max": int(mi["inner_iters"].max()),
        "all_converged": bool((mi["status"] == 0).all()),
        "seconds_batched": round(mi["seconds"], 4),
        "solves_per_s_batched": round(k / mi["seconds"], 3),
        "inner_products": "FP32 block layout",
        "speedup_vs_fp64_batched": round(out["fp64"]["seconds_batched"] / mi["seconds"], 3),
        "speedup_vs_fp64_one_at_a_time": round(out["fp64"]["seconds_one_at_a_time_est"]
                                               / mi["seconds"], 3),
        "max_true_rel_residual_cpu_sample": tr,
        "max_rel_error_vs_fp64_cpu": cpu_err(X)}
    dB.free(); dX.free(); invd32.free(); A32.val.free(); A.free()
    ctx.close()
    return out


def bench_spmm_products(hflmp, rp, ci, v, B, k, reps=20):
    """Only the block products of the batched leg: Y = A X for the 64
    right-hand sides, FP64 and FP32, through the CSR SpMM and through the
    row-block layout (bit-identity checked)."""
    N, nnz = len(rp) - 1, len(ci)
    ctx = hflmp.Context(0)
    out = {"workload": "2D Q2 Poisson 512^2, Y = A X for %d right-hand sides" % k}
    for dt, name in ((np.float64, "fp64"), (np.float32, "fp32")):
        vb = np.dtype(dt).itemsize
        A = hflmp.CsrMatrix(ctx, rp, ci, v.astype(dt))
        dX, dY = ctx.upload(B.astype(dt).ravel()), ctx.empty(N * k, dt)
        t_mm = time_launches(ctx, lambda: hflmp.spmm(ctx, A, dX, dY, k), reps)
        Y0 = dY.download()
        P = hflmp.SpmmPlan(ctx, A, k)
        t_plan = time_launches(ctx, lambda: P.spmm(A, dX, dY, k), reps)
        nb = spmm_bytes(N, nnz, k, vb)
        out[name] = {"bytes": nb, "csr_us": round(t_mm * 1e6, 2),
                     "csr_frac_hbm": round(nb / t_mm / 1e9 / HBM_PEAK_GBS, 4),
                     "row_block_us": round(t_plan * 1e6, 2),
                     "row_block_frac_hbm": round(nb / t_plan / 1e9 / HBM_PEAK_GBS, 4),
                     "bit_identical": bool(np.array_equal(dY.download(), Y0)), **P.info()}
        P.free(); dX.free(); dY.free(); A.free()
    ctx.close()
    return out


def serve_stream(nreq, seed):
    """The serving mix: 70% 2D Q1 with n (cells per side) log-uniform in
    [16, 1024], 30% 3D Q1 with n log-uniform in [16, 96] (n rounded), each
    with its own log-normal diffusivity (sigma 1) and uniform RHS seed, and
    unit-rate exponential inter-arrival gaps (scaled by 1/rate later)."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(nreq):
        dim, lo, hi = (2, 16, 1024) if rng.uniform() < 0.7 else (3, 16, 96)
        n = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        specs.append((dim, n, seed * 100003 + 2 * k, seed * 100003 + 2 * k + 1))
    gaps = rng.exponential(1.0, nreq)
    return specs, gaps


def make_request(hflmp, spec):
    dim, n, ks, bs = spec
    kap = hflmp.fe_kappa(dim, n, 1.0, ks)
    rp, ci, v = hflmp.fe_system(1, dim, n, kap)
    return rp, ci, v, hflmp.fe_rhs(1, dim, n, bs)


class RequestFeed:
    """The stream's requests generated on host threads shortly before they
    are submitted (at most `ahead` past the last one taken), so the whole
    stream is never held in memory at once."""

    def __init__(self, hflmp, specs, threads=8, ahead=48):
        from concurrent.futures import ThreadPoolExecutor
        self.hflmp, self.specs, self.ahead = hflmp, specs, ahead
        self.ex = ThreadPoolExecutor(threads)
        self.futs, self.nxt = {}, 0

    def _fill(self, k):
        while self.nxt < len(self.specs) and self.nxt <= k + self.ahead:
            self.futs[self.nxt] = self.ex.submit(make_request, self.hflmp, self.specs[self.nxt])
            self.nxt += 1

    def get(self, k):
        self._fill(k)
        return self.futs.pop(k).result()

    def warm(self):
        """Generates the first window of requests (the first `ahead`) and
        waits for them: a run timed after this starts with requests ready,
        as a server that has been taking a stream does."""
        self._fill(0)
        for f in list(self.futs.values()):
            f.result()

    def close(self):
        self.ex.shutdown(wait=True)


def fifo_latencies(arrivals, service):
    """Completion - arrival of a one-at-a-time server taking requests in
    arrival order with the measured service times."""
    t, out = 0.0, []
    for a, s in zip(arrivals, service):
        t = max(t, a) + s
        out.append(t - a)
    return np.array(out)


def server_of(hflmp, ctx, cfg):
    return hflmp.Server(ctx, cfg["hbm_budget"], cfg["max_active"], cfg["chunk_iters"],
                        cfg["use_graphs"], cfg.get("mixed_min_nnz", 0),
                        cfg.get("inner_rtol", 1e-2), cfg.get("mixed_max_nnz", 0))


def bench_serve_full_load(hflmp, dim=2, n=255, count=256, graphs=True):
    """The solve server at full load: `count` copies of one FE system (2D
    Q1 n x n by default: 65,536 rows, 586,756 nonzeros) submitted at once
    and solved to 1e-8, in FP64 and in the mixed mode.  Reports the
    grouped chunks' device time (HIP events) per CG iteration and the
    bandwidth it implies, from the iteration's algorithmic bytes per
    system: FP64 10 nnz + 113 N (int16-column SELL-64 SpMV with the row
    length, Ap, p at the row and the gathered p once; the x/r/z update 64 N;
    the p update 24 N), the mixed mode's FP32 inner iteration 6 nnz + 57 N."""
    kap = hflmp.fe_kappa(dim, n, 1.0, 7)
    rp, ci, v = hflmp.fe_system(1, dim, n, kap)
    b = hflmp.fe_rhs(1, dim, n, 8)
    N, nnz = len(rp) - 1, len(ci)
    ctx = hflmp.Context(0)
    res = {"systems": count, "system": "%dD Q1 %d^%d: %d rows, %d nonzeros" % (dim, n, dim, N, nnz),
           "rows_in_all": count * N}
    for name, prec, per_iter in (("fp64", 64, 10 * nnz + 113 * N),
                                 ("mixed", 32, 6 * nnz + 57 * N)):
        sv = hflmp.Server(ctx, 32 << 30, 256, 16, graphs, 0, 1e-2, 0)
        try:
            t0 = time.perf_counter()
            tickets = [sv.submit(rp, ci, v, b, rtol=1e-8, precision=prec) for _ in range(count)]
            left, iters, ok = set(tickets), 0, True
            while left:
                if sv.step():
                    for t in list(left):
                        if sv.poll(t)["state"] == 2:
                            _, info = sv.take(t)
                            iters = max(iters, info["iters"])
                            ok = ok and info["status"] == 0
                            left.discard(t)
            wall = time.perf_counter() - t0
            dev = sv.stats()["chunk_device_seconds"]
        finally:
            sv.close()
        res[name] = {"iterations": iters, "converged": ok, "seconds": round(wall, 4),
                     "device_seconds_in_chunks": round(dev, 4),
                     "device_us_per_iteration": round(dev / iters * 1e6, 2),
                     "GBps": round(count * per_iter * iters / dev / 1e9, 1),
                     "bytes_per_iteration_per_system": per_iter}
    ctx.close()
    return res


def run_server_stream(hflmp, ctx, feed_specs, arrivals, cfg, rtol, precision, keep,
                      max_queued=32):
    """The stream through one server, its requests generated as it runs
    (RequestFeed).  arrivals None: each request is submitted as soon as
    fewer than `max_queued` wait for admission (the server never idles for
    lack of work: its capacity); else at its arrival time (latency from
    arrival to the result's return).  Keeps x only for the request indices
    in `keep`; returns (per-request info, kept x by index, latencies,
    makespan, counters)."""
    feed = RequestFeed(hflmp, feed_specs)
    feed.warm()
    sv = server_of(hflmp, ctx, cfg)
    try:
        n = len(feed_specs)
        infos, lat, xs = [None] * n, np.zeros(n), {}
        by_ticket = {}
        nxt = done = 0
        t_feed = t_submit = t_step = 0.0
        t0 = time.perf_counter()
        while nxt < n or by_ticket:
            now = time.perf_counter() - t0
            while nxt < n and ((arrivals is None and sv.stats()["queued"] < max_queued)
                               or (arrivals is not None and arrivals[nxt] <= now)):
                tw = time.perf_counter()
                r = feed.get(nxt)
                tg = time.perf_counter()
                by_ticket[sv.submit(*r, rtol=rtol, precision=precision)] = nxt
                t_feed += tg - tw
                t_submit += time.perf_counter() - tg
                nxt += 1
                now = time.perf_counter() - t0
            if not by_ticket:
                time.sleep(max(0.0, min(arrivals[nxt] - now, 0.005)))
                continue
            ts = time.perf_counter()
            stepped = sv.step()
            t_step += time.perf_counter() - ts
            if stepped:
                now = time.perf_counter() - t0
                for t in list(by_ticket):
                    if sv.poll(t)["state"] == 2:
                        k = by_ticket.pop(t)
                        x, info = sv.take(t)
                        infos[k] = info
                        lat[k] = now - (0.0 if arrivals is None else arrivals[k])
                        if k in keep:
                            xs[k] = x
                        done += 1
                        if done % 1000 == 0:
                            progress("serving (%s): %d of %d requests done"
                                     % ("capacity" if arrivals is None else "latency",
                                        done, n))
        makespan = time.perf_counter() - t0
        stats = sv.stats()
        stats["host_seconds"] = {"waiting_for_generated_requests": round(t_feed, 3),
                                 "submit": round(t_submit, 3), "step": round(t_step, 3),
                                 "step_admit": round(stats["admit_seconds"], 3),
                                 "step_tables_and_graphs": round(stats["capture_seconds"], 3),
                                 "step_wait_for_kernels": round(stats["wait_seconds"], 3),
                                 "step_wait_for_admission": round(stats["join_seconds"], 3),
                                 "step_retire": round(stats["retire_seconds"], 3)}
        # the device's own time in the grouped chunks (HIP events): the rest
        # of the makespan the device waited for the host
        stats["device_seconds_in_chunks"] = round(stats["chunk_device_seconds"], 3)
    finally:
        sv.close()
        feed.close()
    return infos, xs, lat, makespan, stats


def run_singles(hflmp, ctx, specs, rtol):
    """One at a time, device-resident FP64 Jacobi-CG per request: upload,
    SELL-64 build (int16 columns where they fit), Jacobi, solve, download.
    The requests come from the same generator as the server's (RequestFeed,
    8 host threads working ahead of the solves), so both runs share the
    host with the generation.  Returns the requests too (for parity)."""
    service, out, reqs, solve_s = [], [], [], 0.0
    feed = RequestFeed(hflmp, specs)
    feed.warm()
    t0 = time.perf_counter()
    for k in range(len(specs)):
        ts = time.perf_counter()
        rp, ci, v, b = feed.get(k)
        reqs.append((rp, ci, v, b))
        A = hflmp.CsrMatrix(ctx, rp, ci, v)
        S = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
        invd = hflmp.invdiag(ctx, A)
        db, dx = ctx.upload(b), ctx.zeros(A.nrow, np.float64)
        info = hflmp.pcg_jacobi_sell(ctx, S, invd, db, dx, rtol=rtol, maxit=100000)
        solve_s += info["seconds"]
        x = dx.download()
        S.free(); A.free(); invd.free(); db.free(); dx.free()
        service.append(time.perf_counter() - ts)
        out.append((x, info))
    t_all = time.perf_counter() - t0
    feed.close()
    return service, out, solve_s, t_all, reqs


def bench_serve(hflmp, nreq, seed, rate, ref_style, graphs=True, ntune=160, sample_every=20,
                ref_every=100, load=0.7):
    """C5: a stream of `nreq` FE solves (Jacobi-CG to 1e-8) with Poisson
    arrivals through the solve server.  Requests are generated on host
    threads as the server runs (never the whole stream at once).
      * capacity: every request present as soon as the server can queue it;
      * latency: the same stream submitted at its Poisson arrival times, at
        `load` x the measured capacity (or `rate`);
      * one at a time (device-resident FP64 CG) on every `sample_every`-th
        request of the stream, and its FIFO latencies on that sub-stream at
        the same offered rate; the reference's host-driven CG<LAD> on every
        `ref_every`-th request;
      * parity on the sampled requests: server x vs the single solve and the
        true residual b - A x (computed after the timed runs).
    The server's precision window and step length come from the autotuner
    run on a separate tuning sample."""
    from ba_amd import serve_tune
    from oracle import oracle as O
    rtol = 1e-8
    specs, gaps = serve_stream(nreq, seed)
    tspecs, _ = serve_stream(ntune, seed + 1)
    ctx = hflmp.Context(0)
    # generated on the stream's 8 host threads; held while the tuner runs
    # (160 requests of the mix: about 4.3 GB)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(8) as ex:
        treqs = list(ex.map(lambda sp: make_request(hflmp, sp), tspecs))
    base = {"hbm_budget": 32 << 30, "max_active": 256, "chunk_iters": 16, "use_graphs": graphs,
            "mixed_min_nnz": 0, "inner_rtol": 1e-2, "mixed_max_nnz": 0}
    t0 = time.perf_counter()
    cfg, tune_report = serve_tune.tune(ctx, treqs, base, rtol)
    t_tune = time.perf_counter() - t0
    progress("serving: autotuner done in %.1f s" % t_tune)
    del treqs
    sample = list(range(0, nreq, sample_every))
    keep = set(sample)
    # capacity: the whole stream, the server never short of queued work
    infos, xs, _, makespan, stats = run_server_stream(hflmp, ctx, specs, None, cfg, rtol, 0,
                                                      keep)
    capacity = nreq / makespan
    progress("serving: capacity run %.1f s, %.1f solves/s" % (makespan, capacity))
    # latency under Poisson arrivals
    lam = rate if rate else load * capacity
    arrivals = np.cumsum(gaps) / lam
    linfos, _, lat, mk2, _ = run_server_stream(hflmp, ctx, specs, arrivals, cfg, rtol, 0, set())
    progress("serving: latency run %.1f s" % mk2)
    # one at a time on the sample (requests generated again, identical by seed)
    service, singles, solve_s, t_single, sreqs = run_singles(
        hflmp, ctx, [specs[k] for k in sample], rtol)
    single_rate = len(sample) / t_single
    progress("serving: one at a time on %d requests, %.1f solves/s" % (len(sample), single_rate))
    # FIFO on the sample's own arrivals, rescaled to the same offered rate
    sub_arr = arrivals[sample] / sample_every
    base_lat = fifo_latencies(sub_arr, service)
    # parity on the sample
    err = [float(np.linalg.norm(xs[k] - xsgl) / max(np.linalg.norm(xsgl), 1e-300))
           for k, (xsgl, _) in zip(sample, singles)]
    itd = [abs(infos[k]["iters"] - si["iters"]) for k, (_, si) in zip(sample, singles)
           if infos[k]["precision"] == 64]
    # FP64 requests follow the single solve's arithmetic; mixed ones meet
    # the same tolerance on the true residual by another path
    e64 = [e for e, k in zip(err, sample) if infos[k]["precision"] == 64]
    e32 = [e for e, k in zip(err, sample) if infos[k]["precision"] == 32]
    true_res = [float(np.linalg.norm(b - O.spmv(rp, ci, v, xs[k])) /
                      max(np.linalg.norm(b), 1e-300))
                for k, (rp, ci, v, b) in zip(sample, sreqs)]
    fp64_bytes = sum(si["iters"] * cg_iter_bytes(len(r[0]) - 1, len(r[1]))
                     for (_, si), r in zip(singles, sreqs))
    nmixed = sum(1 for i in infos if i["precision"] == 32)
    res = {"stream": {"requests": nreq, "seed": seed, "rtol": rtol,
                      "mix": "70% 2D Q1 n log-uniform [16,1024], 30% 3D Q1 n log-uniform "
                             "[16,96]; log-normal diffusivity sigma 1; uniform RHS; "
                             "generated on 8 host threads while the server runs (the first 49 "
                             "before the clock starts, for the server and the one-at-a-time "
                             "runs alike)",
                      "arrivals": "Poisson (exponential gaps, seed %d)" % seed},
           "autotune": {"sample": "%d requests of the same mix, seed %d, submitted at once"
                                  % (ntune, seed + 1),
                        "seconds": round(t_tune, 2), "report": tune_report},
           "server": {"config": cfg, "solves_per_s": round(capacity, 2),
                      "makespan_s": round(makespan, 3), "steps": stats["steps"],
                      "graph_captures": stats["graph_captures"],
                      "host_seconds": stats["host_seconds"],
                      "device_seconds_in_chunks": stats["device_seconds_in_chunks"],
                      "peak_active": stats["peak_active"],
                      "peak_arena_in_use_GB": round(stats["peak_arena_in_use"] / 1e9, 3),
                      "mixed_requests": nmixed,
                      "all_converged": all(i["status"] == 0 for i in infos),
                      "iterations_total": int(sum(i["iters"] for i in infos)),
                      "what": "capacity: every request queued as soon as fewer than 32 "
                              "wait for admission"},
           "latency": {"rate_per_s": round(lam, 2), "offered_load": round(lam / capacity, 3),
                       "server_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "server_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                       "server_max_ms": round(float(lat.max()) * 1e3, 2),
                       "server_solves_per_s": round(nreq / mk2, 2),
                       "all_converged": all(i["status"] == 0 for i in linfos),
                       "one_at_a_time_p50_ms": round(float(np.percentile(base_lat, 50)) * 1e3,
                                                     2),
                       "one_at_a_time_p99_ms": round(float(np.percentile(base_lat, 99)) * 1e3,
                                                     2),
                       "note": "server measured live over all %d requests; one at a time: "
                               "FIFO on its measured service times over the sampled "
                               "sub-stream at the same offered rate" % nreq},
           "one_at_a_time_device": {"sample": "every %dth request of the stream (%d), "
                                              "from the same 8-thread generator as the "
                                              "server's" % (sample_every, len(sample)),
                                    "solves_per_s": round(single_rate, 2),
                                    "seconds": round(t_single, 3),
                                    "solve_loop_seconds": round(solve_s, 3),
                                    "effective_GBps_fp64_cg_bytes":
                                        round(fp64_bytes / t_single / 1e9, 1),
                                    "all_converged": all(s["status"] == 0
                                                         for _, s in singles)},
           "parity_on_sample": {
               "fp64_requests": len(e64), "mixed_requests": len(e32),
               "max_rel_diff_x_vs_single_fp64": max(e64) if e64 else None,
               "max_rel_diff_x_vs_single_mixed": max(e32) if e32 else None,
               "max_iter_diff_fp64": max(itd) if itd else None,
               "max_true_rel_residual": max(true_res)}}
    res["speedup_vs_one_at_a_time_device"] = round(capacity / single_rate, 3)
    if ref_style:
        rsample = list(range(0, nreq, ref_every))
        ok, t_ref = True, 0.0
        for k in rsample:
            rp, ci, v, b = make_request(hflmp, specs[k])
            ts = time.perf_counter()
            _, r = hflmp.lmp_solve(rp, ci, v, b, solver=0, precond=1, rtol=rtol, maxit=100000)
            t_ref += time.perf_counter() - ts
            ok = ok and r["status"] == 0
        res["one_at_a_time_reference_style"] = {
            "sample": "every %dth request (%d)" % (ref_every, len(rsample)),
            "solves_per_s": round(len(rsample) / t_ref, 2), "all_converged": ok,
            "what": "CG<LAD>::SolvePrecond with the block-Jacobi preconditioner over "
                    "CoupledMatrix/CoupledVector on the HIP platform (host-driven, one "
                    "synchronising dot per reduction)"}
        res["speedup_vs_one_at_a_time_reference_style"] = round(
            capacity * t_ref / len(rsample), 3)
    ctx.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=128, help="cells per side")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-stencil", action="store_true")
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--batched-spmm-only", action="store_true",
                    help="in the batched leg, only the block products (no solves)")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-seed", type=int, default=20241205)
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="Poisson arrival rate for the latency run (default: 60%% of the "
                         "server's measured capacity)")
    ap.add_argument("--no-serve-reference", action="store_true")
    ap.add_argument("--serve-full-load-only", action="store_true",
                    help="in the serving leg, only the full-load run (256 systems at once)")
    ap.add_argument("--serve-no-graphs", action="store_true",
                    help="run the server's steps as direct launches instead of graph replay")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (hflmp_copy_f64), the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from ba_amd import hflmp
    if args.copy_calibration:
        copy_calibration(hflmp, args.steps)
        return
    n = args.n
    t0 = time.perf_counter()
    rp, ci, v = hflmp.fe_system(1, 3, n)
    b = hflmp.fe_rhs(1, 3, n, 20241202)
    t_gen = time.perf_counter() - t0
    N, nnz = len(rp) - 1, len(ci)
    x = np.random.default_rng(7).uniform(-1, 1, N)

    ctx = hflmp.Context(0)
    A = hflmp.CsrMatrix(ctx, rp, ci, v)
    t0 = time.perf_counter()
    S = hflmp.SellMatrix(ctx, A, 0)
    ctx.sync()
    t_build = time.perf_counter() - t0
    sinfo = S.info()
    dx = ctx.upload(x)
    dy = ctx.empty(N, np.float64)
    nb = spmv_bytes(N, nnz)

    # ---- the timed steps: SELL-64 FP64 SpMV, int32 columns
    for _ in range(args.warmup):
        S.spmv(dx, dy)
    ev0, ev1 = ctx.event(), ctx.event()
    ctx.sync()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        S.spmv(dx, dy)
    ev1.record()
    ctx.sync()
    wall = time.perf_counter() - w0
    per_launch_s = ev1.elapsed_ms(ev0) / 1e3 / args.steps
    gbs = nb / per_launch_s / 1e9
    progress("SpMV timed steps: %.2f us per launch, %.1f GB/s" % (per_launch_s * 1e6, gbs))

    out = {
        "metric": "FP64 SpMV GB/s (CSR algorithmic bytes), 3D Q1 Poisson 128^3",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: structured 3D Q1 FE stiffness, Dirichlet rows, random x",
        "config": {"workload": "3D Q1 Poisson %d^3 cells, FP64 SpMV" % n, "nrow": N,
                   "nnz": nnz, "bytes_per_spmv": nb, "parallelism": "single GPU",
                   "operator_layout": "SELL-64 (int32 columns, P=2) built on the device from "
                                      "the CSR arrays; products bit-identical to the CSR loop",
                   "layout_build_ms": round(t_build * 1e3, 2),
                   "layout_entries": sinfo["entries"],
                   "layout_device_bytes": sinfo["device_bytes"]},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(n, "sell_spmv_f64")},
        "spmv_us": round(per_launch_s * 1e6, 2),
        "gen_seconds": round(t_gen, 3),
    }

    # ---- other layouts of the same product, reported separately
    dz = ctx.empty(N, np.float64)
    t_csr = time_launches(ctx, lambda: hflmp.spmv(ctx, A, dx, dz), max(args.steps // 2, 1))
    S16 = hflmp.SellMatrix(ctx, A, hflmp.SELL_COL16)
    dw = ctx.empty(N, np.float64)
    t_16 = time_launches(ctx, lambda: S16.spmv(dx, dw), max(args.steps // 2, 1))
    out["spmv_csr_tiles"] = {"us": round(t_csr * 1e6, 2), "GBps": round(nb / t_csr / 1e9, 2),
                             "frac": round(nb / t_csr / 1e9 / HBM_PEAK_GBS, 4),
                             "traffic": measured_traffic(n, "csr_spmv_f64"),
                             "note": "row-tile kernel reading the CSR arrays directly"}
    tr16 = measured_traffic(n, "sell_spmv_f64_int16")
    out["spmv_sell_int16"] = {"us": round(t_16 * 1e6, 2), "GBps": round(nb / t_16 / 1e9, 2),
                              "frac": round(nb / t_16 / 1e9 / HBM_PEAK_GBS, 4),
                              "traffic": tr16,
                              "hbm_GBps_measured_traffic": round(tr16 / t_16 / 1e9, 2)
                              if tr16 else None,
                              "note": "compressed (int16 column offsets, 10 B per nonzero; the "
                                      "layout HIP_CSR_lMatrix builds when every column fits); "
                                      "scored with the CSR bytes"}

    if not args.no_parity:
        from oracle import oracle as O
        yref = O.spmv(rp, ci, v, x)
        S.spmv(dx, dy)
        hflmp.spmv(ctx, A, dx, dz)
        S16.spmv(dx, dw)
        y, z, w = dy.download(), dz.download(), dw.download()
        out["parity"] = {"spmv_bit_exact": bool(np.array_equal(y, yref)),
                         "csr_tiles_bit_exact": bool(np.array_equal(z, yref)),
                         "sell_int16_bit_exact": bool(np.array_equal(w, yref)),
                         "max_abs_diff": float(np.max(np.abs(y - yref)))}
    dz.free(); dw.free()

    x_cg = None
    if not args.no_cg:
        out["cg"], x_cg = bench_cg(hflmp, ctx, A, S16, S, b)
        progress("Jacobi-CG leg done")
    S16.free()
    if not args.no_stencil:
        out["matrix_free"] = bench_stencil(hflmp, ctx, n, x, b, max(args.steps // 2, 1),
                                           out.get("cg", {}).get("sell"),
                                           None if args.no_parity else (S, rp, ci, v))
        progress("matrix-free leg done")
    if not args.no_mg:
        out["multigrid"] = bench_mg(hflmp, ctx, n, max(args.steps // 10, 5),
                                    out.get("cg", {}).get("sell"))
        progress("multigrid leg done")
    S.free(); A.free(); dx.free(); dy.free()
    ctx.close()

    if not args.no_assembly:
        out["assembly"] = bench_assembly(hflmp, n)
        progress("assembly leg done")
    if not args.no_mixed:
        out["mixed_precision"] = bench_mixed(hflmp, n)
        progress("mixed-precision leg done")
    if not args.no_batched:
        out["batched"] = bench_batched(hflmp, spmm_only=args.batched_spmm_only)
        progress("batched leg done")
    if not args.no_serve and not args.serve_full_load_only:
        out["serving"] = bench_serve(hflmp, args.serve_requests, args.serve_seed,
                                     args.serve_rate, not args.no_serve_reference,
                                     not args.serve_no_graphs)
        progress("serving leg done")
    if not args.no_serve:
        out["serving_full_load"] = bench_serve_full_load(hflmp, graphs=not args.serve_no_graphs)
        progress("serving at full load done")

    # CG time to 1e-8 on the 128^3 Poisson system, every solver path side
    # by side (each measured in its leg above)
    tt = {}
    if "cg" in out:
        tt["jacobi_cg_sell64_int16_ms"] = round(out["cg"]["sell"]["seconds_to_tol"] * 1e3, 3)
        tt["jacobi_cg_sell64_int32_ms"] = round(
            out["cg"]["sell_int32"]["seconds_to_tol"] * 1e3, 3)
        tt["jacobi_cg_csr_ms"] = round(out["cg"]["csr"]["seconds_to_tol"] * 1e3, 3)
    if "matrix_free" in out:
        tt["jacobi_cg_matrix_free_ms"] = round(out["matrix_free"]["cg"]["seconds_to_tol"] * 1e3, 3)
    if "multigrid" in out:
        mgp = out["multigrid"]["poisson"]
        tt["mg_pcg_sell64_ms"] = round(mgp["seconds_to_tol"] * 1e3, 3)
        tt["mg_pcg_matrix_free_ms"] = round(
            mgp["matrix_free_fine_level"]["seconds_to_tol"] * 1e3, 3)
    if tt:
        out["cg_time_to_1e-8"] = tt
    if not args.no_cpu:
        base, naive = cpu_baseline(rp, ci, v, x, args.cpu_seconds)
        base["cpu"] = cpu_model()
        progress("CPU SpMV baseline done")
        out["cpu_baseline"] = base
        out["cpu_baseline_naive"] = naive
        if "cg" in out:
            # CG time to 1e-8 at 128^3 along the reference's CPU path, and
            # the GPU's int16 SELL-64 solve against it
            g = out["cg"]["sell"]
            ccg, _ = cpu_cg(rp, ci, v, b, 1e-8, x_cg, g["iters"])
            ccg["speedup_gpu_sell_int16_vs_cpu"] = round(ccg["seconds_to_tol"] /
                                                         g["seconds_to_tol"], 2)
            out["cpu_baseline_cg"] = ccg
            tt["cpu_reference_path_ms"] = round(ccg["seconds_to_tol"] * 1e3, 1)
            progress("CPU CG baseline done")
        if not args.no_c1:
            out["c1_2d_q1_256"] = bench_c1(hflmp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
