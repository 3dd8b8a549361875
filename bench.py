#!/usr/bin/env python3
"""Benchmark of the gfx950 lmp platform on the 3D Q1 Poisson system
(128^3 hex cells, 129^3 = 2,146,689 DoFs, 57,066,625 nonzeros, FP64).

A step is one FP64 CSR SpMV y = A x of that matrix with A, x, y resident in
HBM.  `value` is the SpMV throughput in GB/s of algorithmic bytes
(12 nnz + 20 N + 4 per SpMV: int32 col + FP64 val per nonzero, int32
row_ptr, x read once, y written once).  The same run also times
Jacobi-preconditioned CG to a relative residual of 1e-8 on the same system
and the CPU restatement of the reference SpMV on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def spmv_bytes(N, nnz, vb=8):
    return (4 + vb) * nnz + 4 * (N + 1) + 2 * vb * N


def cg_iter_bytes(N, nnz, vb=8):
    # SpMV + 13 N-vector streams of the fused iteration
    return (4 + vb) * nnz + 4 * (N + 1) + 13 * vb * N


def host_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


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


def measured_traffic(n):
    """HBM bytes per SpMV launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (scaled by a 1 GiB copy), recorded in profiles/; None if not measured
    for this size."""
    path = os.path.join(ROOT, "profiles", "spmv_128cube_traffic.json")
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["traffic_bytes"])


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
    dx = ctx.upload(x)
    dy = ctx.empty(N, np.float64)
    nb = spmv_bytes(N, nnz)

    for _ in range(args.warmup):
        hflmp.spmv(ctx, A, dx, dy)
    ev0, ev1 = ctx.event(), ctx.event()
    ctx.sync()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        hflmp.spmv(ctx, A, dx, dy)
    ev1.record()
    ctx.sync()
    wall = time.perf_counter() - w0
    ms = ev1.elapsed_ms(ev0)
    per_launch_s = ms / 1e3 / args.steps
    gbs = nb / per_launch_s / 1e9

    out = {
        "metric": "FP64 CSR SpMV GB/s (algorithmic bytes), 3D Q1 Poisson 128^3",
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
        "config": {"workload": "3D Q1 Poisson %d^3 cells, CSR FP64 SpMV" % n, "nrow": N,
                   "nnz": nnz, "bytes_per_spmv": nb, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(n)},
        "spmv_us": round(per_launch_s * 1e6, 2),
        "gen_seconds": round(t_gen, 3),
    }

    if not args.no_parity:
        from oracle import oracle as O
        hflmp.spmv(ctx, A, dx, dy)
        y = dy.download()
        yref = O.spmv(rp, ci, v, x)
        out["parity"] = {"spmv_bit_exact": bool(np.array_equal(y, yref)),
                         "max_abs_diff": float(np.max(np.abs(y - yref)))}

    if not args.no_cg:
        invd = hflmp.invdiag(ctx, A)
        db, dxs = ctx.upload(b), ctx.zeros(N, np.float64)
        info = hflmp.pcg_jacobi(ctx, A, invd, db, dxs, maxit=20000, rtol=1e-8)  # warm
        dxs.free()
        dxs = ctx.zeros(N, np.float64)
        info = hflmp.pcg_jacobi(ctx, A, invd, db, dxs, maxit=20000, rtol=1e-8)
        it = max(info["iters"], 1)
        out["cg"] = {"precond": "jacobi", "rtol": 1e-8, "iters": info["iters"],
                     "status": hflmp.STATUS.get(info["status"], "?"),
                     "seconds_to_tol": round(info["seconds"], 5),
                     "ms_per_iter": round(info["seconds"] * 1e3 / it, 5),
                     "GBps_per_iter": round(cg_iter_bytes(N, nnz) * it / info["seconds"] / 1e9,
                                            2),
                     "res_final_rel": info["res_final"] / info["res_init"]}
        dxs.free(); db.free(); invd.free()

    if not args.no_cpu:
        base, naive = cpu_baseline(rp, ci, v, x, args.cpu_seconds)
        out["cpu_baseline"] = base
        out["cpu_baseline_naive"] = naive
    print(json.dumps(out), flush=True)
    A.free(); dx.free(); dy.free()
    ctx.close()


if __name__ == "__main__":
    main()
