"""Autotuner for the solve server (hflmp_server_*).

Per size class (rows of the system) it chooses the precision mode -- FP64
CG or the FP32-inner / FP64-refined mode -- by timing the class's systems
from a tuning sample through the server in both modes; the classes where
the mixed mode wins from some size upwards give `mixed_min_rows`.  Then it
chooses the iterations per step (the step is the unit of admission,
retirement and graph replay) among candidates by the sample's throughput.

The tuning sample is solved for real (same tolerance), so a mode that does
not converge for a class is never chosen for it.
"""
import time

import numpy as np

from ba_amd import hflmp

# row-count boundaries of the size classes
CLASS_BOUNDS = (0, 4096, 32768, 262144, 1 << 31)


def size_class(nrow):
    for k in range(len(CLASS_BOUNDS) - 1):
        if CLASS_BOUNDS[k] <= nrow < CLASS_BOUNDS[k + 1]:
            return k
    return len(CLASS_BOUNDS) - 2


def _run(ctx, reqs, cfg, precision, rtol):
    sv = hflmp.Server(ctx, cfg["hbm_budget"], cfg["max_active"], cfg["chunk_iters"],
                      cfg["use_graphs"], cfg.get("mixed_min_rows", 0),
                      cfg.get("inner_rtol", 1e-4))
    try:
        t0 = time.perf_counter()
        tickets = [sv.submit(*r, rtol=rtol, precision=precision) for r in reqs]
        left = set(tickets)
        ok = True
        while left:
            if sv.step():
                for t in list(left):
                    if sv.poll(t)["state"] == 2:
                        _, info = sv.take(t)
                        ok = ok and info["status"] == 0
                        left.discard(t)
        return time.perf_counter() - t0, ok
    finally:
        sv.close()


def tune(ctx, sample, base_cfg, rtol=1e-8, chunks=(8, 16, 32)):
    """sample: list of (rp, ci, val, b).  Returns (config, report)."""
    cfg = dict(base_cfg)
    classes = {}
    for r in sample:
        classes.setdefault(size_class(len(r[0]) - 1), []).append(r)
    report = {"classes": {}}
    mixed_wins = {}
    for k in sorted(classes):
        reqs = classes[k]
        _run(ctx, reqs[:1], cfg, 64, rtol)  # warm the kernels
        t64, ok64 = _run(ctx, reqs, cfg, 64, rtol)
        t32, ok32 = _run(ctx, reqs, cfg, 32, rtol)
        mixed_wins[k] = ok32 and (t32 < t64 or not ok64)
        report["classes"][k] = {"rows": [CLASS_BOUNDS[k], CLASS_BOUNDS[k + 1]],
                                "systems": len(reqs), "fp64_s": round(t64, 4),
                                "mixed_s": round(t32, 4),
                                "choice": "mixed" if mixed_wins[k] else "fp64"}
    # the smallest class from which the mixed mode wins for every larger class
    start = None
    for k in sorted(classes, reverse=True):
        if mixed_wins[k]:
            start = k
        else:
            break
    cfg["mixed_min_rows"] = CLASS_BOUNDS[start] if start is not None else 0
    if cfg["mixed_min_rows"] == 0 and start is not None:
        cfg["mixed_min_rows"] = 1
    best = None
    report["chunks"] = {}
    for c in chunks:
        trial = dict(cfg, chunk_iters=c)
        t, ok = _run(ctx, sample, trial, 0, rtol)
        report["chunks"][c] = round(t, 4)
        if ok and (best is None or t < best[0]):
            best = (t, c)
    if best is not None:
        cfg["chunk_iters"] = best[1]
    report["mixed_min_rows"] = cfg["mixed_min_rows"]
    report["chunk_iters"] = cfg["chunk_iters"]
    return cfg, report
