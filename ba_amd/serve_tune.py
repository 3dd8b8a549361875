"""Autotuner for the solve server (hflmp_server_*).

It chooses the precision window [`mixed_min_nnz`, `mixed_max_nnz`) -- the
systems solved by the FP32-inner / FP64-refined mode rather than FP64 CG --
by timing the whole tuning sample through the server under each candidate
window; with a mixed window, the inner tolerance (how far the FP32 inner
solve drives its residual before the FP64 one replaces it); then the
iterations per step (the step is the unit of admission, retirement and
graph replay) among candidates by the sample's throughput.  The sample must be large enough to load the server: the mixed
mode's gain is a bandwidth gain (fewer bytes per iteration), which a few
systems at a time do not show, since their launches' latency hides it.
Candidate windows start at size-class boundaries (nonzeros of the system,
which the bytes per iteration follow) and have no upper bound: the larger
the system, the more of its time is bandwidth.

The tuning sample is solved for real (same tolerance), so a window under
which a system does not converge is never chosen.
"""
import time

import numpy as np

from ba_amd import hflmp

# nonzero-count boundaries of the size classes
CLASS_BOUNDS = (0, 1 << 16, 1 << 20, 8 << 20, (1 << 31) - 1)


def size_class(nnz):
    for k in range(len(CLASS_BOUNDS) - 1):
        if CLASS_BOUNDS[k] <= nnz < CLASS_BOUNDS[k + 1]:
            return k
    return len(CLASS_BOUNDS) - 2


def _run(ctx, reqs, cfg, precision, rtol):
    sv = hflmp.Server(ctx, cfg["hbm_budget"], cfg["max_active"], cfg["chunk_iters"],
                      cfg["use_graphs"], cfg.get("mixed_min_nnz", 0),
                      cfg.get("inner_rtol", 1e-2), cfg.get("mixed_max_nnz", 0))
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


MARGIN = 0.05  # a chunk length replaces the default only when this much faster
# A window replaces FP64 only when the whole sample gains more than this: it
# adds per-system outer steps to launches shared with every other system.
WINDOW_MARGIN = 0.10
# The inner tolerance of a mixed window changes the same kernels' iteration
# counts, not which kernels run: a smaller gain carries over (stream probes:
# 0.1 against 0.01 took 2.5-3% off the device time, every time).
IRT_MARGIN = 0.02
# candidate windows (lo, hi): FP64 only, then mixed from each class boundary
WINDOWS = ((0, 0),) + tuple((max(b, 1), 0) for b in reversed(CLASS_BOUNDS[:-1]))


def choose_window(times, margin=WINDOW_MARGIN):
    """The window with the shortest time among those whose sample
    converged (times: window -> seconds, None where a system failed), if it
    beats FP64 (window (0, 0)) by more than `margin` or FP64 failed;
    otherwise FP64."""
    ok = {w: t for w, t in times.items() if t is not None}
    if not ok:
        return (0, 0)
    best = min(ok, key=ok.get)
    t64 = ok.get((0, 0))
    if t64 is None or ok[best] < (1 - margin) * t64:
        return best
    return (0, 0)


def _best(ctx, reqs, cfg, precision, rtol):
    (t1, ok1), (t2, ok2) = (_run(ctx, reqs, cfg, precision, rtol),
                            _run(ctx, reqs, cfg, precision, rtol))
    return min(t1, t2), ok1 and ok2


def tune(ctx, sample, base_cfg, rtol=1e-8, chunks=(8, 16, 32), windows=WINDOWS,
         inner_rtols=(1e-1,)):
    """sample: list of (rp, ci, val, b), submitted at once (enough systems
    to load the server).  Returns (config, report).  Every timing is the
    better of two runs; the window departs from FP64 on a gain of more than
    WINDOW_MARGIN, the inner tolerance (with a mixed window) from
    base_cfg's on more than IRT_MARGIN, the chunk length on more than
    MARGIN."""
    cfg = dict(base_cfg)
    _run(ctx, sample[:1], cfg, 64, rtol)  # warm the kernels
    report = {"sample_systems": len(sample), "windows": {}}
    times = {}
    for lo, hi in windows:
        t, ok = _best(ctx, sample, dict(cfg, mixed_min_nnz=lo, mixed_max_nnz=hi), 0, rtol)
        report["windows"]["[%d, %d)" % (lo, hi)] = {"seconds": round(t, 4), "converged": ok}
        times[(lo, hi)] = t if ok else None
    cfg["mixed_min_nnz"], cfg["mixed_max_nnz"] = choose_window(times)
    if cfg["mixed_min_nnz"] > 0:
        # how far the FP32 inner solve drives its residual before the FP64
        # one replaces it: replacing more often costs outer steps and keeps
        # the FP32 recursion closer to the true residual (fewer inner
        # iterations on this mix)
        t_def = times[(cfg["mixed_min_nnz"], cfg["mixed_max_nnz"])]
        report["inner_rtol"] = {str(cfg["inner_rtol"]): round(t_def, 4)}
        for irt in inner_rtols:
            if irt == cfg["inner_rtol"]:
                continue
            t, ok = _best(ctx, sample, dict(cfg, inner_rtol=irt), 0, rtol)
            report["inner_rtol"][str(irt)] = round(t, 4)
            if ok and t < (1 - IRT_MARGIN) * t_def:
                cfg["inner_rtol"], t_def = irt, t
    report["chunks"] = {}
    ctimes = {}
    for c in sorted(set(chunks) | {cfg["chunk_iters"]}):
        t, ok = _best(ctx, sample, dict(cfg, chunk_iters=c), 0, rtol)
        report["chunks"][c] = round(t, 4)
        if ok:
            ctimes[c] = t
    if ctimes:
        c_best = min(ctimes, key=ctimes.get)
        t_def = ctimes.get(cfg["chunk_iters"])
        if t_def is None or ctimes[c_best] < (1 - MARGIN) * t_def:
            cfg["chunk_iters"] = c_best
    report["mixed_min_nnz"] = cfg["mixed_min_nnz"]
    report["mixed_max_nnz"] = cfg["mixed_max_nnz"]
    report["inner_rtol_chosen"] = cfg["inner_rtol"]
    report["chunk_iters"] = cfg["chunk_iters"]
    return cfg, report
