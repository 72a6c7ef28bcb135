"""Autotuner of the window solver's plan options (work-item tile shapes and launch shapes) for one batch on this GPU.

Options tuned (``xba_options`` in include/xba.h):
  slot_tile     keyframes spanned by the narrow work-item tile: landmarks coupling at most this many consecutive
                keyframes go to the narrow assembly kernel (more wavefronts per workgroup, smaller LDS tile), wider
                ones to the wide kernel
  lm_per_chunk  landmarks per work item (0: the solver's rule, landmarks / (4 x compute units))
  big_n         reduced-system size above which a window's Cholesky is tiled over many workgroups (64 x 64 tiles,
                trailing updates on the matrix cores) instead of one workgroup per window
  big_fp64      mixed precision: 1, the tiled windows factored in FP64 (one solve) instead of FP32 with FP64
                refinement (several solves); 2, every window
  grid          persistent grid of the assembly kernels (0: one workgroup per compute unit)

Search: coordinate descent over the options in that order, starting from the solver's defaults, two rounds (the
second revisits each option with the others at their first-round choices).  A candidate is timed as the median of
``repeats`` runs of the first ``iters`` LM iterations replayed from its captured graph (the initial state is
reloaded before each, so every candidate times the same work); a candidate the solver rejects (a tile that does not
fit the LDS) is skipped.  Results are cached in a JSON file keyed by the batch's shape and the device name, so a
second call with the same workload reuses them without timing.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), ".autotune_cache.json")
ORDER = ("slot_tile", "lm_per_chunk", "big_n", "big_fp64", "grid")
DEFAULTS = dict(slot_tile=0, lm_per_chunk=0, big_n=0, big_fp64=0, grid=0)


def candidate_space(nlm, ncu, ns, mixed=True):
    """Candidate values per option for a batch of ``nlm`` landmarks whose reduced systems have sizes ``ns``, on a
    GPU with ``ncu`` compute units, in mixed precision or FP64.  0 is the solver's own default in every list."""
    auto_lpc = max(8, min(4096, int(round(nlm / (4.0 * ncu)))))
    ns = np.asarray(ns)
    big = [0]
    for b in (64, 96, 112, 128, 144, 160, 256, 384, 512, -1):
        # keep thresholds that move at least one window across the tiled / one-workgroup boundary relative to 192
        if (b < 0 and (ns > 192).any()) or (b > 0 and ((ns > b) != (ns > 192)).any()):
            big.append(b)
    return dict(
        slot_tile=[0, 8, 12, 20, 24],
        lm_per_chunk=[0] + sorted({max(8, auto_lpc // 2), min(4096, auto_lpc * 2), min(4096, auto_lpc * 4)}
                                  - {auto_lpc}),
        big_n=big,
        # the factor precision matters only in mixed precision; 1 (the tiled windows in FP64) only when some window
        # can be tiled, 2 (every window in FP64) always
        big_fp64=([0, 1, 2] if (len(big) > 1 or (ns > 192).any()) else [0, 2]) if mixed else [0],
        grid=[0, 2 * ncu],
    )


def search(measure, space, start=None, order=ORDER, rounds=2):
    """Coordinate descent: for each option in ``order``, try every value of ``space[option]`` with the others at
    their current best and keep the fastest; ``rounds`` passes over the options (a plan is timed once).
    ``measure(opts) -> seconds`` (None or an exception: rejected).  Returns (best options, best seconds, trials
    [(opts, seconds or None)])."""
    cur = dict(DEFAULTS if start is None else start)
    trials = []
    cache = {}

    def timed(o):
        key = tuple(sorted(o.items()))
        if key not in cache:
            try:
                t = measure(dict(o))
            except Exception:  # noqa: BLE001 -- a rejected plan is just not a candidate
                t = None
            cache[key] = t
            trials.append((dict(o), t))
        return cache[key]

    best_t = timed(cur)
    if best_t is None:
        raise RuntimeError("the solver's default plan is rejected for this batch")
    for _ in range(rounds):
        for name in order:
            for v in space.get(name, []):
                o = dict(cur, **{name: v})
                t = timed(o)
                if t is not None and t < best_t:
                    cur, best_t = o, t
    return cur, best_t, trials


def workload_key(solver, device_name):
    """The cache key of a batch: its shape (windows, rows, landmarks, keyframe counts, reduced-system sizes,
    precision) and the device."""
    h = hashlib.sha1()
    for a in (solver.win_slot0, solver.win_lm0, solver.win_obs0, solver.win_n0):
        h.update(np.ascontiguousarray(a).tobytes())
    return "%s|prec%d|%s" % (device_name, solver.prec, h.hexdigest()[:16])


def tune(solver, repeats=3, cache_path=DEFAULT_CACHE, use_cache=True, space=None, log=None, iters=1):
    """Pick the plan options for ``solver`` (a ``window.BatchSolver``); leaves the solver re-planned with them and
    returns dict(options, ms, trials, cached, seconds).  A candidate is timed over ``iters`` LM iterations from the
    initial state (the workload a caller replays: 1 for a single LM step, the step's iteration count otherwise);
    ``ms`` is per LM iteration."""
    torch = solver.torch
    dev = solver.dev
    props = torch.cuda.get_device_properties(dev)
    key = workload_key(solver, props.name) + ("" if iters == 1 else "|it%d" % iters)  # (one entry per objective)
    db = {}
    if use_cache and cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
    if use_cache and key in db:
        solver.replan(**db[key]["options"])
        return dict(db[key], cached=True, seconds=0.0)
    if space is None:
        ns = np.diff(solver.win_n0)
        space = candidate_space(solver.nlm, props.multi_processor_count, ns, mixed=solver.prec != 0)
    t0 = time.time()

    def measure(opts):
        solver.replan(**opts)
        solver.upload(1e-4)
        solver.iterate(1)  # graph capture and first replay
        ts = []
        for _ in range(repeats):
            solver.upload(1e-4)
            torch.cuda.synchronize(dev)
            a = time.perf_counter()
            solver.iterate(iters)
            torch.cuda.synchronize(dev)
            ts.append((time.perf_counter() - a) / iters)
        t = float(np.median(ts))
        if log:
            log("autotune %s: %.3f ms" % (opts, t * 1e3))
        return t

    best, best_t, trials = search(measure, space)
    solver.replan(**best)
    rec = dict(options=best, ms=best_t * 1e3,
               trials=[dict(options=o, ms=(None if t is None else t * 1e3)) for o, t in trials])
    if cache_path:
        db[key] = rec
        try:
            with open(cache_path, "w") as f:
                json.dump(db, f, indent=1)
        except OSError:
            pass
    return dict(rec, cached=False, seconds=time.time() - t0)
