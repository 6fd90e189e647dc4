"""Autotuner for operator tile shapes.

The kernels have a few compiled tile shapes (keys per thread of the radix
sort's onesweep tile and its digit width, merged items per find_bounds and
expansion tile, rows per thread of the select pass);
which is fastest depends on the key and value widths, the row counts and
the key distribution.  tune() times a whole plan (a callable that enqueues
it on the context's stream) under each candidate of one knob at a time,
keeps the fastest, and moves to the next knob (coordinate descent).  The
choices live in the context, so a plan captured into a HIP graph afterwards
replays them.  Results never depend on the choice; only time does.
"""
from __future__ import annotations

import json
import os
import time
from typing import Callable, Dict, Iterable, Optional, Tuple

import numpy as np

from . import rfx as R

# knob -> candidates, the built-in default (0) first
KNOBS: Dict[str, Tuple[int, ...]] = {
    "sort.ipt": (0, 8, 12, 16),
    "sort.digit": (0, 8, 9, 10),
    "join.fb_tile": (0, 2048),
    "join.ex_tile": (0, 2048),
    "select.vt": (0, 8, 16, 24, 32),
}


def time_plan(ctx: R.Context, run: Callable[[], None], reps: int = 3) -> float:
    """Median milliseconds of run() on ctx's stream, after one warm run."""
    run()
    ctx.synchronize()
    t = R.Timer(ctx)
    times = []
    for _ in range(reps):
        t.begin()
        run()
        times.append(t.end())
    t.close()
    return float(np.median(times))


def tune(ctx: R.Context, run: Callable[[], None], knobs: Optional[Dict[str, Iterable[int]]] = None,
         reps: int = 3, cache_path: Optional[str] = None, key: Optional[str] = None,
         margin: float = 0.01) -> dict:
    """Pick each knob's fastest value for run(); returns {knob: value} plus
    the timings.  A knob keeps its first candidate (the built-in default)
    unless another is faster by more than `margin` (a fraction of its time),
    so timing noise does not move it.  With cache_path and key, a stored
    choice is applied without timing, and a new one is stored."""
    knobs = dict(KNOBS if knobs is None else knobs)
    if cache_path and key:
        cache = _load(cache_path)
        if key in cache:
            for k, v in cache[key]["choice"].items():
                ctx.tune(k, v)
            return {"choice": dict(cache[key]["choice"]), "timings_ms": {}, "cached": True}
    t0 = time.perf_counter()
    timings: Dict[str, Dict[int, float]] = {}
    choice: Dict[str, int] = {}
    for knob, values in knobs.items():
        res = {}
        values = [int(v) for v in values]
        for v in values:
            ctx.tune(knob, v)
            res[v] = time_plan(ctx, run, reps)
        best = min(res, key=res.get)
        if res[best] > res[values[0]] * (1.0 - margin):
            best = values[0]
        ctx.tune(knob, best)
        choice[knob] = best
        timings[knob] = res
    out = {"choice": choice, "timings_ms": timings, "cached": False,
           "tuning_s": time.perf_counter() - t0}
    if cache_path and key:
        cache = _load(cache_path)
        cache[key] = {"choice": choice}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return out


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}
