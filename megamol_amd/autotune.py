"""Autotuner for the traversal launch and the batched refit/rebuild policy.

Knobs (library options, see include/mmhrt.h):
  tile_w         pixel tile width; a tile is one wavefront of 64 pixels (tile_w x 64/tile_w)
  stack_lds      traversal stack entries kept in LDS per lane (deeper entries spill to HBM)
  blocks_per_cu  persistent workgroups per compute unit
  refill         ballot-refill threshold in active lanes (0 = plain persistent kernel)
  refit_period   batched series: rebuild every k-th timestep, refit in between
  leaf_pairs     nodes over two leaves stored or not (tune_tree: build + trace of a frame)

Every candidate is timed on the device (HIP events around the work, median of `reps` runs)
and the search is coordinate descent from the library defaults: each knob in turn takes its
best value while the others stay fixed, for `rounds` passes. The result is a JSON-able dict.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

SPACE = {
    "tile_w": [4, 8, 16, 32],
    "stack_lds": [8, 16, 24, 32],
    "blocks_per_cu": [2, 4, 8],
    "refill": [0, 16, 32, 48, 56],
}
DEFAULTS = {"tile_w": 8, "stack_lds": 16, "blocks_per_cu": 8, "refill": 0}


def log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def _apply(r, cfg):
    for k, v in cfg.items():
        if k in DEFAULTS:
            r.set_option(k, int(v))


def tune(measure, r, space=None, rounds=2, reps=3, start=None):
    """measure(r) -> milliseconds of one run with the options currently set on r."""
    space = space or SPACE
    cfg = dict(start or DEFAULTS)
    cache = {}

    def cost(c):
        key = tuple(sorted(c.items()))
        if key not in cache:
            _apply(r, c)
            cache[key] = float(np.median([measure(r) for _ in range(reps)]))
        return cache[key]

    best = cost(cfg)
    for rd in range(rounds):
        changed = False
        for knob, values in space.items():
            for v in values:
                cand = dict(cfg)
                cand[knob] = v
                ms = cost(cand)
                if ms < best * 0.995:
                    best, cfg, changed = ms, cand, True
            log(f"round {rd} {knob} -> {cfg[knob]} ({best:.3f} ms)")
        if not changed:
            break
    _apply(r, cfg)
    return dict(cfg), best, {str(dict(k)): v for k, v in cache.items()}


def tune_frame(r, cam, p, rounds=2, reps=3):
    """Tune the launch of one frame (render only; the LBVH is already built)."""
    from . import _lib as L  # noqa: F401

    def measure(rr):
        rr.render_async(cam, p)
        rr.sync()
        return rr.timings()["trace"]

    measure(r)  # warm-up allocations
    return tune(measure, r, rounds=rounds, reps=reps)


def tune_tree(r, cam, p, bits=30, reps=3):
    """leaf_pairs by the time of the frame's build plus its trace: a tree without leaf pairs
    takes longer to build (more nodes written), a tree with them costs the traversal a branch
    per visit, which shows in the coherent sample-lanes kernel. Returns (value, {value: ms})."""
    from . import _lib as L

    res = {}
    for v in (1, 0):
        r.set_option("leaf_pairs", v)
        ms = []
        for _ in range(reps + 1):
            r.build(L.BUILD_REBUILD, bits)
            r.render_async(cam, p)
            r.sync()
            t = r.timings()
            ms.append(t["build"] + t["trace"])
        res[v] = float(np.median(ms[1:]))
        log(f"leaf_pairs {v}: build + trace {res[v]:.3f} ms")
    best = min(res, key=res.get)
    r.set_option("leaf_pairs", best)
    return best, res


def tune_batch(r, jobs, periods=(1, 2, 4, 8, 16, 64), rounds=2, reps=2, use_graph=True):
    """Tune refit_period and the launch knobs for a batched series (set_timesteps already done)."""
    period = {"v": 1}

    def measure(rr):
        _, ms = rr.render_batch(jobs, refit_period=period["v"], use_graph=use_graph)
        return ms

    res = {}
    for k in periods:
        period["v"] = k
        res[k] = float(np.median([measure(r) for _ in range(reps)]))
        log(f"refit_period {k}: {res[k]:.2f} ms")
    period["v"] = min(res, key=res.get)
    cfg, best, cache = tune(measure, r, rounds=rounds, reps=reps)
    cfg["refit_period"] = period["v"]
    return cfg, best, {"refit_period_ms": res, "launch": cache}


def save(path, cfg, best_ms, meta=None):
    with open(path, "w") as f:
        json.dump({"config": cfg, "ms": best_ms, "meta": meta or {}, "time": time.time()}, f, indent=1)
