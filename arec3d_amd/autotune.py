"""Autotuner of the forward projector's launch plan (no counterpart in the
reference).

The forward projector's results do not depend on its plan; its speed does.
A plan is (detector rows per lane, minimum workgroup count): rows per lane
sets how many rows one wavefront carries (1, 2 or 4: one, two or four FP32
per lane per column read), the workgroup count sets how finely each angle's
detector is split into segments.  :func:`autotune` times every allowed plan
on the geometry with HIP events and keeps the fastest; results are cached in
a JSON file keyed by (method, nx, ny, na) so a shape is timed once.
"""
import json
import os

import numpy as np

ROWS_PER_LANE = (4, 2, 1)
MIN_GROUPS = (512, 1024, 2048, 4096)


def default_cache():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "arec3d_amd", "plans.json")


def _key(g):
    return f"{g.method}:{g.nx}:{g.ny}:{g.na}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(ctx, g, cache=None, reps=3, log=None):
    """Pick and set the fastest forward plan for geometry g; returns
    (rows_per_lane, min_groups, ms)."""
    path = cache or default_cache()
    table = _load(path) if path else {}
    k = _key(g)
    if k in table:
        r, m, ms = table[k]
        g.plan(r, m)
        return r, m, ms
    x = ctx.to_device(np.ones(g.vol_shape, np.float32))
    p = ctx.empty(g.proj_shape)
    best = None
    for r in ROWS_PER_LANE:
        if g.ny % r:
            continue
        for m in MIN_GROUPS:
            g.plan(r, m)
            g.time_operator(0, x, p, 1)  # warm
            ms = g.time_operator(0, x, p, reps)
            if log:
                log(f"plan rows/lane {r} groups>={m}: {ms:.3f} ms")
            if best is None or ms < best[2]:
                best = (r, m, ms)
    g.plan(best[0], best[1])
    if path:
        table = _load(path)
        table[k] = list(best)
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass
    return best
