"""Autotuner of the forward projector's design and launch plan and of the
back projector's tile shape (no counterpart in the reference).

The forward projector's results do not depend on its plan and agree between
designs to FP32 summation order; its speed depends on both.  Design 1 (the
detector sweep) has plans (detector rows per lane, minimum workgroup count):
rows per lane sets how many rows one wavefront carries (1, 2 or 4: one, two
or four FP32 per lane per column read), the workgroup count how finely each
angle's detector is split into segments.  Designs 2 (depth strips), 3
(depth strips walked by angle pairs) and 4 (staged strips: groups of 8 angles
through LDS-staged column chunks) have only rows per lane; their tables are
built while timed and released again when another design wins.
The back projector's tile shape (detector rows per workgroup x angles staged
per LDS pass) does not change its results at all (each voxel sums its angles
in the same order); short row counts waste fewer lanes with 16-row tiles.
:func:`autotune` times every candidate on the geometry with HIP events and
keeps the fastest; results are cached in a JSON file keyed by
(method, nx, ny, na) so a shape is timed once.
"""
import json
import os

import numpy as np

import arec3d_amd as A

ROWS_PER_LANE = (4, 2, 1)
MIN_GROUPS = (512, 1024, 2048, 4096)
STRIP_ROWS_PER_LANE = (4, 2)
PAIR_ROWS_PER_LANE = (2, 1)
STAGED_ROWS_PER_LANE = (4, 2, 1)
DESIGNS = (1, 2, 3, 4)
BACK_TILES = ((32, 2), (16, 4), (64, 2))


def default_cache():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "arec3d_amd", "plans.json")


def _key(g):
    return f"designs1-4,back:{g.method}:{g.nx}:{g.ny}:{g.na}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def candidates(g, designs=DESIGNS):
    """(design, rows_per_lane, min_groups) plans allowed for geometry g."""
    out = []
    if 1 in designs:
        out += [(1, r, m) for r in ROWS_PER_LANE if g.ny % r == 0 for m in MIN_GROUPS]
    if 2 in designs:
        out += [(2, r, 0) for r in STRIP_ROWS_PER_LANE if g.ny % r == 0]
    if 3 in designs:
        out += [(3, r, 0) for r in PAIR_ROWS_PER_LANE if g.ny % r == 0]
    if 4 in designs:
        out += [(4, r, 0) for r in STAGED_ROWS_PER_LANE]  # any row count (unaligned: row loads)
    return out


def apply(g, plan):
    design, r, m = plan
    g.set_forward(design, r, m)


def autotune(ctx, g, cache=None, reps=3, log=None, designs=DESIGNS, back=True):
    """Pick and set the fastest forward design and plan for geometry g and
    (back=True) the fastest back-projector tile; returns (design,
    rows_per_lane, min_groups, forward ms, (tile rows, tile angles), back
    ms)."""
    path = cache or default_cache()
    table = _load(path) if path else {}
    k = _key(g)
    if k in table and (table[k][4] is not None or not back):
        d, r, m, ms, tile, bms = table[k]
        apply(g, (d, r, m))
        if back and tile is not None:
            g.set_back(*tile)
        return d, r, m, ms, tuple(tile) if tile else None, bms
    x = ctx.to_device(np.ones(g.vol_shape, np.float32))
    p = ctx.empty(g.proj_shape)
    best = None
    for plan in candidates(g, designs):
        try:
            apply(g, plan)
        except A.ArecError as e:  # e.g. the depth-strip table does not fit
            if log:
                log(f"design {plan[0]} rows/lane {plan[1]}: skipped ({e})")
            continue
        g.time_operator(0, x, p, 1)  # warm
        ms = g.time_operator(0, x, p, reps)
        if log:
            log(f"design {plan[0]} rows/lane {plan[1]} groups>={plan[2]}: {ms:.3f} ms")
        if best is None or ms < best[3]:
            best = (*plan, ms)
    apply(g, best[:3])
    tile, bms = None, None
    if back:
        p.upload(np.ones(g.proj_shape, np.float32))
        for t in BACK_TILES:
            g.set_back(*t)
            g.time_operator(1, p, x, 1)  # warm
            ms = g.time_operator(1, p, x, reps)
            if log:
                log(f"back tile {t[0]} rows x {t[1]} angles: {ms:.3f} ms")
            if bms is None or ms < bms:
                tile, bms = t, ms
        g.set_back(*tile)
    result = (*best, tile, bms)
    if path:
        table = _load(path)
        table[k] = list(result)
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass
    return result
