"""Launch choices for a batch of uneven 2D grids (pskel_batch_*).

The cost model prices a grid alone.  A small grid (64..2048 a side) swept
alone is launch-bound, a few microseconds per sweep whatever its kernel, so
the model's per-grid picks favour shallow launches; inside a batch's
persistent launches that floor is shared and only the work counts.  So the
family-wide choices are measured on the batch itself, the batch-level
counterpart of the exhaustive search that judges the model: life at 1, 3, 5 or
7 generations per launch, f32 at 1 or 3 sweeps in LDS tiles or 3 through the
register march.  Every grid's tile at the chosen depth comes from the model.
The timings run on scratch grids of the same shapes (the kernels' speed does
not depend on the values), so the caller's grids are left as they are.

The reference tunes one stencil at a time with GAlib
(include/PSkelStencil.hpp:2014-2074); it has no batched launch.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .. import _lib as L
from .. import kernels as K

MARCH = "m"      # the f32 family through the register march (3 sweeps per launch)
MARCH_SEG = 32   # rows per march segment: 16, 32 and 64 measured on 256 grids, 32 fastest

# grid kind -> (storage, tuner query kind)
KINDS = {"jacobi": ("f32", "jacobi2d"), "jacobi2d": ("f32", "jacobi2d"), "diffusion9": ("f32", "diffusion9"),
         "life": ("u8", "life")}

Shape = Tuple[int, int, str]  # (height, width, kind)


def stencil_for(kind: str) -> K.Stencil:
    return {"life": K.life, "diffusion9": K.diffusion9}.get(kind, K.jacobi2d)()


def candidates(kind: str) -> List[L.LaunchCfg]:
    """Per-grid shapes the batch kernels run: register tiles of 1, 2, 4 or 8
    rows per lane at depth 1; f32 LDS tiles of 16 or 32 rows at depth 3; life
    bit tiles at 3, 5 or 7 generations."""

    def cfg(ty, tz, k):
        c = L.LaunchCfg()
        c.variant = L.VARIANT_TUNED
        c.tile_y, c.tile_z, c.time_tile = ty, tz, k
        return c

    out = [cfg(ry, 4, 1) for ry in (1, 2, 4, 8)]
    if kind == "life":
        out += [cfg(32, 0, k) for k in (3, 5, 7)]
    else:
        out += [cfg(cy, 0, 3) for cy in (16, 32)]
    return out


def query(h: int, w: int, kind: str, sweeps: int):
    from . import Query

    storage, qkind = KINDS[kind]
    return Query((h, w), storage, qkind, sweeps)


def model_choices(shapes: Sequence[Shape], sweeps: int, life_depth=None, f32_depth=None,
                  model=None) -> Optional[List[L.LaunchCfg]]:
    """Per grid: the model's best batch shape among those of the given family
    depth (None: the model's pick over every depth; MARCH: depth 3, whose LDS
    tile stays the shape of a grid the march cannot take).  None without a
    model."""
    from . import default_model, features

    model = model or default_model()
    if model is None:
        return None
    out = []
    for h, w, kind in shapes:
        want = life_depth if kind == "life" else (3 if f32_depth == MARCH else f32_depth)
        cands = [c for c in candidates(kind) if want is None or max(1, int(c.time_tile)) == want]
        pred = model.predict(features(query(h, w, kind, sweeps), cands))
        out.append(cands[int(np.argmin(pred))])
    return out


def march_segments(shapes: Sequence[Shape], f32_depth) -> List[int]:
    """Per grid: rows per march segment (0: not marched) when the f32 family
    runs the register march: every f32 grid over 8 cells wide."""
    return [MARCH_SEG if f32_depth == MARCH and kind != "life" and w > 8 else 0 for h, w, kind in shapes]


def plan(shapes: Sequence[Shape], sweeps: int, life_depth=None, f32_depth=None, model=None) -> Optional[Dict]:
    """rows, depths and march segments of one family choice (the arguments of
    api.Batch)."""
    picks = model_choices(shapes, sweeps, life_depth, f32_depth, model)
    if picks is None:
        return None
    return {"rows": [int(c.tile_y) for c in picks], "depths": [max(1, int(c.time_tile)) for c in picks],
            "march": march_segments(shapes, f32_depth), "picks": picks}


FAMILY_CHOICES = [(None, None)] + [(ld, fd) for ld in (1, 3, 5, 7) for fd in (1, 3, MARCH)]


def search(shapes: Sequence[Shape], sweeps: int, reps: int = 2, model=None,
           timer: Optional[Callable[[Dict], float]] = None) -> Optional[Dict]:
    """Times the model's own picks and every family choice (life 1/3/5/7 x f32
    1, 3, 3 through the march) on the batch and returns the fastest plan with
    its ms per run of `sweeps` sweeps and the table of all of them (None
    without a model).  timer(plan) -> ms replaces the timing on scratch grids
    (a caller timing on its own grids)."""
    from . import field_for
    from .. import api

    plans = [(ld, fd, plan(shapes, sweeps, ld, fd, model)) for ld, fd in FAMILY_CHOICES]
    if plans[0][2] is None:
        return None
    grids, stencils = [], []
    if timer is None:
        for h, w, kind in shapes:
            q = query(h, w, kind, sweeps)
            g = api.Grid((h, w), q.storage)
            g.upload_pair(field_for(q))
            grids.append(g)
            stencils.append(stencil_for(kind))

        def timer(p):
            with api.Batch(grids, stencils, p["rows"], p["depths"], p["march"]) as b:
                b.run(sweeps, use_graph=True)  # capture and warm-up
                return min(b.run(sweeps, use_graph=True)["elapsed_ms"] for _ in range(reps))

    table, best = [], None
    try:
        for ld, fd, p in plans:
            ms = float(timer(p))
            table.append({"life_depth": "model" if ld is None else ld,
                          "f32_depth": {None: "model", MARCH: "3 (register march)"}.get(fd, fd),
                          "ms_per_run": round(ms, 3)})
            if best is None or ms < best[0]:
                best = (ms, ld, fd, p)
    finally:
        for g in grids:
            g.close()
    ms, ld, fd, p = best
    return dict(p, life_depth=ld, f32_depth=fd, ms_per_run=ms, table=table)
