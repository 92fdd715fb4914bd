"""Benchmark of the iterative stencil hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload jacobi3d|life|jacobi2d|diffusion9]

A step is one runIterativeGPU of the workload: the configured number of sweeps
over the whole grid, with both buffers already resident in HBM when the timed
region starts (uploads happen before it).  Prints ONE JSON line.

value        cell-updates/s of the headline workload (tuned kernels; the
             Jacobi-3D workloads reduce the FP64 residual every 50 sweeps
             inside the timed region)
roofline     the dominant kernel: the HBM bytes a launch must move (one read
             and one write of the grid, 2 x storage bytes per cell, whatever
             number of sweeps it fuses) over its average launch time, from HIP
             events on the stream the kernels run on; `effective` counts every
             fused sweep's read and write (passes the peak under temporal
             tiling); `traffic` is rocprofv3's measured bytes per launch
cpu_baseline the CPU oracle (OpenMP restatement of runIterativeCPU) on a
             bounded sample of the same workload (and of Game of Life), on
             every host processor OpenMP sees
workloads    the other measured forms (plain vs tuned, GoL, 16-bit storage)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# CPU baseline threads: one per core, packed (read when the OpenMP runtime of
# the CPU oracle loads, which happens after this)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Jacobi-3D 512^3 x1000 reduces the FP64 residual ||u^k - u^(k-1)||_2 on the
# device every 50 sweeps (20 per step), inside the timed region.
RES3D = 50

WORKLOADS = {
    # name: (shape, storage, stencil, sweeps per step, generator)
    "jacobi3d": dict(shape=(512, 512, 512), storage="f32", sweeps=1000, residual_every=RES3D,
                     label="jacobi3d-7pt 512^3 f32 x1000, FP64 residual every 50"),
    "jacobi3d_single": dict(shape=(512, 512, 512), storage="f32", sweeps=1000, residual_every=RES3D,
                            label="jacobi3d-7pt 512^3 f32 x1000, FP64 residual every 50, one sweep per launch"),
    "jacobi3d_pair": dict(shape=(512, 512, 512), storage="f32", sweeps=1000, residual_every=RES3D,
                          label="jacobi3d-7pt 512^3 f32 x1000, FP64 residual every 50, two sweeps per launch"),
    "jacobi3d_bf16": dict(shape=(512, 512, 512), storage="bf16", sweeps=1000, residual_every=RES3D,
                          label="jacobi3d-7pt 512^3 bf16-storage (stochastic rounding) x1000, FP64 residual every 50"),
    "jacobi3d_bf16_single": dict(shape=(512, 512, 512), storage="bf16", sweeps=1000, residual_every=RES3D,
                                 label="jacobi3d-7pt 512^3 bf16-storage (stochastic rounding) x1000, "
                                       "FP64 residual every 50, one sweep per launch"),
    "jacobi3d_f16": dict(shape=(512, 512, 512), storage="f16", sweeps=1000, residual_every=RES3D,
                         label="jacobi3d-7pt 512^3 f16-storage (stochastic rounding) x1000, FP64 residual every 50"),
    "life": dict(shape=(8194, 8194), storage="u8", sweeps=500, label="life B3/S23 8192^2 u8 x500"),
    "jacobi2d": dict(shape=(2048, 2048), storage="f32", sweeps=100, label="jacobi2d-5pt 2048^2 f32 x100"),
    "diffusion9": dict(shape=(4096, 4096), storage="f32", sweeps=1000, label="diffusion9 4096^2 f32 x1000"),
    # field arguments: the grid plus two wind fields read every sweep
    "cloudsim": dict(shape=(4096, 4096), storage="f32", sweeps=1000, label="cloudsim 4096^2 f32 + 2 wind fields x1000",
                     bytes_per_cu=16),
}

# Launch shapes of the workloads the tuner does not choose for: one and two
# sweeps per launch (the storage comparison at equal depth), CLOUDSIM (field
# arguments).  The other tuned rows take pskel_tuner_choose's pick (these
# shapes are only the base of a --tile / --block / --time-tile override).
TUNED = {
    # three fused sweeps per launch (f32): 16-row x 512-cell tiles of 2 x 4
    # waves, segment length chosen from the device's resident blocks
    "jacobi3d": dict(tile=(4, 4, 0), block=(512, 2), time_tile=3),
    # two fused sweeps per launch: 18-row x 512-cell tiles of 2 x 6 waves of 3
    # rows, the launch the tuner picks for bf16 and f16 storage at 512^3
    "jacobi3d_pair": dict(tile=(4, 3, 0), block=(768, 2), time_tile=2),
    "jacobi3d_bf16": dict(tile=(4, 3, 0), block=(768, 2), time_tile=2),
    "jacobi3d_f16": dict(tile=(4, 3, 0), block=(768, 2), time_tile=2),
    # one sweep per launch: the storage comparison through the same kernel
    "jacobi3d_single": dict(tile=(4, 1, 128)),
    "jacobi3d_bf16_single": dict(tile=(2, 2, 32)),
    # 8 generations per launch on bit-packed tiles of 32 rows per lane
    "life": dict(tile=(0, 32, 0), time_tile=8),
    "jacobi2d": dict(tile=(0, 8, 0), time_tile=4),
    "diffusion9": dict(tile=(0, 16, 0), time_tile=4),
    "cloudsim": dict(tile=(0, 1, 2)),
}


def stencil_of(name):
    from pskel_amd import kernels as K

    if name.startswith("jacobi3d"):
        return K.jacobi3d()
    return {"life": K.life, "jacobi2d": K.jacobi2d, "diffusion9": K.diffusion9, "cloudsim": K.cloudsim}[name]()


def field_of(name):
    from pskel_amd import inputs

    if name.startswith("jacobi3d"):
        return inputs.jacobi3d_field(512, seed=4)
    if name == "life":
        return inputs.life_field(8192, seed=2)
    if name == "jacobi2d":
        return inputs.jacobi2d_field(2048, 2048, seed=1)
    if name == "cloudsim":
        return inputs.cloudsim_fields(4096, 4096, seed=6)
    return inputs.blobs_field(4096, seed=3)


def store_offset_of(name):
    # 16-bit storage holds u - offset with the offset set automatically to the
    # mean of the uploaded field (about 0.5 here): over 1000 sweeps of a 512^3
    # grid most cells stay near it (the faces' pull towards 1.0 reaches a few
    # tens of cells), so the stored deviations are small and keep more
    # absolute precision
    return "auto" if name.startswith(("jacobi3d_bf16", "jacobi3d_f16")) else 0.0


def rounding_of(name):
    # 16-bit stores round stochastically: nearest rounding stalls the slow
    # relaxation once updates drop below half an ulp
    return "stochastic" if name.startswith(("jacobi3d_bf16", "jacobi3d_f16")) else "nearest"


# ---------------------------------------------------------------------------
# the tuner's choices and their regret (measured, this device)
# ---------------------------------------------------------------------------
TUNER_WORKLOADS = ("jacobi3d", "jacobi3d_bf16", "jacobi3d_f16", "life", "jacobi2d", "diffusion9")
TUNER_QUERY = {"jacobi3d": "jacobi3d", "jacobi3d_bf16": "jacobi3d", "jacobi3d_f16": "jacobi3d", "life": "life",
               "jacobi2d": "jacobi2d", "diffusion9": "diffusion9"}


def tuner_pick(name):
    """pskel_tuner_choose's launch for the workload (the shipped cost model)."""
    import ctypes as C

    from pskel_amd import _lib as L
    from pskel_amd import tuner as T

    w = WORKLOADS[name]
    q = T.Query(w["shape"], w["storage"], TUNER_QUERY[name], w["sweeps"])
    cq, d = q.c_query(), stencil_of(name).desc()
    out = L.LaunchCfg()
    L.check(L.lib().pskel_tuner_choose(None, C.byref(cq), C.byref(d), C.byref(out)))
    return out


def tuner_regret(name, pick, sweeps=None, reps=3):
    """Exhaustive search over the workload's candidates on its own grid
    (pskel_tuner_search, `sweeps` sweeps per timing, median of `reps`), and the
    tuner's pick timed the same way: regret = pick / best - 1.  By default 24
    sweeps per timing, 240 on grids of at most 4M cells (whose sweeps take a
    few microseconds: fewer would time mostly launch jitter)."""
    import ctypes as C

    from pskel_amd import _lib as L
    from pskel_amd import api
    from pskel_amd import tuner as T

    w = WORKLOADS[name]
    if sweeps is None:
        sweeps = 240 if int(np.prod(w["shape"])) <= (1 << 22) else 24
    d = stencil_of(name).desc()
    with api.Grid(w["shape"], w["storage"], store_offset=store_offset_of(name), rounding=rounding_of(name),
                  seed=17) as g:
        g.upload_pair(field_of(name))
        best, ms, nm = L.LaunchCfg(), C.c_double(), C.c_int32()
        L.check(L.lib().pskel_tuner_search(g._h, C.byref(d), w["sweeps"], L.SEARCH_EXHAUSTIVE, sweeps, reps,
                                           C.byref(best), C.byref(ms), C.byref(nm)))
        tp = C.c_double()
        L.check(L.lib().pskel_tuner_measure(g._h, C.byref(d), sweeps, C.byref(pick), reps, C.byref(tp)))
    return {"pick": T.cfg_dict(pick), "pick_ms_per_sweep": tp.value, "best": T.cfg_dict(best),
            "best_ms_per_sweep": ms.value, "candidates_timed": nm.value,
            "regret": tp.value / ms.value - 1.0}


def measure(name, variant, steps, warmup, field=None, run_kw=None):
    """Run `warmup` + `steps` runIterativeGPU steps on a resident grid.
    Returns (cell-updates/s from the host wall clock, per-step details)."""
    from pskel_amd import api

    w = WORKLOADS[name]
    kw = dict(run_kw or {})
    if variant == "tuned" and not kw:
        if name in TUNER_WORKLOADS:
            from pskel_amd import tuner as T

            kw = T.cfg_dict(tuner_pick(name))
        else:
            kw = dict(TUNED[name])
    variant = kw.pop("variant", variant)  # the tuner may pick the plain kernel
    if w.get("residual_every"):
        kw["residual_every"] = w["residual_every"]
    f = field_of(name) if field is None else field
    s = stencil_of(name)
    winds = []
    if name == "cloudsim":
        f, wx, wy = f
        winds = [api.Field(wx), api.Field(wy)]
    with api.Grid(w["shape"], w["storage"], store_offset=store_offset_of(name), rounding=rounding_of(name),
                  seed=17) as g:
        for slot, fld in enumerate(winds):
            g.set_field(slot, fld)
        g.upload_pair(f)
        for _ in range(warmup):
            g.run(s, w["sweeps"], variant=variant, **kw)
        api.synchronize()
        t0 = time.perf_counter()
        dev_ms = 0.0
        launches = 0
        residuals = 0
        for _ in range(steps):
            st = g.run(s, w["sweeps"], variant=variant, **kw)
            dev_ms += st["elapsed_ms"]
            launches += st["launches"]
            residuals += len(st["residuals"])
        api.synchronize()
        wall = time.perf_counter() - t0
    cells = int(np.prod(w["shape"]))
    cu = cells * w["sweeps"] * steps
    for fld in winds:
        fld.close()
    elem = {"f32": 4, "bf16": 2, "f16": 2, "u8": 1}[w["storage"]]
    bpc = w.get("bytes_per_cu", 2 * elem)
    return {
        "workload": w["label"],
        "variant": variant,
        "launch": {k: v for k, v in kw.items()},
        "cell_updates_per_s": cu / wall,
        "ms_per_step": wall * 1e3 / steps,
        "device_ms_per_step": dev_ms / steps,
        "launches_per_step": launches // max(steps, 1),
        "avg_launch_ms": dev_ms / max(launches, 1),
        "sweeps_per_launch": w["sweeps"] * steps / max(launches, 1),
        "algorithmic_bytes_per_cu": bpc,
        "effective_roofline_frac": (cu / wall) * bpc / (HBM_PEAK_GBS * 1e9),
        "residuals_per_step": residuals // max(steps, 1),
        "cells": cells,
        "elem_bytes": elem,
    }


def hbm_frac(r):
    """HBM fraction of a measured workload: the bytes one launch must move at
    least (one read and one write of the grid, plus one read of any field
    argument, whatever its temporal depth) over the average launch time,
    against the 8 TB/s peak.  Equal to the effective figure for one sweep per
    launch."""
    return r["algorithmic_bytes_per_cu"] * r["cells"] / (r["avg_launch_ms"] * 1e-3) / (HBM_PEAK_GBS * 1e9)


def measure_dropin(sweeps=1000, every=50):
    """Jacobi-3D 512^3 x1000 through the C++ drop-in headers: a PSkel program
    (tests/cpp/drive.cpp built by g++) calling Stencil3D::runIterativeGPU(N)
    with no block sizes (the tuner's launch), FP64 residual every 50 sweeps;
    the second of two calls is reported.  gpu_ms: device time of the sweeps
    (HIP events); wall_ms: the whole call, both host arrays uploaded and the
    result downloaded over PCIe as the reference's runIterativeGPU does."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "cpp", "bin", "drive_gpu_gxx")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, PSKEL_DRIVE_TIMES="1", PSKEL_DRIVE_RESIDUAL=str(every), PSKEL_DRIVE_REPEAT="2")
    p = subprocess.run([exe, "jacobi3d", "default", "512", "512", "512", str(sweeps), "gen:4", "gen:4", "-"],
                       capture_output=True, text=True, env=env, timeout=900)
    if p.returncode != 0:
        return {"error": p.stderr[-500:]}
    times = re.findall(r"gpu_ms ([0-9.eE+-]+) wall_ms ([0-9.eE+-]+)", p.stderr)
    res = [l for l in p.stderr.splitlines() if l.startswith("residuals")]
    gpu_ms, wall_ms = (float(x) for x in times[-1])
    cu = 512 ** 3 * sweeps
    return {"workload": f"jacobi3d-7pt 512^3 f32 x{sweeps}, FP64 residual every {every}, "
                        "Stencil3D::runIterativeGPU(N) through the C++ drop-in (g++ build)",
            "gpu_ms": gpu_ms, "wall_ms": wall_ms, "cell_updates_per_s_device": cu / (gpu_ms * 1e-3),
            "cell_updates_per_s_with_pcie": cu / (wall_ms * 1e-3),
            "residuals": len(res[-1].split()) - 1 if res else 0}


def measure_transfers(shape=(512, 512, 512), reps=3):
    """pskel_grid_upload / pskel_grid_download of a 512^3 f32 host array (the
    drop-in's per-call copies) into f32 and bf16 storage: best of `reps` after
    one warm call (which allocates the pinned staging), GB/s of host bytes."""
    from pskel_amd.api import Grid

    h = np.random.default_rng(0).standard_normal(shape, dtype=np.float32)
    out = {}
    for st in ("f32", "bf16"):
        with Grid(shape, storage=st) as g:
            up, down = [], []
            for _ in range(reps + 1):
                t0 = time.perf_counter()
                g.upload(0, h)
                t1 = time.perf_counter()
                g.download(np.float32)
                t2 = time.perf_counter()
                up.append(t1 - t0)
                down.append(t2 - t1)
            u, d = min(up[1:]), min(down[1:])
            out[st] = {"upload_ms": round(u * 1e3, 2), "download_ms": round(d * 1e3, 2),
                       "upload_gbs": round(h.nbytes / u / 1e9, 1), "download_gbs": round(h.nbytes / d / 1e9, 1)}
    out["host_bytes"] = h.nbytes
    return out


def host_search(f, s, sweeps):
    """(slab height, sweeps per pass) from pskel_tuner_host_search."""
    import ctypes as C

    from pskel_amd import _lib as L

    gd = L.GridDesc()
    gd.ndim = 3
    gd.storage = L.DTYPE_F32
    gd.height, gd.width, gd.depth = f.shape
    d = s.desc()
    tile, inner, ms = C.c_size_t(), C.c_size_t(), C.c_double()
    L.check(L.lib().pskel_tuner_host_search(C.byref(gd), f.ctypes.data, L.DTYPE_F32, C.byref(d), sweeps, 0,
                                            C.byref(tile), C.byref(inner), C.byref(ms)))
    return tile.value, inner.value


def measure_host_tiled(sweeps=64, tile=None, inner=None, pcie_duplex_gbs=48.6):
    """Jacobi-3D 512^3 f32 out of core: the grid stays in host memory, slabs of
    `tile` h-planes plus an `inner`-plane halo visit the device for `inner`
    sweeps at a time, (tile, inner) from the library's out-of-core search
    unless given.  The time is the library call's own (its host copies in and
    out included).  The bound is the PCIe traffic of the passes at the duplex
    rate one MI355X link sustains (48.6 GB/s each way, pinned buffers, H2D and
    D2H at once)."""
    from pskel_amd import api

    f = field_of("jacobi3d")
    s = stencil_of("jacobi3d")
    searched = tile is None
    if searched:
        tile, inner = host_search(f, s, sweeps)
    api.run_host_tiled(s, f, inner, tile, inner)  # warm-up
    _, st = api.run_host_tiled(s, f, sweeps, tile, inner)
    wall = st["elapsed_ms"] / 1e3
    cu = f.size * sweeps
    H = f.shape[0]
    plane = f[0].nbytes
    up = down = 0
    for done in range(0, sweeps, inner):
        halo = min(inner, sweeps - done)  # range 1
        for h0 in range(0, H, tile):
            up += (min(H, h0 + tile + halo) - max(0, h0 - halo)) * plane
            down += (min(H, h0 + tile) - h0) * plane
    bound = max(up, down) / (pcie_duplex_gbs * 1e9)
    return {
        "workload": f"jacobi3d-7pt 512^3 f32 host-staged, slabs of {tile} planes, {inner} sweeps per pass, "
                    f"{sweeps} sweeps",
        "cell_updates_per_s": cu / wall,
        "ms_per_sweep": wall * 1e3 / sweeps,
        "launches": st["launches"],
        "slab_planes_and_inner_from": "pskel_tuner_host_search" if searched else "given",
        "pcie_bytes": {"h2d": up, "d2h": down},
        "pcie_bound_ms": bound * 1e3,
        "pcie_bound_frac": bound / wall,
    }


def batch_fields(seed=5):
    """The 256-grid workload: sides log-uniform in [64, 2048] per axis, half
    Jacobi f32 (interior uniform [0,1), zero ring), half life u8 (p = 0.3)."""
    from pskel_amd import inputs

    shapes = inputs.batch_shapes(256, seed=seed)
    fields = []
    for i, (h, w, kind) in enumerate(shapes):
        if kind == "life":
            fields.append((inputs.uniform(1000 + i, 0, h * w).reshape(h, w) < np.float32(0.3)).astype(np.uint8))
        else:
            fields.append(inputs.jacobi2d_field(h, w, seed=1000 + i))
    return shapes, fields


# The batch's per-grid shapes and family-wide choices (life depth; f32 depth in
# LDS tiles or through the register march) come from pskel_amd.tuner.batch.


def batch_regret_sample(shapes, picks, sweeps, nsample=16, reps=3):
    """Regret of the model's batch choices on a sample of the grids: each
    sampled grid alone, every batch-supported shape timed (pskel_tuner_measure)."""
    import ctypes as C

    from pskel_amd import _lib as L
    from pskel_amd import api
    from pskel_amd import tuner as T
    from pskel_amd.tuner import batch as TB

    idx = np.linspace(0, len(shapes) - 1, nsample).astype(int)
    rows = []
    for i in idx:
        h, w, kind = shapes[i]
        q = TB.query(h, w, kind, sweeps)
        s = TB.stencil_for(kind)
        d = s.desc()
        with api.Grid((h, w), q.storage) as g:
            g.upload_pair(T.field_for(q))
            t = []
            for c in TB.candidates(kind):
                ms = C.c_double()
                L.check(L.lib().pskel_tuner_measure(g._h, C.byref(d), 24, C.byref(c), reps, C.byref(ms)))
                t.append(ms.value)
        keys = [T.cfg_key(c) for c in TB.candidates(kind)]
        tp = t[keys.index(T.cfg_key(picks[i]))]
        rows.append({"grid": [h, w, kind], "pick": T.cfg_dict(picks[i]), "regret": tp / min(t) - 1.0,
                     "depth1_best_over_best": min(t[:4]) / min(t)})
    return rows


def measure_batch(steps, warmup, sweeps=200, compare=True, family_depths=None):
    """256 uneven grids, `sweeps` sweeps per step: persistent kernels per
    launch, captured in one hipGraph, with every grid's launch shape from the
    cost model at the family choices measured on the batch (tuner.batch.search,
    timed here on the workload's own grids: life depth, f32 in LDS tiles or
    through the register march); beside it the model's own picks, the same
    batch at depth 1 (register tiles, the model's rows per lane) and the same
    work launched grid by grid (one runIterativeGPU per grid, the reference's
    way)."""
    from pskel_amd import api
    from pskel_amd import kernels as K
    from pskel_amd.tuner import batch as TB

    shapes, fields = batch_fields()
    grids, stencils = [], []
    for (h, w, kind), f in zip(shapes, fields):
        g = api.Grid((h, w), "u8" if kind == "life" else "f32")
        g.upload_pair(f)
        grids.append(g)
        stencils.append(K.life() if kind == "life" else K.jacobi2d())
    cells = sum(h * w for h, w, _ in shapes)
    nbytes = sum(h * w * (1 if k == "life" else 4) * 2 for h, w, k in shapes)

    def timed(rows_, depths_, march_=None, steps_=steps, warmup_=warmup):
        # the grids restart from the same fields, so every configuration times the same work
        for g, f in zip(grids, fields):
            g.upload_pair(f)
        with api.Batch(grids, stencils, rows_, depths_, march_) as b:
            for _ in range(warmup_):
                b.run(sweeps, use_graph=True)
            api.synchronize()
            t0 = time.perf_counter()
            dev = 0.0
            launches = 0
            for _ in range(steps_):
                st = b.run(sweeps, use_graph=True)
                dev += st["elapsed_ms"]
                launches = st["launches"]
            api.synchronize()
            return time.perf_counter() - t0, dev, launches, b.info()

    search = None
    march, f32_choice = None, None
    model_picks = TB.model_choices(shapes, sweeps)
    if model_picks is None:
        how = "default (depth 1, 4 rows per lane)"
        rows, depths = [4] * len(shapes), [1] * len(shapes)
        rows1 = rows
    else:
        if family_depths is None:
            how = "cost model tiles at the family choices measured on the batch"
            chosen = TB.search(shapes, sweeps,
                               timer=lambda p: timed(p["rows"], p["depths"], p["march"], 2, 1)[0] * 1e3 / 2)
            f32_choice = chosen["f32_depth"]
            search = [{"life_depth": r["life_depth"], "f32_depth": r["f32_depth"], "ms_per_step": r["ms_per_run"]}
                      for r in chosen["table"]]
        else:
            how = f"cost model tiles at life depth {family_depths[0]}, f32 depth {family_depths[1]}"
            f32_choice = family_depths[1]
            chosen = TB.plan(shapes, sweeps, *family_depths)
        rows, depths, march = chosen["rows"], chosen["depths"], chosen["march"]
        # depth 1 beside it: the model's best register tile per grid
        rows1 = TB.plan(shapes, sweeps, 1, 1)["rows"]

    try:
        wall, dev, launches, info = timed(rows, depths, march)
        if compare:
            wall1, _, launches1, _ = timed(rows1, [1] * len(shapes))
            # separate launches: every grid on its own, tuned kernel, no graph
            t1 = time.perf_counter()
            for g, s in zip(grids, stencils):
                g.run(s, sweeps, variant="tuned")
            api.synchronize()
            sep = time.perf_counter() - t1
    finally:
        for g in grids:
            g.close()
    out = {
        "workload": f"batch 256 grids (sides 64..2048) jacobi f32 + life u8 x{sweeps}, one hipGraph",
        "cell_updates_per_s": cells * sweeps * steps / wall,
        "ms_per_step": wall * 1e3 / steps,
        "device_ms_per_step": dev / steps,
        "launch_steps": launches,
        "hbm_roofline_frac_per_sweep_bytes": nbytes * sweeps * steps / wall / (HBM_PEAK_GBS * 1e9),
        "depths": {str(k): depths.count(k) for k in sorted(set(depths))},
        "f32_family": ("register march, 3 sweeps per launch, %d-row segments" % TB.MARCH_SEG
                       if f32_choice == TB.MARCH else "LDS temporal tiles" if 3 in depths else "register tiles"),
        "march_grids": sum(1 for m in (march or []) if m),
        "items": info["items"],
        "persistent_blocks": info["blocks"],
        "launch_shapes_from": how,
        "cells": cells,
    }
    if search is not None:
        out["depth_search"] = search
        model_ms = next(r["ms_per_step"] for r in search if r["life_depth"] == "model")
        out["model_picks_ms_per_step"] = model_ms
        out["speedup_vs_model_picks"] = model_ms / (wall * 1e3 / steps)
        # the model's picks against the best of the batch-level search
        out["model_regret_vs_batch_search"] = model_ms / min(r["ms_per_step"] for r in search) - 1.0
    if not compare:
        return out
    out.update({"depth1_ms_per_step": wall1 * 1e3 / steps, "depth1_launch_steps": launches1,
                "speedup_vs_depth1": wall1 / wall, "separate_launches_ms": sep * 1e3,
                "speedup_vs_separate": sep / (wall / steps)})
    if model_picks is not None:
        # the model's per-grid picks against each grid's exhaustive best, each grid timed alone
        reg = batch_regret_sample(shapes, model_picks, sweeps)
        out["regret_sample"] = {"grids": len(reg), "mean": float(np.mean([r["regret"] for r in reg])),
                                "max": float(np.max([r["regret"] for r in reg])), "per_grid": reg}
    return out


def accuracy_3d(sweeps=1000, every=50, n=512):
    """Jacobi-3D n^3 x sweeps against the FP64 CPU oracle: f32 storage, and bf16
    and f16 storage of u - offset with stochastic rounding and the automatic
    offset (the mean of the uploaded field; f32 arithmetic, tuned kernels),
    plus three bf16 controls (nearest rounding; offset 0.75; no offset).
    Reports the max absolute and relative-L2 error of each GPU result and the
    FP64 residuals the device reduces every `every` sweeps next to the
    oracle's."""
    import oracle
    from pskel_amd import api
    from pskel_amd import inputs
    from pskel_amd import kernels as K

    f64 = inputs.jacobi3d_field(n, seed=4, dtype=np.float64)
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or oracle.num_procs(), 16)
    t0 = time.perf_counter()
    want, want_res = oracle.run_3d(K.jacobi3d(), f64, sweeps, threads=threads, residual_every=every)
    oracle_s = time.perf_counter() - t0
    f32 = f64.astype(np.float32)
    norm = float(np.linalg.norm(want))
    out = {"oracle": f"FP64 runIterativeCPU restatement, {threads} threads, {oracle_s:.0f} s",
           "grid": f"{n}^3 x {sweeps}", "residual_every": every}
    for name, storage, off, rnd in (("f32", "f32", 0.0, "nearest"), ("bf16", "bf16", "auto", "stochastic"),
                                    ("f16", "f16", "auto", "stochastic"), ("bf16_nearest", "bf16", "auto", "nearest"),
                                    ("bf16_offset_0.75", "bf16", 0.75, "stochastic"),
                                    ("bf16_no_offset", "bf16", 0.0, "stochastic")):
        kw = dict(TUNED["jacobi3d" if storage == "f32" else "jacobi3d_" + storage])
        with api.Grid(f32.shape, storage, store_offset=off, rounding=rnd, seed=17) as g:
            g.upload_pair(f32)
            off = g.info()["store_offset"] if storage != "f32" else 0.0
            st = g.run(K.jacobi3d(), sweeps, variant="tuned", residual_every=every, **kw)
            got = g.download(np.float64)
        d = got - want
        res = np.array(st["residuals"])
        out[name] = {
            "storage_offset": off,
            "rounding": rnd,
            "max_abs_err": float(np.abs(d).max()),
            "rel_l2_err": float(np.linalg.norm(d) / norm),
            "residual_first_last": [float(res[0]), float(res[-1])] if len(res) else None,
            "residual_rel_diff_vs_fp64_max": float(np.max(np.abs(res - want_res) / want_res)) if len(res) else None,
        }
    out["fp64_residual_first_last"] = [float(want_res[0]), float(want_res[-1])]
    return out


# CPUs this process may run on, read at import: once OpenMP binds its threads
# (OMP_PROC_BIND=close) the main thread's own mask shrinks to one place
START_CPUS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name, seconds_target=12.0):
    """The oracle (runIterativeCPU restatement, OpenMP) on a bounded sample:
    the full grid for as many sweeps as fit about `seconds_target`, on every
    processor OpenMP sees (omp_get_num_procs), threads bound close to cores
    (OMP_PROC_BIND / OMP_PLACES are set before the OpenMP runtime loads)."""
    import oracle

    w = WORKLOADS[name]
    f = field_of(name)
    s = stencil_of(name)
    threads = oracle.num_procs()
    # calibrate with 2 sweeps (the second run: the first pays for first-touch
    # page faults), then time the sample
    run = (lambda n: oracle.run_3d(s, f, n, threads=threads)) if len(w["shape"]) == 3 else \
        (lambda n: oracle.run_2d(s, f, n, threads=threads))
    run(2)
    run(2)
    per = oracle.last_seconds() / 2
    n = max(2, min(w["sweeps"], int(seconds_target / max(per, 1e-6))))
    n -= n % 2
    run(n)
    secs = oracle.last_seconds()
    cells = int(np.prod(w["shape"]))
    kind = {"f32": "f32", "u8": "u8"}.get(w["storage"], w["storage"])
    return {
        "value": cells * n / secs,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{w['label'].split(' x')[0]}, {n} of {w['sweeps']} sweeps ({secs:.1f} s, {kind} OpenMP)",
        "cpu_model": cpu_model(),
        "omp": {"threads": threads, "proc_bind": os.environ.get("OMP_PROC_BIND"),
                "places": os.environ.get("OMP_PLACES"),
                "affinity_cpus": START_CPUS},
    }


# rocprofv3 summary of the dominant kernel of each (workload, variant) (kernel
# trace + PMC traffic), for the launch shape the tuner picks
PROFILES = {
    ("jacobi3d", "tuned"): "k3d_march2_depth3_jacobi3d_512_f32.txt",
    ("jacobi3d", "plain"): "k3d_plain_jacobi3d_512_f32.txt",
    ("jacobi3d_pair", "tuned"): "k3d_march2_jacobi3d_512_f32.txt",
    ("jacobi3d_single", "tuned"): "k3d_march_jacobi3d_512_f32.txt",
    ("jacobi3d_bf16", "tuned"): "k3d_march2_jacobi3d_512_bf16_stochastic.txt",
    ("jacobi3d_bf16_single", "tuned"): "k3d_march_jacobi3d_512_bf16_stochastic.txt",
    ("life", "tuned"): "k_life_bits_life_8192_u8.txt",
    ("jacobi2d", "tuned"): "k2d_march_jacobi2d_2048_f32.txt",
    ("diffusion9", "tuned"): "k2d_march_diffusion9_4096_f32.txt",
}


def kernel_name(name, variant, launch):
    """The kernel a workload's launch runs (tuned variant)."""
    if variant != "tuned":
        return None
    k = launch.get("time_tile", 1)
    tile = tuple(launch.get("tile", (0, 0, 0)))
    if name.startswith("jacobi3d"):
        return f"k3d_march2 (depth {k})" if k >= 2 else "k3d_march"
    if name == "life":
        return f"k_life_bits (depth {k})" if k >= 2 else "k_life_tuned"
    if name in ("jacobi2d", "diffusion9"):
        if k >= 2:
            return f"k2d_march (depth {k})" if tile[0] == 4 else f"k2d_ttile (depth {k})"
        return "k2d_tuned"
    return None


def traffic_from_profiles(name, variant="tuned"):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factor)."""
    if (name, variant) not in PROFILES:
        return None
    p = os.path.join(ROOT, "profiles", PROFILES[(name, variant)])
    if not os.path.exists(p):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read(), re.M)
    return float(m.group(1)) if m else None


def copy_calibration(mib, reps):
    """Plain 16-byte-vector copy of `mib` MiB, `reps` times: the yardstick that
    converts FETCH_SIZE / WRITE_SIZE readings to bytes, and the achievable
    copy bandwidth of this device."""
    import ctypes as C

    from pskel_amd import _lib as L

    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    L.check(L.lib().pskel_device_alloc(n, C.byref(a)))
    L.check(L.lib().pskel_device_alloc(n, C.byref(b)))
    try:
        ms = C.c_double()
        L.check(L.lib().pskel_device_copy(b, a, n, None, C.byref(ms)))  # warm
        times = []
        for _ in range(reps):
            L.check(L.lib().pskel_device_copy(b, a, n, None, C.byref(ms)))
            times.append(ms.value)
    finally:
        L.lib().pskel_device_free(a)
        L.lib().pskel_device_free(b)
    best = min(times)
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "reps": reps,
                      "best_ms": best, "copy_GBps": 2 * n / (best * 1e-3) / 1e9}))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="jacobi3d", choices=sorted(WORKLOADS))
    ap.add_argument("--variant", default="tuned", choices=["tuned", "plain"])
    ap.add_argument("--tile", default="", help="tuned tile x,y,z")
    ap.add_argument("--block", default="", help="tuned block x,y")
    ap.add_argument("--time-tile", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64-oracle error report")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run the plain copy kernel on two buffers of this size (PMC calibration)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 256-grid batch with the model's launch shapes (no comparisons)")
    ap.add_argument("--batch-depths", default="",
                    help="with --batch-only: life,f32 family depths, f32 'm' for the register march "
                         "(no depth search; a profiling target)")
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a grid would shard "
              "into h-slabs with one-plane halo exchanges); run with --gpus 1", file=sys.stderr)
        return 2

    from pskel_amd import _lib, api

    if _lib.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    if args.copy_mib:
        return copy_calibration(args.copy_mib, max(args.steps, 1))
    if args.batch_only:
        # the batch alone (a profiling target): the depth search, or the given family depths
        from pskel_amd.tuner import batch as TB

        depths = None
        if args.batch_depths:
            life_d, f32_d = args.batch_depths.split(",")
            depths = (int(life_d), f32_d if f32_d == TB.MARCH else int(f32_d))
        print(json.dumps(measure_batch(max(args.steps, 1), args.warmup, compare=False, family_depths=depths)))
        return 0

    run_kw = None
    if args.tile or args.time_tile or args.block:
        run_kw = dict(TUNED[args.workload])
        if args.tile:
            run_kw["tile"] = tuple(int(x) for x in args.tile.split(","))
        if args.block:
            run_kw["block"] = tuple(int(x) for x in args.block.split(","))
        if args.time_tile:
            run_kw["time_tile"] = args.time_tile

    head = measure(args.workload, args.variant, args.steps, args.warmup, run_kw=run_kw)
    head_frac = hbm_frac(head)
    print(f"# headline {head['workload']} {head['variant']}: {head['cell_updates_per_s']:.4g} CU/s, "
          f"{head_frac*100:.1f}% of HBM ({head['effective_roofline_frac']*100:.0f}% effective), "
          f"{head['residuals_per_step']} residuals per step", file=sys.stderr, flush=True)

    extra = {}
    if not args.no_extra:
        plan = [("jacobi3d", "plain"), ("jacobi3d_single", "tuned"), ("jacobi3d_pair", "tuned"), ("jacobi3d_bf16", "tuned"),
                ("jacobi3d_bf16_single", "tuned"), ("jacobi3d_f16", "tuned"),
                ("life", "tuned"), ("life", "plain"), ("jacobi2d", "tuned"), ("jacobi2d", "plain"),
                ("diffusion9", "tuned"), ("diffusion9", "plain"), ("cloudsim", "tuned"), ("cloudsim", "plain")]
        for nm, var in plan:
            if (nm, var) == (args.workload, args.variant):
                continue
            r = measure(nm, var, max(1, min(args.steps, 2)), 1)
            extra[f"{nm}:{var}"] = {k: r[k] for k in ("workload", "cell_updates_per_s", "ms_per_step",
                                                       "avg_launch_ms", "sweeps_per_launch",
                                                       "effective_roofline_frac", "residuals_per_step", "launch")}
            extra[f"{nm}:{var}"]["hbm_roofline_frac"] = hbm_frac(r)
            tr = traffic_from_profiles(nm, var)
            if tr:  # HBM bytes per launch the committed profile measured
                extra[f"{nm}:{var}"]["traffic_bytes_per_launch"] = tr
                extra[f"{nm}:{var}"]["traffic_gbs"] = round(tr / (r["avg_launch_ms"] * 1e-3) / 1e9, 1)
            print(f"# {nm} {var}: {r['cell_updates_per_s']:.4g} CU/s, {hbm_frac(r)*100:.1f}% of HBM "
                  f"({r['effective_roofline_frac']*100:.0f}% effective)", file=sys.stderr, flush=True)
        rh = measure_host_tiled()
        extra["jacobi3d:host_staged"] = rh
        print(f"# jacobi3d host-staged: {rh['cell_updates_per_s']:.4g} CU/s, {rh['ms_per_sweep']:.2f} ms/sweep",
              file=sys.stderr, flush=True)
        rb = measure_batch(max(1, min(args.steps, 3)), 1)
        extra["batch256:graph"] = rb
        print(f"# batch256: {rb['cell_updates_per_s']:.4g} CU/s, {rb['ms_per_step']:.2f} ms/step, "
              f"{rb['speedup_vs_depth1']:.2f}x vs depth 1, {rb['speedup_vs_separate']:.1f}x vs separate launches, "
              f"depths {rb['depths']}, f32 {rb['f32_family']}, model's own picks {rb.get('model_picks_ms_per_step', float('nan')):.2f} ms/step "
              f"(regret {rb.get('model_regret_vs_batch_search', float('nan')):.3f} vs the batch search), "
              f"per-grid sample regret {rb.get('regret_sample', {}).get('mean', float('nan')):.3f}",
              file=sys.stderr, flush=True)
        rd = measure_dropin()
        extra["jacobi3d:dropin"] = rd
        if rd and "gpu_ms" in rd:
            print(f"# jacobi3d drop-in runIterativeGPU(1000): {rd['cell_updates_per_s_device']:.4g} CU/s device, "
                  f"{rd['cell_updates_per_s_with_pcie']:.4g} CU/s with PCIe", file=sys.stderr, flush=True)
        rt = measure_transfers()
        extra["jacobi3d:host_transfers"] = rt
        print("# host transfers 512^3 f32 host array: " + ", ".join(
            f"{st} up {rt[st]['upload_gbs']} GB/s down {rt[st]['download_gbs']} GB/s" for st in ("f32", "bf16")),
              file=sys.stderr, flush=True)
        # the tuner: its pick for each workload against exhaustive search
        tuner = {}
        for nm in TUNER_WORKLOADS:
            tuner[nm] = tuner_regret(nm, tuner_pick(nm))
            print(f"# tuner {nm}: pick {tuner[nm]['pick']} regret {tuner[nm]['regret']:.3f} "
                  f"(best {tuner[nm]['best']})", file=sys.stderr, flush=True)
        extra["tuner_regret_vs_exhaustive"] = tuner

    # HBM bytes one launch must move: one read and one write of the grid (plus
    # field reads), whatever the number of sweeps it fuses
    launch_ms = head["avg_launch_ms"]
    bytes_per_launch = head["algorithmic_bytes_per_cu"] * head["cells"]
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    eff_bytes = bytes_per_launch * head["sweeps_per_launch"]
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(args.workload, args.variant),
        "profile": (f"profiles/{PROFILES[(args.workload, args.variant)]}"
                    if (args.workload, args.variant) in PROFILES else None),
        "kernel": kernel_name(args.workload, args.variant, head["launch"]),
        "sweeps_per_launch": round(head["sweeps_per_launch"], 3),
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "algorithmic_bytes": (f"{head['algorithmic_bytes_per_cu']} B per cell (one read + one write of the grid) "
                              f"x {head['cells']} cells per launch, at any temporal depth"),
        "avg_launch_ms": round(launch_ms, 5),
        # the per-sweep figure: every sweep's read + write counted, as if each
        # sweep went to HBM; under temporal tiling it passes the peak
        "effective": {"achieved": round(eff_bytes / (launch_ms * 1e-3) / 1e9, 1),
                      "frac": round(eff_bytes / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "bytes_per_launch": eff_bytes},
    }
    # the HBM bytes a launch actually moved (rocprofv3 FETCH_SIZE + WRITE_SIZE)
    if roofline["traffic"]:
        roofline["traffic_gbs"] = round(roofline["traffic"] / (launch_ms * 1e-3) / 1e9, 1)
        roofline["traffic_frac"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 4)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.workload)
        print(f"# cpu baseline {args.workload}: {cpu['value']:.4g} CU/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)
        if args.workload != "life" and not args.no_extra:
            # the metric names Game of Life too
            cpu["life"] = cpu_baseline("life")
            print(f"# cpu baseline life: {cpu['life']['value']:.4g} CU/s", file=sys.stderr, flush=True)
    acc = None
    if not args.no_accuracy and args.workload.startswith("jacobi3d") and not args.no_extra:
        acc = accuracy_3d()
        print(f"# accuracy vs FP64: " + ", ".join(f"{k} rel-L2 {v['rel_l2_err']:.2e}" for k, v in acc.items()
                                                  if isinstance(v, dict)), file=sys.stderr, flush=True)

    info = api.device_info(0)
    line = {
        "metric": "cell-updates/s (+% of HBM roofline), Jacobi3D & GoL, tuned vs plain",
        "value": head["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "bf16": "bf16-storage/f32", "f16": "f16-storage/f32", "u8": "u8"}[
            WORKLOADS[args.workload]["storage"]],
        "data": "synthetic (counter-based splitmix64 fields)",
        "config": {"workload": head["workload"], "variant": head["variant"], "launch": head["launch"],
                   "hbm_roofline_frac": round(head_frac, 4),
                   "residuals_per_step": head["residuals_per_step"], "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "workloads": extra,
        "accuracy_vs_fp64": acc,
        # bf16 storage (stochastic rounding) against f32, each through its
        # fastest path (bf16: two fused sweeps per launch; f32: three)
        "bf16_storage_speedup_vs_f32": (round(extra["jacobi3d_bf16:tuned"]["cell_updates_per_s"] /
                                              head["cell_updates_per_s"], 3)
                                        if (args.workload, args.variant) == ("jacobi3d", "tuned")
                                        and "jacobi3d_bf16:tuned" in extra else None),
        # storage alone: bf16 and f32 through the same fused-pair launch
        "bf16_storage_speedup_vs_f32_same_launch": (
            round(extra["jacobi3d_bf16:tuned"]["cell_updates_per_s"] /
                  extra["jacobi3d_pair:tuned"]["cell_updates_per_s"], 3)
            if "jacobi3d_bf16:tuned" in extra and "jacobi3d_pair:tuned" in extra
            and all(extra["jacobi3d_bf16:tuned"]["launch"].get(k) == extra["jacobi3d_pair:tuned"]["launch"].get(k)
                    for k in ("tile", "block", "time_tile")) else None),
        # storage alone: bf16 and f32 through the same one-sweep-per-launch kernel
        "bf16_storage_speedup_vs_f32_one_sweep_per_launch": (
            round(extra["jacobi3d_bf16_single:tuned"]["cell_updates_per_s"] /
                  extra["jacobi3d_single:tuned"]["cell_updates_per_s"], 3)
            if "jacobi3d_bf16_single:tuned" in extra and "jacobi3d_single:tuned" in extra else None),
        # fused sweeps against one sweep per launch, f32: the pair, and the
        # headline's launch
        "fused_pair_speedup_vs_single_sweep": (round(extra["jacobi3d_pair:tuned"]["cell_updates_per_s"] /
                                                     extra["jacobi3d_single:tuned"]["cell_updates_per_s"], 3)
                                               if "jacobi3d_pair:tuned" in extra
                                               and "jacobi3d_single:tuned" in extra else None),
        "headline_speedup_vs_single_sweep": (round(head["cell_updates_per_s"] /
                                                   extra["jacobi3d_single:tuned"]["cell_updates_per_s"], 3)
                                             if (args.workload, args.variant) == ("jacobi3d", "tuned")
                                             and "jacobi3d_single:tuned" in extra else None),
        "device": {"arch": info["name"], "cus": info["cus"]},
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
