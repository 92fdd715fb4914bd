"""Launch tuner: learned cost model, exhaustive and genetic search.

Replaces runIterativeAutoGPU's GAlib search (include/PSkelStencil.hpp:1865-2076),
which picked tile sizes and sub-iteration counts by maximising an analytic
objective and never timed anything.  Here:

* candidates: the launch configurations of the kernel family a stencil maps to
  (pskel_tuner_candidates: register-tile shapes, temporal depths, 3D marching
  tiles, plain block shapes);
* measure: device time per sweep of a candidate on a resident grid
  (pskel_tuner_measure);
* cost model: an MLP (trained by the library's own HIP kernels,
  pskel_mlp_train) predicting log(ms per sweep) from pskel_tuner_features of
  (grid dims, mask radius, storage type, iteration count, kernel family,
  candidate);
* choose: argmin of the model's prediction over the candidates;
* exhaustive / GA search over measured times: the fallback and the yardstick
  of the model's regret.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import random
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .. import _lib as L
from .. import inputs
from .. import kernels as K

DATA_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data")
SHIPPED_MODEL = os.path.join(DATA_DIR, "costmodel.bin")
# PSKEL_TUNER_MODEL names another model file ("none": no model), as it does
# for the library's own pskel_tuner_choose
_ENV_MODEL = os.environ.get("PSKEL_TUNER_MODEL")
DEFAULT_MODEL = SHIPPED_MODEL if _ENV_MODEL is None else ("" if _ENV_MODEL in ("", "none") else _ENV_MODEL)

STORAGE = {"u8": L.DTYPE_U8, "f32": L.DTYPE_F32, "bf16": L.DTYPE_BF16, "f16": L.DTYPE_F16}


@dataclass(frozen=True)
class Query:
    """One tuning problem: grid shape (h, w[, d]), storage, stencil, iterations."""
    shape: Tuple[int, ...]
    storage: str
    stencil_name: str
    iterations: int

    def stencil(self) -> K.Stencil:
        return {"jacobi2d": K.jacobi2d, "diffusion9": K.diffusion9, "life": K.life, "jacobi3d": K.jacobi3d}[
            self.stencil_name]()

    def c_query(self) -> L.TuneQuery:
        q = L.TuneQuery()
        q.ndim = len(self.shape)
        q.storage = STORAGE[self.storage]
        q.height = self.shape[0]
        q.width = self.shape[1]
        q.depth = self.shape[2] if len(self.shape) == 3 else 1
        q.iterations = self.iterations
        return q


def candidates(q: Query) -> List[L.LaunchCfg]:
    cq, d = q.c_query(), q.stencil().desc()
    buf = (L.LaunchCfg * L.TUNER_MAX_CANDIDATES)()
    n = C.c_size_t()
    L.check(L.lib().pskel_tuner_candidates(C.byref(cq), C.byref(d), buf, L.TUNER_MAX_CANDIDATES, C.byref(n)))
    return [buf[i] for i in range(min(n.value, L.TUNER_MAX_CANDIDATES))]


def features(q: Query, cfgs: Sequence[L.LaunchCfg]) -> np.ndarray:
    cq, d = q.c_query(), q.stencil().desc()
    out = np.zeros((len(cfgs), L.TUNER_NFEAT), np.float32)
    for i, c in enumerate(cfgs):
        L.check(L.lib().pskel_tuner_features(C.byref(cq), C.byref(d), C.byref(c),
                                             out[i].ctypes.data_as(C.POINTER(C.c_float))))
    return out


def cfg_dict(c: L.LaunchCfg) -> Dict:
    """Keyword arguments of api.Grid.run for a launch configuration."""
    return {"variant": "tuned" if c.variant == L.VARIANT_TUNED else "plain", "block": (c.block_x, c.block_y),
            "tile": (c.tile_x, c.tile_y, c.tile_z), "time_tile": max(1, c.time_tile)}


def cfg_key(c: L.LaunchCfg) -> Tuple:
    return (c.variant, c.block_x, c.block_y, c.tile_x, c.tile_y, c.tile_z, c.time_tile)


# ---------------------------------------------------------------------------
# measurement
# ---------------------------------------------------------------------------
def field_for(q: Query, seed: int = 0) -> np.ndarray:
    if len(q.shape) == 3:
        h, w, d = q.shape
        a = inputs.uniform(seed + 7, 0, h * w * d).reshape(q.shape)
        return a
    h, w = q.shape
    if q.stencil_name == "life":
        return (inputs.uniform(seed + 3, 0, h * w).reshape(h, w) < 0.3).astype(np.uint8)
    return inputs.uniform(seed + 5, 0, h * w).reshape(h, w)


def sweeps_for(q: Query, target_cells: float = 2.0e8) -> int:
    """Sweeps per timing: enough work (about target_cells cell-updates) to
    average out launch jitter, at least 12 (a multiple of 12 so every temporal
    depth 1..4 divides it)."""
    cells = float(np.prod(q.shape))
    n = int(max(12, min(240, target_cells / cells)))
    return n - n % 12 or 12


class Measurer:
    """Times candidates of one query on one resident grid."""

    def __init__(self, q: Query, sweeps: Optional[int] = None, reps: int = 3):
        from ..api import Grid

        self.q = q
        self.sweeps = sweeps or sweeps_for(q)
        self.reps = reps
        self.grid = Grid(q.shape, q.storage)
        self.grid.upload_pair(field_for(q))
        self.desc = q.stencil().desc()
        self.cache: Dict[Tuple, float] = {}

    def __call__(self, c: L.LaunchCfg) -> float:
        k = cfg_key(c)
        if k not in self.cache:
            ms = C.c_double()
            L.check(L.lib().pskel_tuner_measure(self.grid._h, C.byref(self.desc), self.sweeps, C.byref(c),
                                                self.reps, C.byref(ms)))
            self.cache[k] = ms.value
        return self.cache[k]

    def close(self):
        self.grid.close()


# ---------------------------------------------------------------------------
# search
# ---------------------------------------------------------------------------
def exhaustive(q: Query, measurer: Optional[Measurer] = None) -> Tuple[L.LaunchCfg, float, List[float]]:
    own = measurer is None
    m = measurer or Measurer(q)
    try:
        cands = candidates(q)
        times = [m(c) for c in cands]
        i = int(np.argmin(times))
        return cands[i], times[i], times
    finally:
        if own:
            m.close()


def genetic(q: Query, measurer: Measurer, popsize: int = 16, generations: int = 8, pmut: float = 0.01,
            pcross: float = 0.6, seed: int = 1) -> Tuple[L.LaunchCfg, float, int]:
    """Binary-genome GA over the candidate list (the reference's GAlib set-up:
    simple GA, bit-flip mutation pmut, single-point crossover pcross, fitness
    from sigma-truncation scaling) but scoring measured time per sweep.
    Returns (best config, its time, distinct candidates measured)."""
    cands = candidates(q)
    nbits = max(1, math.ceil(math.log2(len(cands))))
    rng = random.Random(seed)

    def decode(g):
        return int("".join(map(str, g)), 2) % len(cands)

    def cost(g):
        return measurer(cands[decode(g)])

    pop = [[rng.randint(0, 1) for _ in range(nbits)] for _ in range(popsize)]
    best_g = min(pop, key=cost)
    for _ in range(generations):
        costs = np.array([cost(g) for g in pop])
        score = -costs
        mu, sd = score.mean(), score.std()
        fit = np.maximum(score - (mu - 2.0 * sd), 0.0) + 1e-12  # sigma truncation
        cum = np.cumsum(fit / fit.sum())

        def pick():
            return pop[min(int(np.searchsorted(cum, rng.random())), len(pop) - 1)]

        nxt = []
        while len(nxt) < popsize:
            a, b = pick(), pick()
            a, b = list(a), list(b)
            if rng.random() < pcross and nbits > 1:
                cut = rng.randint(1, nbits - 1)
                a, b = a[:cut] + b[cut:], b[:cut] + a[cut:]
            for g in (a, b):
                for k in range(nbits):
                    if rng.random() < pmut:
                        g[k] ^= 1
            nxt.extend([a, b])
        pop = nxt[:popsize]
        cand = min(pop, key=cost)
        if cost(cand) < cost(best_g):
            best_g = cand
    return cands[decode(best_g)], cost(best_g), len(measurer.cache)


# ---------------------------------------------------------------------------
# cost model
# ---------------------------------------------------------------------------
class CostModel:
    """MLP over tuner features, predicting log(ms per sweep); trained and
    evaluated by the library's HIP kernels."""

    def __init__(self, sizes=(L.TUNER_NFEAT, 64, 64, 1), seed: int = 1, handle=None):
        self.sizes = list(sizes)
        if handle is None:
            h = C.c_void_p()
            arr = (C.c_int32 * len(self.sizes))(*self.sizes)
            L.check(L.lib().pskel_mlp_create(arr, len(self.sizes) - 1, seed, 0, C.byref(h)))
            handle = h
        self._h = handle

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL) -> "CostModel":
        h = C.c_void_p()
        L.check(L.lib().pskel_mlp_load(path.encode(), 0, C.byref(h)))
        return cls(handle=h)

    def save(self, path: str = DEFAULT_MODEL):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        L.check(L.lib().pskel_mlp_save(self._h, path.encode()))

    @staticmethod
    def normalization(X: np.ndarray, y: np.ndarray):
        """Per-feature mean and spread of a training table, and of its targets."""
        mean = X.mean(0).astype(np.float32)
        sd = X.std(0).astype(np.float32)
        sd[sd < 1e-6] = 1.0
        return mean, sd, float(y.mean()), float(max(y.std(), 1e-6))

    def fit(self, X: np.ndarray, y: np.ndarray, epochs: int = 400, batch: int = 128, lr: float = 2e-3,
            seed: int = 1) -> float:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        mean, sd, ym, ys = self.normalization(X, y)
        F = C.POINTER(C.c_float)
        L.check(L.lib().pskel_mlp_set_normalization(self._h, mean.ctypes.data_as(F), sd.ctypes.data_as(F),
                                                    ym, ys))
        loss = C.c_float()
        L.check(L.lib().pskel_mlp_train(self._h, X.ctypes.data_as(F), y.ctypes.data_as(F), len(y), epochs, batch,
                                        lr, seed, C.byref(loss)))
        return loss.value

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(len(X), np.float32)
        F = C.POINTER(C.c_float)
        L.check(L.lib().pskel_mlp_predict(self._h, X.ctypes.data_as(F), len(X), out.ctypes.data_as(F)))
        return out

    def choose(self, q: Query) -> L.LaunchCfg:
        cq, d = q.c_query(), q.stencil().desc()
        out = L.LaunchCfg()
        L.check(L.lib().pskel_tuner_choose(self._h, C.byref(cq), C.byref(d), C.byref(out)))
        return out

    def install(self):
        """Make this the model pskel_tuner_choose(NULL, ...) uses."""
        L.check(L.lib().pskel_tuner_set_default_model(self._h))

    def params(self) -> np.ndarray:
        n = C.c_size_t()
        L.check(L.lib().pskel_mlp_num_params(self._h, C.byref(n)))
        p = np.zeros(n.value, np.float32)
        L.check(L.lib().pskel_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), n.value))
        return p

    def close(self):
        if getattr(self, "_h", None):
            L.lib().pskel_tuner_set_default_model(None)
            L.lib().pskel_mlp_destroy(self._h)
            self._h = None


def pack_ensemble(members: Sequence[CostModel], mean: np.ndarray, sd: np.ndarray, y_mean: float,
                  y_sd: float) -> CostModel:
    """One MLP whose output is the mean of the members' outputs: the members'
    hidden layers side by side (block-diagonal weights), the output layer's
    weights and bias divided by the member count.  Members share the input
    and output normalisation (trained on the same table) and the layer sizes
    (in, h1, ..., 1); the result has layer sizes (in, M*h1, ..., 1), so it is an
    ordinary model file for pskel_tuner_choose."""
    big, flat = pack_params([m.params() for m in members], members[0].sizes)
    model = CostModel(sizes=big)
    F = C.POINTER(C.c_float)
    mean = np.ascontiguousarray(mean, np.float32)
    sd = np.ascontiguousarray(sd, np.float32)
    L.check(L.lib().pskel_mlp_set_normalization(model._h, mean.ctypes.data_as(F), sd.ctypes.data_as(F),
                                                float(y_mean), float(y_sd)))
    L.check(L.lib().pskel_mlp_set_params(model._h, flat.ctypes.data_as(F), flat.size))
    return model


def pack_params(ps: Sequence[np.ndarray], sizes: Sequence[int]) -> Tuple[List[int], np.ndarray]:
    """Flat parameters of the averaged ensemble (see pack_ensemble) from the
    members' flat parameters, each laid out layer by layer as W (out x in,
    row-major) then b."""
    M = len(ps)
    nl = len(sizes) - 1
    big = [sizes[0]] + [M * h for h in sizes[1:-1]] + [1]
    out = []
    offs = [0]
    for l in range(nl):
        offs.append(offs[-1] + sizes[l + 1] * sizes[l] + sizes[l + 1])
    for l in range(nl):
        W = np.zeros((big[l + 1], big[l]), np.float32)
        b = np.zeros(big[l + 1], np.float32)
        for j, p in enumerate(ps):
            w = p[offs[l]:offs[l] + sizes[l + 1] * sizes[l]].reshape(sizes[l + 1], sizes[l])
            bb = p[offs[l] + sizes[l + 1] * sizes[l]:offs[l + 1]]
            if l == nl - 1:
                W[:, j * sizes[l]:(j + 1) * sizes[l]] = w / M
                b += bb / M
            elif l == 0:
                W[j * sizes[1]:(j + 1) * sizes[1], :] = w
                b[j * sizes[1]:(j + 1) * sizes[1]] = bb
            else:
                W[j * sizes[l + 1]:(j + 1) * sizes[l + 1], j * sizes[l]:(j + 1) * sizes[l]] = w
                b[j * sizes[l + 1]:(j + 1) * sizes[l + 1]] = bb
        out += [W.ravel(), b]
    return big, np.ascontiguousarray(np.concatenate(out), np.float32)


def heuristic_choice(q: Query) -> L.LaunchCfg:
    cq, d = q.c_query(), q.stencil().desc()
    out = L.LaunchCfg()
    L.check(L.lib().pskel_tuner_heuristic(C.byref(cq), C.byref(d), C.byref(out)))
    return out


# ---------------------------------------------------------------------------
# data collection
# ---------------------------------------------------------------------------
def random_queries(n: int, seed: int, kinds=("jacobi2d", "diffusion9", "life", "jacobi3d")) -> List[Query]:
    """Random shapes: 2D sides log-uniform in [64, 4096] for half the queries
    of each 2D kind and in [1024, 8448] for the other half (at most 2^27
    cells: 8192^2 boards with their ring included), where bit-packed life
    tiles and deep temporal tiles pay; 3D sides
    log-uniform in [32, 512] for half the 3D queries and in [192, 640] for the
    other half, where the HBM-bound kernels part ways (at most 2^27 cells);
    iterations log-uniform in [10, 1000]."""
    rng = random.Random(seed)
    out = []
    n3 = 0
    n2 = {}
    for i in range(n):
        kind = kinds[i % len(kinds)]
        it = int(round(math.exp(rng.uniform(math.log(10), math.log(1000)))))
        if kind == "jacobi3d":
            lo, hi = (32, 512) if n3 % 2 == 0 else (192, 640)
            n3 += 1
            while True:
                s = tuple(int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(3))
                if np.prod(s) <= 2 ** 27:
                    break
            st = rng.choice(["f32", "bf16", "f16"])
            out.append(Query(s, st, kind, it))
        else:
            k2 = n2.get(kind, 0)
            n2[kind] = k2 + 1
            lo, hi = (64, 4096) if k2 % 2 == 0 else (1024, 8448)
            while True:
                s = tuple(int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(2))
                if np.prod(s) <= 2 ** 27:
                    break
            out.append(Query(s, "u8" if kind == "life" else "f32", kind, it))
    return out


def collect(queries: Sequence[Query], reps: int = 3, log=None) -> List[Dict]:
    """Exhaustive timings of every candidate of every query."""
    rows = []
    for qi, q in enumerate(queries):
        m = Measurer(q, reps=reps)
        try:
            cands = candidates(q)
            for ci, c in enumerate(cands):
                rows.append({"q": q, "ci": ci, "ms": m(c)})
        finally:
            m.close()
        if log:
            log(f"collected {qi + 1}/{len(queries)} {q.stencil_name} {q.shape} {q.storage}: "
                f"{len(cands)} candidates, best {min(r['ms'] for r in rows[-len(cands):]):.4f} ms/sweep")
    return rows


def rows_to_json(rows: Sequence[Dict]) -> List[Dict]:
    """Timing rows as JSON, each candidate named by its launch configuration
    (cfg_key), so a saved table stays valid when the candidate list grows."""
    out = []
    cands: Dict[Query, List[L.LaunchCfg]] = {}
    for r in rows:
        q = r["q"]
        if q not in cands:
            cands[q] = candidates(q)
        out.append({"shape": list(q.shape), "storage": q.storage, "stencil": q.stencil_name,
                    "iterations": q.iterations, "cfg": list(cfg_key(cands[q][r["ci"]])), "ms": r["ms"]})
    return out


def rows_from_json(items: Sequence[Dict]) -> Tuple[List[Dict], List[Tuple[Query, L.LaunchCfg]]]:
    """Rows of a saved table whose configuration is still a candidate, and the
    (query, candidate) pairs the table lacks."""
    rows, seen = [], {}
    for x in items:
        q = Query(tuple(x["shape"]), x["storage"], x["stencil"], x["iterations"])
        if q not in seen:
            seen[q] = ({cfg_key(c): i for i, c in enumerate(candidates(q))}, set())
        index, have = seen[q]
        k = tuple(x["cfg"])
        if k in index:
            rows.append({"q": q, "ci": index[k], "ms": x["ms"]})
            have.add(k)
    missing = []
    for q, (index, have) in seen.items():
        cands = candidates(q)
        missing += [(q, cands[i]) for k, i in index.items() if k not in have]
    return rows, missing


def dataset(rows: Sequence[Dict]) -> Tuple[np.ndarray, np.ndarray]:
    by_q: Dict[Query, List[Dict]] = {}
    for r in rows:
        by_q.setdefault(r["q"], []).append(r)
    Xs, ys = [], []
    for q, rs in by_q.items():
        cands = candidates(q)
        f = features(q, cands)
        for r in rs:
            Xs.append(f[r["ci"]])
            ys.append(math.log(r["ms"]))
    return np.stack(Xs), np.array(ys, np.float32)


def regret(model: CostModel, rows: Sequence[Dict]) -> List[Dict]:
    """Per query: time of the model's choice over the exhaustive best, from the
    measured table (no re-timing)."""
    by_q: Dict[Query, Dict[int, float]] = {}
    for r in rows:
        by_q.setdefault(r["q"], {})[r["ci"]] = r["ms"]
    out = []
    for q, t in by_q.items():
        cands = candidates(q)
        keys = [cfg_key(c) for c in cands]
        pick = model.choose(q)
        heur = heuristic_choice(q)
        best = min(t.values())
        ti = t[keys.index(cfg_key(pick))]
        th = t[keys.index(cfg_key(heur))] if cfg_key(heur) in keys else float("nan")
        plain = [i for i, c in enumerate(cands) if c.variant == L.VARIANT_PLAIN]
        out.append({"q": q, "best_ms": best, "model_ms": ti, "heuristic_ms": th, "regret": ti / best - 1.0,
                    "heuristic_regret": th / best - 1.0,
                    "plain_ms": t[plain[0]] if plain else float("nan")})
    return out


# ---------------------------------------------------------------------------
# runIterativeAutoGPU
# ---------------------------------------------------------------------------
_installed: Optional[CostModel] = None


def default_model() -> Optional[CostModel]:
    """The shipped cost model (pskel_amd/data/costmodel.bin), installed as the
    library's default on first use; None if the file is absent."""
    global _installed
    if _installed is None and os.path.exists(DEFAULT_MODEL):
        _installed = CostModel.load(DEFAULT_MODEL)
        _installed.install()
    return _installed


def choose_config(stencil: K.Stencil, shape, storage: str, iterations: int) -> Dict:
    default_model()
    q = L.TuneQuery()
    q.ndim = len(shape)
    q.storage = STORAGE[storage]
    q.height = shape[0]
    q.width = shape[1]
    q.depth = shape[2] if len(shape) == 3 else 1
    q.iterations = iterations
    d = stencil.desc()
    out = L.LaunchCfg()
    L.check(L.lib().pskel_tuner_choose(None, C.byref(q), C.byref(d), C.byref(out)))
    return cfg_dict(out)
