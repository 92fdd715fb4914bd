"""The learned launch tuner on the GPU: collection, training by the device
kernels, choice, regret against exhaustive search, the gain over the
library's fixed heuristic, and the C ABI's measured search fallback."""
import os

import numpy as np
import pytest

from pskel_amd import _lib as L
from pskel_amd import tuner as T

pytestmark = pytest.mark.gpu


def test_measure_and_exhaustive(gpu):
    q = T.Query((300, 1100), "f32", "jacobi2d", 48)
    best, t, times = T.exhaustive(q)
    assert len(times) == len(T.candidates(q)) and t == min(times) and t > 0


def test_small_model_trains_and_chooses(gpu):
    qs = T.random_queries(6, 11, kinds=("jacobi2d", "life"))
    rows = T.collect(qs, reps=1)
    X, y = T.dataset(rows)
    m = T.CostModel(seed=3)
    try:
        loss = m.fit(X, y, epochs=150)
        assert np.isfinite(loss)
        pred = m.predict(X)
        # fitted on its training set: rank correlation with the measured times
        r = np.corrcoef(np.argsort(np.argsort(pred)), np.argsort(np.argsort(y)))[0, 1]
        assert r > 0.6, r
        for q in qs:
            c = m.choose(q)
            assert T.cfg_key(c) in {T.cfg_key(x) for x in T.candidates(q)}
        reg = T.regret(m, rows)
        assert all(x["regret"] >= 0 for x in reg)
    finally:
        m.close()


def test_genetic_search_finds_a_good_candidate(gpu):
    q = T.Query((1500, 1500), "u8", "life", 60)
    m = T.Measurer(q, reps=1)
    try:
        best, tb, _ = T.exhaustive(q, m)
        cfg, t, nmeas = T.genetic(q, m, popsize=12, generations=6, seed=5)
        assert t <= 1.5 * tb and nmeas >= 1
    finally:
        m.close()


@pytest.mark.skipif(not os.path.exists(T.DEFAULT_MODEL), reason="no shipped cost model")
def test_shipped_model_regret_on_the_workloads(gpu):
    """The shipped model on the benchmark workloads' shapes (not its training
    set): its mean regret against the exhaustive best is at most 10% and no
    query's exceeds 20%.  A query whose table shows more than 10% is timed
    again, the pick against the table's best, both with more repetitions,
    before it counts: the small grids are launch-bound, where one median of
    three can be off by 10%."""
    m = T.CostModel.load()
    try:
        qs = [T.Query((2048, 2048), "f32", "jacobi2d", 100), T.Query((8194, 8194), "u8", "life", 500),
              T.Query((4096, 4096), "f32", "diffusion9", 1000), T.Query((512, 512, 512), "f32", "jacobi3d", 1000),
              T.Query((512, 512, 512), "bf16", "jacobi3d", 1000), T.Query((700, 333), "f32", "jacobi2d", 40),
              T.Query((96, 200, 400), "bf16", "jacobi3d", 100)]
        rows = T.collect(qs, reps=3)
        regs = []
        for x in T.regret(m, rows):
            reg = x["regret"]
            if reg > 0.10:
                q = x["q"]
                t = {r["ci"]: r["ms"] for r in rows if r["q"] == q}
                cands = T.candidates(q)
                best = cands[min(t, key=t.get)]
                pick = m.choose(q)
                again = T.Measurer(q, reps=9)
                try:  # best, pick, best again (the measurer caches by configuration)
                    tb = again(best)
                    again.cache.clear()
                    tp = again(pick)
                    again.cache.clear()
                    tb = min(tb, again(best))
                finally:
                    again.close()
                reg = tp / tb - 1.0
            assert reg <= 0.20, (x["q"], x["regret"], reg)
            regs.append(reg)
        assert np.mean(regs) <= 0.10, regs
    finally:
        m.close()


@pytest.mark.skipif(not os.path.exists(T.DEFAULT_MODEL), reason="no shipped cost model")
def test_shipped_model_gain_over_heuristic_on_fresh_queries(gpu):
    """Random queries from a seed neither the training nor the held-out set
    used: the model's choices, over the exhaustive table, against the fixed
    heuristic the library falls back to (queries over 10% timed again).  The
    model's mean regret stays under 10% and its choices are at least as fast
    as the heuristic's overall (geometric mean, within 2% for timing noise)."""
    m = T.CostModel.load()
    try:
        qs = T.random_queries(12, 4242)
        rows = T.collect(qs, reps=3)
        reg = T.regret(m, rows)
        # A query whose model or heuristic pick reads more than 10% above the
        # best is timed again (pick, heuristic and best, more repetitions;
        # each keeps its faster reading): small grids sweep in a microsecond
        # or two, where one table reading is mostly launch jitter.
        for x in reg:
            if max(x["regret"], x["heuristic_regret"]) <= 0.10:
                continue
            q = x["q"]
            cands = T.candidates(q)
            by_ci = {r["ci"]: r["ms"] for r in rows if r["q"] == q}
            best_c = cands[min(by_ci, key=by_ci.get)]
            again = T.Measurer(q, reps=7)
            try:
                for key, c in (("best_ms", best_c), ("model_ms", m.choose(q)), ("heuristic_ms", T.heuristic_choice(q))):
                    again.cache.clear()
                    x[key] = min(x[key], again(c))
            finally:
                again.close()
            x["best_ms"] = min(x["best_ms"], x["model_ms"], x["heuristic_ms"])
            x["regret"] = x["model_ms"] / x["best_ms"] - 1.0
            x["heuristic_regret"] = x["heuristic_ms"] / x["best_ms"] - 1.0
        mean_model = float(np.mean([x["regret"] for x in reg]))
        gain = float(np.exp(np.mean([np.log(x["heuristic_ms"] / x["model_ms"]) for x in reg])))
        print(f"model mean regret {mean_model:.4f}, heuristic mean regret "
              f"{np.mean([x['heuristic_regret'] for x in reg]):.4f}, model gain over heuristic {gain:.4f}")
        assert mean_model <= 0.10, mean_model
        assert gain >= 0.98, gain
    finally:
        m.close()


def _grid_for(q):
    from pskel_amd import api

    g = api.Grid(q.shape, q.storage)
    g.upload_pair(T.field_for(q))
    return g


def test_library_search_exhaustive_and_genetic(gpu):
    """pskel_tuner_search (the C ABI's own fallback): exhaustive returns the
    fastest candidate it timed, the GA times at most every candidate and lands
    near it, and a repeated search is answered from the process cache."""
    import ctypes as C

    q = T.Query((600, 900), "f32", "diffusion9", 48)
    d = q.stencil().desc()
    n_cands = len(T.candidates(q))
    keys = {T.cfg_key(c) for c in T.candidates(q)}
    out, ms, nm = L.LaunchCfg(), C.c_double(), C.c_int32()
    g = _grid_for(q)
    try:
        L.check(L.lib().pskel_tuner_search(g._h, C.byref(d), 48, L.SEARCH_EXHAUSTIVE, 24, 3, C.byref(out),
                                           C.byref(ms), C.byref(nm)))
        best, t_ex = T.cfg_key(out), ms.value
        assert best in keys and nm.value == n_cands and t_ex > 0
        L.check(L.lib().pskel_tuner_search(g._h, C.byref(d), 48, L.SEARCH_EXHAUSTIVE, 24, 3, C.byref(out),
                                           C.byref(ms), C.byref(nm)))
        assert nm.value == 0 and T.cfg_key(out) == best  # cached
        L.check(L.lib().pskel_tuner_search(g._h, C.byref(d), 48, L.SEARCH_GENETIC, 24, 3, C.byref(out),
                                           C.byref(ms), C.byref(nm)))
        assert T.cfg_key(out) in keys and 1 <= nm.value <= n_cands
        assert ms.value <= 1.5 * t_ex
    finally:
        g.close()


def test_dropin_auto_without_model_searches(gpu, monkeypatch):
    """PSKEL_TUNER_MODEL=none: the drop-in's runIterativeAutoGPU and
    runIterativeGPU(N) time the candidates on the grid, start again from the
    host state and give the oracle's result."""
    import dropin
    import oracle
    from pskel_amd import inputs
    from pskel_amd import kernels as K

    monkeypatch.setenv("PSKEL_TUNER_MODEL", "none")
    a3 = inputs.jacobi3d_field(40, seed=61)
    a2 = inputs.jacobi2d_field(150, 260, seed=62)
    for method in ("auto", "default"):
        rc, got, err = dropin.run("drive_gpu_gxx", "jacobi3d", method, a3, 9)
        assert rc == 0, err
        assert np.array_equal(got.view(np.uint32), oracle.run_3d(K.jacobi3d(), a3, 9)[0].view(np.uint32))
        rc, got, err = dropin.run("drive_gpu_gxx", "jacobi2d", method, a2, 9)
        assert rc == 0, err
        assert np.array_equal(got.view(np.uint32), oracle.run_2d(K.jacobi2d(), a2, 9).view(np.uint32))


def test_auto_gpu_uses_tuner(gpu):
    """Stencil2D.runIterativeAutoGPU: tuner-chosen launch, same result."""
    import oracle
    from pskel_amd import api, inputs
    from pskel_amd import kernels as K

    a = inputs.jacobi2d_field(257, 399, seed=3)
    inp, out = api.Array2D(399, 257), api.Array2D(399, 257)
    inp.hostArray[...] = a
    out.hostArray[...] = a
    st = api.Stencil2D(inp, out, api.Mask2D(4), K.jacobi2d())
    st.runIterativeAutoGPU(17)
    assert np.array_equal(out.hostArray, oracle.run_2d(K.jacobi2d(), a, 17))


def test_model_of_another_width_is_refused(gpu):
    """A model whose input width is not the tuner's feature count is refused
    with an error, not read at the wrong stride; the heuristic alone still
    answers."""
    q = T.Query((64, 64, 64), "f32", "jacobi3d", 10)
    m = T.CostModel(sizes=(L.TUNER_NFEAT - 2, 8, 1))
    try:
        with pytest.raises(L.PSkelError):
            m.choose(q)
    finally:
        m.close()
    h = T.heuristic_choice(q)
    assert T.cfg_key(h) in {T.cfg_key(c) for c in T.candidates(q)}
