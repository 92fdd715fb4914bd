"""The learned launch tuner on the GPU: collection, training by the device
kernels, choice, regret, and the gain over the plain kernel."""
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
def test_shipped_model_regret_and_gain(gpu):
    """The shipped model on fresh queries (not its training set): its choice is
    within 25% of the exhaustive best on every query and clearly faster than
    the plain kernel overall."""
    m = T.CostModel.load()
    try:
        qs = [T.Query((2048, 2048), "f32", "jacobi2d", 100), T.Query((8194, 8194), "u8", "life", 500),
              T.Query((4096, 4096), "f32", "diffusion9", 1000), T.Query((512, 512, 512), "f32", "jacobi3d", 1000),
              T.Query((700, 333), "f32", "jacobi2d", 40), T.Query((96, 200, 400), "bf16", "jacobi3d", 100)]
        rows = T.collect(qs, reps=2)
        reg = T.regret(m, rows)
        for x in reg:
            assert x["regret"] <= 0.25, (x["q"], x["regret"])
        gain = np.exp(np.mean([np.log(x["plain_ms"] / x["model_ms"]) for x in reg]))
        assert gain > 1.5, gain
    finally:
        m.close()


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
