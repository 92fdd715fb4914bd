"""Device residency semantics: resume, graphs, ping-pong parity, errors."""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K
from pskel_amd._lib import PSkelError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", [{"variant": "plain"}, {"variant": "tuned"}, {"variant": "tuned", "time_tile": 3},
                                {"variant": "tuned", "time_tile": 4}])
def test_resume_equals_single_run(gpu, kw):
    """run(k); run(n-k) is bit-identical to run(n) (checkpoint/resume)."""
    a = inputs.jacobi2d_field(200, 300, seed=1)
    b = a.copy()
    b[0, :] = 1.0  # distinct rings make parity mistakes visible
    with api.Grid(a.shape, "f32") as g:
        g.upload(0, a)
        g.upload(1, b)
        g.run(K.jacobi2d(), 7, **kw)
        g.run(K.jacobi2d(), 6, **kw)
        got = g.download()
    assert np.array_equal(got, oracle.run_2d(K.jacobi2d(), a, 13, out=b))


@pytest.mark.parametrize("kw", [{"variant": "plain"}, {"variant": "tuned"}, {"variant": "tuned", "time_tile": 2}])
def test_graph_replay_matches_eager(gpu, kw):
    """A captured hipGraph replays the same launches: every replay (with eager
    work and synchronisation between replays) equals the oracle."""
    a = inputs.jacobi2d_field(160, 260, seed=2)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        for rep in range(1, 5):
            g.run(K.jacobi2d(), 10, use_graph=True, **kw)
            g.run(K.jacobi2d(), 1, variant="plain")  # eager work between replays
            api.synchronize()
            got = g.download()
            assert np.array_equal(got, oracle.run_2d(K.jacobi2d(), a, 11 * rep)), rep


def test_graph_residuals_3d(gpu):
    a = inputs.jacobi3d_field(48, seed=3)
    want_u, want_r = oracle.run_3d(K.jacobi3d(), a, 20, residual_every=5)
    with api.Grid(a.shape, "f32") as g:
        for rep in range(3):
            g.upload_pair(a)
            st = g.run(K.jacobi3d(), 20, variant="tuned", residual_every=5, use_graph=True)
            np.testing.assert_allclose(st["residuals"], want_r, rtol=1e-12)
            assert np.array_equal(g.download(), want_u), rep


def test_previous_state_and_residual(gpu):
    a = inputs.jacobi3d_field(32, seed=4)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        g.run(K.jacobi3d(), 3)
        cur, prev = g.download(), g.download_previous()
        r = g.residual()
    want2, _ = oracle.run_3d(K.jacobi3d(), a, 2)
    want3, _ = oracle.run_3d(K.jacobi3d(), a, 3)
    assert np.array_equal(cur, want3) and np.array_equal(prev, want2)
    assert abs(r - np.linalg.norm(want3.astype(np.float64) - want2)) <= 1e-12 * max(r, 1e-30)


def test_zero_iterations_is_identity(gpu):
    a = inputs.jacobi2d_field(30, 40, seed=5)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a, np.zeros_like(a))
        st = g.run(K.jacobi2d(), 0)
        assert st["launches"] == 0
        assert np.array_equal(g.download(), a)


def test_pskel_mirror_leaves_input(gpu):
    """Stencil2D.runIterativeGPU: input untouched, result in output
    (include/PSkelStencil.hpp:1078-1090)."""
    inp = api.Array2D(50, 40)
    out = api.Array2D(50, 40)
    inp.hostArray[...] = inputs.jacobi2d_field(40, 50, seed=6)
    out.hostArray[...] = inp.hostArray
    before = inp.hostArray.copy()
    m = api.Mask2D(4)
    for i, (dh, dw) in enumerate(K.J5_OFFSETS):
        m.set(i, dh, dw)
    st = api.Stencil2D(inp, out, m, api.StencilArgs(K.SCALED_SUM, coeff=0.25))
    st.runIterativeGPU(9, 32, 4)
    assert np.array_equal(inp.hostArray, before)
    assert np.array_equal(out.hostArray, oracle.run_2d(K.jacobi2d(), before, 9))
    out.hostArray[...] = before
    st.runIterativeTilingGPU(9, 50, 40, 1, innerIterations=3)
    assert np.array_equal(out.hostArray, oracle.run_2d(K.jacobi2d(), before, 9))


def test_errors_are_reported(gpu):
    a = inputs.jacobi2d_field(30, 40, seed=7)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        with pytest.raises(PSkelError):
            g.run(K.life(), 1)  # life on a float grid
        bad = K.jacobi2d()
        bad.radius = 0
        bad.offsets = [(-2, 0), (1, 0), (0, -1), (0, 1)]
        bad.radius = 1  # range smaller than the farthest offset
        with pytest.raises(PSkelError):
            g.run(bad, 1)
        with pytest.raises(PSkelError):
            g.run(K.jacobi2d(), 1, variant="plain", block=(64, 64))
        with pytest.raises(PSkelError):
            g.run(K.jacobi2d(), 1, variant="tuned", time_tile=9)
