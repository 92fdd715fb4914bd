"""GPU parity of field arguments and the CLOUDSIM stencil against the CPU oracle.

Fields (Args / Args2D / Args3D) live in HBM in a reversed layout; their host
round trip must be exact.  CLOUDSIM sweeps (plain and tuned) are bit-exact
against the oracle: same expression, same order, no contraction.
"""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K
from pskel_amd._lib import PSkelError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(17,), (5, 7), (300, 1000), (3, 4, 5), (6, 33, 260)])
def test_field_round_trip(gpu, shape):
    host = inputs.uniform(9, 0, int(np.prod(shape))).reshape(shape)
    with api.Field(host) as f:
        assert np.array_equal(f.download(), host)
        host2 = host * 2
        f.upload(host2)
        assert np.array_equal(f.download(), host2)


def _run(shape, steps, variant, seed=1, dt=0.01, **kw):
    t, wx, wy = inputs.cloudsim_fields(*shape, seed=seed)
    with api.Grid(shape, "f32") as g, api.Field(wx) as fx, api.Field(wy) as fy:
        g.set_field(0, fx)
        g.set_field(1, fy)
        g.upload_pair(t)
        g.run(K.cloudsim(dt), steps, variant=variant, **kw)
        got = g.download()
    return got, oracle.run_2d_cloud(t, wx, wy, dt, steps)


@pytest.mark.parametrize("shape", [(3, 3), (4, 9), (33, 70), (130, 257), (257, 1031), (512, 512)])
@pytest.mark.parametrize("variant", ["plain", "tuned"])
def test_cloud_parity(gpu, shape, variant):
    for steps in (0, 1, 6):
        got, want = _run(shape, steps, variant)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (shape, variant, steps)


@pytest.mark.parametrize("tile", [(0, 1, 1), (0, 2, 4), (0, 4, 2), (0, 8, 1), (0, 8, 4)])
def test_cloud_tuned_tiles(gpu, tile):
    got, want = _run((300, 777), 5, "tuned", seed=2, tile=tile)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), tile


def test_cloud_plain_blocks_and_graph(gpu):
    for kw in ({"block": (32, 8)}, {"block": (256, 1)}, {"use_graph": True}):
        got, want = _run((99, 201), 7, "plain", seed=3, **kw)
        assert np.array_equal(got, want), kw
    got, want = _run((99, 201), 7, "tuned", seed=3, use_graph=True)
    assert np.array_equal(got, want)


def test_cloud_field_reupload_between_runs(gpu):
    """A field re-uploaded between runs is what the next run reads (also when
    the launch sequence is a cached graph)."""
    shape = (40, 64)
    t, wx, wy = inputs.cloudsim_fields(*shape, seed=4)
    s = K.cloudsim(0.02)
    with api.Grid(shape, "f32") as g, api.Field(wx) as fx, api.Field(wy) as fy:
        g.set_field(0, fx)
        g.set_field(1, fy)
        g.upload_pair(t)
        g.run(s, 3, variant="tuned", use_graph=True)
        mid = oracle.run_2d_cloud(t, wx, wy, 0.02, 3)
        assert np.array_equal(g.download(), mid)
        wx2 = -wx
        fx.upload(wx2)
        g.run(s, 3, variant="tuned", use_graph=True)
        assert np.array_equal(g.download(), oracle.run_2d_cloud(mid, wx2, wy, 0.02, 3))


def test_cloud_residuals(gpu):
    shape = (128, 300)
    t, wx, wy = inputs.cloudsim_fields(*shape, seed=5)
    with api.Grid(shape, "f32") as g, api.Field(wx) as fx, api.Field(wy) as fy:
        g.set_field(0, fx)
        g.set_field(1, fy)
        g.upload_pair(t)
        st = g.run(K.cloudsim(0.01), 10, variant="tuned", residual_every=5)
    prev = oracle.run_2d_cloud(t, wx, wy, 0.01, 4)
    cur = oracle.run_2d_cloud(t, wx, wy, 0.01, 5)
    want5 = np.sqrt(np.sum((cur.astype(np.float64) - prev.astype(np.float64)) ** 2))
    assert st["residuals"][0] == pytest.approx(want5, rel=1e-12)


def test_cloud_rejects_bad_setups(gpu):
    shape = (20, 30)
    t, wx, wy = inputs.cloudsim_fields(*shape, seed=6)
    with api.Grid(shape, "f32") as g, api.Field(wx) as fx, api.Field(wy[:10]) as small:
        g.upload_pair(t)
        with pytest.raises(PSkelError):
            g.run(K.cloudsim(), 1)              # no fields attached
        g.set_field(0, fx)
        g.set_field(1, small)
        with pytest.raises(PSkelError):
            g.run(K.cloudsim(), 1)              # field too small for the grid
        with pytest.raises(PSkelError):
            g.set_field(4, fx)
    with api.Grid(shape, "u8") as g8:
        with pytest.raises(PSkelError):
            g8.run(K.cloudsim(), 1)
