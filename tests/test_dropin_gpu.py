"""The drop-in headers' GPU methods on the MI355X against the oracle, bit-exact.

drive_gpu_gxx: a g++ build -- stock kernels run through libpskel_amd.so.
drive_gpu_hip: a hipcc build -- additionally a user stencilKernel with no
descriptor runs on the device through the generic one-lane-per-cell path.
"""
import numpy as np
import pytest

import oracle
from pskel_amd import inputs
from pskel_amd import kernels as K

import dropin

pytestmark = pytest.mark.gpu

CASES = [
    ("jacobi2d", K.jacobi2d, lambda: inputs.jacobi2d_field(130, 300, seed=1)),
    ("diffusion9", K.diffusion9, lambda: inputs.blobs_field(300, seed=3, nblobs=5, sigma=20.0)[:210, :290].copy()),
    ("life", K.life, lambda: inputs.life_field(260, seed=4)),
    ("jacobi3d", K.jacobi3d, lambda: inputs.jacobi3d_field(40, seed=5)),
]


def _want(stencil, a, iters):
    if a.ndim == 3:
        return oracle.run_3d(stencil(), a, iters)[0]
    return oracle.run_2d(stencil(), a, iters)


@pytest.mark.parametrize("binary", ["drive_gpu_gxx", "drive_gpu_hip"])
@pytest.mark.parametrize("kernel,stencil,field", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("method", ["gpu", "default", "pyramid", "tiling", "auto", "hosttiling", "hostauto",
                                    "partition", "partition_small"])
def test_dropin_gpu_methods(gpu, binary, kernel, stencil, field, method):
    a = field()
    for iters in (1, 6, 11):
        rc, got, err = dropin.run(binary, kernel, method, a, iters)
        assert rc == 0, err
        assert np.array_equal(got.view(np.uint8), _want(stencil, a, iters).view(np.uint8)), (binary, method, iters)


@pytest.mark.parametrize("kernel,stencil,field", [
    ("user2d", K.jacobi2d, lambda: inputs.jacobi2d_field(130, 300, seed=11)),
    ("user3d", K.jacobi3d, lambda: inputs.jacobi3d_field(36, seed=12)),
], ids=["user2d", "user3d"])
def test_dropin_user_kernel_on_device(gpu, kernel, stencil, field):
    """A user-written stencilKernel (no descriptor) compiled by hipcc for gfx950,
    alone and split with the CPU (runIterativePartition)."""
    a = field()
    for iters in (1, 4, 9):
        for method in ("gpu", "partition"):
            rc, got, err = dropin.run("drive_gpu_hip", kernel, method, a, iters)
            assert rc == 0, err
            assert np.array_equal(got.view(np.uint8), _want(stencil, a, iters).view(np.uint8)), (iters, method)


def test_dropin_gxx_rejects_user_kernel(gpu):
    """In a g++ build a kernel with no descriptor cannot run on the device: the
    method reports PSKEL_ERR_UNSUPPORTED instead of running anything else."""
    a = inputs.jacobi2d_field(64, 64, seed=13)
    rc, _, err = dropin.run("drive_gpu_gxx", "user2d", "gpu", a, 3)
    assert rc == 10 + 5 and "unsupported" in err


@pytest.mark.parametrize("binary", ["drive_gpu_gxx", "drive_gpu_hip"])
@pytest.mark.parametrize("method", ["gpu", "pyramid", "tiling", "auto"])
def test_dropin_cloudsim_fields_on_device(gpu, binary, method):
    """CloudsimArgs: its wind fields (Args2D) are uploaded to HBM by the run
    method and attached to the grid; the library's CLOUDSIM kernels read them."""
    t, wx, wy = inputs.cloudsim_fields(150, 333, seed=9)
    for iters in (1, 6):
        rc, got, err = dropin.run(binary, "cloudsim", method, t, iters, wind=(wx, wy))
        assert rc == 0, err
        want = oracle.run_2d_cloud(t, wx, wy, 0.01, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (binary, method, iters)


def test_dropin_user_kernel_reads_fields_on_device(gpu):
    """A user Args type with Args2D members, compiled by hipcc: the kernel reads
    the fields' device copies (args.f(h, w) -> HBM)."""
    t, wx, wy = inputs.cloudsim_fields(130, 270, seed=10)
    for iters in (1, 5):
        rc, got, err = dropin.run("drive_gpu_hip", "usercloud", "gpu", t, iters, wind=(wx, wy))
        assert rc == 0, err
        want = oracle.run_2d_cloud(t, wx, wy, 0.01, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), iters



@pytest.mark.parametrize("kernel,shape", [("map1d", (1, 5000)), ("map2d", (300, 517)), ("map3d", (20, 31, 130))])
def test_dropin_map_on_device(gpu, kernel, shape):
    """A user mapKernel compiled by hipcc runs on the device over every cell;
    in a g++ build the GPU method reports PSKEL_ERR_UNSUPPORTED."""
    a = inputs.uniform(22, 0, int(np.prod(shape))).reshape(shape)
    arr = a  # map1d: a (1, n) host array, read as Array<float>(n)
    for iters in (0, 1, 6):
        rc, got, err = dropin.run("drive_gpu_hip", kernel, "gpu", arr, iters)
        assert rc == 0, err
        want = arr.copy()
        for _ in range(iters):
            want = (np.float32(0.5) * want * want + np.float32(0.125)).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (kernel, iters)
    rc, _, err = dropin.run("drive_gpu_gxx", kernel, "gpu", arr, 2)
    assert rc == 10 + 5 and "unsupported" in err


def _residuals(err):
    line = [l for l in err.splitlines() if l.startswith("residuals")][-1]
    return np.array([float(x) for x in line.split()[1:]])


@pytest.mark.parametrize("binary", ["drive_gpu_gxx", "drive_gpu_hip"])
@pytest.mark.parametrize("method", ["default", "auto", "tiling", "pyramid"])
def test_dropin_residual_interval_1(gpu, binary, method, monkeypatch):
    """setResidualInterval(1) with the tuned methods (temporal depths 2 in 3D,
    3 or 4 in 2D): every sweep's FP64 residual, and the result unchanged."""
    monkeypatch.setenv("PSKEL_DRIVE_RESIDUAL", "1")
    a = inputs.jacobi3d_field(40, seed=25)
    iters = 7
    rc, got, err = dropin.run(binary, "jacobi3d", method, a, iters)
    assert rc == 0, err
    want, want_res = oracle.run_3d(K.jacobi3d(), a, iters, residual_every=1)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), method
    np.testing.assert_allclose(_residuals(err), want_res, rtol=1e-12, atol=0)
    b = inputs.jacobi2d_field(130, 300, seed=26)
    rc, got, err = dropin.run(binary, "jacobi2d", method, b, iters)
    assert rc == 0, err
    assert np.array_equal(got.view(np.uint8), oracle.run_2d(K.jacobi2d(), b, iters).view(np.uint8)), method
    want_res = [np.linalg.norm(oracle.run_2d(K.jacobi2d(), b, n).astype(np.float64) -
                               oracle.run_2d(K.jacobi2d(), b, n - 1).astype(np.float64)) for n in range(1, iters + 1)]
    np.testing.assert_allclose(_residuals(err), want_res, rtol=1e-12, atol=0)


def test_dropin_user_kernel_residual_interval(gpu, monkeypatch):
    """A user stencilKernel (no descriptor) on the device: one residual every
    n sweeps, as for the library's kernels."""
    monkeypatch.setenv("PSKEL_DRIVE_RESIDUAL", "3")
    a = inputs.jacobi3d_field(36, seed=27)
    rc, got, err = dropin.run("drive_gpu_hip", "user3d", "gpu", a, 10)
    assert rc == 0, err
    want, want_res = oracle.run_3d(K.jacobi3d(), a, 10, residual_every=3)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    np.testing.assert_allclose(_residuals(err), want_res, rtol=1e-12, atol=0)


def test_dropin_host_staged_refuses_residuals(gpu, monkeypatch):
    """Host-staged passes reduce no residuals: an interval that was asked for
    is refused instead of silently returning none."""
    monkeypatch.setenv("PSKEL_DRIVE_RESIDUAL", "2")
    a = inputs.jacobi3d_field(30, seed=28)
    rc, _, err = dropin.run("drive_gpu_gxx", "jacobi3d", "hosttiling", a, 4)
    assert rc == 10 + 5 and "unsupported" in err
