"""The drop-in headers' CPU methods (runIterativeCPU / runIterativeSequential,
compiled by g++ from include/PSkel.h) against the oracle, bit-exact."""
import numpy as np
import pytest

import oracle
from pskel_amd import inputs
from pskel_amd import kernels as K

import dropin

CASES = [
    ("jacobi2d", K.jacobi2d, lambda: inputs.jacobi2d_field(37, 53, seed=1)),
    ("user2d", K.jacobi2d, lambda: inputs.jacobi2d_field(37, 53, seed=2)),
    ("diffusion9", K.diffusion9, lambda: inputs.blobs_field(64, seed=3, nblobs=3, sigma=6.0)[:41, :57].copy()),
    ("life", K.life, lambda: inputs.life_field(45, seed=4)),
    ("jacobi3d", K.jacobi3d, lambda: inputs.jacobi3d_field(19, seed=5)),
    ("user3d", K.jacobi3d, lambda: inputs.jacobi3d_field(17, seed=6)),
]


@pytest.mark.parametrize("kernel,stencil,field", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("method", ["cpu", "seq"])
def test_dropin_cpu_matches_oracle(kernel, stencil, field, method):
    a = field()
    for iters in (0, 1, 4, 7):
        rc, got, err = dropin.run("drive_cpu", kernel, method, a, iters)
        assert rc == 0, err
        if a.ndim == 3:
            want, _ = oracle.run_3d(stencil(), a, iters)
        else:
            want = oracle.run_2d(stencil(), a, iters)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (kernel, method, iters)


def test_dropin_cpu_distinct_output_ring():
    """The reference's ping-pong: the result keeps the ring of the buffer it ends in."""
    a = inputs.jacobi2d_field(20, 30, seed=7)
    b = a.copy()
    b[0, :] = 9.0
    for iters in (1, 2, 3):
        rc, got, err = dropin.run("drive_cpu", "jacobi2d", "cpu", a, iters, out0=b)
        assert rc == 0, err
        assert np.array_equal(got, oracle.run_2d(K.jacobi2d(), a, iters, out=b))


def test_dropin_gxx_gpu_build_reports_missing_device_or_descriptor():
    """Without a device the g++ GPU build fails cleanly with a status (no crash);
    the user kernel without a descriptor is unsupported in a g++ build."""
    a = inputs.jacobi2d_field(16, 16, seed=8)
    rc, _, err = dropin.run("drive_gpu_gxx", "user2d", "gpu", a, 1)
    assert rc != 0 and "GPUassert" in err


@pytest.mark.parametrize("kernel", ["cloudsim", "usercloud"])
@pytest.mark.parametrize("method", ["cpu", "seq"])
def test_dropin_cpu_cloudsim_fields(kernel, method):
    """CloudsimArgs (stock kernel) and a user Args type holding Args2D fields:
    the CPU methods read the fields through Args2D's host map."""
    t, wx, wy = inputs.cloudsim_fields(23, 41, seed=8)
    for iters in (0, 1, 5):
        rc, got, err = dropin.run("drive_cpu", kernel, method, t, iters, wind=(wx, wy))
        assert rc == 0, err
        want = oracle.run_2d_cloud(t, wx, wy, 0.01, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (kernel, method, iters)



@pytest.mark.parametrize("kernel,shape", [("map1d", (1, 1000)), ("map2d", (37, 53)), ("map3d", (9, 11, 13))])
@pytest.mark.parametrize("method", ["cpu", "seq"])
def test_dropin_map_cpu(kernel, shape, method):
    """Map / Map2D / Map3D (include/PSkelMap.h): every cell mapped, ping-pong
    through a clone of the input, result in the output.  The user's map is
    x -> 0.5*x*x + 0.125, applied iteratively (a plain f32 reference of it)."""
    a = inputs.uniform(21, 0, int(np.prod(shape))).reshape(shape)
    arr = a  # map1d: a (1, n) host array, read as Array<float>(n)
    for iters in (0, 1, 4, 7):
        rc, got, err = dropin.run("drive_cpu", kernel, method, arr, iters)
        assert rc == 0, err
        want = arr.copy()
        for _ in range(iters):
            want = (np.float32(0.5) * want * want + np.float32(0.125)).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (kernel, method, iters)


def test_dropin_cpu_under_sanitizers():
    """The headers' CPU methods and StencilTiling::tile built with
    AddressSanitizer + UndefinedBehaviorSanitizer (g++, host code): every
    kernel family, both CPU methods, field arguments and Map, on small
    ragged grids, with results equal to the oracle's."""
    import os
    import subprocess

    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    os.environ.update({k: env[k] for k in ("ASAN_OPTIONS", "UBSAN_OPTIONS")})
    for kernel, stencil, field in CASES:
        a = field()
        for method in ("cpu", "seq"):
            rc, got, err = dropin.run("drive_cpu_asan", kernel, method, a, 3)
            assert rc == 0 and "runtime error" not in err, err[-3000:]
            want = oracle.run_3d(stencil(), a, 3)[0] if a.ndim == 3 else oracle.run_2d(stencil(), a, 3)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (kernel, method)
    t, wx, wy = inputs.cloudsim_fields(13, 21, seed=9)
    rc, got, err = dropin.run("drive_cpu_asan", "usercloud", "cpu", t, 2, wind=(wx, wy))
    assert rc == 0 and "runtime error" not in err, err[-3000:]
    rc, got, err = dropin.run("drive_cpu_asan", "map3d", "cpu", inputs.uniform(3, 0, 99).reshape(3, 3, 11), 2)
    assert rc == 0 and "runtime error" not in err, err[-3000:]
    p = subprocess.run([dropin.driver("tile_geometry_asan")], input="100 50 30 1 3 2 0 0 10 10 10\n7 5 3 2 2 0 4 1 9 9 9\n",
                       capture_output=True, text=True, env=env, timeout=120)
    assert p.returncode == 0 and "runtime error" not in p.stderr, p.stderr[-3000:]
