"""The cost-model MLP's device kernels against the FP64 CPU checker."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import oracle
from pskel_amd import _lib as L

pytestmark = pytest.mark.gpu

F = C.POINTER(C.c_float)


def fp(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(F)


def make(sizes, seed=1):
    h = C.c_void_p()
    arr = (C.c_int32 * len(sizes))(*sizes)
    L.check(L.lib().pskel_mlp_create(arr, len(sizes) - 1, seed, 0, C.byref(h)))
    n = C.c_size_t()
    L.check(L.lib().pskel_mlp_num_params(h, C.byref(n)))
    return h, n.value


def params(h, n):
    p = np.zeros(n, np.float32)
    L.check(L.lib().pskel_mlp_get_params(h, p.ctypes.data_as(F), n))
    return p


@pytest.mark.parametrize("sizes", [[5, 7, 1], [6, 16, 12, 1], [24, 64, 64, 1]])
def test_loss_and_gradient_match_fp64_checker(gpu, sizes):
    rng = np.random.default_rng(sum(sizes))
    h, n = make(sizes, seed=3)
    try:
        x = rng.standard_normal((37, sizes[0])).astype(np.float32)
        y = rng.standard_normal(37).astype(np.float32)
        loss = C.c_float()
        g = np.zeros(n, np.float32)
        L.check(L.lib().pskel_mlp_loss_grad(h, fp(x), fp(y), 37, C.byref(loss), g.ctypes.data_as(F)))
        p = params(h, n).astype(np.float64)
        want_loss, want_g = oracle.mlp_loss_grad(sizes, p, x.astype(np.float64), y.astype(np.float64))
        assert abs(loss.value - want_loss) <= 1e-5 * abs(want_loss)
        np.testing.assert_allclose(g, want_g, rtol=2e-4, atol=2e-6 * np.abs(want_g).max())
    finally:
        L.lib().pskel_mlp_destroy(h)


def test_normalisation_is_applied(gpu):
    sizes = [4, 8, 1]
    rng = np.random.default_rng(5)
    h, n = make(sizes, seed=4)
    try:
        mean = np.array([1.0, -2.0, 0.5, 3.0], np.float32)
        sd = np.array([2.0, 0.5, 1.0, 4.0], np.float32)
        L.check(L.lib().pskel_mlp_set_normalization(h, fp(mean), fp(sd), 1.5, 2.0))
        x = rng.standard_normal((20, 4)).astype(np.float32) * 3
        y = rng.standard_normal(20).astype(np.float32)
        loss = C.c_float()
        g = np.zeros(n, np.float32)
        L.check(L.lib().pskel_mlp_loss_grad(h, fp(x), fp(y), 20, C.byref(loss), g.ctypes.data_as(F)))
        xs = (x.astype(np.float64) - mean) / sd
        ys = (y.astype(np.float64) - 1.5) / 2.0
        p = params(h, n).astype(np.float64)
        want_loss, want_g = oracle.mlp_loss_grad(sizes, p, xs, ys)
        assert abs(loss.value - want_loss) <= 1e-5 * abs(want_loss)
        np.testing.assert_allclose(g, want_g, rtol=2e-4, atol=1e-6)
        pred = np.zeros(20, np.float32)
        L.check(L.lib().pskel_mlp_predict(h, fp(x), 20, pred.ctypes.data_as(F)))
        np.testing.assert_allclose(pred, oracle.mlp_predict(sizes, p, xs) * 2.0 + 1.5, rtol=1e-5, atol=1e-5)
    finally:
        L.lib().pskel_mlp_destroy(h)


def test_training_reduces_loss_and_round_trips(gpu):
    """Fit a smooth function; the trained model predicts it and survives save/load."""
    rng = np.random.default_rng(6)
    x = rng.uniform(-2, 2, (2000, 3)).astype(np.float32)
    y = (np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - 0.3 * x[:, 2]).astype(np.float32)
    h, n = make([3, 32, 32, 1], seed=7)
    try:
        before = C.c_float()
        L.check(L.lib().pskel_mlp_loss_grad(h, fp(x), fp(y), 2000, C.byref(before), None))
        last = C.c_float()
        L.check(L.lib().pskel_mlp_train(h, fp(x), fp(y), 2000, 200, 128, 3e-3, 11, C.byref(last)))
        after = C.c_float()
        L.check(L.lib().pskel_mlp_loss_grad(h, fp(x), fp(y), 2000, C.byref(after), None))
        assert after.value < 0.02 * before.value, (before.value, after.value)
        pred = np.zeros(2000, np.float32)
        L.check(L.lib().pskel_mlp_predict(h, fp(x), 2000, pred.ctypes.data_as(F)))
        assert np.sqrt(np.mean((pred - y) ** 2)) < 0.1
        with tempfile.TemporaryDirectory() as t:
            path = os.path.join(t, "m.bin").encode()
            L.check(L.lib().pskel_mlp_save(h, path))
            h2 = C.c_void_p()
            L.check(L.lib().pskel_mlp_load(path, 0, C.byref(h2)))
            pred2 = np.zeros(2000, np.float32)
            L.check(L.lib().pskel_mlp_predict(h2, fp(x), 2000, pred2.ctypes.data_as(F)))
            L.lib().pskel_mlp_destroy(h2)
            assert np.array_equal(pred, pred2)
    finally:
        L.lib().pskel_mlp_destroy(h)


def test_training_is_deterministic(gpu):
    rng = np.random.default_rng(8)
    x = rng.standard_normal((300, 5)).astype(np.float32)
    y = x.sum(1).astype(np.float32)
    outs = []
    for _ in range(2):
        h, n = make([5, 16, 1], seed=9)
        L.check(L.lib().pskel_mlp_train(h, fp(x), fp(y), 300, 5, 64, 1e-2, 3, None))
        outs.append(params(h, n))
        L.lib().pskel_mlp_destroy(h)
    assert np.array_equal(outs[0], outs[1])
