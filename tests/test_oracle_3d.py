"""Pins the CPU oracle's 3D runners, storage conversions and FP64 MLP."""
import math

import numpy as np
import pytest

import oracle
from pskel_amd import inputs
from pskel_amd import kernels as K


@pytest.mark.parametrize("H,W,D,a,b,c", [(17, 17, 17, 1, 1, 1), (13, 21, 30, 1, 2, 1), (25, 11, 19, 2, 1, 3)])
def test_jacobi3d_separable_mode(H, W, D, a, b, c):
    """u = 1 + sin(a pi h/(H-1)) sin(b pi w/(W-1)) sin(c pi d/(D-1)) with faces at 1:
    the mode decays by (cos(a pi/(H-1)) + cos(b pi/(W-1)) + cos(c pi/(D-1)))/3 per
    sweep.  Distinct dims and wave numbers catch any axis mix-up."""
    h = np.arange(H)[:, None, None]
    w = np.arange(W)[None, :, None]
    d = np.arange(D)[None, None, :]
    mode = np.sin(a * math.pi * h / (H - 1)) * np.sin(b * math.pi * w / (W - 1)) * np.sin(c * math.pi * d / (D - 1))
    lam = (math.cos(a * math.pi / (H - 1)) + math.cos(b * math.pi / (W - 1)) + math.cos(c * math.pi / (D - 1))) / 3
    u = 1.0 + mode
    got, res = oracle.run_3d(K.jacobi3d(), u, 6, residual_every=1)
    np.testing.assert_allclose(got, 1.0 + mode * lam ** 6, atol=1e-13, rtol=0)
    # residual closed form: ||u^k - u^(k-1)|| = (1 - lam) lam^(k-1) ||mode||
    want = (1 - lam) * lam ** np.arange(6) * np.linalg.norm(mode)
    np.testing.assert_allclose(res, np.abs(want), rtol=1e-9)


def test_jacobi3d_ones_fixed_point_exact_f32():
    u = np.ones((9, 10, 11), np.float32)
    got, res = oracle.run_3d(K.jacobi3d(), u, 4, residual_every=2)
    assert np.array_equal(got, u) and np.all(res == 0)


def test_jacobi3d_faces_copied():
    u = inputs.jacobi3d_field(12, seed=2)
    got, _ = oracle.run_3d(K.jacobi3d(), u, 3)
    for sl in (np.s_[0], np.s_[-1], np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1]):
        assert np.all(got[sl] == 1.0)


def test_halo_boundary_closed_form():
    """Mask3D::get semantics: outside the grid reads the halo value.  With u = 0 and
    halo 1, one weighted sweep leaves on each cell the sum of the weights of its
    outside neighbours: face cells 1 outside point, edges 2, corners 3."""
    wts = [0.05, 0.11, 0.13, 0.17, 0.19, 0.23]
    s = K.Stencil(K.WEIGHTED_SUM, 3, list(K.J7_OFFSETS), weights=wts, boundary=K.BOUNDARY_HALO, halo_value=1.0)
    H, W, D = 5, 6, 7
    got, _ = oracle.run_3d(s, np.zeros((H, W, D), np.float64), 1)
    wf = s.f32_weights().astype(np.float64)
    want = np.zeros((H, W, D))
    want[0] += wf[0]      # h-1 outside
    want[-1] += wf[1]     # h+1 outside
    want[:, 0] += wf[2]
    want[:, -1] += wf[3]
    want[:, :, 0] += wf[4]
    want[:, :, -1] += wf[5]
    np.testing.assert_allclose(got, want, atol=1e-15)


def test_jacobi3d_flip_symmetry_exact():
    """((hm + hp) + ...) is symmetric in h: flipping h commutes bit-exactly."""
    u = inputs.jacobi3d_field(11, seed=3)
    a, _ = oracle.run_3d(K.jacobi3d(), u[::-1].copy(), 3)
    b, _ = oracle.run_3d(K.jacobi3d(), u, 3)
    assert np.array_equal(a, b[::-1])


def test_3d_f32_matches_brute_force():
    rng = np.random.default_rng(4)
    u = rng.random((5, 6, 7), dtype=np.float32)
    c = np.float32(1.0 / 6.0)
    cur = u.copy()
    for _ in range(3):
        nxt = cur.copy()
        t = cur[:-2, 1:-1, 1:-1] + cur[2:, 1:-1, 1:-1]
        t = t + cur[1:-1, :-2, 1:-1]
        t = t + cur[1:-1, 2:, 1:-1]
        t = t + cur[1:-1, 1:-1, :-2]
        t = t + cur[1:-1, 1:-1, 2:]
        nxt[1:-1, 1:-1, 1:-1] = c * t
        cur = nxt
    got, _ = oracle.run_3d(K.jacobi3d(), u, 3)
    assert np.array_equal(got, cur)


# ---------------------------------------------------------------------------
# 16-bit storage
# ---------------------------------------------------------------------------
def test_f16_encode_matches_numpy():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(3000).astype(np.float32) * 10,
                        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e-8, 6.0e-8, 3e-5, -7e-6, 2.0 ** -25,
                                  3 * 2.0 ** -26, np.inf, -np.inf], np.float32)])
    got = oracle.encode16(x, oracle.STORE_F16)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    # and decode is exact
    assert np.array_equal(oracle.decode16(want, oracle.STORE_F16), want.view(np.float16).astype(np.float32))


def test_bf16_encode_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.standard_normal(3000).astype(np.float32) * 3,
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 3.0e38, 1e-40],
                                 np.float32)])
    got = oracle.encode16(x, oracle.STORE_BF16)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(got, want)


def test_bf16_ties_to_even():
    # 1 + 2^-8 is halfway between 1.0 (0x3f80) and 1.0078125 (0x3f81): even -> 0x3f80
    assert oracle.encode16(np.float32(1 + 2 ** -8), oracle.STORE_BF16) == 0x3F80
    # 1 + 3*2^-8 is halfway between 0x3f81 and 0x3f82: even -> 0x3f82
    assert oracle.encode16(np.float32(1 + 3 * 2 ** -8), oracle.STORE_BF16) == 0x3F82


@pytest.mark.parametrize("store", [oracle.STORE_BF16, oracle.STORE_F16])
def test_r16_runner_equals_f32_on_exact_data(store):
    """Data on a coarse dyadic grid: the f32 sums and products are exact and the
    results representable in 16 bits, so one reduced-storage sweep must equal the
    f32 sweep.  Pins that the r16 runner evaluates the same expression."""
    s = K.Stencil(K.SCALED_SUM, 3, list(K.J7_OFFSETS)[:4], coeff=0.25, boundary=K.BOUNDARY_COPY)
    rng = np.random.default_rng(7)
    u = (rng.integers(0, 16, (6, 7, 8)) / 16.0).astype(np.float32)
    w = oracle.encode16(u, store)
    assert np.array_equal(oracle.decode16(w, store), u)
    got_w, _ = oracle.run_3d_r16(s, store, w, 1)
    want, _ = oracle.run_3d(s, u, 1)
    assert np.array_equal(oracle.decode16(got_w, store), want)


def test_r16_offset_storage():
    """Storage of (u - 0.5): decoded + offset equals the offset-free result when
    everything is exact."""
    s = K.Stencil(K.SCALED_SUM, 3, list(K.J7_OFFSETS)[:4], coeff=0.25, boundary=K.BOUNDARY_COPY)
    rng = np.random.default_rng(8)
    u = (rng.integers(0, 16, (6, 7, 8)) / 16.0).astype(np.float32)
    w0 = oracle.encode16(u, oracle.STORE_BF16, 0.0)
    w5 = oracle.encode16(u, oracle.STORE_BF16, 0.5)
    r0, _ = oracle.run_3d_r16(s, oracle.STORE_BF16, w0, 1, store_offset=0.0)
    r5, _ = oracle.run_3d_r16(s, oracle.STORE_BF16, w5, 1, store_offset=0.5)
    assert np.array_equal(oracle.decode16(r0, oracle.STORE_BF16, 0.0), oracle.decode16(r5, oracle.STORE_BF16, 0.5))


def test_r16_stays_close_to_f64():
    u = inputs.jacobi3d_field(20, seed=9).astype(np.float64)
    want, _ = oracle.run_3d(K.jacobi3d(), u, 10)
    for store, tol in ((oracle.STORE_BF16, 4e-3), (oracle.STORE_F16, 5e-4)):
        w = oracle.encode16(u.astype(np.float32), store)
        got_w, _ = oracle.run_3d_r16(K.jacobi3d(), store, w, 10)
        got = oracle.decode16(got_w, store).astype(np.float64)
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert rel < tol, (store, rel)


# ---------------------------------------------------------------------------
# MLP (cost model checker)
# ---------------------------------------------------------------------------
def test_mlp_gradient_finite_differences():
    sizes = [5, 7, 6, 1]
    rng = np.random.default_rng(10)
    n_par = sum(sizes[i + 1] * sizes[i] + sizes[i + 1] for i in range(len(sizes) - 1))
    p = rng.standard_normal(n_par) * 0.5
    x = rng.standard_normal((9, 5))
    y = rng.standard_normal(9)
    loss, g = oracle.mlp_loss_grad(sizes, p, x, y)
    eps = 1e-6
    for i in range(0, n_par, 3):
        pp = p.copy()
        pm = p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (oracle.mlp_loss_grad(sizes, pp, x, y, False)[0] - oracle.mlp_loss_grad(sizes, pm, x, y, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), i
    pred = oracle.mlp_predict(sizes, p, x)
    assert abs(np.mean((pred - y) ** 2) - loss) < 1e-14


def test_mlp_linear_closed_form():
    """One linear layer: y_hat = W x + b, gradient 2/n sum (y_hat - y) [x, 1]."""
    rng = np.random.default_rng(11)
    p = rng.standard_normal(4)
    x = rng.standard_normal((6, 3))
    y = rng.standard_normal(6)
    loss, g = oracle.mlp_loss_grad([3, 1], p, x, y)
    r = x @ p[:3] + p[3] - y
    assert abs(loss - np.mean(r ** 2)) < 1e-14
    np.testing.assert_allclose(g[:3], 2 * (r[:, None] * x).mean(0), atol=1e-14)
    np.testing.assert_allclose(g[3], 2 * r.mean(), atol=1e-14)
