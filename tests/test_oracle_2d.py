"""Pins the CPU oracle's 2D runners (f32/f64 stencils, u8 life).

The reference ships no golden vectors or known-answer tests for this path
(its test/ holds timing logs only), so each check below comes from what the
mathematics fixes, chosen so that a plausible mistake -- a swapped axis, a
wrong sign, a changed association order, an off-by-one interior, a ping-pong
parity slip -- fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle
from pskel_amd import inputs
from pskel_amd import kernels as K


# ---------------------------------------------------------------------------
# Jacobi 5-point
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("H,W,p,q", [(33, 33, 1, 1), (21, 40, 1, 3), (40, 21, 2, 1), (17, 65, 3, 5)])
def test_jacobi_eigenmode_decay(H, W, p, q):
    """Zero-ring discrete Laplacian mode sin(p*pi*h/(H-1)) sin(q*pi*w/(W-1)) decays by
    exactly (cos(p*pi/(H-1)) + cos(q*pi/(W-1)))/2 per sweep.  A non-square grid with
    p != q catches a transposed index map."""
    h = np.arange(H)[:, None]
    w = np.arange(W)[None, :]
    mode = np.sin(p * math.pi * h / (H - 1)) * np.sin(q * math.pi * w / (W - 1))
    lam = (math.cos(p * math.pi / (H - 1)) + math.cos(q * math.pi / (W - 1))) / 2
    for n in (1, 2, 7):
        got = oracle.run_2d(K.jacobi2d(), mode, n)
        np.testing.assert_allclose(got, mode * lam ** n, atol=1e-13, rtol=0)


def test_jacobi_fixed_point_exact_f32():
    """A constant field with a matching ring is a fixed point, exactly in f32
    (0.375 + ... sums are exact, 0.25 * 1.5 is exact)."""
    a = np.full((19, 23), 0.375, np.float32)
    assert np.array_equal(oracle.run_2d(K.jacobi2d(), a, 5), a)


def test_jacobi_linearity_f64():
    x = inputs.jacobi2d_field(30, 41, seed=1).astype(np.float64)
    y = inputs.jacobi2d_field(30, 41, seed=2).astype(np.float64)
    a, b = 0.3, -1.7
    lhs = oracle.run_2d(K.jacobi2d(), a * x + b * y, 6)
    rhs = a * oracle.run_2d(K.jacobi2d(), x, 6) + b * oracle.run_2d(K.jacobi2d(), y, 6)
    np.testing.assert_allclose(lhs, rhs, atol=1e-13)


def test_jacobi_flip_h_commutes_exactly_f32():
    """The sum order ((n + s) + w) + e is symmetric under h -> H-1-h (n and s swap
    inside one commutative add), so flipping rows commutes with the f32 sweep
    bit for bit.  An order that splits n and s (e.g. ((n + w) + s) + e) fails."""
    a = inputs.jacobi2d_field(37, 29, seed=3)
    got = oracle.run_2d(K.jacobi2d(), a[::-1].copy(), 4)
    assert np.array_equal(got, oracle.run_2d(K.jacobi2d(), a, 4)[::-1])


def test_jacobi_brute_force_tiny():
    """Independent numpy restatement (vectorised slices) on tiny grids."""
    rng = np.random.default_rng(0)
    for H, W in [(3, 3), (4, 7), (9, 5)]:
        a = rng.random((H, W), dtype=np.float32)
        b = a.copy()
        cur, nxt = a.copy(), b.copy()
        for _ in range(5):
            n = cur[:-2, 1:-1]
            s = cur[2:, 1:-1]
            wv = cur[1:-1, :-2]
            e = cur[1:-1, 2:]
            nxt[1:-1, 1:-1] = np.float32(0.25) * (((n + s) + wv) + e)
            cur, nxt = nxt, cur
        # after 5 (odd) sweeps the result is the buffer that started as `output`
        assert np.array_equal(oracle.run_2d(K.jacobi2d(), a, 5, out=b), cur)


# ---------------------------------------------------------------------------
# runIterativeCPU ping-pong semantics (include/PSkelStencil.hpp:880-938)
# ---------------------------------------------------------------------------
def test_pingpong_rings_and_clobbering():
    a = inputs.jacobi2d_field(12, 14, seed=4)
    b = a.copy()
    b[0, :] = 5.0  # output's ring differs from input's
    for n in range(0, 6):
        inp, out = a.copy(), b.copy()
        oracle.run_2d_inplace(K.jacobi2d(), inp, out, n)
        # the result buffer keeps the ring of the buffer it lives in:
        # odd n -> output's ring, even n -> input's ring copied into output
        want_ring = 5.0 if n % 2 else 0.0
        assert np.all(out[0, :] == want_ring), n
        if n % 2 == 0:
            assert np.array_equal(inp, out)  # output.hostMemCopy(input)
        if n == 0:
            assert np.array_equal(out, a)


def test_interior_bounds_with_explicit_range():
    """Mask range 2 leaves two rings untouched."""
    s = K.jacobi2d()
    s.radius = 2
    a = inputs.jacobi2d_field(10, 11, seed=5) + np.float32(1)
    got = oracle.run_2d(s, a, 1, out=np.zeros_like(a))
    assert np.all(got[:2] == 0) and np.all(got[-2:] == 0) and np.all(got[:, :2] == 0) and np.all(got[:, -2:] == 0)
    assert np.all(got[2:-2, 2:-2] != 0)


def test_range_is_chebyshev_radius():
    s = K.Stencil(K.WEIGHTED_SUM, 2, [(0, -3), (2, 1), (-1, 0)], weights=[1, 1, 1])
    assert oracle.range_of(s) == 3
    s3 = K.Stencil(K.SCALED_SUM, 3, [(0, 0, -2), (1, 1, 1)], coeff=1.0)
    assert oracle.range_of(s3) == 2


# ---------------------------------------------------------------------------
# 9-point diffusion / advection
# ---------------------------------------------------------------------------
def test_diffusion9_weights_closed_form():
    w = np.array(K.diffusion9_weights(), np.float64).reshape(3, 3)  # [dh+1, dw+1]
    assert abs(w.sum() - 1.0) < 1e-15
    dh = np.array([-1, 0, 1])[:, None]
    dw = np.array([-1, 0, 1])[None, :]
    # first moment of the weights: -(drift per sweep)
    assert abs((w * dw).sum() - (-0.06)) < 1e-15
    assert abs((w * dh).sum() - 0.03) < 1e-15
    assert w[1, 1] == 0.5 and abs(w[1, 0] - 0.12) < 1e-15 and abs(w[0, 1] - 0.09) < 1e-15


def test_diffusion9_mass_and_drift():
    """One sweep of a blob far from the ring: mass scales by sum(w) and the centroid
    moves by -sum_k w_k o_k (w: +0.06, h: -0.03).  A transposed offset or a sign
    error moves it elsewhere."""
    s = K.diffusion9()
    H, W = 80, 90
    h = np.arange(H)[:, None]
    w = np.arange(W)[None, :]
    blob = np.exp(-((h - 37.0) ** 2 + (w - 46.0) ** 2) / (2 * 5.0 ** 2))
    got = oracle.run_2d(s, blob, 1)
    wf = s.f32_weights().astype(np.float64).reshape(3, 3)
    m0, m1 = blob.sum(), got.sum()
    assert abs(m1 / m0 - wf.sum()) < 1e-12
    dw_drift = (got * w).sum() / m1 - (blob * w).sum() / m0
    dh_drift = (got * h).sum() / m1 - (blob * h).sum() / m0
    ow = np.array([-1, 0, 1])[None, :]
    oh = np.array([-1, 0, 1])[:, None]
    assert abs(dw_drift - (-(wf * ow).sum() / wf.sum())) < 1e-9
    assert abs(dh_drift - (-(wf * oh).sum() / wf.sum())) < 1e-9
    assert abs(dw_drift - 0.06) < 1e-6 and abs(dh_drift + 0.03) < 1e-6


def test_weighted_brute_force_tiny():
    s = K.diffusion9()
    wf = s.f32_weights()
    rng = np.random.default_rng(1)
    a = rng.random((6, 8), dtype=np.float32)
    cur = a.copy()
    nxt = a.copy()
    for _ in range(3):
        for hh in range(1, 5):
            for ww in range(1, 7):
                acc = np.float32(wf[0] * cur[hh - 1, ww - 1])
                for k, (dh, dw) in enumerate(K.BOX9_OFFSETS[1:], start=1):
                    acc = np.float32(acc + np.float32(wf[k] * cur[hh + dh, ww + dw]))
                nxt[hh, ww] = acc
        cur, nxt = nxt, cur
    assert np.array_equal(oracle.run_2d(s, a, 3), cur)


# ---------------------------------------------------------------------------
# Game of Life
# ---------------------------------------------------------------------------
def _board(n, cells):
    b = np.zeros((n, n), np.uint8)
    for h, w in cells:
        b[h, w] = 1
    return b


def test_life_blinker_period_two():
    b = _board(7, [(3, 2), (3, 3), (3, 4)])
    one = oracle.run_2d(K.life(), b, 1)
    assert np.array_equal(one, _board(7, [(2, 3), (3, 3), (4, 3)]))
    assert np.array_equal(oracle.run_2d(K.life(), b, 2), b)


def test_life_still_lifes():
    block = _board(6, [(2, 2), (2, 3), (3, 2), (3, 3)])
    beehive = _board(8, [(3, 2), (2, 3), (2, 4), (3, 5), (4, 3), (4, 4)])
    for b in (block, beehive):
        assert np.array_equal(oracle.run_2d(K.life(), b, 5), b)


def test_life_glider_translates():
    b = _board(14, [(1, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
    got = oracle.run_2d(K.life(), b, 8)
    want = np.roll(np.roll(b, 2, 0), 2, 1)
    assert np.array_equal(got, want)
    assert got.sum() == 5


def test_life_brute_force_numpy():
    """Independent count by shifted slices with a dead boundary."""
    rng = np.random.default_rng(2)
    b = (rng.random((40, 53)) < 0.35).astype(np.uint8)
    b[0, :] = b[-1, :] = 0
    b[:, 0] = b[:, -1] = 0
    cur = b.copy()
    for _ in range(6):
        p = np.pad(cur, 1)
        n = sum(p[1 + dh:1 + dh + cur.shape[0], 1 + dw:1 + dw + cur.shape[1]].astype(int)
                for dh in (-1, 0, 1) for dw in (-1, 0, 1) if (dh, dw) != (0, 0))
        nxt = ((n == 3) | ((n == 2) & (cur == 1))).astype(np.uint8)
        nxt[0, :] = nxt[-1, :] = 0
        nxt[:, 0] = nxt[:, -1] = 0
        cur = nxt
    assert np.array_equal(oracle.run_2d(K.life(), b, 6), cur)


def test_life_padded_board_equals_dead_boundary():
    """The padded n x n board (dead ring) evolves like an n x n board whose
    out-of-bounds neighbours are dead: the ring never changes."""
    b = inputs.life_field(60, seed=3)
    got = oracle.run_2d(K.life(), b, 9)
    assert got[0].sum() == got[-1].sum() == got[:, 0].sum() == got[:, -1].sum() == 0


# ---------------------------------------------------------------------------
# CLOUDSIM (wind advection + conduction with field arguments)
# ---------------------------------------------------------------------------
def _winds(H, W, vx=0.0, vy=0.0):
    """Args2D host arrays (shape (W, H), [y, x] with x = h, y = w) of constant wind."""
    return np.full((W, H), vx, np.float32), np.full((W, H), vy, np.float32)


def test_cloud_constant_field_is_fixed_point():
    H, W = 9, 11
    u = np.full((H, W), 3.25, np.float32)
    for vx, vy in ((0.0, 0.0), (0.7, -0.3), (-1.0, 1.0)):
        wx, wy = _winds(H, W, vx, vy)
        assert np.array_equal(oracle.run_2d_cloud(u, wx, wy, 0.3, 4), u)


@pytest.mark.parametrize("p,q", [(1, 1), (2, 3)])
def test_cloud_conduction_eigenmode_decay(p, q):
    """Zero wind: out = c - 0.0243 * 0.25 * dt * (4c - sum of neighbours), so the
    zero-ring mode sin(p pi h/(H-1)) sin(q pi w/(W-1)) decays by exactly
    1 - 0.006075 dt (4 - 2 cos a - 2 cos b) per sweep."""
    H, W, dt, n = 33, 41, 4.0, 6
    h = np.arange(H)[:, None]
    w = np.arange(W)[None, :]
    mode = np.sin(p * math.pi * h / (H - 1)) * np.sin(q * math.pi * w / (W - 1))
    lam = 1 - 0.0243 * 0.25 * dt * (4 - 2 * math.cos(p * math.pi / (H - 1)) - 2 * math.cos(q * math.pi / (W - 1)))
    wx, wy = _winds(H, W)
    got = oracle.run_2d_cloud(mode.astype(np.float32), wx, wy, dt, n).astype(np.float64)
    want = mode * lam ** n
    assert np.max(np.abs(got - want)) < 1e-5


@pytest.mark.parametrize("vx,vy,axis,sign", [(0.5, 0.0, 1, 1), (-0.5, 0.0, 1, -1), (0.0, 0.5, 0, 1),
                                             (0.0, -0.5, 0, -1)])
def test_cloud_advection_of_a_ramp(vx, vy, axis, sign):
    """A linear ramp (slope 0.25 along w or h) has zero conduction; the wind
    term compares each cell with its upwind-side neighbour (w + xf or h + yf,
    xf = sign of the wind), so one sweep adds exactly 10 * 0.25 * 0.5 * dt with
    the sign of the wind.  A wrong side or sign changes the increment."""
    H, W, dt = 7, 9, 0.125
    hh, ww = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    u = (1.0 + 0.25 * (ww if axis == 1 else hh)).astype(np.float32)
    wx, wy = _winds(H, W, vx, vy)
    got = oracle.run_2d_cloud(u, wx, wy, dt, 1)
    want = u.copy()
    want[1:-1, 1:-1] += np.float32(sign * 10 * 0.25 * 0.5 * dt)
    assert np.array_equal(got, want)


def test_cloud_field_index_map():
    """The kernel reads wind(x = h, y = w) through Args2D's (x, y) -> y*width + x
    (include/PSkelArgs.hpp:116-123): on a non-square grid a wind set only at
    host element [y=w0, x=h0] moves only cell (h0, w0)."""
    H, W, dt = 6, 10, 0.125
    hh, ww = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    u = (1.0 + 0.25 * ww).astype(np.float32)
    wx, wy = _winds(H, W)
    h0, w0 = 2, 7
    wx[w0, h0] = 0.5
    got = oracle.run_2d_cloud(u, wx, wy, dt, 1)
    diff = np.argwhere(got != u)
    assert diff.tolist() == [[h0, w0]]
    assert got[h0, w0] == u[h0, w0] + np.float32(0.15625)


def test_cloud_ring_untouched_and_parity():
    t, wx, wy = inputs.cloudsim_fields(12, 17, seed=3)
    r1 = oracle.run_2d_cloud(t, wx, wy, 0.01, 1)
    r2 = oracle.run_2d_cloud(r1, wx, wy, 0.01, 1)
    assert np.array_equal(oracle.run_2d_cloud(t, wx, wy, 0.01, 2), r2)
    for r in (r1, r2):
        assert np.array_equal(r[0], t[0]) and np.array_equal(r[-1], t[-1])
        assert np.array_equal(r[:, 0], t[:, 0]) and np.array_equal(r[:, -1], t[:, -1])
    assert np.array_equal(oracle.run_2d_cloud(t, wx, wy, 0.01, 0), t)
