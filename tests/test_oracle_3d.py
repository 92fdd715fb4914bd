"""Pins the CPU oracle's 3D runners, storage conversions and FP64 MLP."""
import math
import os

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


def _ulp16(v, store):
    """Spacing of the 16-bit format at |v| (normal range; f16 subnormals: 2^-24)."""
    mbits = 7 if store == oracle.STORE_BF16 else 10
    e = np.floor(np.log2(np.maximum(np.abs(v), 1e-30)))
    if store == oracle.STORE_F16:
        e = np.maximum(e, -14)
    return 2.0 ** (e - mbits)


def _offset_cases():
    j7 = list(K.J7_OFFSETS)
    box = [(dh, dw, dd) for dh in (-1, 0, 1) for dw in (-1, 0, 1) for dd in (-1, 0, 1) if abs(dh) + abs(dw) + abs(dd) <= 1]
    return [
        ("jacobi7", K.jacobi3d()),
        # coeff * n = 1.2: the folded offset term is 0.2 * offset
        ("scaled_1.2", K.Stencil(K.SCALED_SUM, 3, j7, coeff=0.2, boundary=K.BOUNDARY_COPY)),
        # every cell updated, points outside read the halo value
        ("scaled_halo", K.Stencil(K.SCALED_SUM, 3, j7, coeff=0.15, boundary=K.BOUNDARY_HALO, halo_value=2.0)),
        # weights summing to 0.9: the weighted path decodes u = x + offset
        ("weighted", K.Stencil(K.WEIGHTED_SUM, 3, box, weights=[0.1, 0.15, 0.2, 0.05, 0.1, 0.2, 0.1],
                               boundary=K.BOUNDARY_HALO, halo_value=0.5)),
    ]


@pytest.mark.parametrize("store", [oracle.STORE_BF16, oracle.STORE_F16])
@pytest.mark.parametrize("offset", [0.0, 0.5, -0.75, 3.0])
@pytest.mark.parametrize("case", range(4), ids=["jacobi7", "scaled_1.2", "scaled_halo", "weighted"])
def test_r16_sweep_within_rounding_of_f64(store, offset, case):
    """One 16-bit sweep, decoded (x + offset), equals the FP64 sweep of the same
    stored state within half a 16-bit ulp of the stored result plus f32
    rounding, whatever the offset.  Folding the offset into
    fma(coeff, sum x_i, (coeff * n - 1) * offset) must not change the value: a
    wrong sign or a missing term is off by about 2 * 0.2 * offset (coeff * n =
    1.2), a halo read without the offset by 0.15 * offset, a weighted sum with
    a folded term by 0.1 * offset."""
    name, s = _offset_cases()[case]
    rng = np.random.default_rng(11 + case)
    u = rng.uniform(0.0, 1.0, (9, 10, 11)).astype(np.float32)
    w = oracle.encode16(u, store, offset)
    # the stored state, exactly: x + offset in f64
    x = (oracle.decode16(w, store, 0.0)).astype(np.float64)
    u64 = x + float(np.float32(offset))
    want, _ = oracle.run_3d(s, u64, 1)
    got_w, _ = oracle.run_3d_r16(s, store, w, 1, store_offset=offset)
    got = oracle.decode16(got_w, store, 0.0).astype(np.float64) + float(np.float32(offset))
    bound = 0.5 * _ulp16(want - offset, store) + 2.0 ** -20 * (np.abs(want) + abs(offset) + 1.0)
    err = np.abs(got - want)
    assert np.all(err <= bound), (name, offset, float(err.max()), float(bound[np.argmax(err - bound)]))


def test_offset_term_closed_form():
    """c0 = (coeff * n - 1) * offset, in double, rounded once."""
    f = oracle.lib().oracle_scaled_sum_offset_term
    import ctypes as C
    f.argtypes = [C.c_float, C.c_int, C.c_float]
    f.restype = C.c_float
    assert f(0.2, 6, 0.5) == np.float32((float(np.float32(0.2)) * 6.0 - 1.0) * 0.5)
    assert abs(f(0.2, 6, 0.5) - 0.1) < 1e-7
    assert f(0.25, 4, 3.0) == 0.0
    assert f(np.float32(1.0 / 6.0), 6, 0.5) == np.float32((float(np.float32(1.0 / 6.0)) * 6.0 - 1.0) * 0.5)
    assert f(0.5, 6, 0.0) == 0.0


def test_r16_stays_close_to_f64():
    u = inputs.jacobi3d_field(20, seed=9).astype(np.float64)
    want, _ = oracle.run_3d(K.jacobi3d(), u, 10)
    for store, tol in ((oracle.STORE_BF16, 4e-3), (oracle.STORE_F16, 5e-4)):
        w = oracle.encode16(u.astype(np.float32), store)
        got_w, _ = oracle.run_3d_r16(K.jacobi3d(), store, w, 10)
        got = oracle.decode16(got_w, store).astype(np.float64)
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert rel < tol, (store, rel)


def _bf16_val(w):
    return float(oracle.decode16(np.array([w], np.uint16), oracle.STORE_BF16)[0])


def _f16_val(w):
    return float(np.array([w], np.uint16).view(np.float16).astype(np.float64)[0])


@pytest.mark.parametrize("x", [1.0 + 2 ** -9, -3.14159, 0.3, 1e-20, -7.5e-39, 65504.0, 3.3895e38])
def test_bf16_stochastic_exact_probability(x):
    """Over all 2^16 values of the random word's top half, the result rounds the
    magnitude up exactly `dropped` times (P(up) = dropped fraction of an ulp) and
    otherwise truncates, so the mean of the decoded results equals x exactly
    (finite results).  The low half of the word is ignored."""
    xf = np.float32(x)
    bits = int(xf.view(np.uint32))
    dropped = bits & 0xFFFF
    down = bits >> 16
    vals = [oracle.encode16_sr(xf, oracle.STORE_BF16, k << 16) for k in range(1 << 16)]
    assert sum(1 for v in vals if v != down) == dropped
    assert all(v in (down, down + 1) for v in vals)
    if all((v & 0x7F80) != 0x7F80 for v in vals):
        assert sum(_bf16_val(v) for v in vals) / (1 << 16) == float(xf)
    assert oracle.encode16_sr(xf, oracle.STORE_BF16, 0x1234FFFF) == oracle.encode16_sr(xf, oracle.STORE_BF16, 0x12340000)


@pytest.mark.parametrize("x", [1.0 + 2 ** -12, -0.3, 1000.123, 6.2e-5, -2.0 ** -14 * 1.5, 65500.0])
def test_f16_stochastic_exact_probability_normal(x):
    """Normal f16 range: the top 13 bits of the random word decide.  Over all
    2^13 of them the two bracketing f16 values occur with the exact weights that
    make the mean equal x; a word of 0 truncates."""
    xf = np.float32(x)
    vals = [oracle.encode16_sr(xf, oracle.STORE_F16, k << 19) for k in range(1 << 13)]
    assert len(set(vals)) <= 2
    assert sum(_f16_val(v) for v in vals) / (1 << 13) == float(xf)
    t = _f16_val(oracle.encode16_sr(xf, oracle.STORE_F16, 0))
    assert abs(t) <= abs(float(xf)) < abs(t) + abs(float(xf)) * 2 ** -9
    assert oracle.encode16_sr(xf, oracle.STORE_F16, (77 << 19) | 0x7FFFF) == oracle.encode16_sr(xf, oracle.STORE_F16, 77 << 19)


def test_f16_stochastic_overflow():
    """Above the largest f16 (65504) the round-up neighbour is 65536, which is inf."""
    for x, frac in ((65504.0, 0.0), (65510.0, 6 / 32), (65519.0, 15 / 32), (65535.0, 31 / 32)):
        vals = [oracle.encode16_sr(np.float32(x), oracle.STORE_F16, k << 19) for k in range(1 << 13)]
        assert set(vals) <= {0x7BFF, 0x7C00}
        assert sum(v == 0x7C00 for v in vals) == round(frac * 8192)


@pytest.mark.parametrize("x", [3.3e-6, -1.0e-7, 2.0 ** -20 + 2.0 ** -30, 5.0e-8, 2.0 ** -14 - 2.0 ** -30, 1e-30, 0.0])
def test_f16_stochastic_subnormal_mean(x):
    """Below 2^-14 the f16 spacing is 2^-24: the result is one of the two
    multiples of 2^-24 around |x|, and over all 2^13 random words its mean is
    |x| truncated to a multiple of 2^-37 (the dropped fraction counts to 13
    bits, as in the normal range)."""
    xf = np.float32(x)
    vals = [oracle.encode16_sr(xf, oracle.STORE_F16, k << 19) for k in range(1 << 13)]
    ax = abs(float(xf))
    lo = np.floor(ax * 2.0 ** 24) * 2.0 ** -24
    assert {abs(_f16_val(v)) for v in vals} <= {lo, lo + 2.0 ** -24}
    assert sum(abs(_f16_val(v)) for v in vals) / (1 << 13) == np.floor(ax * 2.0 ** 37) * 2.0 ** -37
    assert all((v & 0x8000) == (0x8000 if np.signbit(xf) else 0) for v in vals)


def test_stochastic_nonfinite():
    inf, ninf = np.float32(np.inf), np.float32(-np.inf)
    for r in (0, 0xFFFFFFFF, 0x80000000):
        assert oracle.encode16_sr(inf, oracle.STORE_BF16, r) == 0x7F80
        assert oracle.encode16_sr(ninf, oracle.STORE_BF16, r) == 0xFF80
        assert oracle.encode16_sr(inf, oracle.STORE_F16, r) == 0x7C00
        assert oracle.encode16_sr(ninf, oracle.STORE_F16, r) == 0xFC00
        nan = np.float32(np.nan)
        assert oracle.encode16_sr(nan, oracle.STORE_BF16, r) & 0x7FC0 == 0x7FC0
        assert oracle.encode16_sr(nan, oracle.STORE_F16, r) & 0x7E00 == 0x7E00


def test_stochastic_matches_recorded_gfx950_conversions():
    """4096 (value, random word) pairs with the results of gfx950's
    v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32 recorded on the device
    (tests/golden/cvt_sr_probe.hip): special values, every bit pattern class,
    the f16 range and its subnormal band."""
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "cvt_sr_gfx950.npz"))
    x = f["x_bits"].view(np.float32)
    r = f["r"]
    nan = np.isnan(x)
    for store, want in ((oracle.STORE_BF16, f["bf16"]), (oracle.STORE_F16, f["f16"])):
        got = np.array([oracle.encode16_sr(v, store, int(k)) for v, k in zip(x, r)], np.uint16)
        # NaN payloads pass through a double on the ctypes call: compare class only
        assert np.array_equal(got[~nan], want[~nan]), store
        q = 0x7FC0 if store == oracle.STORE_BF16 else 0x7E00
        assert np.all((got[nan] & q) == q) and np.all((want[nan] & q) == q)


def test_sr_bits_statistics():
    """Random words A(w*D + d) ^ B(t, h): the top 16 bits close to uniform per
    bit along a plane, a column and the sweeps of one cell; pairwise
    decorrelated between neighbouring cells of a plane, of a column, between
    consecutive sweeps of one cell and between seeds.  Words that share one of
    the two terms are related by the XOR of the other two (two planes of a
    sweep, or one plane in two sweeps, by B ^ B' over every cell; two cells by
    A ^ A' over every plane and sweep), and that relating word is itself
    uniform over the pairs: B(t, h) ^ B(t', h) over t, A(c) ^ A(c + 1) over
    c."""
    n = 1 << 14
    top = lambda a: (a >> np.uint64(16)).astype(np.float64) / 2 ** 16
    bits = lambda a: np.array([((a >> np.uint64(k)) & np.uint64(1)).mean() for k in range(16, 32)])
    plane = np.array([oracle.sr_bits(11, 0, 5, c) for c in range(n)], np.uint64)
    column = np.array([oracle.sr_bits(11, 0, h, 77) for h in range(n)], np.uint64)
    sweeps = np.array([oracle.sr_bits(11, t, 5, 77) for t in range(n)], np.uint64)
    seed2 = np.array([oracle.sr_bits(12, 0, 5, c) for c in range(n)], np.uint64)
    lim5 = 5 * 0.5 / np.sqrt(n)
    for a in (plane, column, sweeps):
        assert np.all(np.abs(bits(a) - 0.5) < lim5)
    lim = 5 / np.sqrt(n)
    fp, fc, fs = top(plane), top(column), top(sweeps)
    assert abs(np.corrcoef(fp[:-1], fp[1:])[0, 1]) < lim
    assert abs(np.corrcoef(fc[:-1], fc[1:])[0, 1]) < lim
    assert abs(np.corrcoef(fs[:-1], fs[1:])[0, 1]) < lim
    assert abs(np.corrcoef(fp, top(seed2))[0, 1]) < lim
    h6 = np.array([oracle.sr_bits(11, t, 6, 77) for t in range(n)], np.uint64)
    assert abs(np.corrcoef(fs, top(h6))[0, 1]) < lim
    # the relating words: uniform over sweeps (one plane, sweeps t and t + 1)
    # and over cells (neighbouring cells of one plane)
    assert np.all(np.abs(bits(sweeps[:-1] ^ sweeps[1:]) - 0.5) < lim5)
    assert np.all(np.abs(bits(plane[:-1] ^ plane[1:]) - 0.5) < lim5)
    # the split itself: A depends on the seed and cell only, B on the seed,
    # sweep and plane only
    a = lambda seed, t, h, c: oracle.sr_bits(seed, t, h, c)
    for t, h, c, c2 in ((0, 5, 77, 78), (9, 0, 3, 1000), (1 << 33, 7, 0, 12)):
        assert a(11, t, h, c) ^ a(11, t, h, c2) == a(11, t + 1, h + 3, c) ^ a(11, t + 1, h + 3, c2)
        assert a(11, t, h, c) ^ a(11, t + 1, h, c) == a(11, t, h, c2) ^ a(11, t + 1, h, c2)
        assert a(11, t, h, c) ^ a(12, t, h, c) != a(11, t, h, c2) ^ a(12, t, h, c2)


def test_r16_stochastic_sweep_is_unbiased():
    """One sweep of a field whose h-planes alternate between 1 + 2^-8 and 1
    (exact in both formats at offset 1.0): the interior updates are 1 + 2^-8 *
    {2/3, 1/3} in f32, inexact in 16 bits.  Nearest rounding errs by the same
    third of an ulp in every cell; stochastic rounding lands on one of the two
    neighbours of every f32 value and averages back to it within the sampling
    error."""
    n = 40
    u = np.ones((n, n, n), np.float32)
    u[0::2] += np.float32(2 ** -8)
    # the exact sweep with the f32 coefficient (in f64): 16-bit results round
    # the deviation x = u - 1 once, at a resolution far below f32's near 1.0
    cf = float(np.float32(1.0 / 6.0))
    s64 = K.Stencil(K.SCALED_SUM, 3, list(K.J7_OFFSETS), coeff=cf, coeff_f64=cf, boundary=K.BOUNDARY_COPY)
    want = oracle.run_3d(s64, u.astype(np.float64), 1)[0][1:-1, 1:-1, 1:-1]
    for store, mbits in ((oracle.STORE_BF16, 7), (oracle.STORE_F16, 10)):
        w = oracle.encode16(u, store, 1.0)
        assert np.array_equal(oracle.decode16(w, store, 1.0), u)
        ulp = 2.0 ** (np.floor(np.log2(want - 1.0)) - mbits)      # per-cell spacing of (v - 1)
        sigma = 0.5 * np.sqrt(np.mean(ulp ** 2) / want.size)
        bias = {}
        for mode in (oracle.ROUND_NEAREST, oracle.ROUND_STOCHASTIC):
            got_w, _ = oracle.run_3d_r16(K.jacobi3d(), store, w, 1, store_offset=1.0, rounding=mode, seed=3)
            err = oracle.decode16(got_w, store, 1.0)[1:-1, 1:-1, 1:-1].astype(np.float64) - want
            assert np.all(np.abs(err) < ulp), (store, mode)
            bias[mode] = err.mean()
        assert abs(bias[oracle.ROUND_STOCHASTIC]) < 6 * sigma, (store, bias, sigma)
        assert abs(bias[oracle.ROUND_NEAREST]) > 20 * sigma, (store, bias, sigma)


def test_r16_stochastic_reproducible_and_resumable():
    u = inputs.jacobi3d_field(16, seed=2)
    w = oracle.encode16(u, oracle.STORE_BF16, 0.5)
    full, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, w, 7, store_offset=0.5,
                                rounding=oracle.ROUND_STOCHASTIC, seed=5)
    again, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, w, 7, store_offset=0.5,
                                 rounding=oracle.ROUND_STOCHASTIC, seed=5)
    assert np.array_equal(full, again)
    part, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, w, 3, store_offset=0.5,
                                rounding=oracle.ROUND_STOCHASTIC, seed=5)
    rest, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, part, 4, store_offset=0.5,
                                rounding=oracle.ROUND_STOCHASTIC, seed=5, first_sweep=3)
    assert np.array_equal(rest, full)
    other, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, w, 7, store_offset=0.5,
                                 rounding=oracle.ROUND_STOCHASTIC, seed=6)
    assert not np.array_equal(other, full)


def test_r16_stochastic_beats_nearest_vs_f64():
    """48^3 x 300 sweeps from a uniform field with unit faces: nearest bf16
    stagnates (error grows with the sweep count), stochastic bf16 stays under
    1e-3 relative L2 of the FP64 run."""
    u = inputs.jacobi3d_field(48, seed=9).astype(np.float64)
    want, _ = oracle.run_3d(K.jacobi3d(), u, 300)
    rel = {}
    for mode in (oracle.ROUND_NEAREST, oracle.ROUND_STOCHASTIC):
        w = oracle.encode16(u.astype(np.float32), oracle.STORE_BF16, 0.5)
        got_w, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, w, 300, store_offset=0.5,
                                     rounding=mode, seed=7)
        got = oracle.decode16(got_w, oracle.STORE_BF16, 0.5).astype(np.float64)
        rel[mode] = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel[oracle.ROUND_STOCHASTIC] < 1e-3, rel
    assert rel[oracle.ROUND_NEAREST] > 2.5 * rel[oracle.ROUND_STOCHASTIC], rel


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
