"""GPU parity of the TV-L1 path against the CPU oracle.

The kernels are written to the reference's arithmetic (no FMA contraction,
double where the reference promotes), so every stage is compared BIT FOR BIT
with the oracle; the full solve additionally must reproduce the oracle's
per-(scale, warping) iteration counts.  The stated bar of the north star,
max |du|, |dv| <= 1e-3 px, is asserted as well (implied by bit equality).
"""
import numpy as np
import pytest

import oracle as O
from amdflow import TV_L1, OpticalFlowError, synth
from amdflow import debug as D

pytestmark = pytest.mark.gpu

TOL_PX = 1e-3


def rng_img(rng, h, w, lo=0.0, hi=255.0):
    return (rng.random((h, w)) * (hi - lo) + lo).astype(np.float32)


def test_hypot_matches_glibc():
    rng = np.random.default_rng(0)
    n = 1 << 20
    a = (rng.standard_normal(n) * rng.choice([1e-3, 1.0, 30.0, 1e4], n)).astype(np.float32)
    b = (rng.standard_normal(n) * rng.choice([1e-3, 1.0, 30.0, 1e4], n)).astype(np.float32)
    bits = rng.integers(0, 2**32, n // 4, dtype=np.uint64).astype(np.uint32).view(np.float32)
    bits2 = rng.integers(0, 2**32, n // 4, dtype=np.uint64).astype(np.uint32).view(np.float32)
    ok = np.isfinite(bits) & np.isfinite(bits2)
    a = np.concatenate([a, bits[ok]]); b = np.concatenate([b, bits2[ok]])
    with np.errstate(over="ignore"):   # random bit patterns include overflowing pairs
        ref = np.hypot(a, b)   # numpy float32 hypot -> C hypotf
    got = D.hypot(a, b)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_hypot_at_float_rounding_boundaries():
    """Pairs whose hypot is exactly a midpoint between two floats (Pythagorean
    triples with an odd 25-bit hypotenuse, scaled by powers of two) and the
    same pairs moved by one ulp: where rounding the double square root to
    float must break the tie to even, as glibc's hypotf does."""
    rng = np.random.default_rng(7)
    u = rng.integers(2900, 4097, 200000)
    v = rng.integers(1, 2900, 200000)
    A = u * u - v * v
    B = 2 * u * v
    M = u * u + v * v
    keep = (A > 0) & (A < 2 ** 24) & (B < 2 ** 24) & (M >= 2 ** 24) & (M < 2 ** 25) & (M % 2 == 1)
    A, B = A[keep], B[keep]
    k = rng.integers(-60, 61, A.size).astype(np.float64)
    a = (A * 2.0 ** k).astype(np.float32)
    b = (B * 2.0 ** k).astype(np.float32)
    a = np.concatenate([a, np.nextafter(a, np.float32(np.inf)), np.nextafter(a, np.float32(0)), b])
    b = np.concatenate([b, b, b, a[:b.size]])
    assert a.size > 20000
    ref = np.hypot(a, b)
    got = D.hypot(a, b)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("h,w", [(37, 53), (120, 68), (135, 240), (6, 7), (1080, 1920)])
@pytest.mark.parametrize("sigma", [float(np.float32(0.8)), float(np.float32(1.0)),
                                   float(np.float32(0.6) * np.sqrt(3.0))])
def test_gaussian_bitexact(h, w, sigma):
    rng = np.random.default_rng(h * w)
    img = rng_img(rng, h, w)
    assert np.array_equal(D.gaussian(img, sigma), O.gaussian(img, sigma))


@pytest.mark.parametrize("factor", [0.5, 0.6])
@pytest.mark.parametrize("h,w", [(68, 120), (61, 97), (1080, 1920), (7, 9), (13, 300)])
def test_zoom_out_bitexact(h, w, factor):
    """Factor 1/2 runs the tiled blur-and-decimate kernel, other factors the
    two passes and the bicubic resampling."""
    rng = np.random.default_rng(7)
    img = rng_img(rng, h, w)
    assert np.array_equal(D.zoom_out(img, factor), O.zoom_out(img, factor))


@pytest.mark.parametrize("src,dst", [((120, 68), (240, 135)), ((17, 13), (33, 25)),
                                     ((960, 540), (1920, 1080)),
                                     # factors 1.33 and 4, 1 (integer taps) and below 1
                                     ((300, 200), (400, 267)), ((50, 30), (200, 120)),
                                     ((70, 9), (70, 9)), ((400, 300), (130, 90))])
def test_zoom_in_bitexact(src, dst):
    rng = np.random.default_rng(3)
    img = (rng.standard_normal((src[1], src[0])) * 4).astype(np.float32)
    got = D.zoom_in(img, dst[0], dst[1], 2.0)
    want = np.float32(2.0) * O.zoom_in(img, dst[0], dst[1])
    assert np.array_equal(got, want)


def test_centered_gradient_bitexact():
    rng = np.random.default_rng(5)
    img = rng_img(rng, 77, 131)
    gx, gy = D.centered_gradient(img)
    ox, oy = O.centered_gradient(img)
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy)


@pytest.mark.parametrize("lo,hi", [(0.0, 255.0), (-40.0, 300.0), (12.0, 13.0)])
def test_normalize_bitexact(lo, hi):
    rng = np.random.default_rng(11)
    a, b = rng_img(rng, 50, 70, lo, hi), rng_img(rng, 50, 70, lo, hi)
    ga, gb = D.normalize(a, b)
    oa, ob = O.image_normalization(a, b)
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)


def test_normalize_constant_frames_is_noop():
    a = np.full((20, 30), 7.0, np.float32)
    ga, gb = D.normalize(a, a)
    assert np.array_equal(ga, a) and np.array_equal(gb, a)


@pytest.mark.parametrize("scale", [0.5, 3.0, 12.0])
def test_warp_setup_bitexact(scale):
    rng = np.random.default_rng(int(scale * 10))
    h, w = 90, 140
    I0, I1 = rng_img(rng, h, w), rng_img(rng, h, w)
    I1x, I1y = O.centered_gradient(I1)
    u1 = (rng.standard_normal((h, w)) * scale).astype(np.float32)
    u2 = (rng.standard_normal((h, w)) * scale).astype(np.float32)
    got = D.warp_setup(I0, I1, I1x, I1y, u1, u2)
    _, wx, wy, rc = O.tvl1_warp_setup(I0, I1, I1x, I1y, u1, u2)
    for g, o in zip(got, (wx, wy, rc)):
        assert np.array_equal(g, o)


def _random_state(rng, h, w):
    I0, I1 = rng_img(rng, h, w), rng_img(rng, h, w)
    I1x, I1y = O.centered_gradient(O.gaussian(I1, 0.8))
    u1 = (rng.standard_normal((h, w)) * 0.7).astype(np.float32)
    u2 = (rng.standard_normal((h, w)) * 0.7).astype(np.float32)
    _, wx, wy, rc = O.tvl1_warp_setup(I0, O.gaussian(I1, 0.8), I1x, I1y, u1, u2)
    p = [(rng.standard_normal((h, w)) * 0.3).astype(np.float32) for _ in range(4)]
    return wx, wy, rc, u1, u2, p


@pytest.mark.parametrize("K,R", [(1, 2), (2, 2), (1, 4), (2, 4), (4, 4), (1, 8), (2, 8),
                                 (4, 8), (8, 8), (1, 0), (2, 0), (4, 0)])
@pytest.mark.parametrize("h,w", [(45, 70), (130, 257)])
def test_iterations_bitexact(K, R, h, w):
    rng = np.random.default_rng(K * 100 + R + h)
    wx, wy, rc, u1, u2, p = _random_state(rng, h, w)
    iters = 7
    (gu1, gu2, g11, g12, g21, g22), gerr = D.tvl1_iterations(wx, wy, rc, u1, u2, *p, iters, K, R)
    st = (u1, u2, *p)
    errs = []
    for _ in range(iters):
        st, e = O.tvl1_iteration(wx, wy, rc, *st)
        errs.append(e)
    for g, o in zip((gu1, gu2, g11, g12, g21, g22), st):
        assert np.array_equal(g, o)
    np.testing.assert_allclose(gerr, np.array(errs), rtol=1e-12, atol=0)


def _solve_both(q0, q1, **kw):
    h, w = q0.shape
    with TV_L1(w, h, **kw) as tv:
        tv.runDualTVL1Multiscale(q0, q1)
        u = tv.getU()
        it = tv.iterations()
    p = O.default_params(**{("lambda_" if k == "lambda_" else k): v for k, v in kw.items()})
    uo, ito = O.tvl1(q0, q1, p)
    return u, it, uo, ito


@pytest.mark.parametrize("w,h,cap,seed", [(160, 120, 3.0, 1), (97, 61, 2.0, 2),
                                          (256, 192, 6.0, 3), (320, 240, 8.0, 4)])
def test_tvl1_pair_matches_oracle(w, h, cap, seed):
    I0, I1, _, _ = synth.make_pair(w, h, cap, seed, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    u, it, uo, ito = _solve_both(q0, q1)
    assert np.array_equal(it, ito), (it, ito)
    assert np.max(np.abs(u - uo)) <= TOL_PX
    assert np.array_equal(u, uo)


def test_tvl1_float_input_matches_oracle():
    I0, I1, _, _ = synth.make_pair(200, 150, 4.0, 9, bumps=True)
    with TV_L1(200, 150) as tv:
        tv.runDualTVL1Multiscale(I0, I1)
        u = tv.getU()
        it = tv.iterations()
    uo, ito = O.tvl1(I0, I1)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_tvl1_nondefault_params_match_oracle():
    I0, I1, _, _ = synth.make_pair(180, 140, 5.0, 21)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    kw = dict(tau=0.2, lambda_=0.3, theta=0.25, nscales=3, zfactor=0.6, warps=3, epsilon=0.02,
              max_iterations=40)
    u, it, uo, ito = _solve_both(q0, q1, **kw)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_tvl1_iteration_cap_reached():
    I0, I1, _, _ = synth.make_pair(128, 96, 4.0, 5)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    u, it, uo, ito = _solve_both(q0, q1, epsilon=1e-6, max_iterations=7)
    assert (it == 7).all() and np.array_equal(it, ito)
    assert np.array_equal(u, uo)


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("K,R", [(1, 2), (2, 2), (1, 4), (2, 4), (4, 4), (1, 8), (2, 8), (4, 8),
                                 (8, 8)])
def test_tvl1_launch_shapes_agree(K, R, persistent):
    _shapes_agree(K, R, persistent)


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("K", [1, 2, 4])
def test_tvl1_square_tiles_agree(K, persistent):
    """Square 16x16 tiles (rows_per_wave 0) at every scale."""
    _shapes_agree(K, 0, persistent)


@pytest.mark.parametrize("K", [2, 4])
def test_tvl1_square_persistent_1080p(K):
    """Persistent square tiles at every scale of a 1920x1080 pair: one
    workgroup per tile is more than the device holds at the finest scale, so
    workgroups take several tiles per chunk."""
    I0, I1, _, _ = synth.make_pair(1920, 1080, 20.0, 2024, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    with TV_L1(1920, 1080, nscales=5) as tv:
        tv.set_launch_rule(0, K, 0, True)
        tv.runDualTVL1Multiscale(q0, q1)
        u, it = tv.getU(), tv.iterations()
    uo, ito = O.tvl1(q0, q1, O.default_params(nscales=5))
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def _shapes_agree(K, R, persistent):
    I0, I1, _, _ = synth.make_pair(240, 180, 5.0, 13, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    with TV_L1(240, 180) as tv:
        tv.set_launch_rule(0, K, R, persistent)
        tv.runDualTVL1Multiscale(q0, q1)
        u, it = tv.getU(), tv.iterations()
    uo, ito = O.tvl1(q0, q1)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


@pytest.mark.parametrize("K,R", [(2, 8), (4, 4), (1, 2), (2, 4)])
def test_tvl1_persistent_everywhere_1080p(K, R):
    """The whole 1920x1080 solve with persistent warpings at every scale
    (at 1920x1080 each workgroup loops over several tiles; (1,2) and (2,4)
    would fit 8 workgroups per CU, the grid is capped at 4)."""
    cfg = synth.CONFIGS["tvl1_1920x1080"]
    I0, I1, _, _ = synth.make_pair(cfg["width"], cfg["height"], cfg["cap"], 2024, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    with TV_L1(cfg["width"], cfg["height"], nscales=5) as tv:
        tv.set_launch_rule(0, K, R, True)
        tv.runDualTVL1Multiscale(q0, q1)
        u, it = tv.getU(), tv.iterations()
    uo, ito = O.tvl1(q0, q1, O.default_params(nscales=5))
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_tvl1_identical_frames_zero_flow():
    I0, _, _, _ = synth.make_pair(300, 200, 1.0, 8)
    q = synth.quantize(I0)
    with TV_L1(300, 200) as tv:
        tv.runDualTVL1Multiscale(q, q)
        assert np.all(tv.getU() == 0.0)


def test_tvl1_constant_frames_zero_flow():
    q = np.full((64, 80), 100, np.uint8)
    with TV_L1(80, 64) as tv:
        tv.runDualTVL1Multiscale(q, q)
        assert np.all(tv.getU() == 0.0)


def test_tvl1_streaming_push():
    frames = [synth.make_pair(160, 120, 2.0, 30 + i)[0] for i in range(3)]
    with TV_L1(160, 120) as tv:
        tv.runDualTVL1Multiscale(frames[0])
        assert np.all(tv.getU() == 0.0)
        for k in (1, 2):
            tv.runDualTVL1Multiscale(frames[k])
            uo, _ = O.tvl1(frames[k - 1], frames[k])
            assert np.array_equal(tv.getU(), uo)


def test_tvl1_repeat_is_deterministic():
    I0, I1, _, _ = synth.make_pair(200, 160, 4.0, 17, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    with TV_L1(200, 160) as tv:
        tv.runDualTVL1Multiscale(q0, q1)
        a = tv.getU()
        tv.runDualTVL1Multiscale(q0, q1)
        assert np.array_equal(a, tv.getU())


def test_tvl1_1080p_matches_oracle():
    """The 1920x1080 workload (5 scales, 20 px motion, noise sigma 1)."""
    cfg = synth.CONFIGS["tvl1_1920x1080"]
    I0, I1, fu, fv = synth.make_pair(cfg["width"], cfg["height"], cfg["cap"], 2024,
                                     bumps=cfg["bumps"])
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    u, it, uo, ito = _solve_both(q0, q1, nscales=5)
    assert it.shape == (5, 5)
    assert np.array_equal(it, ito), (it, ito)
    d = np.max(np.abs(u - uo))
    assert d <= TOL_PX
    assert np.array_equal(u, uo)


def test_sigma_too_large_reported():
    with pytest.raises(OpticalFlowError) as e:
        TV_L1(4, 40, nscales=1)
    assert e.value.status == 2


@pytest.mark.parametrize("w,h,kw", [
    (16, 12, {}),                                  # one scale by the nscales rule
    (300, 48, {}),                                 # 5 scales down to 19x3, thin image
    (48, 300, {}),                                 # the same, transposed
    (65, 33, {}),                                  # odd sizes, 3 scales
    (96, 64, dict(nscales=1, warps=1)),            # one scale, one warping
    (130, 70, dict(zfactor=0.75, nscales=6, warps=2)),
])
def test_tvl1_edge_shapes_match_oracle(w, h, kw):
    """Thin, tiny and odd frames, and single-scale / single-warping solves:
    same flow and iteration counts as the oracle."""
    I0, I1, _, _ = synth.make_pair(w, h, 2.0, 50 + w + h, bumps=False)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    u, it, uo, ito = _solve_both(q0, q1, **kw)
    assert np.array_equal(it, ito), (it, ito)
    assert np.array_equal(u, uo)
