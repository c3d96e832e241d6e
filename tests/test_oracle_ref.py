"""The oracle's restatement of the bicubic sampler and zoom_in against the
reference's own sources (bicubic_interpolation.cpp, zoom.cpp) compiled into
oracle/_ref/libref.so.  Bit-for-bit equality is required.  Skipped when the
reference build is absent (it needs the reference tree at build time)."""
import numpy as np
import pytest

import oracle as O

ref = O.ref
pytestmark = pytest.mark.skipif(not ref.available, reason="oracle/_ref not built")


def test_bicubic_at_random_points_bitexact():
    rng = np.random.default_rng(0)
    for (h, w) in [(5, 7), (31, 17), (64, 96)]:
        img = (rng.random((h, w)) * 255).astype(np.float32)
        pts = np.concatenate([
            rng.uniform(-3, w + 3, 4000), rng.uniform(-3, h + 3, 4000)]).reshape(2, -1).T
        for border in (True, False):
            for x, y in pts:
                x, y = np.float32(x), np.float32(y)
                a = O.bicubic_at(img, x, y, border)
                b = ref.bicubic_at(img, x, y, border)
                assert np.float32(a).tobytes() == np.float32(b).tobytes(), (x, y, border)


def test_bicubic_at_integer_and_edge_coordinates_bitexact():
    rng = np.random.default_rng(1)
    img = (rng.standard_normal((12, 15)) * 30).astype(np.float32)
    coords = [np.float32(v) for v in
              (-1.0, -0.5, -0.0, 0.0, 0.25, 1.0, 1.5, 13.0, 13.99, 14.0, 14.5, 15.0, 11.0, 10.5)]
    for x in coords:
        for y in coords:
            for border in (True, False):
                a, b = O.bicubic_at(img, x, y, border), ref.bicubic_at(img, x, y, border)
                assert np.float32(a).tobytes() == np.float32(b).tobytes(), (x, y, border)


@pytest.mark.parametrize("scale", [0.3, 2.0, 9.0])
def test_bicubic_warp_bitexact(scale):
    rng = np.random.default_rng(int(scale * 7))
    h, w = 47, 83
    img = (rng.random((h, w)) * 255).astype(np.float32)
    u = (rng.standard_normal((h, w)) * scale).astype(np.float32)
    v = (rng.standard_normal((h, w)) * scale).astype(np.float32)
    for border in (True, False):
        assert np.array_equal(O.bicubic_warp(img, u, v, border), ref.bicubic_warp(img, u, v, border))


@pytest.mark.parametrize("src,dst", [((120, 68), (240, 135)), ((17, 13), (33, 25)),
                                     ((30, 17), (60, 34)), ((240, 135), (480, 270))])
def test_zoom_in_bitexact(src, dst):
    rng = np.random.default_rng(5)
    img = (rng.standard_normal((src[1], src[0])) * 3).astype(np.float32)
    assert np.array_equal(O.zoom_in(img, *dst), ref.zoom_in(img, *dst))


def test_zoom_size_matches():
    for n in [1080, 540, 270, 135, 68, 1920, 960, 480, 240, 120, 7, 5, 3, 2, 1, 333]:
        for f in [0.5, 0.6, 0.75, 0.9]:
            assert O.zoom_size(n, n + 1, f) == ref.zoom_size(n, n + 1, f)
