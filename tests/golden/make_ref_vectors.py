"""Writes tests/golden/ref_bicubic_zoom.npz: input/output vectors of the
reference's own bicubic sampler, bicubic warp, zoom_in and zoom_size
(src/tvl1/sequential/bicubic_interpolation/bicubic_interpolation.cpp,
zoom/zoom.cpp), produced by oracle/_ref/libref.so, which `make -C oracle ref`
compiles from those sources where the reference tree is present.

The file holds data only (seeded inputs and the reference's outputs), so the
checks that use it run wherever the reference tree is absent.

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "ref_bicubic_zoom.npz")

# bicubic_at: image shapes (h, w) and random points per image, plus a grid of
# integer, half-integer and border coordinates
AT_SHAPES = [(5, 7), (31, 17), (64, 96)]
AT_POINTS = 1500
# bicubic_warp: one image, flows of three magnitudes (px)
WARP_SHAPE = (29, 41)
WARP_SCALES = [0.3, 2.0, 9.0]
# zoom_in: (src w, h) -> (dst w, h), as between consecutive scales at zfactor 0.5
ZOOM_CASES = [((30, 17), (60, 34)), ((17, 13), (33, 25)), ((60, 34), (120, 68))]
ZOOM_SIZES = [1080, 540, 270, 135, 68, 1920, 960, 480, 240, 120, 436, 1024, 7, 5, 3, 2, 1, 333]
ZOOM_FACTORS = [0.5, 0.6, 0.75, 0.9]


def main() -> None:
    ref = O.ref
    if not ref.available:
        sys.exit("oracle/_ref/libref.so is not built: run `make -C oracle ref` with the reference tree present")
    rng = np.random.default_rng(20240607)
    data = {}
    for k, (h, w) in enumerate(AT_SHAPES):
        img = (rng.random((h, w)) * 255).astype(np.float32)
        pts = np.stack([rng.uniform(-3, w + 3, AT_POINTS), rng.uniform(-3, h + 3, AT_POINTS)], 1)
        edge = np.array([-1.0, -0.5, 0.0, 0.25, 1.0, 1.5, w - 2.0, w - 1.5, w - 1.0, w - 0.01, w * 1.0,
                         h - 2.0, h - 1.0, h - 0.5, h * 1.0], np.float64)
        gx, gy = np.meshgrid(edge, edge)
        pts = np.concatenate([pts, np.stack([gx.ravel(), gy.ravel()], 1)]).astype(np.float32)
        for b in (0, 1):
            data[f"at{k}_out_b{b}"] = np.array(
                [ref.bicubic_at(img, np.float32(x), np.float32(y), bool(b)) for x, y in pts], np.float32)
        data[f"at{k}_img"] = img
        data[f"at{k}_pts"] = pts
    h, w = WARP_SHAPE
    img = (rng.random((h, w)) * 255).astype(np.float32)
    data["warp_img"] = img
    for k, sc in enumerate(WARP_SCALES):
        u = (rng.standard_normal((h, w)) * sc).astype(np.float32)
        v = (rng.standard_normal((h, w)) * sc).astype(np.float32)
        data[f"warp{k}_u"], data[f"warp{k}_v"] = u, v
        for b in (0, 1):
            data[f"warp{k}_out_b{b}"] = ref.bicubic_warp(img, u, v, bool(b))
    for k, ((sw, sh), (dw, dh)) in enumerate(ZOOM_CASES):
        for c in (0, 1):   # two flow components
            src = (rng.standard_normal((sh, sw)) * 3).astype(np.float32)
            data[f"zoom{k}_in{c}"] = src
            data[f"zoom{k}_out{c}"] = ref.zoom_in(src, dw, dh)
    data["zoom_cases"] = np.array([[sw, sh, dw, dh] for (sw, sh), (dw, dh) in ZOOM_CASES], np.int32)
    sizes = [[n, n + 1, int(f * 100), *ref.zoom_size(n, n + 1, f)] for n in ZOOM_SIZES for f in ZOOM_FACTORS]
    data["zoom_size"] = np.array(sizes, np.int32)   # nx, ny, factor*100, nxx, nyy
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
