"""NumPy restatement of the reference application's flow rendering
(App::flowToColor, src/tvl1/sequential/main.cpp:93-126) for checking the
GPU kernel.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product library.

The reference converts with OpenCV (absent here): cv::Mat::convertTo to
CV_8U rounds to nearest even and saturates, and cv::cvtColor(HSV2BGR) on
8-bit data scales the hue by 2 degrees and applies the hexcone HSV -> RGB
model.  Both are restated from their definitions in float32; OpenCV's own
bits are not available to pin against ("parity unpinned" for the last
rounding step of the colour conversion).
"""
from __future__ import annotations

import numpy as np


def _sat_u8(x: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(x), 0, 255).astype(np.uint8)


def hsv_to_bgr(H, S, V) -> np.ndarray:
    f32 = np.float32
    h6 = H.astype(f32) * f32(6.0 / 180.0)
    s = S.astype(f32) * f32(1.0 / 255.0)
    v = V.astype(f32) * f32(1.0 / 255.0)
    sector = np.floor(h6)
    f = (h6 - sector).astype(f32)
    k = sector.astype(np.int64) % 6
    one = f32(1.0)
    p = v * (one - s)
    q = v * (one - s * f)
    t = v * (one - s * (one - f))
    r = np.choose(k, [v, q, p, p, t, v])
    g = np.choose(k, [t, v, v, q, p, p])
    b = np.choose(k, [p, p, t, v, v, q])
    c = f32(255.0)
    return np.stack([_sat_u8(b * c), _sat_u8(g * c), _sat_u8(r * c)], -1)


def flow_to_color(U: np.ndarray, width: int, height: int) -> np.ndarray:
    f32 = np.float32
    U = np.asarray(U, f32)
    n = width * height
    u, v = U[:n], U[n:2 * n]
    mag = np.sqrt(u * u + v * v)
    val = np.minimum(mag * f32(10.0), f32(255.0))
    hue = np.arctan2(v, u).astype(f32) * f32(180.0 / np.pi / 2.0)
    H, V = _sat_u8(hue), _sat_u8(val)
    S = np.full_like(H, 255)
    return hsv_to_bgr(H, S, V).reshape(height, width, 3)
