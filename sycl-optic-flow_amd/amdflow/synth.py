"""Seeded synthetic frame pairs of known motion.

frame0: uniform noise, Gaussian blur sigma=2, affinely rescaled to [0, 255].
flow:   sum of 4 random 2-D sinusoids (periods 64-256 px, random orientation
        and phase) plus, optionally, 3-6 Gaussian bumps (radius 40-200 px)
        of random displacement; scaled so that max |flow| equals `cap`.
frame1: frame0 sampled bilinearly at x - flow(x) (clamp to edge) plus
        N(0, sigma_noise) noise.  A solver that samples I1 at x + u (as the
        TV-L1 warp does, bicubic_interpolation.cpp:155-160) recovers u ~ flow.
`quantize` rounds and clamps to uint8, the reference pair API's input type.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def make_pair(width: int, height: int, cap: float, seed: int, bumps: bool = False,
              noise: float = 1.0):
    rng = np.random.default_rng(seed)
    base = rng.random((height, width))
    base = ndimage.gaussian_filter(base, 2.0, mode="reflect")
    base = (base - base.min()) / max(base.max() - base.min(), 1e-12) * 255.0

    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    fu = np.zeros((height, width))
    fv = np.zeros((height, width))
    for comp in (fu, fv):
        for _ in range(4):
            period = rng.uniform(64.0, 256.0)
            ang = rng.uniform(0.0, 2 * np.pi)
            phase = rng.uniform(0.0, 2 * np.pi)
            amp = rng.uniform(0.5, 1.0)
            comp += amp * np.sin(2 * np.pi * (np.cos(ang) * xx + np.sin(ang) * yy) / period + phase)
    if bumps:
        for _ in range(int(rng.integers(3, 7))):
            cx, cy = rng.uniform(0, width), rng.uniform(0, height)
            r = rng.uniform(40.0, 200.0)
            du, dv = rng.uniform(-1.5, 1.5, size=2)
            g = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
            fu += du * g
            fv += dv * g
    mag = np.sqrt(fu * fu + fv * fv).max()
    s = cap / mag if mag > 0 else 0.0
    fu *= s
    fv *= s

    coords = np.stack([yy - fv, xx - fu])
    f1 = ndimage.map_coordinates(base, coords, order=1, mode="nearest")
    f1 = f1 + rng.normal(0.0, noise, size=f1.shape)
    return base.astype(np.float32), f1.astype(np.float32), fu.astype(np.float32), fv.astype(np.float32)


def quantize(img: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


# Workloads of the benchmark configurations (sizes and motion caps).
CONFIGS = {
    "hs_640x480": dict(width=640, height=480, cap=2.0, bumps=False),
    "lk_1280x720": dict(width=1280, height=720, cap=8.0, bumps=False),
    "tvl1_1920x1080": dict(width=1920, height=1080, cap=20.0, bumps=True),
}


def draw_params(rng: np.random.Generator, width: int, height: int, bumps: bool = True):
    """Random motion parameters in the layout of of_synth_pair: per flow
    component 4 x (period, angle, phase, amplitude), then 3-6 bumps of
    (cx, cy, radius, du, dv) when `bumps`.  Returns (float32[32 + 5*nb], nb)."""
    p = []
    for _ in range(2):
        for _ in range(4):
            p += [rng.uniform(64.0, 256.0), rng.uniform(0.0, 2 * np.pi), rng.uniform(0.0, 2 * np.pi),
                  rng.uniform(0.5, 1.0)]
    nb = int(rng.integers(3, 7)) if bumps else 0
    for _ in range(nb):
        p += [rng.uniform(0, width), rng.uniform(0, height), rng.uniform(40.0, 200.0),
              *rng.uniform(-1.5, 1.5, size=2)]
    return np.asarray(p, np.float32), nb


class DevicePairs:
    """Synthetic u8 pairs generated on the GPU (of_synth_pair): same
    definition as make_pair, its own random stream.  Buffers are torch
    tensors on the current device, reused across calls."""

    def __init__(self, width: int, height: int, cap: float, bumps: bool = True, noise: float = 1.0):
        import torch

        from . import _lib
        self._lib = _lib
        _lib.of_synth_pair.restype = __import__("ctypes").c_int
        self.width, self.height, self.cap, self.bumps, self.noise = width, height, cap, bumps, noise
        n = width * height
        self.I0 = torch.empty(n, dtype=torch.uint8, device="cuda")
        self.I1 = torch.empty(n, dtype=torch.uint8, device="cuda")
        self.flow = torch.empty(2 * n, dtype=torch.float32, device="cuda")
        self._scratch = torch.empty(2 * n + 4, dtype=torch.float32, device="cuda")

    def generate(self, seed: int, stream: int = 0):
        """Fills (I0, I1, flow) for `seed`; asynchronous on `stream`."""
        import ctypes as C

        from . import _check
        rng = np.random.default_rng(seed)
        params, nb = draw_params(rng, self.width, self.height, self.bumps)
        _check(self._lib.of_synth_pair(
            C.c_int(self.width), C.c_int(self.height), C.c_float(self.cap), C.c_float(self.noise),
            C.c_ulonglong(seed), params.ctypes.data_as(C.c_void_p), C.c_int(nb),
            C.c_void_p(self.I0.data_ptr()), C.c_void_p(self.I1.data_ptr()),
            C.c_void_p(self.flow.data_ptr()), C.c_void_p(self._scratch.data_ptr()),
            C.c_void_p(stream or None)), "synth_pair")
        return self.I0, self.I1, self.flow
