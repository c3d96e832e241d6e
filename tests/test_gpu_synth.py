"""The GPU pair generator (training and measurement inputs): deterministic
per seed, motion scaled to the cap, and frame1 a warp of frame0."""
import numpy as np
import pytest

from amdflow import synth

pytestmark = pytest.mark.gpu


def test_generator_deterministic_and_capped():
    import torch
    g = synth.DevicePairs(320, 240, 8.0, bumps=True)
    a0, a1, f = [t.cpu().numpy().copy() for t in g.generate(123)]
    b0, b1, _ = [t.cpu().numpy().copy() for t in g.generate(123)]
    c0, _, _ = [t.cpu().numpy().copy() for t in g.generate(124)]
    torch.cuda.synchronize()
    assert np.array_equal(a0, b0) and np.array_equal(a1, b1)
    assert not np.array_equal(a0, c0)
    n = 320 * 240
    mag = np.hypot(f[:n], f[n:])
    assert abs(mag.max() - 8.0) < 1e-3
    assert a0.min() <= 1 and a0.max() >= 254          # rescaled to the full range


def test_frame1_is_warped_frame0():
    """Zero motion: frame1 = frame0 + N(0, 1) before rounding."""
    import torch
    g = synth.DevicePairs(256, 192, 0.0, bumps=False, noise=1.0)
    a0, a1, f = [t.cpu().numpy().astype(np.float64) for t in g.generate(7)]
    torch.cuda.synchronize()
    assert not f.any()
    d = a1 - a0
    assert abs(d.mean()) < 0.05 and 0.9 < d.std() < 1.2
