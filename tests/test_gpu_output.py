"""GPU flow rendering (flowToColor of src/tvl1/sequential/main.cpp:93-126)
against the numpy restatement: identical except where atan2f's last bit
moves the rounded hue (rare), and then by one hue step."""
import numpy as np
import pytest

import flow_color as FC
from amdflow import io

pytestmark = pytest.mark.gpu


def test_flow_to_color_matches_oracle():
    rng = np.random.default_rng(0)
    w, h = 333, 211
    U = (rng.normal(size=2 * w * h) * rng.choice([0.5, 5.0, 40.0], 2 * w * h)).astype(np.float32)
    U[:5] = 0.0
    U[w * h:w * h + 5] = 0.0
    got = io.flow_to_color(U, w, h).astype(int)
    ref = FC.flow_to_color(U, w, h).astype(int)
    diff = np.abs(got - ref).max(-1)
    assert (diff > 0).mean() <= 1e-3
    # one hue step moves a channel by at most 255 * 6 / 180 (+1 rounding)
    assert diff.max() <= 10


def test_flow_to_color_device_matches_host():
    import torch
    rng = np.random.default_rng(1)
    w, h = 64, 48
    U = rng.normal(size=2 * w * h).astype(np.float32) * 10
    dU = torch.from_numpy(U).cuda()
    dB = torch.empty(3 * w * h, dtype=torch.uint8, device="cuda")
    io.flow_to_color_device(dU.data_ptr(), w, h, dB.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dB.cpu().numpy().reshape(h, w, 3), io.flow_to_color(U, w, h))
