"""Pins the float64 network oracle used to check the GPU warm-start kernels:
its convolution against scipy's correlate, its gradients against central
finite differences, and the bfloat16 rounding against exact cases."""
import numpy as np
import pytest
from scipy import ndimage

import cnn_oracle as N


def test_bf16_round_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0, np.inf],
                 np.float32)
    r = N.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 -> 1 + 2^-6
    assert r.tolist() == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, np.inf]


def test_conv_layer_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    B, H, W = 2, 7, 9
    X = rng.normal(size=(B, H, W, 5))
    params = N.init_params(0, seed=3)
    (W0, b0), _ = N.unflatten(params, 0)
    acts = N.forward(params, X, 0)
    for b in range(B):
        for n in range(4):
            ref = np.full((H, W), b0[n])
            for c in range(5):
                k = W0[n, :, c].reshape(3, 3)
                ref += ndimage.correlate(X[b, :, :, c], k, mode="constant", cval=0.0)
            np.testing.assert_allclose(acts[0][b, :, :, n], np.maximum(ref, 0.0), rtol=1e-12,
                                       atol=1e-12)


@pytest.mark.parametrize("hidden", [0, 1])
def test_gradients_match_finite_differences(hidden):
    rng = np.random.default_rng(7 + hidden)
    B, H, W = 2, 5, 6
    X = rng.normal(size=(B, H, W, 5))
    T = rng.normal(size=(B, 2, H, W))
    params = N.init_params(hidden, seed=11).astype(np.float64)
    # larger head weights so that the loss depends on every layer visibly
    loss, g = N.loss_and_grads(params, X, T, hidden)
    idx = rng.choice(len(params), 60, replace=False)
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        l1, _ = N.loss_and_grads(p1, X, T, hidden)
        l2, _ = N.loss_and_grads(p2, X, T, hidden)
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_bf16_emulation_close_to_float64():
    rng = np.random.default_rng(5)
    X = rng.normal(size=(1, 6, 6, 5))
    T = rng.normal(size=(1, 2, 6, 6))
    params = N.init_params(1, seed=2)
    l64, g64 = N.loss_and_grads(params, X, T, 1)
    l16, g16 = N.loss_and_grads(params, X, T, 1, bf16=True)
    assert abs(l16 - l64) <= 2e-2 * l64
    assert np.linalg.norm(g16 - g64) <= 3e-2 * np.linalg.norm(g64)
