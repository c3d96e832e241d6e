"""The GPU kernels against the same closed forms as the oracle (tests/pins.py):
the TV-L1 dual update, u update and error term in the iteration kernel
(several launch shapes), the stopping rule in the one-scale solver, and the
HS pyramid steps."""
import numpy as np
import pytest

import oracle as O
import pins
from amdflow import synth

pytestmark = pytest.mark.gpu

ITER_CHECKS = ["t17_one_step", "t17_dual_feasible", "t16_u_update", "t16_error_term"]
OTHER_CHECKS = ["t16_stop_threshold", "h6_downscale_ramp", "h6_upscale_constant",
                "h6_upscale_ramp"]


@pytest.mark.parametrize("K,R", [(2, 8), (4, 4), (1, 2), (4, 0)])
@pytest.mark.parametrize("name", ITER_CHECKS)
def test_gpu_iteration_closed_form(name, K, R):
    pins.CHECKS[name](pins.GpuBackend(K, R))


@pytest.mark.parametrize("name", OTHER_CHECKS)
def test_gpu_closed_form(name):
    pins.CHECKS[name](pins.GpuBackend())


@pytest.mark.parametrize("w,h,seed", [(120, 68, 1), (240, 135, 2), (97, 61, 3)])
def test_gpu_one_scale_matches_oracle(w, h, seed):
    """One scale of the GPU solver (gradient, 5 warpings, stopping rule) on
    synthetic float frames equals the oracle's dualTVL1 bit for bit, with the
    same iteration count per warping."""
    I0, I1, _, _ = synth.make_pair(w, h, 2.0, seed=seed, bumps=True)
    I0 = O.gaussian(I0, 0.8)
    I1 = O.gaussian(I1, 0.8)
    z = np.zeros((h, w), np.float32)
    args = dict(tau=0.25, lambda_=0.15, theta=0.3, warps=5, epsilon=0.01, max_iterations=300)
    gu1, gu2, git = pins.GpuBackend().dual(I0, I1, z, z, **args)
    ou1, ou2, oit = pins.OracleBackend().dual(I0, I1, z, z, **args)
    assert list(git) == list(oit)
    assert np.array_equal(gu1, ou1) and np.array_equal(gu2, ou2)
