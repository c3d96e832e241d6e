"""Frames holding NaN or infinities: the GPU gives the oracle's bits.

Through the float entry points a frame can hold non-finite values.  The
oracle's behaviour on them is pinned on the CPU (test_oracle_nonfinite.py):
the normalisation bounds follow the BLAS scan (a NaN first pixel turns
every pixel NaN, a NaN elsewhere is not picked), a NaN stopping error ends
a warping, and samplers clamp every tap of a non-finite coordinate.  Here
the GPU solvers must match it bit for bit, NaN positions included, and
every call must return normally."""
import numpy as np
import pytest

import oracle as O
from amdflow import TV_L1, HornSchunck, lk_pyramidal, synth

pytestmark = pytest.mark.gpu

CASES = {
    "nan_inside": lambda a, b: (a.__setitem__((40, 50), np.nan), b),
    "nan_first": lambda a, b: (a.__setitem__((0, 0), np.nan), b),
    "inf_inside": lambda a, b: (a.__setitem__((10, 10), np.inf), b),
    "ninf_first_of_second": lambda a, b: (a, b.__setitem__((0, 0), -np.inf)),
    "nan_both": lambda a, b: (a.__setitem__((0, 0), np.nan), b.__setitem__((30, 20), np.nan)),
}


def _frames(case, w=96, h=80, seed=7):
    I0, I1, _, _ = synth.make_pair(w, h, 3.0, seed, bumps=True)
    a, b = I0.astype(np.float32).copy(), I1.astype(np.float32).copy()
    CASES[case](a, b)
    return a, b


@pytest.mark.parametrize("case", list(CASES))
def test_tvl1_float_frames(case):
    a, b = _frames(case)
    with TV_L1(96, 80, nscales=3) as tv:
        tv.runDualTVL1Multiscale(a, b)
        u, it = tv.getU().copy(), tv.iterations()
    uo, ito = O.tvl1(a, b, O.default_params(nscales=3))
    print(case, "iterations", it.sum(), "NaN share", float(np.isnan(u).mean()), flush=True)
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo, equal_nan=True)


@pytest.mark.parametrize("case", list(CASES))
def test_hs_float_frames(case):
    a, b = _frames(case)
    with HornSchunck(2, 96, 96, 80) as hs:
        u, v = hs.ComputeFlow(a, b, 96, 80, 96, 15.0, 2, 2, 30)
    uo, vo = O.hs_compute(a, b, 15.0, 2, 2, 30)
    assert np.array_equal(np.reshape(u, (80, 96)), uo, equal_nan=True)
    assert np.array_equal(np.reshape(v, (80, 96)), vo, equal_nan=True)


@pytest.mark.parametrize("case", list(CASES))
def test_lk_pyramidal_float_frames(case):
    a, b = _frames(case)
    u, v = lk_pyramidal(a, b, levels=3, zf=0.5, sigma=1.0, window=7, iters=5, det_eps=1e-6)
    uo, vo = O.lk_pyr(a, b, levels=3, zf=0.5, sigma=1.0, win=7, iters=5, det_eps=1e-6)
    assert np.array_equal(u, uo, equal_nan=True)
    assert np.array_equal(v, vo, equal_nan=True)
