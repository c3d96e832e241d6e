"""The oracle's TV-L1 dual update, u update, stopping rule and the HS
pyramid steps against closed forms (tests/pins.py), and the stopping rule's
accumulation order against the reference's float sum."""
import numpy as np
import pytest

import oracle as O
import pins
from amdflow import synth


@pytest.mark.parametrize("name", list(pins.CHECKS))
def test_oracle_closed_form(name):
    pins.CHECKS[name](pins.OracleBackend())


def _solve(q0, q1, mode):
    O.set_error_mode(mode)
    try:
        return O.tvl1(q0, q1)
    finally:
        O.set_error_mode(O.ERROR_DOUBLE_ROWS)


@pytest.mark.parametrize("w,h,cap,seed", [(320, 240, 4.0, 11), (640, 480, 6.0, 12)])
def test_stopping_sum_order_against_reference_float_sum(w, h, cap, seed):
    """The GPU (and the default oracle) sum the stopping error in double in a
    fixed order; the reference keeps a float (tvl1.cpp:252-265).  Replaying
    the reference's float sum in index order (its one-thread order) must give
    the same stopping decision at every warping of these solves, hence the
    same flow.  Where the two sums straddle epsilon^2 they would differ: that
    is the residual, unpinned part of the rule, and the bench reports the
    count at the full sizes."""
    I0, I1, _, _ = synth.make_pair(w, h, cap, seed=seed, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    ud, itd = _solve(q0, q1, O.ERROR_DOUBLE_ROWS)
    uf, itf = _solve(q0, q1, O.ERROR_FLOAT_SEQUENTIAL)
    assert np.array_equal(itd, itf), f"stopping decisions differ: {np.argwhere(itd != itf)}"
    assert np.array_equal(ud, uf)


def test_float_mode_replays_the_float_sum():
    """In float mode the returned sum is the float accumulation of the same
    per-pixel terms (checked here against numpy's float32 cumulative sum)."""
    rng = np.random.default_rng(3)
    h, w = 17, 23
    wx, wy, rc = [(rng.standard_normal((h, w)) * s).astype(np.float32) for s in (3, 3, 20)]
    u1, u2 = [(rng.standard_normal((h, w)) * 2).astype(np.float32) for _ in range(2)]
    p = [(rng.standard_normal((h, w)) * 0.2).astype(np.float32) for _ in range(4)]
    O.set_error_mode(O.ERROR_FLOAT_SEQUENTIAL)
    try:
        (n1, n2, *_), ef = O.tvl1_iteration(wx, wy, rc, u1, u2, *p)
    finally:
        O.set_error_mode(O.ERROR_DOUBLE_ROWS)
    (m1, m2, *_), ed = O.tvl1_iteration(wx, wy, rc, u1, u2, *p)
    assert np.array_equal(n1, m1) and np.array_equal(n2, m2)
    d1, d2 = (n1 - u1).ravel(), (n2 - u2).ravel()
    terms = d1 * d1 + d2 * d2
    acc = np.float32(0)
    for t in terms:
        acc = np.float32(acc + t)
    assert ef == float(acc)
    assert abs(ed - ef) <= 1e-5 * ed
