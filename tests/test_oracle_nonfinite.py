"""The CPU oracle on non-finite input, and the oracle under the sanitizers.

The reference picks the normalisation bounds with cblas_isamax/isamin
(tvl1.cpp:302-329).  The published BLAS scan keeps the first element and
replaces it only on a strictly larger (smaller) |x|: a NaN first element is
therefore kept, a NaN anywhere else is never picked.  These tests pin the
oracle to that scan with numpy's own NaN-aware argmax, pin the bicubic
sampler's handling of non-finite coordinates (C leaves the float-to-int
conversion undefined there; the oracle and the GPU both saturate, NaN giving
0), and run every oracle entry point under AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/sanitize_main.c)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from amdflow import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _blas_pick(x, largest):
    """Index chosen by the BLAS scan: the first element unless a later one is
    strictly larger (smaller) in |x|; a NaN first element stays."""
    ax = np.abs(x.ravel().astype(np.float64))
    if np.isnan(ax[0]):
        return 0
    ax[np.isnan(ax)] = -np.inf if largest else np.inf
    return int(np.argmax(ax) if largest else np.argmin(ax))


def _normalised(I0, I1):
    a0, a1 = I0.ravel()[_blas_pick(I0, True)], I1.ravel()[_blas_pick(I1, True)]
    b0, b1 = I0.ravel()[_blas_pick(I0, False)], I1.ravel()[_blas_pick(I1, False)]
    mx = a1 if a0 < a1 else a0
    mn = b1 if b1 < b0 else b0
    den = np.float32(mx - mn)
    if den <= 0:
        return I0.copy(), I1.copy()
    f = lambda x: (255.0 * (x - np.float32(mn)).astype(np.float64) / np.float64(den)).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        return f(I0), f(I1)


@pytest.mark.parametrize("case", ["nan_first", "nan_inside", "nan_both", "inf_inside", "ninf_first",
                                  "all_nan"])
def test_normalisation_follows_the_blas_scan(case):
    rng = np.random.default_rng(3)
    I0 = rng.uniform(-40, 300, (23, 31)).astype(np.float32)
    I1 = rng.uniform(-40, 300, (23, 31)).astype(np.float32)
    if case == "nan_first":
        I0.flat[0] = np.nan
    elif case == "nan_inside":
        I0.flat[100] = np.nan
        I1.flat[7] = np.nan
    elif case == "nan_both":
        I0.flat[0] = np.nan
        I1.flat[0] = np.nan
        I1.flat[5] = np.nan
    elif case == "inf_inside":
        I1.flat[50] = np.inf
    elif case == "ninf_first":
        I0.flat[0] = -np.inf
    else:
        I0[:] = np.nan
    got = O.image_normalization(I0, I1)
    want = _normalised(I0, I1)
    for g, w in zip(got, want):
        assert np.array_equal(g, w, equal_nan=True)


def test_bicubic_nonfinite_coordinates():
    img = np.arange(12 * 9, dtype=np.float32).reshape(9, 12)
    for bad in (np.nan, np.inf, -np.inf, 1e30, -1e30):
        for uu, vv in ((bad, 4.0), (4.0, bad), (bad, bad)):
            # some tap is clamped: border_out gives 0
            assert O.bicubic_at(img, uu, vv, True) == 0.0
            if np.isnan(uu) or np.isnan(vv):
                assert np.isnan(O.bicubic_at(img, uu, vv, False))
    # bilinear: a NaN coordinate gives NaN weights; coordinates beyond the
    # int range saturate and clamp to the edge column
    assert np.isnan(O.bilinear_at(img, np.nan, 3.0))
    assert O.bilinear_at(img, 1e30, 2.0) == img[2, -1]
    assert O.bilinear_at(img, -1e30, 2.0) == img[2, 0]


def test_nan_first_pixel_stops_every_warping_at_once():
    """A NaN first pixel makes the normalised frame NaN everywhere; the
    stopping test error > eps^2 is false for a NaN error, so every warping
    stops after its first iteration (tvl1.cpp:216, 252-265)."""
    I0, I1, _, _ = synth.make_pair(64, 48, 2.0, 9)
    a = I0.astype(np.float32)
    a.flat[0] = np.nan
    u, it = O.tvl1(a, I1.astype(np.float32), O.default_params(nscales=3))
    assert (it == 1).all()
    assert np.isnan(u).all()


def test_infinite_pixel_flattens_both_frames():
    """max |x| = inf: den = inf, every finite pixel of both frames
    normalises to 0 and the flow of two flat frames is 0 after one
    iteration per warping."""
    I0, I1, _, _ = synth.make_pair(64, 48, 2.0, 10)
    a = I0.astype(np.float32)
    a[10, 10] = np.inf
    u, it = O.tvl1(a, I1.astype(np.float32), O.default_params(nscales=3))
    assert (it == 1).all()
    assert np.all(u == 0)


def test_oracle_under_address_and_ub_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "oracle_san")], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all calls returned" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
