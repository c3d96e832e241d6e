"""Closed-form checks of the TV-L1 dual update (T17), u update and stopping
rule (T16), and the Horn-Schunck pyramid steps (H6).

Every expected value below is derived by hand from the algorithm's
definition, not from any implementation:
  tvl1.cpp:252-265  u = v + theta*div(p); error = sum(du1^2 + du2^2) / N;
                    iterate while error > epsilon^2
  tvl1.cpp:272-291  g = 1 + (tau/theta)*|grad u|; p = (p + (tau/theta)*grad u) / g
  downscaleKernel.cuh:27-40   2x2 box mean
  upscaleKernel.cuh:26-40     x2 upscale sampled at ((ix - 0.5)/2, (iy - 0.5)/2),
                              times the scale factor
The inputs are chosen so that every value involved is exact in float32 (small
integers, powers of two, one correctly rounded division), so the checks
demand bit equality.

The same checks run on three backends: the CPU oracle (tests/test_oracle_pins.py),
the GPU kernels through the debug entry points (tests/test_gpu_pins.py), and
copies of the oracle with a planted mistake (tests/test_oracle_mutations.py),
each of which must fail at least one check.  Run as a script, this module
applies every check to the oracle library named by AMDFLOW_ORACLE_LIB and
prints a JSON object {check name: "ok" | failure message}.
"""
from __future__ import annotations

import json
import sys

import numpy as np

F = np.float32


# ----------------------------------------------------------------- backends --

class OracleBackend:
    """The CPU oracle (oracle/oracle.py over liboracle.so)."""

    def __init__(self):
        import oracle as O
        self.O = O

    def iterations(self, wx, wy, rc, u1, u2, p11, p12, p21, p22, n, tau, lambda_, theta):
        st = (u1, u2, p11, p12, p21, p22)
        errs = []
        for _ in range(n):
            st, e = self.O.tvl1_iteration(wx, wy, rc, *st, tau=tau, lambda_=lambda_, theta=theta)
            errs.append(e)
        return st, np.array(errs)

    def dual(self, I0, I1, u1, u2, tau, lambda_, theta, warps, epsilon, max_iterations):
        p = self.O.default_params(tau=tau, lambda_=lambda_, theta=theta, warps=warps,
                                  epsilon=epsilon, max_iterations=max_iterations)
        return self.O.dual_tvl1(I0, I1, u1, u2, p)

    def hs_downscale(self, src, nw, nh):
        return self.O.hs_downscale(src, nw, nh)

    def hs_upscale(self, src, nw, nh, scale):
        return self.O.hs_upscale(src, nw, nh, scale)


class GpuBackend:
    """The GPU kernels through include/amdflow_debug.h, with launch shape
    (K iterations per launch, R rows per wave) for the iteration kernel."""

    def __init__(self, K=2, R=8):
        from amdflow import TVL1Params
        from amdflow import debug as D
        self.D, self.P, self.K, self.R = D, TVL1Params, K, R

    def iterations(self, wx, wy, rc, u1, u2, p11, p12, p21, p22, n, tau, lambda_, theta):
        return self.D.tvl1_iterations(wx, wy, rc, u1, u2, p11, p12, p21, p22, n, self.K, self.R,
                                      tau=tau, lambda_=lambda_, theta=theta)

    def dual(self, I0, I1, u1, u2, tau, lambda_, theta, warps, epsilon, max_iterations):
        p = self.P(tau, lambda_, theta, 1, 0.5, warps, epsilon, max_iterations)
        return self.D.tvl1_dual(I0, I1, u1, u2, p)

    def hs_downscale(self, src, nw, nh):
        return self.D.hs_downscale(src, nw, nh)

    def hs_upscale(self, src, nw, nh, scale):
        return self.D.hs_upscale(src, nw, nh, scale)


# -------------------------------------------------------------------- T17 --

def _ramps(h, w):
    y, x = np.mgrid[0:h, 0:w]
    return F(3) * x.astype(F) + F(4) * y.astype(F), F(5) * x.astype(F) + F(12) * y.astype(F)


def check_t17_one_step(be, h=40, w=48):
    """One iteration from p = 0 with a neutral data term (grad I1w = 0 and
    rho_c = 0, so v = u and u does not move).  With tau/theta = 1/2 and the
    ramps u1 = 3x + 4y, u2 = 5x + 12y (forward differences (3,4) and (5,12),
    0 across the last column/row), Chambolle's projected step gives
    p = (1/2) grad u / (1 + (1/2)|grad u|): 1.5/3.5, 2/3.5, 2.5/7.5, 6/7.5
    inside."""
    u1, u2 = _ramps(h, w)
    z = np.zeros((h, w), F)
    (n1, n2, p11, p12, p21, p22), err = be.iterations(z, z, z, u1, u2, z, z, z, z, 1,
                                                      tau=0.25, lambda_=0.15, theta=0.5)
    assert np.array_equal(n1, u1) and np.array_equal(n2, u2), "u moved under a neutral data term"
    assert err[0] == 0.0, f"error sum {err[0]} for an unchanged u"
    gx1 = np.full((h, w), F(3)); gx1[:, -1] = 0
    gy1 = np.full((h, w), F(4)); gy1[-1, :] = 0
    gx2 = np.full((h, w), F(5)); gx2[:, -1] = 0
    gy2 = np.full((h, w), F(12)); gy2[-1, :] = 0
    half = F(0.5)
    # |grad u| is 5, 13 (or 3, 4, 5, 12, 0 on the last column/row): exact
    g1 = F(1) + half * np.sqrt(gx1.astype(np.float64) ** 2 + gy1.astype(np.float64) ** 2).astype(F)
    g2 = F(1) + half * np.sqrt(gx2.astype(np.float64) ** 2 + gy2.astype(np.float64) ** 2).astype(F)
    want = ((half * gx1) / g1, (half * gy1) / g1, (half * gx2) / g2, (half * gy2) / g2)
    assert want[0][5, 5] == F(1.5) / F(3.5) and want[3][5, 5] == F(6) / F(7.5)
    for name, got, exp in zip(("p11", "p12", "p21", "p22"), (p11, p12, p21, p22), want):
        bad = np.argwhere(got != exp)
        assert bad.size == 0, (f"{name} differs from the closed form at {len(bad)} pixels, "
                               f"first {tuple(bad[0])}: {got[tuple(bad[0])]} vs {exp[tuple(bad[0])]}")


def check_t17_dual_feasible(be, h=37, w=53, iters=25, seed=5):
    """The dual step keeps |(p11,p12)| <= 1 and |(p21,p22)| <= 1 (it maps
    the unit disc into itself): after many iterations on random data with
    widely varying gradients, no pixel may leave the disc (one float
    rounding of slack)."""
    rng = np.random.default_rng(seed)
    wx = (rng.standard_normal((h, w)) * 20).astype(F)
    wy = (rng.standard_normal((h, w)) * 0.05).astype(F)
    rc = (rng.standard_normal((h, w)) * 30).astype(F)
    u1 = (rng.standard_normal((h, w)) * 40).astype(F)
    u2 = (rng.standard_normal((h, w)) * 0.01).astype(F)
    z = np.zeros((h, w), F)
    (_, _, p11, p12, p21, p22), _ = be.iterations(wx, wy, rc, u1, u2, z, z, z, z, iters,
                                                  tau=0.25, lambda_=0.15, theta=0.3)
    for a, b, name in ((p11, p12, "p1"), (p21, p22, "p2")):
        m = float(np.max(np.hypot(a.astype(np.float64), b.astype(np.float64))))
        assert m <= 1.0 + 2e-7, f"|{name}| reaches {m} > 1"


# -------------------------------------------------------------------- T16 --

def check_t16_u_update(be, h=24, w=40):
    """u = v + theta*div(p): with a neutral data term (v = u = 0) and
    p11 = x/64, p12 = p21 = p22 = 0, div p1 = 1/64 in the interior columns,
    so u1 becomes theta/64 there (theta = 1/2: exactly 1/128) and u2 stays 0."""
    y, x = np.mgrid[0:h, 0:w]
    p11 = (x / 64.0).astype(F)
    z = np.zeros((h, w), F)
    (u1, u2, *_), err = be.iterations(z, z, z, z, z, p11, z, z, z, 1,
                                      tau=0.25, lambda_=0.15, theta=0.5)
    inner = u1[:, 1:-1]
    assert np.all(inner == F(1) / F(128)), f"u1 inside {np.unique(inner)[:4]}, want 1/128"
    assert np.all(u2 == 0)
    # the error sum of this step: (1/128)^2 per interior pixel plus the
    # border columns' theta*div terms (div there: p11[0] = 0 and -p11[w-2])
    last = F(0.5) * (-(F(w - 2) / F(64)))
    want = h * ((w - 2) * (1.0 / 128) ** 2 + float(last) ** 2)
    assert err[0] == want, f"error sum {err[0]} vs {want}"


def check_t16_error_term(be, h=30, w=44):
    """With grad I1w = (1, 0) and rho = 10 >> lambda*theta, TH moves u1 by
    exactly -lambda*theta at every pixel in each of the first two iterations
    (grad u stays 0, so p stays 0): each iteration's error sum is
    N * (lambda*theta)^2 with lambda*theta rounded to float."""
    one = np.ones((h, w), F)
    z = np.zeros((h, w), F)
    rc = np.full((h, w), F(10))
    (u1, u2, p11, *_), err = be.iterations(one, z, rc, z, z, z, z, z, z, 2,
                                           tau=0.25, lambda_=0.15, theta=0.3)
    lt = F(0.15) * F(0.3)
    assert np.all(u1 == -(lt + lt)) and np.all(u2 == 0), "u1 did not move by lambda*theta twice"
    assert np.all(p11 == 0)
    want = float(h * w) * float(lt * lt)
    assert list(err) == [want, want], f"error sums {list(err)} vs {want}"


def check_t16_stop_threshold(be, h=36, w=44, max_iterations=12):
    """The stopping rule of one warping divides the error sum by N and
    compares it with epsilon^2.  Frames: I1 = x (a ramp), I0 = x + 100.  The
    bicubic warp of u = 0 keeps the samples whose 4x4 taps lie inside, x in
    [1, w-3] and y in [1, h-3], and zeroes the rest (border_out); there
    grad I1w = (1, 0) and rho = -100, so the first iteration moves u1 by
    +lambda*theta; outside grad I1w = 0 and nothing moves.  Hence
    error_1 = (w-3)(h-3) (lambda*theta)^2 / (w h), and with epsilon just
    above sqrt(error_1) the warping stops after 1 iteration, just below it
    goes on to a second one (later decisions depend on the TV term at the
    edge of that region, so only "more than one" is fixed)."""
    y, x = np.mgrid[0:h, 0:w]
    I1 = x.astype(F)
    I0 = I1 + F(100)
    z = np.zeros((h, w), F)
    lt = F(0.15) * F(0.3)
    e1 = (w - 3) * (h - 3) * float(lt * lt) / (w * h)
    for scale, stop in ((1.0 + 1e-3, True), (1.0 - 1e-3, False)):
        eps = F(np.sqrt(e1) * scale)
        # the margin is far beyond both float and double rounding of e1
        assert (float(eps * eps) > e1) == stop
        u1, u2, iters = be.dual(I0, I1, z, z, tau=0.25, lambda_=0.15, theta=0.3, warps=1,
                                epsilon=float(eps), max_iterations=max_iterations)
        n = int(iters[0])
        assert (n == 1) if stop else (n >= 2), \
            f"epsilon {float(eps)}: {n} iterations, want {'1' if stop else 'more than 1'}"
        if stop:
            inside = np.zeros((h, w), bool)
            inside[1:h - 2, 1:w - 2] = True
            assert np.all(u1[inside] == lt) and np.all(u1[~inside] == 0) and np.all(u2 == 0)


# --------------------------------------------------------------------- H6 --

def check_h6_downscale_ramp(be, h=20, w=26):
    """The 2x2 box mean of f = 4x + 8y at (2ix..2ix+1, 2iy..2iy+1) is
    4(2ix + 1/2) + 8(2iy + 1/2) = 8ix + 16iy + 6."""
    y, x = np.mgrid[0:h, 0:w]
    src = (F(4) * x + F(8) * y).astype(F)
    nw, nh = w // 2, h // 2
    out = be.hs_downscale(src, nw, nh)
    iy, ix = np.mgrid[0:nh, 0:nw]
    want = (8 * ix + 16 * iy + 6).astype(F)
    assert np.array_equal(out, want), f"max deviation {np.max(np.abs(out - want))}"


def check_h6_upscale_constant(be, h=10, w=12):
    """Upscaling a constant field c gives c * scale everywhere."""
    src = np.full((h, w), F(1.75))
    out = be.hs_upscale(src, 2 * w, 2 * h, 1.5)
    assert np.all(out == F(2.625)), f"values {np.unique(out)[:4]}, want 2.625"


def check_h6_upscale_ramp(be, h=10, w=12):
    """Upscaling f = 4x + 8y samples it at x = (ix - 1/2)/2: 2ix - 1 (and
    4iy - 2 for the y part) inside, clamped to the first and last coarse
    columns/rows at ix = 0 and ix = 2w-1; times the scale 2."""
    y, x = np.mgrid[0:h, 0:w]
    src = (F(4) * x + F(8) * y).astype(F)
    out = be.hs_upscale(src, 2 * w, 2 * h, 2.0)
    ix = np.arange(2 * w)
    X = 2.0 * ix - 1
    X[0], X[-1] = 0, 4 * (w - 1)
    iy = np.arange(2 * h)
    Y = 4.0 * iy - 2
    Y[0], Y[-1] = 0, 8 * (h - 1)
    want = ((Y[:, None] + X[None, :]) * 2).astype(F)
    assert np.array_equal(out, want), f"max deviation {np.max(np.abs(out - want))}"


CHECKS = {
    "t17_one_step": check_t17_one_step,
    "t17_dual_feasible": check_t17_dual_feasible,
    "t16_u_update": check_t16_u_update,
    "t16_error_term": check_t16_error_term,
    "t16_stop_threshold": check_t16_stop_threshold,
    "h6_downscale_ramp": check_h6_downscale_ramp,
    "h6_upscale_constant": check_h6_upscale_constant,
    "h6_upscale_ramp": check_h6_upscale_ramp,
}


def run_all(be) -> dict:
    res = {}
    for name, fn in CHECKS.items():
        try:
            fn(be)
            res[name] = "ok"
        except AssertionError as e:
            res[name] = f"fail: {e}"
    return res


if __name__ == "__main__":
    print(json.dumps(run_all(OracleBackend())))
    sys.exit(0)
