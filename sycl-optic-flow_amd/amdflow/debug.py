"""Bindings of the per-kernel test entry points (include/amdflow_debug.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _lib

_vp = C.c_void_p

_SIGS = {
    "of_dbg_gaussian": [_vp, _vp, C.c_int, C.c_int, C.c_double],
    "of_dbg_zoom_out": [_vp, _vp, C.c_int, C.c_int, C.c_float],
    "of_dbg_zoom_in": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float],
    "of_dbg_centered_gradient": [_vp, _vp, _vp, C.c_int, C.c_int],
    "of_dbg_normalize": [_vp, _vp, C.c_int, C.c_int],
    "of_dbg_warp_setup": [_vp] * 9 + [C.c_int, C.c_int],
    "of_dbg_tvl1_iterations": [_vp] * 9 + [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                           C.c_int, C.c_int, C.c_int, _vp],
    "of_dbg_hypot": [_vp, _vp, _vp, C.c_size_t],
    "of_dbg_persist_force_not_resident": [C.c_int],
    "of_dbg_fail_alloc_after": [C.c_longlong],
}
for _n, _a in _SIGS.items():
    _f = getattr(_lib, _n)
    _f.restype = C.c_int
    _f.argtypes = _a


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data


def zoom_size(nx, ny, factor):
    return (int(float(np.float32(np.float32(nx) * np.float32(factor))) + 0.5),
            int(float(np.float32(np.float32(ny) * np.float32(factor))) + 0.5))


def gaussian(img, sigma):
    a = _f32(img)
    h, w = a.shape
    out = np.empty_like(a)
    _check(_lib.of_dbg_gaussian(_p(a), _p(out), w, h, sigma), "dbg_gaussian")
    return out


def zoom_out(img, factor=0.5):
    a = _f32(img)
    h, w = a.shape
    nxx, nyy = zoom_size(w, h, factor)
    out = np.empty((nyy, nxx), np.float32)
    _check(_lib.of_dbg_zoom_out(_p(a), _p(out), w, h, factor), "dbg_zoom_out")
    return out


def zoom_in(img, nxx, nyy, scale=1.0):
    a = _f32(img)
    h, w = a.shape
    out = np.empty((nyy, nxx), np.float32)
    _check(_lib.of_dbg_zoom_in(_p(a), _p(out), w, h, nxx, nyy, scale), "dbg_zoom_in")
    return out


def centered_gradient(img):
    a = _f32(img)
    h, w = a.shape
    dx, dy = np.empty_like(a), np.empty_like(a)
    _check(_lib.of_dbg_centered_gradient(_p(a), _p(dx), _p(dy), w, h), "dbg_centered_gradient")
    return dx, dy


def normalize(I0, I1):
    a, b = _f32(I0).copy(), _f32(I1).copy()
    h, w = a.shape
    _check(_lib.of_dbg_normalize(_p(a), _p(b), w, h), "dbg_normalize")
    return a, b


def warp_setup(I0, I1, I1x, I1y, u1, u2):
    ins = [_f32(x) for x in (I0, I1, I1x, I1y, u1, u2)]
    h, w = ins[0].shape
    outs = [np.empty_like(ins[0]) for _ in range(3)]
    _check(_lib.of_dbg_warp_setup(*[_p(x) for x in ins], *[_p(o) for o in outs], w, h),
           "dbg_warp_setup")
    return tuple(outs)  # I1wx, I1wy, rho_c


def tvl1_iterations(wx, wy, rho_c, u1, u2, p11, p12, p21, p22, iters, K=2, R=8, tau=0.25,
                    lambda_=0.15, theta=0.3):
    wx, wy, rho_c = map(_f32, (wx, wy, rho_c))
    st = [_f32(x).copy() for x in (u1, u2, p11, p12, p21, p22)]
    h, w = wx.shape
    err = np.zeros(iters, np.float64)
    _check(_lib.of_dbg_tvl1_iterations(_p(wx), _p(wy), _p(rho_c), *[_p(x) for x in st], w, h,
                                       tau, lambda_, theta, iters, K, R, _p(err)),
           "dbg_tvl1_iterations")
    return tuple(st), err


def hypot(a, b):
    a, b = _f32(a), _f32(b)
    out = np.empty_like(a)
    _check(_lib.of_dbg_hypot(_p(a), _p(b), _p(out), a.size), "dbg_hypot")
    return out


_lib.of_dbg_hs_jacobi.restype = C.c_int
_lib.of_dbg_hs_jacobi.argtypes = [_vp] * 5 + [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                              C.c_int, _vp, _vp]
_lib.of_dbg_hs_warp_derivatives.restype = C.c_int
_lib.of_dbg_hs_warp_derivatives.argtypes = [_vp] * 4 + [C.c_int, C.c_int] + [_vp] * 4


def hs_jacobi(du, dv, Ix, Iy, Iz, alpha, iters, K=4, R=4):
    du, dv, Ix, Iy, Iz = map(_f32, (du, dv, Ix, Iy, Iz))
    h, w = du.shape
    o1, o2 = np.empty_like(du), np.empty_like(dv)
    _check(_lib.of_dbg_hs_jacobi(_p(du), _p(dv), _p(Ix), _p(Iy), _p(Iz), w, h, alpha, iters, K, R,
                                 _p(o1), _p(o2)), "dbg_hs_jacobi")
    return o1, o2


def hs_warp_derivatives(I0, I1, u, v):
    I0, I1, u, v = map(_f32, (I0, I1, u, v))
    h, w = I0.shape
    outs = [np.empty_like(I0) for _ in range(4)]
    _check(_lib.of_dbg_hs_warp_derivatives(_p(I0), _p(I1), _p(u), _p(v), w, h,
                                           *[_p(o) for o in outs]), "dbg_hs_warp_derivatives")
    return tuple(outs)  # I1w, Ix, Iy, Iz


_lib.of_dbg_hs_downscale.restype = C.c_int
_lib.of_dbg_hs_downscale.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int]
_lib.of_dbg_hs_upscale.restype = C.c_int
_lib.of_dbg_hs_upscale.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_float]


def hs_downscale(src, nw, nh):
    """2x2 box average of the HS pyramid (dense planes)."""
    a = _f32(src)
    h, w = a.shape
    out = np.empty((nh, nw), np.float32)
    _check(_lib.of_dbg_hs_downscale(_p(a), w, h, _p(out), nw, nh), "dbg_hs_downscale")
    return out


def hs_upscale(src, nw, nh, scale):
    """x2 flow upscale of the HS pyramid times `scale` (dense planes)."""
    a = _f32(src)
    h, w = a.shape
    out = np.empty((nh, nw), np.float32)
    _check(_lib.of_dbg_hs_upscale(_p(a), w, h, _p(out), nw, nh, scale), "dbg_hs_upscale")
    return out


def tvl1_dual(I0, I1, u1, u2, params):
    """One scale of the GPU solver (warpings under the stopping rule) on float
    frames as given; `params` is an amdflow.TVL1Params.  Returns (u1, u2,
    iteration count per warping)."""
    from . import TVL1Params
    I0, I1 = _f32(I0), _f32(I1)
    a, b = _f32(u1).copy(), _f32(u2).copy()
    h, w = I0.shape
    iters = np.zeros(params.warps, np.int32)
    _lib.of_dbg_tvl1_dual.restype = C.c_int
    _lib.of_dbg_tvl1_dual.argtypes = [_vp] * 4 + [C.c_int, C.c_int, C.POINTER(TVL1Params), _vp]
    _check(_lib.of_dbg_tvl1_dual(_p(I0), _p(I1), _p(a), _p(b), w, h, C.byref(params), _p(iters)),
           "dbg_tvl1_dual")
    return a, b, iters


def persist_force_not_resident(on: bool) -> None:
    """While on, launches of the persistent warping kernel report that their
    grid was not co-resident, so solves take their host-paced repeat."""
    _check(_lib.of_dbg_persist_force_not_resident(1 if on else 0), "dbg_persist_force_not_resident")


_lib.of_dbg_alloc_attempts.restype = C.c_longlong
_lib.of_dbg_alloc_attempts.argtypes = []
_lib.of_dbg_live_allocs.restype = C.c_longlong
_lib.of_dbg_live_allocs.argtypes = []


def fail_alloc_after(n: int) -> None:
    """The first n device or pinned host allocations from now succeed, every
    later one fails with OF_ERR_ALLOC (n < 0: none fails); restarts the
    attempt count."""
    _check(_lib.of_dbg_fail_alloc_after(int(n)), "dbg_fail_alloc_after")


def alloc_attempts() -> int:
    return int(_lib.of_dbg_alloc_attempts())


def live_allocs() -> int:
    return int(_lib.of_dbg_live_allocs())
