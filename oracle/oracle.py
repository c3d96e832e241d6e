"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference's own
bicubic/zoom sources built into _ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# AMDFLOW_ORACLE_LIB names another build of oracle.c (the mutation tests
# load deliberately broken copies through it)
LIB = os.environ.get("AMDFLOW_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")


def build() -> None:
    """Compile liboracle.so and, when the reference tree is present, _ref/."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _load():
    if not os.path.exists(LIB) and not os.environ.get("AMDFLOW_ORACLE_LIB"):
        build()
    return C.CDLL(LIB)


_lib = _load()
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class Params(C.Structure):
    _fields_ = [
        ("tau", C.c_float), ("lambda_", C.c_float), ("theta", C.c_float),
        ("nscales", C.c_int), ("zfactor", C.c_float), ("warps", C.c_int),
        ("epsilon", C.c_float), ("max_iterations", C.c_int),
    ]


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("orc_tvl1_default_params", None, [C.POINTER(Params)])
_sig("orc_set_threads", None, [C.c_int])
_sig("orc_get_threads", C.c_int, [])
_sig("orc_set_error_mode", None, [C.c_int])

ERROR_DOUBLE_ROWS, ERROR_FLOAT_SEQUENTIAL = 0, 1


def set_error_mode(mode: int) -> None:
    """How TV-L1's stopping error is summed: ERROR_DOUBLE_ROWS (default, the
    GPU's order) or ERROR_FLOAT_SEQUENTIAL (the reference's float sum in
    index order, as with one OpenMP thread)."""
    _lib.orc_set_error_mode(int(mode))


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's loops (its results do not depend on it)."""
    _lib.orc_set_threads(int(n))


def get_threads() -> int:
    return int(_lib.orc_get_threads())
_sig("orc_tvl1_nscales", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float])
_sig("orc_zoom_size", None, [C.c_int, C.c_int, C.c_float, _ip, _ip])
_sig("orc_zoom_sigma", C.c_float, [C.c_float])
_sig("orc_gaussian_taps", C.c_int, [C.c_double, C.POINTER(C.c_double), C.c_int])
_sig("orc_gaussian", C.c_int, [_vp, C.c_int, C.c_int, C.c_double])
_sig("orc_bicubic_at", C.c_float, [_vp, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int])
_sig("orc_bicubic_warp", None, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int])
_sig("orc_zoom_out", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_float, _vp])
_sig("orc_zoom_in", None, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int])
_sig("orc_centered_gradient", None, [_vp, _vp, _vp, C.c_int, C.c_int])
_sig("orc_forward_gradient", None, [_vp, _vp, _vp, C.c_int, C.c_int])
_sig("orc_divergence", None, [_vp, _vp, _vp, C.c_int, C.c_int])
_sig("orc_image_normalization", None, [_vp, _vp, _vp, _vp, C.c_size_t])
_sig("orc_tvl1_th", None, [_vp, _vp, _vp, _vp, _vp, C.c_float, _vp, _vp, C.c_size_t])
_sig("orc_tvl1_warp_setup", None, [_vp] * 10 + [C.c_int, C.c_int])
_sig("orc_tvl1_iteration", C.c_double, [_vp] * 9 + [C.c_int, C.c_int, C.c_float, C.c_float,
                                                    C.c_float, _vp])
_sig("orc_dual_tvl1", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _ip])
_sig("orc_tvl1_run_u8", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _vp, _ip, _ip])
_sig("orc_tvl1_run_f32", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _vp, _ip, _ip])
_sig("orc_tvl1_run_init_u8", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _vp, C.c_int,
                                       _vp, _ip, _ip])
_sig("orc_tvl1_run_schedule_u8", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _vp,
                                           C.c_int, _vp, _vp, _vp, _ip, _ip])
_sig("orc_tvl1_energy", C.c_double, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_float,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)])
_sig("orc_bilinear_at", C.c_float, [_vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float])
_sig("orc_hs_downscale", None, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int])
_sig("orc_hs_upscale", None, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                              C.c_float])
_sig("orc_hs_warp", None, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp])
_sig("orc_hs_derivatives", None, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp])
_sig("orc_hs_jacobi", None, [_vp] * 5 + [C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp])
_sig("orc_hs_compute", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                 C.c_int, C.c_int, _vp, _vp])
_sig("orc_lk_filter", C.c_int, [C.c_int, C.c_int, _vp])
_sig("orc_lk_single", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                                _vp])
_sig("orc_lk_pyr_compute", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_int, C.c_int, C.c_float, _vp, _vp])


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def default_params(**kw) -> Params:
    p = Params()
    _lib.orc_tvl1_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


def nscales(w, h, nscales=100, zfactor=0.5) -> int:
    return _lib.orc_tvl1_nscales(w, h, nscales, zfactor)


def zoom_size(nx, ny, factor):
    a, b = C.c_int(), C.c_int()
    _lib.orc_zoom_size(nx, ny, factor, C.byref(a), C.byref(b))
    return a.value, b.value


def zoom_sigma(factor) -> float:
    return _lib.orc_zoom_sigma(factor)


def gaussian_taps(sigma: float) -> np.ndarray:
    B = (C.c_double * 64)()
    n = _lib.orc_gaussian_taps(sigma, B, 64)
    return np.array(B[:n])


def gaussian(img, sigma: float) -> np.ndarray:
    a = _f32(img).copy()
    h, w = a.shape
    st = _lib.orc_gaussian(_p(a), w, h, sigma)
    if st:
        raise ValueError(f"gaussian: status {st}")
    return a


def bicubic_at(img, uu, vv, border_out: bool) -> float:
    a = _f32(img)
    h, w = a.shape
    return _lib.orc_bicubic_at(_p(a), uu, vv, w, h, int(border_out))


def bicubic_warp(img, u, v, border_out: bool = True) -> np.ndarray:
    a, u, v = _f32(img), _f32(u), _f32(v)
    h, w = a.shape
    out = np.empty_like(a)
    _lib.orc_bicubic_warp(_p(a), _p(u), _p(v), _p(out), w, h, int(border_out))
    return out


def zoom_out(img, factor=0.5) -> np.ndarray:
    a = _f32(img)
    h, w = a.shape
    nxx, nyy = zoom_size(w, h, factor)
    out = np.empty((nyy, nxx), np.float32)
    tmp = np.empty_like(a)
    st = _lib.orc_zoom_out(_p(a), _p(out), w, h, factor, _p(tmp))
    if st:
        raise ValueError(f"zoom_out: status {st}")
    return out


def zoom_in(img, nxx, nyy) -> np.ndarray:
    a = _f32(img)
    h, w = a.shape
    out = np.empty((nyy, nxx), np.float32)
    _lib.orc_zoom_in(_p(a), _p(out), w, h, nxx, nyy)
    return out


def centered_gradient(img):
    a = _f32(img)
    h, w = a.shape
    dx, dy = np.empty_like(a), np.empty_like(a)
    _lib.orc_centered_gradient(_p(a), _p(dx), _p(dy), w, h)
    return dx, dy


def forward_gradient(img):
    a = _f32(img)
    h, w = a.shape
    dx, dy = np.empty_like(a), np.empty_like(a)
    _lib.orc_forward_gradient(_p(a), _p(dx), _p(dy), w, h)
    return dx, dy


def divergence(v1, v2):
    v1, v2 = _f32(v1), _f32(v2)
    h, w = v1.shape
    d = np.empty_like(v1)
    _lib.orc_divergence(_p(v1), _p(v2), _p(d), w, h)
    return d


def image_normalization(I0, I1):
    a, b = _f32(I0).copy(), _f32(I1).copy()
    _lib.orc_image_normalization(_p(a), _p(b), _p(a), _p(b), a.size)
    return a, b


def tvl1_th(rho_c, wx, wy, u1, u2, l_t):
    rho_c, wx, wy, u1, u2 = map(_f32, (rho_c, wx, wy, u1, u2))
    v1, v2 = np.empty_like(u1), np.empty_like(u2)
    _lib.orc_tvl1_th(_p(rho_c), _p(wx), _p(wy), _p(u1), _p(u2), l_t, _p(v1), _p(v2), u1.size)
    return v1, v2


def tvl1_warp_setup(I0, I1, I1x, I1y, u1, u2):
    I0, I1, I1x, I1y, u1, u2 = map(_f32, (I0, I1, I1x, I1y, u1, u2))
    h, w = I0.shape
    outs = [np.empty_like(I0) for _ in range(4)]
    _lib.orc_tvl1_warp_setup(_p(I0), _p(I1), _p(I1x), _p(I1y), _p(u1), _p(u2),
                             *[_p(o) for o in outs], w, h)
    return tuple(outs)  # I1w, I1wx, I1wy, rho_c


def tvl1_iteration(wx, wy, rho_c, u1, u2, p11, p12, p21, p22, tau=0.25, lambda_=0.15, theta=0.3):
    """One primal-dual iteration; returns (new state tuple, error sum)."""
    wx, wy, rho_c = map(_f32, (wx, wy, rho_c))
    st = [_f32(x).copy() for x in (u1, u2, p11, p12, p21, p22)]
    h, w = wx.shape
    scratch = np.empty(6 * w * h, np.float32)
    e = _lib.orc_tvl1_iteration(_p(wx), _p(wy), _p(rho_c), *[_p(x) for x in st], w, h,
                                tau, lambda_, theta, _p(scratch))
    return tuple(st), e


def dual_tvl1(I0, I1, u1, u2, params: Params | None = None):
    I0, I1 = _f32(I0), _f32(I1)
    u1, u2 = _f32(u1).copy(), _f32(u2).copy()
    h, w = I0.shape
    p = params or default_params()
    iters = np.zeros(p.warps, np.int32)
    st = _lib.orc_dual_tvl1(_p(I0), _p(I1), _p(u1), _p(u2), w, h, C.byref(p),
                            iters.ctypes.data_as(_ip))
    if st:
        raise ValueError(f"dual_tvl1: status {st}")
    return u1, u2, iters


def tvl1(I0, I1, params: Params | None = None):
    """Full multiscale TV-L1 of a pair (uint8 or float32 frames).
    Returns (u planar [u1|u2] float32, iteration counts [nscales, warps])."""
    p = params or default_params()
    I0 = np.asarray(I0)
    h, w = I0.shape
    ns = nscales(w, h, p.nscales, p.zfactor)
    u = np.zeros(2 * w * h, np.float32)
    iters = np.zeros(ns * p.warps, np.int32)
    nso = C.c_int()
    if I0.dtype == np.uint8:
        a, b = np.ascontiguousarray(I0), np.ascontiguousarray(I1, dtype=np.uint8)
        st = _lib.orc_tvl1_run_u8(_p(a), _p(b), w, h, C.byref(p), _p(u),
                                  iters.ctypes.data_as(_ip), C.byref(nso))
    else:
        a, b = _f32(I0), _f32(I1)
        st = _lib.orc_tvl1_run_f32(_p(a), _p(b), w, h, C.byref(p), _p(u),
                                   iters.ctypes.data_as(_ip), C.byref(nso))
    if st:
        raise ValueError(f"tvl1: status {st}")
    return u, iters.reshape(ns, p.warps)


def tvl1_init(I0, I1, uinit, params: Params | None = None, first: int = -1):
    """Multiscale TV-L1 of a uint8 pair solving scales first .. 0 (-1: all,
    as the reference) from the initial flow `uinit` at scale `first` (planar
    [u1|u2]; None = zero, the reference's start).  Returns (u, iteration
    counts [nscales, warps], zero for unsolved scales)."""
    p = params or default_params()
    I0 = np.ascontiguousarray(I0, dtype=np.uint8)
    h, w = I0.shape
    ns = nscales(w, h, p.nscales, p.zfactor)
    u = np.zeros(2 * w * h, np.float32)
    iters = np.zeros(ns * p.warps, np.int32)
    nso = C.c_int()
    b = np.ascontiguousarray(I1, dtype=np.uint8)
    ui = None if uinit is None else _f32(uinit)
    st = _lib.orc_tvl1_run_init_u8(_p(I0), _p(b), w, h, C.byref(p), None if ui is None else _p(ui),
                                   int(first), _p(u), iters.ctypes.data_as(_ip), C.byref(nso))
    if st:
        raise ValueError(f"tvl1_init: status {st}")
    return u, iters.reshape(ns, p.warps)


def tvl1_schedule(I0, I1, warps, max_iterations, uinit=None, params: Params | None = None,
                  first: int = -1):
    """tvl1_init with per-scale limits: warps[s] warpings and
    max_iterations[s] iterations per warping at scale s (finest first; 0 =
    the parameters' value).  Returns (u, iteration counts [nscales, warps])."""
    p = params or default_params()
    I0 = np.ascontiguousarray(I0, dtype=np.uint8)
    h, w = I0.shape
    ns = nscales(w, h, p.nscales, p.zfactor)
    u = np.zeros(2 * w * h, np.float32)
    iters = np.zeros(ns * p.warps, np.int32)
    ws = np.zeros(64, np.int32)
    ms = np.zeros(64, np.int32)
    ws[:len(warps)] = warps
    ms[:len(max_iterations)] = max_iterations
    nso = C.c_int()
    b = np.ascontiguousarray(I1, dtype=np.uint8)
    ui = None if uinit is None else _f32(uinit)
    st = _lib.orc_tvl1_run_schedule_u8(_p(I0), _p(b), w, h, C.byref(p),
                                       None if ui is None else _p(ui), first, _p(ws), _p(ms), _p(u),
                                       iters.ctypes.data_as(_ip), C.byref(nso))
    if st:
        raise ValueError(f"tvl1_schedule: status {st}")
    return u, iters.reshape(ns, p.warps)


def tvl1_energy(I0, I1, u, lambda_=0.15):
    """(E, data sum, TV sum) of the planar flow u on the scale images I0, I1."""
    a, b = _f32(I0), _f32(I1)
    h, w = a.shape
    uu = _f32(u)
    d, t = C.c_double(), C.c_double()
    e = _lib.orc_tvl1_energy(_p(a), _p(b), _p(uu), _p(uu[h * w:]), w, h, lambda_, C.byref(d),
                             C.byref(t))
    return e, d.value, t.value


def bilinear_at(img, x, y) -> float:
    a = _f32(img)
    h, w = a.shape
    return _lib.orc_bilinear_at(_p(a), w, h, w, x, y)


def hs_compute(I0, I1, alpha=15.0, nLevels=1, nWarpIters=1, nSolverIters=100, stride=None):
    I0, I1 = _f32(I0), _f32(I1)
    h, w = I0.shape[0], I0.shape[1]
    s = stride or w
    if s != w:
        a = np.zeros((h, s), np.float32); a[:, :w] = I0; I0 = a
        b = np.zeros((h, s), np.float32); b[:, :w] = I1; I1 = b
    u = np.zeros((h, s), np.float32)
    v = np.zeros((h, s), np.float32)
    st = _lib.orc_hs_compute(_p(I0), _p(I1), w, h, s, alpha, nLevels, nWarpIters, nSolverIters,
                             _p(u), _p(v))
    if st:
        raise ValueError(f"hs_compute: status {st}")
    return u, v


def hs_downscale(src, nw, nh):
    """2x2 box average of the HS pyramid (dense planes)."""
    a = _f32(src)
    h, w = a.shape
    out = np.empty((nh, nw), np.float32)
    _lib.orc_hs_downscale(_p(a), w, h, w, _p(out), nw, nh, nw)
    return out


def hs_upscale(src, nw, nh, scale):
    """x2 flow upscale of the HS pyramid times `scale` (dense planes)."""
    a = _f32(src)
    h, w = a.shape
    out = np.empty((nh, nw), np.float32)
    _lib.orc_hs_upscale(_p(a), w, h, w, _p(out), nw, nh, nw, scale)
    return out


def hs_jacobi(du, dv, Ix, Iy, Iz, alpha):
    du, dv, Ix, Iy, Iz = map(_f32, (du, dv, Ix, Iy, Iz))
    h, w = du.shape
    o1, o2 = np.empty_like(du), np.empty_like(dv)
    _lib.orc_hs_jacobi(_p(du), _p(dv), _p(Ix), _p(Iy), _p(Iz), w, h, w, alpha, _p(o1), _p(o2))
    return o1, o2


def hs_derivatives(I0, I1w):
    I0, I1w = _f32(I0), _f32(I1w)
    h, w = I0.shape
    Ix, Iy, Iz = np.empty_like(I0), np.empty_like(I0), np.empty_like(I0)
    _lib.orc_hs_derivatives(_p(I0), _p(I1w), w, h, w, _p(Ix), _p(Iy), _p(Iz))
    return Ix, Iy, Iz


def lk_filter(size: int, temporal: bool):
    f = np.zeros(16, np.float32)
    ok = _lib.orc_lk_filter(size, int(temporal), _p(f))
    return f[:size] if ok else None


def lk_single(frames_u8, iframe, spac=3, win=5):
    fr = np.ascontiguousarray(frames_u8, dtype=np.uint8)
    T, h, w = fr.shape
    vx, vy = np.zeros((h, w), np.float32), np.zeros((h, w), np.float32)
    st = _lib.orc_lk_single(_p(fr), T, iframe, w, h, spac, win, _p(vx), _p(vy))
    if st:
        raise ValueError(f"lk_single: status {st}")
    return vx, vy


def lk_pyr(I0, I1, levels=3, zf=0.5, sigma=1.0, win=7, iters=5, det_eps=1e-6):
    I0, I1 = _f32(I0), _f32(I1)
    h, w = I0.shape
    u, v = np.zeros((h, w), np.float32), np.zeros((h, w), np.float32)
    st = _lib.orc_lk_pyr_compute(_p(I0), _p(I1), w, h, levels, zf, sigma, win, iters, det_eps,
                                 _p(u), _p(v))
    if st:
        raise ValueError(f"lk_pyr: status {st}")
    return u, v


# ---- FP64 oracle (oracle64.c): the same algorithms in double ---------------

_dp = C.POINTER(C.c_double)
_sig("o64_gaussian", C.c_int, [_vp, C.c_int, C.c_int, C.c_double])
_sig("o64_bicubic_at", C.c_double, [_vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int])
_sig("o64_tvl1_iteration", C.c_double, [_vp] * 9 + [C.c_int, C.c_int, C.c_double, C.c_double,
                                                    C.c_double, _vp])
_sig("o64_dual_tvl1", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _ip])
_sig("o64_tvl1_run_f32", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(Params), _vp, _ip, _ip])
_sig("o64_hs_downscale", None, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int])
_sig("o64_hs_upscale", None, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_double])
_sig("o64_hs_jacobi", None, [_vp] * 5 + [C.c_int, C.c_int, C.c_double, _vp, _vp])
_sig("o64_hs_compute", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_int, C.c_int, _vp, _vp])
_sig("o64_lk_pyr_compute", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_double, _vp, _vp])


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def gaussian64(img, sigma: float) -> np.ndarray:
    a = _f64(img).copy()
    h, w = a.shape
    st = _lib.o64_gaussian(_p(a), w, h, sigma)
    if st:
        raise ValueError(f"gaussian64: status {st}")
    return a


def bicubic_at64(img, uu, vv, border_out: bool) -> float:
    a = _f64(img)
    h, w = a.shape
    return _lib.o64_bicubic_at(_p(a), uu, vv, w, h, int(border_out))


def tvl1_iteration64(wx, wy, rho_c, u1, u2, p11, p12, p21, p22, tau=0.25, lambda_=0.15,
                     theta=0.3):
    """One primal-dual iteration in double; returns the new (u1, u2, p11,
    p12, p21, p22) and the error sum (du1^2 + du2^2)."""
    wx, wy, rho_c = _f64(wx), _f64(wy), _f64(rho_c)
    st = [_f64(x).copy() for x in (u1, u2, p11, p12, p21, p22)]
    h, w = wx.shape
    scratch = np.empty(4 * w * h, np.float64)
    e = _lib.o64_tvl1_iteration(_p(wx), _p(wy), _p(rho_c), *[_p(x) for x in st], w, h, tau,
                                lambda_, theta, _p(scratch))
    return tuple(st), e


def dual_tvl1_64(I0, I1, u1, u2, params: Params | None = None):
    """One scale of TV-L1 in double from (u1, u2); returns (u1, u2, iters)."""
    I0, I1 = _f64(I0), _f64(I1)
    u1, u2 = _f64(u1).copy(), _f64(u2).copy()
    h, w = I0.shape
    p = params or default_params()
    iters = np.zeros(p.warps, np.int32)
    st = _lib.o64_dual_tvl1(_p(I0), _p(I1), _p(u1), _p(u2), w, h, C.byref(p),
                            iters.ctypes.data_as(_ip))
    if st:
        raise ValueError(f"dual_tvl1_64: status {st}")
    return u1, u2, iters


def tvl1_64(I0, I1, params: Params | None = None):
    """Multiscale TV-L1 in double of a pair of frames (uint8 or float32, read
    exactly).  Returns (u planar [u1|u2] float64, iteration counts
    [nscales, warps])."""
    p = params or default_params()
    a, b = _f32(I0), _f32(I1)
    h, w = a.shape
    ns = nscales(w, h, p.nscales, p.zfactor)
    u = np.zeros(2 * w * h, np.float64)
    iters = np.zeros(ns * p.warps, np.int32)
    nso = C.c_int()
    st = _lib.o64_tvl1_run_f32(_p(a), _p(b), w, h, C.byref(p), _p(u), iters.ctypes.data_as(_ip),
                               C.byref(nso))
    if st:
        raise ValueError(f"tvl1_64: status {st}")
    return u, iters.reshape(ns, p.warps)


def hs_downscale64(src, nw, nh):
    a = _f64(src)
    h, w = a.shape
    out = np.empty((nh, nw), np.float64)
    _lib.o64_hs_downscale(_p(a), w, h, _p(out), nw, nh)
    return out


def hs_upscale64(src, nw, nh, scale):
    a = _f64(src)
    h, w = a.shape
    out = np.empty((nh, nw), np.float64)
    _lib.o64_hs_upscale(_p(a), w, h, _p(out), nw, nh, scale)
    return out


def hs_jacobi64(du, dv, Ix, Iy, Iz, alpha):
    du, dv, Ix, Iy, Iz = (_f64(x) for x in (du, dv, Ix, Iy, Iz))
    h, w = du.shape
    o1, o2 = np.empty_like(du), np.empty_like(dv)
    _lib.o64_hs_jacobi(_p(du), _p(dv), _p(Ix), _p(Iy), _p(Iz), w, h, alpha, _p(o1), _p(o2))
    return o1, o2


def hs_compute64(I0, I1, alpha=15.0, nLevels=1, nWarpIters=1, nSolverIters=100):
    I0, I1 = _f32(I0), _f32(I1)
    h, w = I0.shape
    u, v = np.zeros((h, w), np.float64), np.zeros((h, w), np.float64)
    st = _lib.o64_hs_compute(_p(I0), _p(I1), w, h, w, alpha, nLevels, nWarpIters, nSolverIters,
                             _p(u), _p(v))
    if st:
        raise ValueError(f"hs_compute64: status {st}")
    return u, v


def lk_pyr64(I0, I1, levels=3, zf=0.5, sigma=1.0, win=7, iters=5, det_eps=1e-6):
    I0, I1 = _f32(I0), _f32(I1)
    h, w = I0.shape
    u, v = np.zeros((h, w), np.float64), np.zeros((h, w), np.float64)
    st = _lib.o64_lk_pyr_compute(_p(I0), _p(I1), w, h, levels, zf, sigma, win, iters, det_eps,
                                 _p(u), _p(v))
    if st:
        raise ValueError(f"lk_pyr64: status {st}")
    return u, v


def flow_error(u, u64) -> dict:
    """Distance of a float flow from the FP64 oracle's: max, mean and 99th
    percentile of |u - u64| over all components, in pixels."""
    d = np.abs(np.asarray(u, np.float64).ravel() - np.asarray(u64, np.float64).ravel())
    return {"max": float(d.max()), "mean": float(d.mean()), "p99": float(np.percentile(d, 99))}


# ---- the reference's own sources (oracle/_ref), when built -----------------

class Ref:
    """Functions compiled from the reference tree; None if _ref is not built."""

    def __init__(self):
        self.lib = None
        if os.path.exists(REF_LIB):
            lib = C.CDLL(REF_LIB)
            lib.ref_bicubic_at.restype = C.c_float
            lib.ref_bicubic_at.argtypes = [_vp, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]
            lib.ref_bicubic_warp.restype = None
            lib.ref_bicubic_warp.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
            lib.ref_zoom_in.restype = None
            lib.ref_zoom_in.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int]
            lib.ref_zoom_size.restype = None
            lib.ref_zoom_size.argtypes = [C.c_int, C.c_int, C.c_float, _ip, _ip]
            self.lib = lib

    @property
    def available(self) -> bool:
        return self.lib is not None

    def bicubic_at(self, img, uu, vv, border_out):
        a = _f32(img)
        h, w = a.shape
        return self.lib.ref_bicubic_at(_p(a), uu, vv, w, h, int(border_out))

    def bicubic_warp(self, img, u, v, border_out=True):
        a, u, v = _f32(img), _f32(u), _f32(v)
        h, w = a.shape
        out = np.empty_like(a)
        self.lib.ref_bicubic_warp(_p(a), _p(u), _p(v), _p(out), w, h, int(border_out))
        return out

    def zoom_in(self, img, nxx, nyy):
        a = _f32(img)
        h, w = a.shape
        out = np.empty((nyy, nxx), np.float32)
        self.lib.ref_zoom_in(_p(a), _p(out), w, h, nxx, nyy)
        return out

    def zoom_size(self, nx, ny, factor):
        a, b = C.c_int(), C.c_int()
        self.lib.ref_zoom_size(nx, ny, factor, C.byref(a), C.byref(b))
        return a.value, b.value


ref = Ref()
