"""Batched TV-L1 binding (of_tvl1_batch_*): many device-resident frame pairs
of mixed sizes solved concurrently."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import TVL1Params, _check, _lib, _exiting

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_lib.of_tvl1_batch_create.restype = C.c_int
_lib.of_tvl1_batch_create.argtypes = [C.POINTER(TVL1Params), C.c_int, C.POINTER(_vp)]
_lib.of_tvl1_batch_destroy.restype = None
_lib.of_tvl1_batch_destroy.argtypes = [_vp]
_lib.of_tvl1_batch_set_shape.restype = C.c_int
_lib.of_tvl1_batch_set_shape.argtypes = [_vp, C.c_longlong, C.c_int, C.c_int]
_lib.of_tvl1_batch_set_mode.restype = C.c_int
_lib.of_tvl1_batch_set_mode.argtypes = [_vp, C.c_int]
_lib.of_tvl1_batch_run_u8.restype = C.c_int
_lib.of_tvl1_batch_run_u8.argtypes = [_vp, C.c_int, _ip, _ip, C.POINTER(_vp), C.POINTER(_vp),
                                      C.POINTER(_vp), _ip]


_lib.of_tvl1_batch_autotune.restype = C.c_int
_lib.of_tvl1_batch_autotune.argtypes = [_vp, C.c_int, _ip, _ip, C.POINTER(_vp), C.POINTER(_vp),
                                        C.POINTER(_vp), C.c_int, _ip, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
TUNE_SHAPES = [(4, 8), (2, 8), (8, 8), (4, 4), (2, 4)]

_lib.of_tvl1_run_batch.restype = C.c_int
_lib.of_tvl1_run_batch.argtypes = [C.POINTER(TVL1Params), C.c_int, _ip, _ip, C.POINTER(_vp),
                                   C.POINTER(_vp), C.POINTER(_vp), _ip]

MODES = {"lockstep": 0, "workers": 1}


def run_batch(widths, heights, d_I0, d_I1, d_U, tau=0.25, lambda_=0.15, theta=0.3,
              nscales=100, zfactor=0.5, warps=5, epsilon=0.01, max_iterations=300):
    """One-call batch (of_tvl1_run_batch): device frames and flows as integer
    pointers; returns each pair's total iteration count."""
    n = len(widths)
    p = TVL1Params(tau, lambda_, theta, nscales, zfactor, warps, epsilon, max_iterations)
    w = np.ascontiguousarray(widths, np.int32)
    h = np.ascontiguousarray(heights, np.int32)
    it = np.zeros(n, np.int32)
    _check(_lib.of_tvl1_run_batch(C.byref(p), n, w.ctypes.data_as(_ip), h.ctypes.data_as(_ip),
                                  (_vp * n)(*d_I0), (_vp * n)(*d_I1), (_vp * n)(*d_U),
                                  it.ctypes.data_as(_ip)), "run_batch")
    return it   # OF_BATCH_LOCKSTEP, OF_BATCH_WORKERS


class TVL1Batch:
    """mode "lockstep" (default): every launch covers one scale of all pairs
    of a lane, `streams` lanes run concurrently; "workers": `streams` host
    threads solve pairs one by one."""

    def __init__(self, streams: int = 4, tau=0.25, lambda_=0.15, theta=0.3, nscales=100,
                 zfactor=0.5, warps=5, epsilon=0.01, max_iterations=300, mode: str = "lockstep"):
        p = TVL1Params(tau, lambda_, theta, nscales, zfactor, warps, epsilon, max_iterations)
        h = _vp()
        _check(_lib.of_tvl1_batch_create(C.byref(p), int(streams), C.byref(h)), "batch_create")
        self._h = h
        self.set_mode(mode)

    def set_mode(self, mode: str) -> None:
        _check(_lib.of_tvl1_batch_set_mode(self._h, MODES[mode]), "batch_set_mode")
        self.mode = mode

    def set_shape(self, min_pixels: int, iters_per_launch: int, rows_per_wave: int) -> None:
        _check(_lib.of_tvl1_batch_set_shape(self._h, int(min_pixels), int(iters_per_launch),
                                            int(rows_per_wave)), "batch_set_shape")

    def autotune(self, widths, heights, d_I0, d_I1, d_U, reps: int = 2):
        """of_tvl1_batch_autotune on this batch: installs and returns the
        fastest shapes ((K, R) at scales of >= 200000 pixels, (K, R) below),
        the uniform shapes' ms per batch (dict over TUNE_SHAPES) and the
        chosen variant's ms."""
        n = len(widths)
        w = np.ascontiguousarray(widths, np.int32)
        h = np.ascontiguousarray(heights, np.int32)
        chosen = np.zeros(4, np.int32)
        table = np.zeros(len(TUNE_SHAPES), np.float64)
        best = C.c_double()
        _check(_lib.of_tvl1_batch_autotune(self._h, n, w.ctypes.data_as(_ip), h.ctypes.data_as(_ip),
                                           (_vp * n)(*d_I0), (_vp * n)(*d_I1), (_vp * n)(*d_U),
                                           int(reps), chosen.ctypes.data_as(_ip),
                                           table.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(best)), "batch_autotune")
        c = [int(v) for v in chosen]
        return ((c[0], c[1]), (c[2], c[3])), dict(zip(TUNE_SHAPES, table.tolist())), best.value

    def run_device_u8(self, widths, heights, d_I0, d_I1, d_U):
        """d_I0/d_I1/d_U: sequences of device pointers (ints).  Returns the
        total iteration count of each pair."""
        n = len(widths)
        w = np.ascontiguousarray(widths, np.int32)
        h = np.ascontiguousarray(heights, np.int32)
        a0 = (_vp * n)(*d_I0)
        a1 = (_vp * n)(*d_I1)
        au = (_vp * n)(*d_U)
        it = np.zeros(n, np.int32)
        _check(_lib.of_tvl1_batch_run_u8(self._h, n, w.ctypes.data_as(_ip), h.ctypes.data_as(_ip),
                                         a0, a1, au, it.ctypes.data_as(_ip)), "batch_run_u8")
        return it

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.of_tvl1_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:   # at interpreter exit this module's globals may be gone already
            if not _exiting():
                self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
