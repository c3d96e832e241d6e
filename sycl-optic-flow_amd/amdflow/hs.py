"""Horn-Schunck binding; mirrors initFlow / ComputeFlow / deleteFlow_mem of
src/hs/cuda/flow.cuh:19-33 (the reference keeps global state; here a handle)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _lib, _exiting

_vp = C.c_void_p
_lib.of_hs_set_jacobi_shape.restype = C.c_int
_lib.of_hs_set_jacobi_shape.argtypes = [_vp, C.c_int, C.c_int]


class HornSchunck:
    """HornSchunck(nLevels, stride, width, height) == initFlow(...)."""

    def __init__(self, nLevels: int, stride: int, width: int, height: int):
        self.nLevels, self.stride, self.width, self.height = nLevels, stride, width, height
        h = _vp()
        _check(_lib.of_hs_create(nLevels, stride, width, height, C.byref(h)), "initFlow")
        self._h = h

    def set_jacobi_shape(self, sweeps_per_launch: int, rows_per_wave: int) -> None:
        _check(_lib.of_hs_set_jacobi_shape(self._h, sweeps_per_launch, rows_per_wave),
               "set_jacobi_shape")

    def ComputeFlow(self, I0, I1, width, height, stride, alpha, nLevels, nWarpIters,
                    nSolverIters, u=None, v=None):
        """Host arrays of stride*height floats; returns (u, v) as (height, stride)."""
        n = stride * height
        a = np.ascontiguousarray(I0, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(I1, dtype=np.float32).reshape(-1)
        if a.size != n or b.size != n:
            raise ValueError(f"frames must hold stride*height = {n} floats")
        uu = np.zeros(n, np.float32) if u is None else u
        vv = np.zeros(n, np.float32) if v is None else v
        _check(_lib.of_hs_compute(self._h, a.ctypes.data, b.ctypes.data, width, height, stride,
                                  alpha, nLevels, nWarpIters, nSolverIters, uu.ctypes.data,
                                  vv.ctypes.data), "ComputeFlow")
        return uu.reshape(height, stride), vv.reshape(height, stride)

    def compute_device(self, d_I0: int, d_I1: int, d_u: int, d_v: int, alpha, nLevels,
                       nWarpIters, nSolverIters, stream: int = 0) -> None:
        _check(_lib.of_hs_compute_device(self._h, d_I0, d_I1, self.width, self.height,
                                         self.stride, alpha, nLevels, nWarpIters, nSolverIters,
                                         d_u, d_v, stream or None), "compute_device")

    def deleteFlow_mem(self) -> None:
        if getattr(self, "_h", None):
            _lib.of_hs_destroy(self._h)
            self._h = None

    close = deleteFlow_mem

    def __del__(self):
        try:   # at interpreter exit this module's globals may be gone already
            if not _exiting():
                self.deleteFlow_mem()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.deleteFlow_mem()
