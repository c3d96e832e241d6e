"""Lucas-Kanade binding; mirrors class LucasKanade of
src/lk/cuda/lucaskanade.cuh:6-35, plus the pyramidal iterative variant."""
from __future__ import annotations

import numpy as np

from . import _check, _lib, _exiting


class LucasKanade:
    """LucasKanade(spac_filt_size, temp_filt_size, window_size, nx, ny)."""

    def __init__(self, spac_filt_size: int, temp_filt_size: int, window_size: int, nx: int,
                 ny: int):
        import ctypes as C
        self.spac, self.temp, self.win, self.nx, self.ny = (spac_filt_size, temp_filt_size,
                                                             window_size, nx, ny)
        h = C.c_void_p()
        _check(_lib.of_lk_create(spac_filt_size, temp_filt_size, window_size, nx, ny,
                                 C.byref(h)), "LucasKanade")
        self._h = h

    def copy_frames_circularbuffer_GPU_wrapper(self, frame, temp_size: int, fr: int,
                                               frame_size: int) -> None:
        a = np.ascontiguousarray(frame, dtype=np.uint8).reshape(-1)
        _check(_lib.of_lk_push_u8(self._h, a.ctypes.data, temp_size, fr, frame_size),
               "copy_frames_circularbuffer_GPU_wrapper")

    def lucas_kanade(self, Vx: np.ndarray, Vy: np.ndarray, iframe: int) -> None:
        """Fills the float32 arrays Vx, Vy (nx*ny) in place, as the reference."""
        for a in (Vx, Vy):
            if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"] or a.size != self.nx * self.ny:
                raise ValueError("Vx, Vy must be contiguous float32 arrays of nx*ny")
        _check(_lib.of_lk_compute(self._h, Vx.ctypes.data, Vy.ctypes.data, iframe), "lucas_kanade")

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.of_lk_destroy(self._h)
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


def lk_pyramidal(I0, I1, levels: int = 3, zf: float = 0.5, sigma: float = 1.0, window: int = 7,
                 iters: int = 5, det_eps: float = 1e-6):
    """Pyramidal iterative LK of a float pair; returns (u, v)."""
    a = np.ascontiguousarray(I0, dtype=np.float32)
    b = np.ascontiguousarray(I1, dtype=np.float32)
    h, w = a.shape
    u = np.zeros((h, w), np.float32)
    v = np.zeros((h, w), np.float32)
    _check(_lib.of_lk_pyr_compute(a.ctypes.data, b.ctypes.data, w, h, levels, zf, sigma, window,
                                  iters, det_eps, u.ctypes.data, v.ctypes.data), "lk_pyramidal")
    return u, v


def lk_pyramidal_device(d_I0: int, d_I1: int, width: int, height: int, d_u: int, d_v: int,
                        levels: int = 3, zf: float = 0.5, sigma: float = 1.0, window: int = 7,
                        iters: int = 5, det_eps: float = 1e-6, stream: int = 0) -> None:
    _check(_lib.of_lk_pyr_compute_device(d_I0, d_I1, width, height, levels, zf, sigma, window,
                                         iters, det_eps, d_u, d_v, stream or None),
           "lk_pyramidal_device")
