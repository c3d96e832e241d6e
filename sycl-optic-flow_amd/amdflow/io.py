"""Output formats of a flow field: the reference application's HSV colour
rendering (of_flow_to_color*) and Middlebury .flo files (of_flo_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _lib

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
for _n, _r, _a in [
    ("of_flow_to_color_device", C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    ("of_flow_to_color", C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    ("of_flo_write", C.c_int, [C.c_char_p, _vp, C.c_int, C.c_int]),
    ("of_flo_read", C.c_int, [C.c_char_p, _vp, C.c_longlong, _ip, _ip]),
]:
    _f = getattr(_lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def flow_to_color(U, width: int, height: int) -> np.ndarray:
    """BGR image [height, width, 3] (uint8) of the planar flow U."""
    u = np.ascontiguousarray(U, np.float32)
    if u.size != 2 * width * height:
        raise ValueError("U must hold 2*width*height floats")
    out = np.empty((height, width, 3), np.uint8)
    _check(_lib.of_flow_to_color(u.ctypes.data, width, height, out.ctypes.data), "flow_to_color")
    return out


def flow_to_color_device(d_U: int, width: int, height: int, d_BGR: int, stream: int = 0) -> None:
    _check(_lib.of_flow_to_color_device(d_U, width, height, d_BGR, stream or None),
           "flow_to_color_device")


def write_flo(path: str, U, width: int, height: int) -> None:
    u = np.ascontiguousarray(U, np.float32)
    if u.size != 2 * width * height:
        raise ValueError("U must hold 2*width*height floats")
    _check(_lib.of_flo_write(path.encode(), u.ctypes.data, width, height), "flo_write")


def read_flo(path: str):
    """Returns (U planar float32, width, height)."""
    w, h = C.c_int(), C.c_int()
    _check(_lib.of_flo_read(path.encode(), None, 0, C.byref(w), C.byref(h)), "flo_read")
    u = np.empty(2 * w.value * h.value, np.float32)
    _check(_lib.of_flo_read(path.encode(), u.ctypes.data, u.size, C.byref(w), C.byref(h)), "flo_read")
    return u, w.value, h.value
