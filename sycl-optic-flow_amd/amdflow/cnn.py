"""Warm-start network binding (of_cnn_*, of_tvl1_warm_run_device_u8).

The network predicts the coarsest-scale TV-L1 flow from that scale's
I0, I1, Ix, Iy, It; see include/amdflow.h for the layer definition.  Device
buffers are passed as integer pointers (e.g. torch's data_ptr()).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _check, _lib, _exiting

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

_SIGS = {
    "of_cnn_create": (C.c_int, [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_vp)]),
    "of_cnn_destroy": (None, [_vp]),
    "of_cnn_num_params": (C.c_longlong, [_vp]),
    "of_cnn_get_params": (C.c_int, [_vp, _vp]),
    "of_cnn_set_params": (C.c_int, [_vp, _vp]),
    "of_cnn_features": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp]),
    "of_cnn_forward": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "of_cnn_predict": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp]),
    "of_cnn_gradients": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _dp]),
    "of_cnn_train_step": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _dp]),
    "of_tvl1_warm_run_device_u8": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _ip,
                                             _dp, _vp]),
    "of_dbg_cnn_mfma": (C.c_int, [_vp, _vp, _vp]),
}
for _n, (_r, _a) in _SIGS.items():
    _f = getattr(_lib, _n)
    _f.restype = _r
    _f.argtypes = _a

HIDDEN = 5          # 6 conv layers of 32 channels, then the 2-channel head
# first solved scale of a warm-started solve: 2 = 480x270 at 1920x1080 (2 warpings:
# 2.5x the plain solve's pairs/s at 0.22 px mean EPE from its flow; seeding 960x540
# instead is 5% faster at 0.48 px)
FIRST_SCALE = 2
CHANNELS = 32
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights",
                       "warmstart_c32_h5.npy")


class WarmStartNet:
    def __init__(self, hidden: int = HIDDEN, channels: int = CHANNELS, seed: int = 0):
        self.hidden, self.channels = int(hidden), int(channels)
        h = _vp()
        _check(_lib.of_cnn_create(self.hidden, self.channels, C.c_ulonglong(seed), C.byref(h)),
               "cnn_create")
        self._h = h

    @classmethod
    def pretrained(cls, path: str = WEIGHTS) -> "WarmStartNet":
        """The network with the trained parameters shipped with the package."""
        w = np.load(path)   # float32 array; no pickles
        net = cls()
        net.set_params(w)
        return net

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.of_cnn_destroy(self._h)
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

    @property
    def handle(self) -> int:
        return self._h.value

    @property
    def num_params(self) -> int:
        return int(_lib.of_cnn_num_params(self._h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        _check(_lib.of_cnn_get_params(self._h, out.ctypes.data), "cnn_get_params")
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        _check(_lib.of_cnn_set_params(self._h, p.ctypes.data), "cnn_set_params")

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.save(path, self.get_params())

    @staticmethod
    def features(d_I0s: int, d_I1s: int, H: int, W: int, d_X: int, stream: int = 0) -> None:
        _check(_lib.of_cnn_features(d_I0s, d_I1s, int(H), int(W), d_X, stream or None), "cnn_features")

    def forward(self, d_X: int, B: int, H: int, W: int, d_Y: int, stream: int = 0) -> None:
        _check(_lib.of_cnn_forward(self._h, d_X, int(B), int(H), int(W), d_Y, stream or None),
               "cnn_forward")

    def predict(self, d_I0s: int, d_I1s: int, H: int, W: int, d_U: int, stream: int = 0) -> None:
        _check(_lib.of_cnn_predict(self._h, d_I0s, d_I1s, int(H), int(W), d_U, stream or None),
               "cnn_predict")

    def gradients(self, d_X: int, d_T: int, B: int, H: int, W: int):
        g = np.empty(self.num_params, np.float32)
        loss = C.c_double()
        _check(_lib.of_cnn_gradients(self._h, d_X, d_T, int(B), int(H), int(W), g.ctypes.data,
                                     C.byref(loss)), "cnn_gradients")
        return loss.value, g

    def train_step(self, d_X: int, d_T: int, B: int, H: int, W: int, lr: float = 1e-3) -> float:
        loss = C.c_double()
        _check(_lib.of_cnn_train_step(self._h, d_X, d_T, int(B), int(H), int(W), C.c_float(lr),
                                      C.byref(loss)), "cnn_train_step")
        return loss.value

    def warm_run(self, tv, d_I0: int, d_I1: int, d_U: int, first_scale: int = FIRST_SCALE,
                 warps: int = 2, max_iterations: int = 100, stream: int = 0):
        """Warm-started TV-L1 of a device u8 pair on solver `tv`: the
        predicted coarsest flow, carried to `first_scale`, seeds the solve of
        scales first_scale .. 0.  Returns (fell_back, energy of the
        prediction, energy of zero flow) at first_scale."""
        fb = C.c_int()
        e = (C.c_double * 2)()
        _check(_lib.of_tvl1_warm_run_device_u8(tv.handle, self._h, d_I0, d_I1, int(first_scale),
                                               int(warps), int(max_iterations), d_U, C.byref(fb), e,
                                               stream or None), "warm_run")
        return bool(fb.value), e[0], e[1]


def mfma_product(A, B) -> np.ndarray:
    """16x16 product through the convolution kernel's MFMA path (test aid)."""
    a = np.ascontiguousarray(A, np.float32)
    b = np.ascontiguousarray(B, np.float32)
    d = np.empty((16, 16), np.float32)
    _check(_lib.of_dbg_cnn_mfma(a.ctypes.data, b.ctypes.data, d.ctypes.data), "dbg_cnn_mfma")
    return d
