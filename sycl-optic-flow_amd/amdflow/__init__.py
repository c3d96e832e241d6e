"""Python binding of libamdflow.so (the MI355X optical-flow library).

The classes mirror the reference's per-algorithm entry points so that code
written against the reference reads the same:

  TV_L1(width, height, tau=0.25, lambda_=0.15, theta=0.3, nscales=100,
        zfactor=0.5, warps=5, epsilon=0.01)       src/tvl1/sequential/tvl1.hpp:12-17
      .runDualTVL1Multiscale(I0, I1)  (uint8 pair) / (float frame, streaming)
      .getU()                         planar [u1 | u2]
  HornSchunck: initFlow / ComputeFlow / deleteFlow_mem   src/hs/cuda/flow.cuh:19-33
  LucasKanade(spac, temp, window, nx, ny)                src/lk/cuda/lucaskanade.cuh:6-35
      .copy_frames_circularbuffer_GPU_wrapper(frame, temp_size, fr, frame_size)
      .lucas_kanade(Vx, Vy, iframe)

Errors raise OpticalFlowError with the library's status code (the reference
throws or terminates).  The library has no CPU fallback: if libamdflow.so is
missing or cannot be loaded, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import sys
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libamdflow.so")


class OpticalFlowError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {status_string(status)} (status {status})")
        self.status = status


def _one_hip_runtime() -> None:
    """PyTorch-ROCm ships its own HIP runtime (NEEDED as "libamdhip64.so",
    soname "libamdhip64.so.7"); libamdflow.so needs "libamdhip64.so.7".  If
    torch is loaded first, the dynamic loader binds libamdflow.so to torch's
    runtime.  In the other order two HIP runtimes end up in the process and
    the second one to initialise finds no device.  So when torch is
    installed it is imported first (AMDFLOW_NO_TORCH=1 skips this)."""
    import importlib.util
    import sys
    if "torch" in sys.modules or os.environ.get("AMDFLOW_NO_TORCH") == "1":
        return
    if importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() "
            "(the HIP library is required, there is no CPU fallback)")
    _one_hip_runtime()
    return C.CDLL(LIB_PATH)


_lib = _load()


class TVL1Params(C.Structure):
    _fields_ = [
        ("tau", C.c_float), ("lambda_", C.c_float), ("theta", C.c_float),
        ("nscales", C.c_int), ("zfactor", C.c_float), ("warps", C.c_int),
        ("epsilon", C.c_float), ("max_iterations", C.c_int),
    ]


class TVL1Stats(C.Structure):
    _fields_ = [
        ("nscales", C.c_int), ("warps", C.c_int),
        ("pixel_iterations", C.c_longlong), ("pixel_warps", C.c_longlong),
        ("pixels_all_scales", C.c_longlong), ("launches", C.c_longlong),
        ("algorithmic_bytes", C.c_double), ("host_launch_seconds", C.c_double),
        ("host_wait_seconds", C.c_double),
    ]


class TVL1KernelProfile(C.Structure):
    _fields_ = [
        ("iters_per_launch", C.c_int), ("rows_per_wave", C.c_int), ("persistent", C.c_int),
        ("launches", C.c_longlong), ("pixel_iterations", C.c_longlong),
        ("milliseconds", C.c_double), ("algorithmic_bytes", C.c_double),
    ]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)

# exported symbols and their ctypes signatures (restype, argtypes)
SIGNATURES = {
    "of_status_string": (C.c_char_p, [C.c_int]),
    "of_build_info": (C.c_char_p, []),
    "of_tvl1_default_params": (None, [C.POINTER(TVL1Params)]),
    "of_tvl1_create": (C.c_int, [C.POINTER(TVL1Params), C.c_int, C.c_int, C.POINTER(_vp)]),
    "of_tvl1_destroy": (None, [_vp]),
    "of_tvl1_nscales": (C.c_int, [_vp]),
    "of_tvl1_scale_size": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "of_tvl1_run_u8": (C.c_int, [_vp, _vp, _vp]),
    "of_tvl1_run_f32": (C.c_int, [_vp, _vp, _vp]),
    "of_tvl1_push_f32": (C.c_int, [_vp, _vp]),
    "of_tvl1_flow": (_fp, [_vp]),
    "of_tvl1_run_device_u8": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "of_tvl1_run_device_f32": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "of_tvl1_iterations": (C.c_int, [_vp, _ip, C.c_int]),
    "of_tvl1_get_stats": (C.c_int, [_vp, C.POINTER(TVL1Stats)]),
    "of_tvl1_set_launch_rule": (C.c_int, [_vp, C.c_longlong, C.c_int, C.c_int, C.c_int]),
    "of_persist_launch_mode": (C.c_int, []),
    "of_set_persist_launch_mode": (C.c_int, [C.c_int]),
    "of_tvl1_autotune": (C.c_int, [_vp, _vp, _vp, C.c_int, _ip, C.POINTER(C.c_double)]),
    "of_tvl1_set_inflight": (C.c_int, [_vp, C.c_int]),
    "of_tvl1_profile_enable": (C.c_int, [_vp, C.c_int]),
    "of_tvl1_profile_get": (C.c_int, [_vp, C.POINTER(TVL1KernelProfile), C.c_int, _ip]),
    "of_tvl1_scale_timing_enable": (C.c_int, [_vp, C.c_int]),
    "of_tvl1_scale_times": (C.c_int, [_vp, C.POINTER(C.c_double), C.c_int, _ip]),
    "of_tvl1_prepare_device_u8": (C.c_int, [_vp, _vp, _vp, _vp]),
    "of_tvl1_scale_images": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp)]),
    "of_tvl1_solve_prepared": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "of_tvl1_solve_prepared_schedule": (C.c_int, [_vp, _vp, C.c_int, _ip, _ip, _vp, _vp]),
    "of_tvl1_set_init_flow": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "of_tvl1_lift_flow": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp, _vp]),
    "of_tvl1_flow_to_scale": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    "of_tvl1_scale_flow": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "of_tvl1_energy": (C.c_int, [_vp, C.c_int, _vp, C.POINTER(C.c_double), _vp]),
    "of_hs_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "of_hs_destroy": (None, [_vp]),
    "of_hs_compute": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                C.c_int, C.c_int, _vp, _vp]),
    "of_hs_compute_device": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                       C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "of_lk_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "of_lk_destroy": (None, [_vp]),
    "of_lk_push_u8": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int]),
    "of_lk_compute": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "of_lk_pyr_compute": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                    C.c_int, C.c_int, C.c_float, _vp, _vp]),
    "of_lk_pyr_compute_device": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                           C.c_float, C.c_int, C.c_int, C.c_float, _vp, _vp, _vp]),
    "of_lk_pyr_compute_device_noalloc": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                                   C.c_float, C.c_int, C.c_int, C.c_float, _vp,
                                                   _vp, _vp]),
}

for _name, (_res, _args) in SIGNATURES.items():
    if hasattr(_lib, _name):
        _f = getattr(_lib, _name)
        _f.restype = _res
        _f.argtypes = _args


def status_string(status: int) -> str:
    return _lib.of_status_string(int(status)).decode()


def build_info() -> str:
    return _lib.of_build_info().decode()


def _exiting() -> bool:
    """True while the interpreter shuts down: module globals (the library
    handle among them) may already be gone and the process releases the
    device memory anyway, so finalisers skip their destroy calls."""
    try:
        return sys.is_finalizing()
    except Exception:   # sys itself already torn down
        return True


def _check(status: int, what: str) -> None:
    if status != 0:
        raise OpticalFlowError(status, what)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _as(a, dtype, shape=None) -> np.ndarray:
    arr = np.ascontiguousarray(a, dtype=dtype)
    if shape is not None and arr.size != int(np.prod(shape)):
        raise ValueError(f"expected {int(np.prod(shape))} elements, got {arr.size}")
    return arr


PERSIST_PLAIN, PERSIST_HOST_PACED = 0, 1


def persist_launch_mode() -> int:
    """Process-wide launch mode of the warpings whose launch rule is
    persistent: PERSIST_PLAIN (one launch per warping with a residency check,
    the default) or PERSIST_HOST_PACED (after a residency failure,
    AMDFLOW_HOST_PACED=1, or set_persist_launch_mode)."""
    return int(_lib.of_persist_launch_mode())


def set_persist_launch_mode(mode: int) -> None:
    _check(_lib.of_set_persist_launch_mode(int(mode)), "set_persist_launch_mode")


def default_tvl1_params() -> TVL1Params:
    p = TVL1Params()
    _lib.of_tvl1_default_params(C.byref(p))
    return p


class TV_L1:
    """TV-L1 solver handle; mirrors class TV_L1 of src/tvl1/sequential/tvl1.hpp:10-41.

    runDualTVL1Multiscale(I0, I1) with two uint8 frames solves the pair
    (tvl1.hpp:17); with one float frame it is the streaming call of the GPU
    variants, I0 being the previous call's frame (src/tvl1/cuda/tvl1.cuh:19).
    """

    def __init__(self, width: int, height: int, tau: float = 0.25, lambda_: float = 0.15,
                 theta: float = 0.3, nscales: int = 100, zfactor: float = 0.5, warps: int = 5,
                 epsilon: float = 0.01, max_iterations: int = 300):
        self.width, self.height = int(width), int(height)
        p = TVL1Params(tau, lambda_, theta, nscales, zfactor, warps, epsilon, max_iterations)
        self.params = p
        h = _vp()
        _check(_lib.of_tvl1_create(C.byref(p), self.width, self.height, C.byref(h)), "TV_L1")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.of_tvl1_destroy(self._h)
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
    def nscales(self) -> int:
        return int(_lib.of_tvl1_nscales(self._h))

    def scale_size(self, s: int) -> tuple[int, int]:
        nx, ny = C.c_int(), C.c_int()
        _check(_lib.of_tvl1_scale_size(self._h, s, C.byref(nx), C.byref(ny)), "scale_size")
        return nx.value, ny.value

    def runDualTVL1Multiscale(self, I0, I1=None) -> None:
        n = self.width * self.height
        if I1 is None:
            f = _as(I0, np.float32, (n,))
            _check(_lib.of_tvl1_push_f32(self._h, _ptr(f)), "runDualTVL1Multiscale")
            return
        if np.asarray(I0).dtype == np.uint8 and np.asarray(I1).dtype == np.uint8:
            a, b = _as(I0, np.uint8, (n,)), _as(I1, np.uint8, (n,))
            _check(_lib.of_tvl1_run_u8(self._h, _ptr(a), _ptr(b)), "runDualTVL1Multiscale")
        else:
            a, b = _as(I0, np.float32, (n,)), _as(I1, np.float32, (n,))
            _check(_lib.of_tvl1_run_f32(self._h, _ptr(a), _ptr(b)), "runDualTVL1Multiscale")

    def getU(self) -> np.ndarray:
        n = self.width * self.height
        p = _lib.of_tvl1_flow(self._h)
        return np.ctypeslib.as_array(p, shape=(2 * n,)).copy()

    def flow(self) -> tuple[np.ndarray, np.ndarray]:
        u = self.getU()
        n = self.width * self.height
        return (u[:n].reshape(self.height, self.width), u[n:].reshape(self.height, self.width))

    def run_device_u8(self, d_I0: int, d_I1: int, d_U: int, stream: int = 0) -> None:
        _check(_lib.of_tvl1_run_device_u8(self._h, d_I0, d_I1, d_U, stream or None),
               "run_device_u8")

    def run_device_f32(self, d_I0: int, d_I1: int, d_U: int, stream: int = 0) -> None:
        _check(_lib.of_tvl1_run_device_f32(self._h, d_I0, d_I1, d_U, stream or None),
               "run_device_f32")

    def iterations(self) -> np.ndarray:
        ns, w = self.nscales, int(self.params.warps)
        out = np.zeros(ns * w, dtype=np.int32)
        _check(_lib.of_tvl1_iterations(self._h, out.ctypes.data_as(_ip), out.size), "iterations")
        return out.reshape(ns, w)

    def stats(self) -> dict:
        st = TVL1Stats()
        _check(_lib.of_tvl1_get_stats(self._h, C.byref(st)), "stats")
        return {f: getattr(st, f) for f, _ in TVL1Stats._fields_}

    def set_launch_rule(self, min_pixels: int, iters_per_launch: int, rows_per_wave: int,
                        persistent: int = 0) -> None:
        """persistent: 0 host-paced, 1 (or True) one launch per warping with
        grid barriers, 2 one launch per warping with neighbour flags."""
        _check(_lib.of_tvl1_set_launch_rule(self._h, int(min_pixels), int(iters_per_launch),
                                            int(rows_per_wave), int(persistent)),
               "set_launch_rule")

    AUTOTUNE_SHAPES = [(1, 2), (2, 2), (1, 4), (2, 4), (4, 4), (1, 8), (2, 8), (4, 8), (8, 8)]
    AUTOTUNE_VARIANTS = ([(k, r, 0) for k, r in AUTOTUNE_SHAPES] + [(k, r, 1) for k, r in AUTOTUNE_SHAPES]
                         + [(1, 0, 0), (2, 0, 0), (4, 0, 0)]      # square 16x16 tiles, host-paced
                         + [(1, 0, 1), (2, 0, 1), (4, 0, 1)]      # and persistent
                         + [(k, r, 2) for k, r in AUTOTUNE_SHAPES]  # neighbour flags
                         + [(1, 0, 2), (2, 0, 2), (4, 0, 2)])

    def autotune(self, d_I0: int, d_I1: int, reps: int = 2):
        """Times every launch variant (K, R, persistent) per scale on a device u8
        pair and keeps the fastest per scale; returns (chosen [(K, R, P)] per
        scale, ms table [36 x nscales])."""
        ns = self.nscales
        chosen = np.zeros(3 * ns, np.int32)
        table = np.zeros(len(self.AUTOTUNE_VARIANTS) * ns, np.float64)
        _check(_lib.of_tvl1_autotune(self._h, d_I0, d_I1, int(reps), chosen.ctypes.data_as(_ip),
                                     table.ctypes.data_as(C.POINTER(C.c_double))), "autotune")
        return ([tuple(int(v) for v in chosen[3 * s:3 * s + 3]) for s in range(ns)],
                table.reshape(len(self.AUTOTUNE_VARIANTS), ns))

    def set_inflight(self, launches: int) -> None:
        _check(_lib.of_tvl1_set_inflight(self._h, int(launches)), "set_inflight")

    def profile(self, enable: bool = True) -> None:
        _check(_lib.of_tvl1_profile_enable(self._h, int(enable)), "profile_enable")

    def profile_results(self) -> list[dict]:
        arr = (TVL1KernelProfile * 16)()
        n = C.c_int()
        _check(_lib.of_tvl1_profile_get(self._h, arr, 16, C.byref(n)), "profile_get")
        return [{f: getattr(arr[i], f) for f, _ in TVL1KernelProfile._fields_}
                for i in range(min(n.value, 16))]

    def scale_timing(self, enable: bool = True) -> None:
        """Per-scale device time of the following solves (resets the sums)."""
        _check(_lib.of_tvl1_scale_timing_enable(self._h, int(enable)), "scale_timing_enable")

    def scale_times(self) -> tuple[list[float], int]:
        """(summed milliseconds per scale, finest first; solves timed)."""
        arr = (C.c_double * 32)()
        n = C.c_int()
        _check(_lib.of_tvl1_scale_times(self._h, arr, 32, C.byref(n)), "scale_times")
        return [arr[s] for s in range(self.nscales)], n.value

    def clear_launch_rules(self) -> None:
        _check(_lib.of_tvl1_set_launch_rule(self._h, -1, 1, 2, 0), "set_launch_rule")

    # two-phase solve (warm starts)
    def prepare_device_u8(self, d_I0: int, d_I1: int, stream: int = 0) -> None:
        _check(_lib.of_tvl1_prepare_device_u8(self._h, d_I0, d_I1, stream or None), "prepare")

    def scale_images(self, s: int) -> tuple[int, int]:
        """Device pointers to the prepared pyramid images of scale s."""
        a, b = _vp(), _vp()
        _check(_lib.of_tvl1_scale_images(self._h, int(s), C.byref(a), C.byref(b)), "scale_images")
        return a.value, b.value

    def set_init_flow(self, level: int, u, v) -> None:
        """Initial flow of the next solve at scale `level`: numpy arrays (host)
        or integer device pointers, that scale's size each; None, None clears."""
        if u is None and v is None:
            _check(_lib.of_tvl1_set_init_flow(self._h, int(level), None, None), "set_init_flow")
            return
        keep = []

        def ptr(x):
            if isinstance(x, int):
                return x
            a = np.ascontiguousarray(x, np.float32)
            keep.append(a)
            return a.ctypes.data
        _check(_lib.of_tvl1_set_init_flow(self._h, int(level), ptr(u), ptr(v)), "set_init_flow")

    def solve_prepared(self, d_U: int, d_Uinit: int = 0, first_scale: int = -1, warps: int = 0,
                       max_iterations: int = 0, stream: int = 0) -> None:
        """Solves scales first_scale .. 0 (-1: all) of the prepared pair from
        the device flow d_Uinit at first_scale (0: zero flow)."""
        _check(_lib.of_tvl1_solve_prepared(self._h, d_Uinit or None, int(first_scale), int(warps),
                                           int(max_iterations), d_U, stream or None), "solve_prepared")

    def solve_prepared_schedule(self, d_U: int, warps, max_iterations, d_Uinit: int = 0,
                                first_scale: int = -1, stream: int = 0) -> None:
        """solve_prepared with per-scale limits: warps[s] warpings and
        max_iterations[s] iterations per warping at scale s (finest first;
        0 = the handle's value)."""
        ns = self.nscales
        w = np.zeros(ns, np.int32)
        m = np.zeros(ns, np.int32)
        w[:len(warps)] = warps
        m[:len(max_iterations)] = max_iterations
        _check(_lib.of_tvl1_solve_prepared_schedule(self._h, d_Uinit or None, int(first_scale),
                                                    w.ctypes.data_as(_ip), m.ctypes.data_as(_ip),
                                                    d_U, stream or None), "solve_prepared_schedule")

    def lift_flow(self, from_scale: int, d_from: int, to_scale: int, d_to: int, stream: int = 0) -> None:
        _check(_lib.of_tvl1_lift_flow(self._h, int(from_scale), d_from, int(to_scale), d_to,
                                      stream or None), "lift_flow")

    def flow_to_scale(self, d_U: int, s: int, d_Us: int, stream: int = 0) -> None:
        _check(_lib.of_tvl1_flow_to_scale(self._h, d_U, int(s), d_Us, stream or None), "flow_to_scale")

    def scale_flow(self, s: int, d_Us: int, stream: int = 0) -> None:
        _check(_lib.of_tvl1_scale_flow(self._h, int(s), d_Us, stream or None), "scale_flow")

    def energy(self, s: int, d_Us: int, stream: int = 0) -> tuple[float, float, float]:
        """(E, data sum, TV sum) of the device flow d_Us at scale s."""
        e = (C.c_double * 3)()
        _check(_lib.of_tvl1_energy(self._h, int(s), d_Us, e, stream or None), "energy")
        return e[0], e[1], e[2]


from .hs import HornSchunck  # noqa: E402
from .lk import LucasKanade, lk_pyramidal, lk_pyramidal_device  # noqa: E402
