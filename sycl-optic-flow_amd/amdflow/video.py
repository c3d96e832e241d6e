"""Frame sources and the reference application's frame loop, without OpenCV.

The reference application (src/tvl1/sequential/main.cpp:129-238, App::run)
reads a video with cv::VideoCapture (main.cpp:72-90), converts each frame to
gray with cv::cvtColor(..., COLOR_BGR2GRAY), solves every pair of
consecutive frames with TV_L1::runDualTVL1Multiscale(I0, I1), renders the
flow with App::flowToColor and, without a display, stops after 100 pairs.
This module keeps that call contract for files a process can read without a
video decoder:

- YUV4MPEG2 (.y4m), 8 bits per sample, any chroma layout: the luma plane is
  the gray frame (a decoder's YUV -> BGR conversion followed by
  COLOR_BGR2GRAY reproduces luma up to the range convention and rounding);
- raw files of back-to-back gray (w*h bytes) or BGR24 (w*h*3 bytes) frames,
  BGR converted with `bgr_to_gray`.

Frames are read through a memory map, one frame at a time."""
from __future__ import annotations

import mmap
import os

import numpy as np

# 8-bit chroma layouts of YUV4MPEG2: bytes of the chroma (and alpha) planes
# per frame as a function of (width, height)
_CHROMA_BYTES = {
    "420jpeg": lambda w, h: 2 * ((w + 1) // 2) * ((h + 1) // 2),
    "420paldv": lambda w, h: 2 * ((w + 1) // 2) * ((h + 1) // 2),
    "420mpeg2": lambda w, h: 2 * ((w + 1) // 2) * ((h + 1) // 2),
    "420": lambda w, h: 2 * ((w + 1) // 2) * ((h + 1) // 2),
    "422": lambda w, h: 2 * ((w + 1) // 2) * h,
    "411": lambda w, h: 2 * ((w + 3) // 4) * h,
    "444": lambda w, h: 2 * w * h,
    "444alpha": lambda w, h: 3 * w * h,
    "mono": lambda w, h: 0,
}


class Y4MReader:
    """Gray frames (the luma plane) of a YUV4MPEG2 file.

    Parses the stream header (W, H, C; other tags are kept in `tags`),
    then the frames: each is "FRAME", optional parameters, a newline, and
    the planes.  Raises ValueError on a malformed header or frame marker, a
    sample depth above 8 bits, and a truncated last frame."""

    def __init__(self, path: str):
        self.path = path
        self._f = open(path, "rb")
        try:
            size = os.fstat(self._f.fileno()).st_size
            if size == 0:
                raise ValueError(f"{path}: empty file")
            self._m = mmap.mmap(self._f.fileno(), 0, access=mmap.ACCESS_READ)
        except Exception:
            self._f.close()
            raise
        end = self._m.find(b"\n")
        if end < 0 or not self._m[:10] == b"YUV4MPEG2 ":
            self.close()
            raise ValueError(f"{path}: not a YUV4MPEG2 stream")
        tags = self._m[10:end].decode("ascii", "replace").split()
        self.tags = {t[0]: t[1:] for t in tags if t}
        try:
            self.width = int(self.tags["W"])
            self.height = int(self.tags["H"])
        except (KeyError, ValueError):
            self.close()
            raise ValueError(f"{path}: header without W and H")
        self.colorspace = self.tags.get("C", "420jpeg")
        if self.colorspace not in _CHROMA_BYTES:
            self.close()
            raise ValueError(f"{path}: colorspace C{self.colorspace} is not an 8-bit layout "
                             f"({', '.join(sorted(_CHROMA_BYTES))})")
        if self.width < 1 or self.height < 1:
            self.close()
            raise ValueError(f"{path}: bad frame size {self.width}x{self.height}")
        self._luma = self.width * self.height
        self._frame_bytes = self._luma + _CHROMA_BYTES[self.colorspace](self.width, self.height)
        self._first = end + 1
        self._offsets = None

    def _scan(self):
        """Offsets of every frame's luma plane."""
        offs, pos, m = [], self._first, self._m
        n = len(m)
        while pos < n:
            if m[pos:pos + 5] != b"FRAME":
                raise ValueError(f"{self.path}: frame marker expected at byte {pos}")
            nl = m.find(b"\n", pos)
            if nl < 0:
                raise ValueError(f"{self.path}: unterminated frame header at byte {pos}")
            data = nl + 1
            if data + self._frame_bytes > n:
                raise ValueError(f"{self.path}: truncated frame {len(offs)}")
            offs.append(data)
            pos = data + self._frame_bytes
        return offs

    def __len__(self) -> int:
        if self._offsets is None:
            self._offsets = self._scan()
        return len(self._offsets)

    def frame(self, i: int) -> np.ndarray:
        """Frame i as a (height, width) uint8 array (a copy)."""
        if self._offsets is None:
            self._offsets = self._scan()
        o = self._offsets[i]
        return np.frombuffer(self._m, np.uint8, self._luma, o).reshape(self.height, self.width).copy()

    def __iter__(self):
        for i in range(len(self)):
            yield self.frame(i)

    def close(self) -> None:
        m = getattr(self, "_m", None)
        if m is not None:
            m.close()
            self._m = None
        if getattr(self, "_f", None):
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def bgr_to_gray(bgr) -> np.ndarray:
    """8-bit BGR (..., 3) to gray as COLOR_BGR2GRAY defines it: the ITU-R
    BT.601 weights 0.299 R + 0.587 G + 0.114 B, in 14-bit fixed point
    (4899, 9617, 1868; they sum to 2^14) with rounding."""
    a = np.asarray(bgr)
    if a.dtype != np.uint8 or a.shape[-1] != 3:
        raise ValueError("bgr_to_gray takes uint8 arrays with 3 channels (B, G, R)")
    b = a[..., 0].astype(np.uint32)
    g = a[..., 1].astype(np.uint32)
    r = a[..., 2].astype(np.uint32)
    return ((r * 4899 + g * 9617 + b * 1868 + 8192) >> 14).astype(np.uint8)


class RawReader:
    """Back-to-back frames of `width` x `height`, gray (1 byte per pixel)
    or BGR24 (3 bytes, converted with bgr_to_gray).  A trailing partial
    frame is an error."""

    def __init__(self, path: str, width: int, height: int, pixel_format: str = "gray"):
        if pixel_format not in ("gray", "bgr24"):
            raise ValueError("pixel_format is 'gray' or 'bgr24'")
        if width < 1 or height < 1:
            raise ValueError(f"bad frame size {width}x{height}")
        self.path, self.width, self.height, self.pixel_format = path, int(width), int(height), pixel_format
        self._bpp = 1 if pixel_format == "gray" else 3
        self._frame_bytes = self.width * self.height * self._bpp
        size = os.path.getsize(path)
        if size % self._frame_bytes:
            raise ValueError(f"{path}: {size} bytes is not a whole number of "
                             f"{self.width}x{self.height} {pixel_format} frames")
        self._n = size // self._frame_bytes
        self._f = open(path, "rb")
        self._m = mmap.mmap(self._f.fileno(), 0, access=mmap.ACCESS_READ) if size else None

    def __len__(self) -> int:
        return self._n

    def frame(self, i: int) -> np.ndarray:
        if not 0 <= i < self._n:
            raise IndexError(i)
        a = np.frombuffer(self._m, np.uint8, self._frame_bytes, i * self._frame_bytes)
        if self._bpp == 1:
            return a.reshape(self.height, self.width).copy()
        return bgr_to_gray(a.reshape(self.height, self.width, 3))

    def __iter__(self):
        for i in range(self._n):
            yield self.frame(i)

    def close(self) -> None:
        if self._m is not None:
            self._m.close()
            self._m = None
        if self._f:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def open_video(path: str, width: int | None = None, height: int | None = None,
               pixel_format: str = "gray"):
    """A Y4MReader for .y4m files, else a RawReader of the given size."""
    if path.lower().endswith(".y4m"):
        return Y4MReader(path)
    if width is None or height is None:
        raise ValueError("raw frames need width and height")
    return RawReader(path, width, height, pixel_format)


def run_sequence(frames, max_pairs: int = 100, render: bool = False, solver=None, **params):
    """App::run's loop over a frame source (an iterable of (H, W) uint8
    frames, e.g. a Y4MReader): one TV-L1 solve per pair of consecutive
    frames, at most `max_pairs` pairs (the reference stops after 100
    without a display).  Yields (pair index, flow U planar float32 [2*H*W],
    BGR rendering or None).  `solver` reuses a TV_L1 handle of the frame
    size; otherwise one is created with `params` (TV_L1's keyword
    arguments) and closed at the end."""
    from . import TV_L1
    from .io import flow_to_color

    it = iter(frames)
    try:
        prev = np.ascontiguousarray(next(it), np.uint8)
    except StopIteration:
        return
    h, w = prev.shape
    own = solver is None
    tv = TV_L1(w, h, **params) if own else solver
    try:
        for k, cur in enumerate(it):
            if k >= max_pairs:
                break
            cur = np.ascontiguousarray(cur, np.uint8)
            if cur.shape != (h, w):
                raise ValueError(f"frame {k + 1} is {cur.shape[1]}x{cur.shape[0]}, expected {w}x{h}")
            tv.runDualTVL1Multiscale(prev, cur)
            u = tv.getU().copy()
            yield k, u, (flow_to_color(u, w, h) if render else None)
            prev = cur
    finally:
        if own:
            tv.close()
