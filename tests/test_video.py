"""Frame sources without OpenCV (amdflow.video): YUV4MPEG2 and raw files,
and the BGR -> gray conversion of the reference application."""
import numpy as np
import pytest

from amdflow import video


def _y4m(path, frames, w, h, colorspace="420jpeg", frame_params=None, extra=b""):
    """Writes a YUV4MPEG2 file: luma from `frames`, chroma planes filled with
    a running byte pattern (so a wrong plane size shifts the next frame)."""
    chroma = video._CHROMA_BYTES[colorspace](w, h)
    hdr = f"YUV4MPEG2 W{w} H{h} F25:1 Ip A1:1" + ("" if colorspace == "420jpeg" else f" C{colorspace}")
    out = [hdr.encode() + extra + b"\n"]
    for k, f in enumerate(frames):
        p = (frame_params or {}).get(k, b"")
        out.append(b"FRAME" + p + b"\n" + f.astype(np.uint8).tobytes() + bytes((k * 7 + i) % 251 for i in range(chroma)))
    path.write_bytes(b"".join(out))


@pytest.mark.parametrize("cs", sorted(video._CHROMA_BYTES))
@pytest.mark.parametrize("w,h", [(7, 5), (16, 9), (1, 1)])
def test_y4m_luma_planes_every_layout(tmp_path, cs, w, h):
    rng = np.random.default_rng(w * 31 + h)
    frames = [rng.integers(0, 256, (h, w)) for _ in range(3)]
    p = tmp_path / "a.y4m"
    _y4m(p, frames, w, h, cs, frame_params={1: b" Ixyz"})
    with video.Y4MReader(str(p)) as r:
        assert (r.width, r.height, len(r), r.colorspace) == (w, h, 3, cs)
        for k, f in enumerate(r):
            assert f.dtype == np.uint8 and f.shape == (h, w)
            assert np.array_equal(f, frames[k])


def test_y4m_chroma_plane_sizes():
    """4:2:0 rounds the half width and height up; 4:1:1 a quarter width."""
    assert video._CHROMA_BYTES["420jpeg"](7, 5) == 2 * 4 * 3
    assert video._CHROMA_BYTES["422"](7, 5) == 2 * 4 * 5
    assert video._CHROMA_BYTES["411"](7, 5) == 2 * 2 * 5
    assert video._CHROMA_BYTES["444alpha"](7, 5) == 3 * 35
    assert video._CHROMA_BYTES["mono"](7, 5) == 0


def test_y4m_errors(tmp_path):
    p = tmp_path / "bad.y4m"
    p.write_bytes(b"RIFF....")
    with pytest.raises(ValueError):
        video.Y4MReader(str(p))
    p.write_bytes(b"YUV4MPEG2 W4 H4 C420p10\n")
    with pytest.raises(ValueError, match="8-bit"):
        video.Y4MReader(str(p))
    p.write_bytes(b"YUV4MPEG2 H4\n")
    with pytest.raises(ValueError):
        video.Y4MReader(str(p))
    f = np.zeros((4, 4), np.uint8)
    _y4m(p, [f, f], 4, 4, "mono")
    p.write_bytes(p.read_bytes()[:-3])   # truncate the last frame
    with video.Y4MReader(str(p)) as r, pytest.raises(ValueError, match="truncated"):
        len(r)
    _y4m(p, [f], 4, 4, "mono")
    p.write_bytes(p.read_bytes() + b"FRAMX\n")
    with video.Y4MReader(str(p)) as r, pytest.raises(ValueError, match="marker"):
        len(r)


def test_bgr_to_gray_known_values():
    """COLOR_BGR2GRAY of the primaries and greys (BT.601 weights)."""
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0],
                    [128, 128, 128], [10, 20, 30]]], np.uint8)   # B, G, R
    assert video.bgr_to_gray(px).tolist() == [[29, 150, 76, 255, 0, 128, 22]]


def test_bgr_to_gray_is_the_rounded_weighted_sum():
    """Against the weights in floating point: within half a level plus the
    fixed point's error (at most 2^-14 * 255 per channel weight)."""
    rng = np.random.default_rng(1)
    bgr = rng.integers(0, 256, (200, 300, 3)).astype(np.uint8)
    y = video.bgr_to_gray(bgr).astype(np.float64)
    f = 0.114 * bgr[..., 0] + 0.587 * bgr[..., 1] + 0.299 * bgr[..., 2]
    assert np.all(np.abs(y - f) <= 0.5 + 0.05)
    assert np.mean(y == np.floor(f + 0.5)) > 0.99
    # grey stays grey
    g = np.repeat(np.arange(256, dtype=np.uint8)[:, None], 3, axis=1)[None]
    assert np.array_equal(video.bgr_to_gray(g)[0], np.arange(256))


def test_raw_reader_gray_and_bgr(tmp_path):
    rng = np.random.default_rng(2)
    g = rng.integers(0, 256, (3, 6, 5)).astype(np.uint8)
    p = tmp_path / "g.raw"
    p.write_bytes(g.tobytes())
    with video.RawReader(str(p), 5, 6) as r:
        assert len(r) == 3 and all(np.array_equal(a, b) for a, b in zip(r, g))
    c = rng.integers(0, 256, (2, 6, 5, 3)).astype(np.uint8)
    p.write_bytes(c.tobytes())
    with video.open_video(str(p), 5, 6, "bgr24") as r:
        assert len(r) == 2
        assert np.array_equal(r.frame(1), video.bgr_to_gray(c[1]))
    p.write_bytes(c.tobytes()[:-1])
    with pytest.raises(ValueError, match="whole number"):
        video.RawReader(str(p), 5, 6, "bgr24")
    with pytest.raises(ValueError):
        video.open_video(str(p))   # raw needs a size
