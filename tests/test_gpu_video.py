"""The reference application's frame loop (App::run,
src/tvl1/sequential/main.cpp:129-238) over a YUV4MPEG2 file: one TV-L1
solve per pair of consecutive frames, each flow equal to the oracle's, the
pair limit, and the colour rendering of App::flowToColor."""
import numpy as np
import pytest

import oracle as O
from amdflow import io, synth, video

pytestmark = pytest.mark.gpu

W, H = 128, 96


def _frames(n=5):
    """A moving pattern: frame k is the base warped by k times a smooth flow."""
    out = []
    for k in range(n):
        I0, I1, _, _ = synth.make_pair(W, H, 1.5 * k + 0.5, 77, bumps=True)
        out.append(synth.quantize(I1 if k else I0))
    return out


def _write_y4m(path, frames):
    chroma = 2 * ((W + 1) // 2) * ((H + 1) // 2)
    data = [f"YUV4MPEG2 W{W} H{H} F30:1 Ip C420mpeg2\n".encode()]
    for f in frames:
        data.append(b"FRAME\n" + f.tobytes() + bytes([128]) * chroma)
    path.write_bytes(b"".join(data))


def test_sequence_flows_match_oracle(tmp_path):
    frames = _frames()
    p = tmp_path / "seq.y4m"
    _write_y4m(p, frames)
    with video.Y4MReader(str(p)) as r:
        got = list(video.run_sequence(r, nscales=4))
    assert [k for k, _, _ in got] == [0, 1, 2, 3]
    for k, u, bgr in got:
        uo, _ = O.tvl1(frames[k], frames[k + 1], O.default_params(nscales=4))
        assert bgr is None
        assert np.array_equal(np.ravel(u), np.ravel(uo)), k


def test_pair_limit_and_rendering(tmp_path):
    frames = _frames(4)
    p = tmp_path / "seq.y4m"
    _write_y4m(p, frames)
    with video.open_video(str(p)) as r:
        got = list(video.run_sequence(r, max_pairs=2, render=True, nscales=3))
    assert len(got) == 2
    for k, u, bgr in got:
        assert bgr.shape == (H, W, 3) and bgr.dtype == np.uint8
        assert np.array_equal(bgr, io.flow_to_color(u, W, H))
