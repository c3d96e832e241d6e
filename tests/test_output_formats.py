"""Output formats: Middlebury .flo round trip and byte layout (host code of
the library, no GPU), and the colour oracle pinned against the textbook
HSV -> RGB conversion (colorsys)."""
import colorsys
import struct

import numpy as np

import flow_color as FC
from amdflow import io


def test_flo_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(0)
    w, h = 7, 5
    U = rng.normal(size=2 * w * h).astype(np.float32)
    p = str(tmp_path / "a.flo")
    io.write_flo(p, U, w, h)
    raw = open(p, "rb").read()
    assert raw[:4] == b"PIEH" and struct.unpack("<ii", raw[4:12]) == (w, h)
    assert len(raw) == 12 + 8 * w * h
    body = np.frombuffer(raw[12:], "<f4").reshape(h, w, 2)
    assert np.array_equal(body[..., 0].ravel(), U[:w * h])
    assert np.array_equal(body[..., 1].ravel(), U[w * h:])
    V, w2, h2 = io.read_flo(p)
    assert (w2, h2) == (w, h) and np.array_equal(V, U)


def test_flo_rejects_bad_tag(tmp_path):
    p = tmp_path / "bad.flo"
    p.write_bytes(b"XXXX" + struct.pack("<ii", 2, 2) + b"\0" * 32)
    try:
        io.read_flo(str(p))
    except Exception as e:
        assert "flo_read" in str(e)
    else:
        raise AssertionError("bad tag accepted")


def test_hsv_oracle_matches_colorsys():
    H = np.arange(180, dtype=np.uint8).repeat(6)
    S = np.tile(np.array([0, 60, 128, 200, 254, 255], np.uint8), 180)
    V = np.tile(np.array([255, 10, 128, 77, 200, 255], np.uint8), 180)
    got = FC.hsv_to_bgr(H, S, V).astype(int)
    for i in range(H.size):
        r, g, b = colorsys.hsv_to_rgb(int(H[i]) * 2 / 360.0, int(S[i]) / 255.0, int(V[i]) / 255.0)
        ref = np.array([b, g, r]) * 255.0
        assert np.max(np.abs(got[i] - ref)) <= 0.5 + 1e-3, (H[i], S[i], V[i], got[i], ref)


def test_flow_color_oracle_directions():
    # unit directions scaled to |flow| = 25.5 -> value 255
    U = np.array([25.5, 0.0, -25.5, 0.0, 0.0, 25.5, 0.0, -25.5], np.float32)   # [u | v]
    img = FC.flow_to_color(U, 4, 1)[0].astype(int)
    assert img[0].tolist() == [0, 0, 255]        # +u: hue 0, red
    # +v: 90 deg -> hue 45 (90 deg): red = 1 - f with f = 45*(6/180) - 1 = 0.50000006 in float
    assert img[1].tolist() == [0, 255, 127]
    # -u: atan2 = 180 deg -> hue 90 (180 deg): cyan;  -v: negative angle saturates to hue 0
    assert img[2].tolist() == [255, 255, 0]
    assert img[3].tolist() == [0, 0, 255]
