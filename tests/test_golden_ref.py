"""Golden vectors of the reference's own bicubic sampler, bicubic warp,
zoom_in and zoom_size (bicubic_interpolation.cpp:17-163, zoom.cpp:14-85),
written by tests/golden/make_ref_vectors.py from the reference sources
compiled into oracle/_ref.  The oracle must reproduce them bit for bit (CPU),
and so must the GPU's zoom_in between scales (of_tvl1_lift_flow = zoom_in
times 1/zfactor = 2, which is exact in float)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import gpu_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_bicubic_zoom.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("border", [0, 1])
def test_oracle_bicubic_at_matches_reference_vectors(g, k, border):
    img, pts = g[f"at{k}_img"], g[f"at{k}_pts"]
    want = g[f"at{k}_out_b{border}"]
    got = np.array([O.bicubic_at(img, x, y, bool(border)) for x, y in pts], np.float32)
    assert np.array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("border", [0, 1])
def test_oracle_bicubic_warp_matches_reference_vectors(g, k, border):
    got = O.bicubic_warp(g["warp_img"], g[f"warp{k}_u"], g[f"warp{k}_v"], bool(border))
    assert np.array_equal(_bits(got), _bits(g[f"warp{k}_out_b{border}"]))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_oracle_zoom_in_matches_reference_vectors(g, k):
    sw, sh, dw, dh = (int(v) for v in g["zoom_cases"][k])
    for c in (0, 1):
        got = O.zoom_in(g[f"zoom{k}_in{c}"], dw, dh)
        assert got.shape == (dh, dw)
        assert np.array_equal(_bits(got), _bits(g[f"zoom{k}_out{c}"]))


def test_oracle_zoom_size_matches_reference_vectors(g):
    for nx, ny, f100, nxx, nyy in g["zoom_size"]:
        assert O.zoom_size(int(nx), int(ny), f100 / 100) == (nxx, nyy), (nx, ny, f100)


def test_vectors_cover_both_border_modes_and_outside_points(g):
    # the border_out=true outputs must hold zeros (taps outside) and the
    # border_out=false outputs none of them at the same points
    out_t, out_f = g["at2_out_b1"], g["at2_out_b0"]
    assert (out_t == 0).sum() > 50 and (out_f[out_t == 0] != 0).any()


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
@pytest.mark.parametrize("k", [0, 1, 2])
def test_gpu_zoom_in_between_scales_matches_reference_vectors(g, k):
    import torch
    from amdflow import TV_L1
    sw, sh, dw, dh = (int(v) for v in g["zoom_cases"][k])
    tv = TV_L1(dw, dh)            # scale 1 of a (dw, dh) frame is (sw, sh)
    assert tv.nscales >= 2 and tv.scale_size(1) == (sw, sh)
    src = np.concatenate([g[f"zoom{k}_in0"].ravel(), g[f"zoom{k}_in1"].ravel()])
    dsrc = torch.from_numpy(src).cuda()
    ddst = torch.full((2 * dw * dh,), np.nan, dtype=torch.float32, device="cuda")
    tv.lift_flow(1, dsrc.data_ptr(), 0, ddst.data_ptr())
    torch.cuda.synchronize()
    got = ddst.cpu().numpy().reshape(2, dh, dw)
    for c in (0, 1):
        want = np.float32(2.0) * g[f"zoom{k}_out{c}"]
        assert np.array_equal(_bits(got[c]), _bits(want)), c
