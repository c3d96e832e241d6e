"""Allocation failures on every path of the library.

A debug hook makes the n-th device or pinned host allocation from now fail
as a real failure would (of_dbg_fail_alloc_after).  For each path the test
first counts the allocations of a clean call, then fails each of them in
turn.  Every such call must report OF_ERR_ALLOC (the Python layer raises),
must give back every block it took (the library counts its live blocks),
and must leave the library usable: a clean call afterwards gives the same
result as before."""
import numpy as np
import pytest

from amdflow import TV_L1, OpticalFlowError, synth
from amdflow import debug as D

pytestmark = pytest.mark.gpu

OF_ERR_ALLOC = 3
W, H = 192, 144


@pytest.fixture(autouse=True)
def _no_injection():
    D.fail_alloc_after(-1)
    yield
    D.fail_alloc_after(-1)


def _pair():
    I0, I1, _, _ = synth.make_pair(W, H, 3.0, 61, bumps=True)
    return I0, I1, synth.quantize(I0), synth.quantize(I1)


def _tvl1():
    _, _, q0, q1 = _pair()
    with TV_L1(W, H, nscales=3) as tv:
        tv.runDualTVL1Multiscale(q0, q1)
        return tv.getU().copy()


def _hs():
    from amdflow import HornSchunck
    I0, I1, _, _ = _pair()
    with HornSchunck(2, W, W, H) as hs:
        u, v = hs.ComputeFlow(I0, I1, W, H, W, 15.0, 2, 1, 20)
    return np.concatenate([np.ravel(u), np.ravel(v)])


def _lk_pyr():
    from amdflow.lk import lk_pyramidal
    I0, I1, _, _ = _pair()
    u, v = lk_pyramidal(I0, I1, levels=3)
    return np.concatenate([u.ravel(), v.ravel()])


def _lk_stream():
    from amdflow.lk import LucasKanade
    frames = [synth.quantize(synth.make_pair(W, H, 1.0, 70 + k)[0]) for k in range(3)]
    Vx = np.zeros(W * H, np.float32)
    Vy = np.zeros(W * H, np.float32)
    with LucasKanade(3, 2, 5, W, H) as lk:
        for fr, f in enumerate(frames):
            lk.copy_frames_circularbuffer_GPU_wrapper(f, 2, fr, W * H)
            if fr >= 1:
                lk.lucas_kanade(Vx, Vy, fr)
    return np.concatenate([Vx, Vy])


def _batch():
    import torch
    from amdflow.batch import TVL1Batch
    sizes = [(W, H), (120, 96), (W, H)]
    d0, d1, dU = [], [], []
    for k, (w, h) in enumerate(sizes):
        I0, I1, _, _ = synth.make_pair(w, h, 3.0, 80 + k, bumps=True)
        d0.append(torch.from_numpy(synth.quantize(I0).reshape(-1)).cuda())
        d1.append(torch.from_numpy(synth.quantize(I1).reshape(-1)).cuda())
        dU.append(torch.zeros(2 * w * h, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    with TVL1Batch(streams=2) as b:
        b.run_device_u8([w for w, _ in sizes], [h for _, h in sizes], [t.data_ptr() for t in d0],
                        [t.data_ptr() for t in d1], [t.data_ptr() for t in dU])
    return np.concatenate([t.cpu().numpy() for t in dU])


def _color():
    from amdflow import io
    rng = np.random.default_rng(5)
    U = rng.normal(0, 3, 2 * W * H).astype(np.float32)
    return io.flow_to_color(U, W, H).ravel().astype(np.float32)


PATHS = {"tvl1": _tvl1, "hs": _hs, "lk_pyramidal": _lk_pyr, "lk_stream": _lk_stream,
         "batch": _batch, "flow_to_color": _color}


@pytest.mark.parametrize("name", list(PATHS))
def test_every_allocation_failure_is_reported_and_released(name):
    path = PATHS[name]
    path()                                   # warm-up: lazily built pools
    live0 = D.live_allocs()
    D.fail_alloc_after(-1)
    want = path()
    attempts = D.alloc_attempts()
    assert D.live_allocs() == live0
    assert attempts > 0, "the path allocated nothing"
    for n in range(attempts):
        D.fail_alloc_after(n)
        with pytest.raises(OpticalFlowError) as e:
            path()
        assert e.value.status == OF_ERR_ALLOC, (name, n)
        D.fail_alloc_after(-1)
        assert D.live_allocs() == live0, (name, n, D.live_allocs(), live0)
    got = path()
    assert np.array_equal(got, want, equal_nan=True)
