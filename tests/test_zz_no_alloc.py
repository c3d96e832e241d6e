"""of_lk_pyr_compute_device_noalloc (the pyramidal LK on the library's fixed
per-GPU scratch, no allocation) against of_lk_pyr_compute_device, called
through the C ABI: bit-identical flow on two calls in a row, the second on a
smaller input (the scratch then still holds the first call's data), both
kernel paths (blocked windows 3-9, three-kernel windows 1 and 11), and the
capacity error for an input beyond the scratch."""
import numpy as np
import pytest

from amdflow import _lib, synth

pytestmark = pytest.mark.gpu

OF_ERR_CAPACITY = 7


def _device_pair(w, h, cap, seed):
    import torch
    I0, I1, _, _ = synth.make_pair(w, h, cap, seed)
    return (torch.from_numpy(I0.reshape(-1).copy()).cuda(),
            torch.from_numpy(I1.reshape(-1).copy()).cuda())


def _solve(fn, d0, d1, w, h, levels, win, iters):
    import torch
    du = torch.full((w * h,), np.nan, dtype=torch.float32, device="cuda")
    dv = torch.full((w * h,), np.nan, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = fn(d0.data_ptr(), d1.data_ptr(), w, h, levels, 0.5, 1.0, win, iters, 1e-6,
            du.data_ptr(), dv.data_ptr(), None)
    assert st == 0, st
    return du.cpu().numpy(), dv.cpu().numpy()


@pytest.mark.parametrize("first,second", [
    ((1280, 720, 8.0, 3, 7, 5), (640, 360, 4.0, 3, 7, 5)),
    ((1920, 1080, 8.0, 4, 9, 3), (301, 173, 4.0, 2, 11, 2)),
    ((997, 611, 6.0, 3, 5, 4), (257, 190, 3.0, 1, 1, 2)),
])
def test_noalloc_matches_allocating_entry_point(first, second):
    print(f"no-alloc LK: {first} then {second}", flush=True)
    for k, (w, h, cap, levels, win, iters) in enumerate((first, second)):
        d0, d1 = _device_pair(w, h, cap, 40 + k)
        ref = _solve(_lib.of_lk_pyr_compute_device, d0, d1, w, h, levels, win, iters)
        got = _solve(_lib.of_lk_pyr_compute_device_noalloc, d0, d1, w, h, levels, win, iters)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (w, h)
        assert np.isfinite(got[0]).all() and np.isfinite(got[1]).all()


def test_noalloc_rejects_input_beyond_scratch():
    import torch
    w, h = 2048, 1080   # 4 * (2 * 737,280 + 15 * 2,211,840) bytes > 128 MiB
    d = torch.zeros(w * h, dtype=torch.float32, device="cuda")
    st = _lib.of_lk_pyr_compute_device_noalloc(d.data_ptr(), d.data_ptr(), w, h, 3, 0.5, 1.0, 7,
                                               5, 1e-6, d.data_ptr(), d.data_ptr(), None)
    assert st == OF_ERR_CAPACITY
    assert "scratch" in _lib.of_status_string(st).decode()
