"""Launch modes of the warpings whose launch rule is persistent.

By default each such warping is one launch of a grid that fits the device at
once; its first grid barrier checks that every workgroup started.  A grid
that did not is reported, the solve is repeated with host-paced launches of
the same shape, and the process stays host-paced.  These tests force that
report from the host (a test hook: every workgroup still arrives and ends
normally), and check that the repeat, and host-paced mode itself, give the
oracle's bits."""
import numpy as np
import pytest

import oracle as O
from amdflow import (PERSIST_HOST_PACED, PERSIST_PLAIN, TV_L1, OpticalFlowError,
                     persist_launch_mode, set_persist_launch_mode, synth)
from amdflow import debug as D

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_mode():
    yield
    D.persist_force_not_resident(False)
    set_persist_launch_mode(PERSIST_PLAIN)


def _solve(w, h, K, R, seed, nscales=5):
    I0, I1, _, _ = synth.make_pair(w, h, 6.0, seed, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    with TV_L1(w, h, nscales=nscales) as tv:
        tv.set_launch_rule(0, K, R, True)
        tv.runDualTVL1Multiscale(q0, q1)
        u, it = tv.getU().copy(), tv.iterations()
    uo, ito = O.tvl1(q0, q1, O.default_params(nscales=nscales))
    return u, it, uo, ito


def test_persistent_mode_is_the_default():
    """No earlier solve of this process found a grid that was not resident."""
    u, it, uo, ito = _solve(320, 240, 2, 8, 41)
    assert np.array_equal(it, ito) and np.array_equal(u, uo)
    assert persist_launch_mode() == PERSIST_PLAIN


def test_mode_setter_rejects_unknown_modes():
    with pytest.raises(OpticalFlowError):
        set_persist_launch_mode(7)


@pytest.mark.parametrize("K,R", [(2, 8), (4, 0)])   # flat and two-level barriers
def test_not_resident_takes_the_host_paced_repeat(K, R, restore_mode):
    set_persist_launch_mode(PERSIST_PLAIN)
    D.persist_force_not_resident(True)
    u, it, uo, ito = _solve(320, 240, K, R, 42)
    assert persist_launch_mode() == PERSIST_HOST_PACED
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_repeat_on_device_buffers(restore_mode):
    """Through the device entry point: the repeat starts from the same
    prepared pyramids, and the caller's flow buffer ends with the oracle's
    bits."""
    import torch
    W, H = 400, 300
    I0, I1, _, _ = synth.make_pair(W, H, 6.0, 44, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    dU = torch.full((2 * W * H,), 7.0, dtype=torch.float32, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    D.persist_force_not_resident(True)
    with TV_L1(W, H, nscales=4) as tv:
        tv.set_launch_rule(0, 4, 4, True)
        tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), dU.data_ptr(), sp)
        torch.cuda.synchronize()
        it = tv.iterations()
    uo, ito = O.tvl1(q0, q1, O.default_params(nscales=4))
    assert persist_launch_mode() == PERSIST_HOST_PACED
    assert np.array_equal(it, ito)
    assert np.array_equal(dU.cpu().numpy(), uo)


@pytest.mark.parametrize("K,R", [(2, 8), (4, 4), (4, 0)])
def test_host_paced_mode_matches_oracle(K, R, restore_mode):
    set_persist_launch_mode(PERSIST_HOST_PACED)
    u, it, uo, ito = _solve(480, 272, K, R, 43)
    assert persist_launch_mode() == PERSIST_HOST_PACED
    assert np.array_equal(it, ito)
    assert np.array_equal(u, uo)


def test_modes_agree_at_1080p(restore_mode):
    """Both modes on the 1920x1080 workload with the tuned shapes."""
    from amdflow import tuning
    cfg = synth.CONFIGS["tvl1_1920x1080"]
    W, H = cfg["width"], cfg["height"]
    I0, I1, _, _ = synth.make_pair(W, H, cfg["cap"], 2024, bumps=True)
    q0, q1 = synth.quantize(I0), synth.quantize(I1)
    out = []
    for mode in (PERSIST_PLAIN, PERSIST_HOST_PACED):
        set_persist_launch_mode(mode)
        with TV_L1(W, H, nscales=5) as tv:
            entry = tuning.lookup(W, H, tv.nscales, 5)
            if entry is not None:
                tuning.apply(tv, [tuple(s) for s in entry["shapes"]])
            tv.runDualTVL1Multiscale(q0, q1)
            out.append((tv.getU().copy(), tv.iterations()))
        assert persist_launch_mode() == mode
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
