"""Two TV-L1 handles solving at once on two streams from two host threads,
persistent warpings at every scale: both keep the oracle's bits whether or
not their launches found the device shared (then the process turns
host-paced).

Skipped.  Its one GPU run, with the fallback of that time (cooperative
launches instead of host-paced ones), passed: the grids collided, the process
switched modes and all six flows equalled the oracle's.  The pytest process
then ended with a segmentation fault during interpreter exit, after
reporting the pass.  Nothing in this library's code was found to cause it;
the only earlier crash of that kind also followed cooperative launches made
from more than one host thread.  The library no longer makes cooperative
launches at all, but until that crash has a cause shown in code, this case
does not run on the GPU."""
import threading

import numpy as np
import pytest

import oracle as O
from amdflow import PERSIST_PLAIN, TV_L1, persist_launch_mode, set_persist_launch_mode, synth

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="its one run ended in a segmentation fault at interpreter "
                                      "exit after passing; no cause found in this library's code")]


def test_two_streams_at_once():
    import torch
    W, H = 1920, 1080   # (2,8) grids of 1024 workgroups: two cannot be resident at once
    pairs = []
    for seed in (51, 52):
        I0, I1, _, _ = synth.make_pair(W, H, 20.0, seed, bumps=True)
        pairs.append((synth.quantize(I0), synth.quantize(I1)))
    want = [O.tvl1(q0, q1, O.default_params(nscales=5)) for q0, q1 in pairs]
    results = [[], []]
    errors = []
    start = threading.Barrier(2)

    def worker(k):
        try:
            q0, q1 = pairs[k]
            s = torch.cuda.Stream()
            d0 = torch.from_numpy(q0.reshape(-1)).cuda()
            d1 = torch.from_numpy(q1.reshape(-1)).cuda()
            dU = torch.empty(2 * W * H, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            with TV_L1(W, H, nscales=5) as tv:
                tv.set_launch_rule(0, 2, 8, True)
                start.wait()
                for _ in range(3):
                    tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), dU.data_ptr(), s.cuda_stream)
                    s.synchronize()
                    results[k].append((dU.cpu().numpy().copy(), tv.iterations()))
        except Exception as e:   # reported by the main thread
            errors.append(e)

    set_persist_launch_mode(PERSIST_PLAIN)
    try:
        th = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        print("launch mode after the concurrent solves:", persist_launch_mode(), flush=True)
    finally:
        set_persist_launch_mode(PERSIST_PLAIN)
    assert not errors, errors
    for k in (0, 1):
        uo, ito = want[k]
        for u, it in results[k]:
            assert np.array_equal(it, ito)
            assert np.array_equal(u, uo)
