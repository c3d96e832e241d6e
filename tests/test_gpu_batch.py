"""The batched TV-L1 solver gives, for every pair of a mixed-size batch, the
same flow as the single-pair solver (and hence the oracle)."""
import numpy as np
import pytest

import oracle as O
from amdflow import synth

pytestmark = pytest.mark.gpu


def test_batch_matches_single_pair_solves():
    import torch
    from amdflow.batch import TVL1Batch

    sizes = [(160, 120), (200, 150), (320, 240), (97, 61), (160, 120), (256, 192), (320, 240)]
    frames = []
    for k, (w, h) in enumerate(sizes):
        I0, I1, _, _ = synth.make_pair(w, h, 3.0, 100 + k, bumps=True)
        frames.append((synth.quantize(I0), synth.quantize(I1)))
    d0 = [torch.from_numpy(a.reshape(-1)).cuda() for a, _ in frames]
    d1 = [torch.from_numpy(b.reshape(-1)).cuda() for _, b in frames]
    dU = [torch.empty(2 * w * h, dtype=torch.float32, device="cuda") for (w, h) in sizes]
    torch.cuda.synchronize()
    with TVL1Batch(streams=4) as b:
        it = b.run_device_u8([w for w, _ in sizes], [h for _, h in sizes],
                             [t.data_ptr() for t in d0], [t.data_ptr() for t in d1],
                             [t.data_ptr() for t in dU])
    for k, (q0, q1) in enumerate(frames):
        uo, ito = O.tvl1(q0, q1)
        assert it[k] == int(ito.sum())
        assert np.array_equal(dU[k].cpu().numpy(), uo), k
