"""The batched TV-L1 solver gives, for every pair of a mixed-size batch, the
same flow and iteration counts as the single-pair solver (and hence the
oracle), in both batch modes."""
import numpy as np
import pytest

import oracle as O
from amdflow import TV_L1, synth

pytestmark = pytest.mark.gpu


def _frames(sizes, cap, seed0):
    import torch
    frames, d0, d1 = [], [], []
    for k, (w, h) in enumerate(sizes):
        I0, I1, _, _ = synth.make_pair(w, h, cap, seed0 + k, bumps=True)
        q0, q1 = synth.quantize(I0), synth.quantize(I1)
        frames.append((q0, q1))
        d0.append(torch.from_numpy(q0.reshape(-1)).cuda())
        d1.append(torch.from_numpy(q1.reshape(-1)).cuda())
    return frames, d0, d1


def _run(b, sizes, d0, d1):
    import torch
    dU = [torch.full((2 * w * h,), float("nan"), dtype=torch.float32, device="cuda") for (w, h) in sizes]
    torch.cuda.synchronize()
    it = b.run_device_u8([w for w, _ in sizes], [h for _, h in sizes],
                         [t.data_ptr() for t in d0], [t.data_ptr() for t in d1],
                         [t.data_ptr() for t in dU])
    return it, [t.cpu().numpy() for t in dU]


@pytest.mark.parametrize("mode,streams", [("lockstep", 1), ("lockstep", 3), ("workers", 4)])
def test_batch_matches_oracle(mode, streams):
    """Lockstep with one lane and with three lanes; the worker mode."""
    from amdflow.batch import TVL1Batch
    # mixed sizes with 3, 4 and 5 scales (default nscales rule), two of one size
    sizes = [(160, 120), (200, 150), (320, 240), (97, 61), (160, 120), (256, 192), (320, 240)]
    frames, d0, d1 = _frames(sizes, 3.0, 100)
    with TVL1Batch(streams=streams, mode=mode) as b:
        it, flows = _run(b, sizes, d0, d1)
    for k, (q0, q1) in enumerate(frames):
        uo, ito = O.tvl1(q0, q1)
        assert it[k] == int(ito.sum()), (mode, k)
        assert np.array_equal(flows[k], uo), (mode, k)


def test_lockstep_matches_single_solves_1080p_mix():
    """A 1920x1080 pair among smaller ones (5 scales each), solved twice by
    one batch object (the second time a subset in another order, reusing
    its pooled handles), against single-pair GPU solves."""
    import torch
    from amdflow.batch import TVL1Batch
    sizes = [(1920, 1080), (320, 240), (1024, 436), (640, 480), (320, 240), (1280, 720)]
    frames, d0, d1 = _frames(sizes, 20.0, 300)
    ref_flows, ref_its = [], []
    for k, (w, h) in enumerate(sizes):
        tv = TV_L1(w, h, nscales=5)
        out = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
        tv.run_device_u8(d0[k].data_ptr(), d1[k].data_ptr(), out.data_ptr())
        ref_flows.append(out.cpu().numpy())
        ref_its.append(int(tv.iterations().sum()))
        tv.close()
    with TVL1Batch(streams=4, nscales=5) as b:
        it, flows = _run(b, sizes, d0, d1)
        for k in range(len(sizes)):
            assert it[k] == ref_its[k], k
            assert np.array_equal(flows[k], ref_flows[k]), k
        order = [4, 0, 2]
        it2, flows2 = _run(b, [sizes[k] for k in order], [d0[k] for k in order],
                           [d1[k] for k in order])
        for j, k in enumerate(order):
            assert it2[j] == ref_its[k], k
            assert np.array_equal(flows2[j], ref_flows[k]), k


@pytest.mark.parametrize("K,R", [(1, 2), (2, 4), (4, 8), (8, 8)])
def test_lockstep_launch_shapes_agree(K, R):
    """Every launch shape gives the oracle's flow (the shape only changes the
    tiling, not the arithmetic)."""
    from amdflow.batch import TVL1Batch
    sizes = [(160, 120), (97, 61), (200, 150)]
    frames, d0, d1 = _frames(sizes, 3.0, 500)
    with TVL1Batch(streams=1) as b:
        b.set_shape(0, K, R)
        it, flows = _run(b, sizes, d0, d1)
    for k, (q0, q1) in enumerate(frames):
        uo, ito = O.tvl1(q0, q1)
        assert it[k] == int(ito.sum()), k
        assert np.array_equal(flows[k], uo), k


@pytest.mark.parametrize("mode", ["lockstep", "workers"])
def test_batch_rejects_unusable_pair_and_stays_usable(mode):
    """A pair too small for the pre-smoothing Gaussian fails the call with an
    error (no launch is left behind), and the same batch object then solves
    a valid batch exactly."""
    from amdflow import OpticalFlowError
    from amdflow.batch import TVL1Batch
    good = [(160, 120), (97, 61)]
    frames, d0, d1 = _frames(good + [(4, 4)], 3.0, 700)
    with TVL1Batch(streams=2, mode=mode) as b:
        with pytest.raises(OpticalFlowError):
            _run(b, good + [(4, 4)], d0, d1)
        it, flows = _run(b, good, d0[:2], d1[:2])
    for k, (q0, q1) in enumerate(frames[:2]):
        uo, ito = O.tvl1(q0, q1)
        assert it[k] == int(ito.sum()), k
        assert np.array_equal(flows[k], uo), k


def test_run_batch_one_call_matches_batch_object():
    import torch
    from amdflow.batch import TVL1Batch, run_batch
    sizes = [(160, 120), (97, 61), (320, 240)]
    frames, d0, d1 = _frames(sizes, 3.0, 900)
    with TVL1Batch(streams=3) as b:
        it, flows = _run(b, sizes, d0, d1)
    dU = [torch.empty(2 * w * h, dtype=torch.float32, device="cuda") for (w, h) in sizes]
    it2 = run_batch([w for w, _ in sizes], [h for _, h in sizes], [t.data_ptr() for t in d0],
                    [t.data_ptr() for t in d1], [t.data_ptr() for t in dU])
    assert np.array_equal(it, it2)
    for k in range(len(sizes)):
        assert np.array_equal(dU[k].cpu().numpy(), flows[k]), k


def test_batch_autotune_installs_a_candidate_and_stays_exact():
    """of_tvl1_batch_autotune times the candidate shapes on the batch,
    installs the fastest, and later runs still match the oracle."""
    from amdflow.batch import TUNE_SHAPES, TVL1Batch
    sizes = [(320, 240), (160, 120), (640, 480), (97, 61), (256, 192)]
    frames, d0, d1 = _frames(sizes, 3.0, 700)
    with TVL1Batch(streams=3) as b:
        import torch
        dU = [torch.empty(2 * w * h, dtype=torch.float32, device="cuda") for (w, h) in sizes]
        (big, small), table, best = b.autotune([w for w, _ in sizes], [h for _, h in sizes],
                                               [t.data_ptr() for t in d0], [t.data_ptr() for t in d1],
                                               [t.data_ptr() for t in dU], reps=1)
        assert big in TUNE_SHAPES and small in TUNE_SHAPES
        assert set(table) == set(TUNE_SHAPES) and all(v > 0 for v in table.values())
        assert 0 < best <= min(table.values())
        it, flows = _run(b, sizes, d0, d1)
    for k, (q0, q1) in enumerate(frames):
        uo, ito = O.tvl1(q0, q1)
        assert it[k] == int(ito.sum()), k
        assert np.array_equal(flows[k], uo), k


def test_lockstep_edge_shapes_match_oracle():
    """Tiny, thin and odd frames (1, 3 and 5 scales) in one lockstep batch."""
    from amdflow.batch import TVL1Batch
    sizes = [(16, 12), (300, 48), (48, 300), (65, 33), (16, 12), (97, 61)]
    frames, d0, d1 = _frames(sizes, 2.0, 900)
    with TVL1Batch(streams=3) as b:
        it, flows = _run(b, sizes, d0, d1)
    for k, (q0, q1) in enumerate(frames):
        uo, ito = O.tvl1(q0, q1)
        assert it[k] == int(ito.sum()), k
        assert np.array_equal(flows[k], uo), k


def test_lockstep_64_pair_mixed_batch_matches_oracle():
    """The batched workload at full size: 64 distinct pairs, sizes drawn
    uniformly from 320x240, 640x480, 1024x436, 1280x720 and 1920x1080, motion
    up to 20 px, 5 scales, default parameters, three lockstep lanes.  Every
    pair's iteration counts per (scale, warping) total and its flow equal the
    oracle's."""
    import time

    from amdflow.batch import TVL1Batch
    all_sizes = [(320, 240), (640, 480), (1024, 436), (1280, 720), (1920, 1080)]
    rng = np.random.default_rng(64)
    sizes = [all_sizes[int(i)] for i in rng.integers(0, len(all_sizes), 64)]
    # frames generated on the GPU (of_synth_pair), one seed per pair
    import torch
    from amdflow.synth import DevicePairs
    gens = {s: DevicePairs(s[0], s[1], 20.0, bumps=True) for s in all_sizes}
    d0, d1 = [], []
    for k, s in enumerate(sizes):
        a, b_, _ = gens[s].generate(5000 + k)
        d0.append(a.clone())
        d1.append(b_.clone())
    torch.cuda.synchronize()
    frames = [(a.cpu().numpy().reshape(h, w), b_.cpu().numpy().reshape(h, w))
              for a, b_, (w, h) in zip(d0, d1, sizes)]
    with TVL1Batch(streams=3, nscales=5) as b:
        it, flows = _run(b, sizes, d0, d1)
    print(f"64-pair batch solved, {int(np.sum(it))} iterations; checking against the oracle",
          flush=True)
    t0 = time.perf_counter()
    p = O.default_params(nscales=5)
    for k, (q0, q1) in enumerate(frames):
        uo, ito = O.tvl1(q0, q1, p)
        assert it[k] == int(ito.sum()), (k, sizes[k], it[k], int(ito.sum()))
        assert np.array_equal(flows[k], uo), (k, sizes[k])
    print(f"oracle: 64 pairs in {time.perf_counter() - t0:.1f} s", flush=True)
