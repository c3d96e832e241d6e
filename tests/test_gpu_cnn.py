"""GPU checks of the learned warm start: the MFMA operand layout, the
network's forward pass and gradients against the float64 oracle (with and
without bfloat16 operand rounding), deterministic training, the two-phase
solver API and the TV-L1 energy against the C oracle, and the warm-started
solve (kept prediction and energy fallback) against the oracle."""
import ctypes as C
import os

import numpy as np
import pytest

import cnn_oracle as NO
import oracle as O
from amdflow import TV_L1, synth

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _d2h(ptr: int, n: int) -> np.ndarray:
    """Copies n floats from a raw device pointer (test aid)."""
    hip = C.CDLL("libamdhip64.so.7")
    out = np.empty(n, np.float32)
    assert hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(4 * n), 2) == 0
    return out


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


# ---------------------------------------------------------------- network


def test_mfma_layout_exact_small_integers():
    from amdflow.cnn import mfma_product
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, (16, 16)).astype(np.float32) / 4
    B = rng.integers(-8, 9, (16, 16)).astype(np.float32) / 8
    assert np.array_equal(mfma_product(A, B), (A.astype(np.float64) @ B).astype(np.float32))


def test_mfma_rounds_operands_to_bf16():
    from amdflow.cnn import mfma_product
    rng = np.random.default_rng(1)
    A = rng.normal(size=(16, 16)).astype(np.float32)
    B = rng.normal(size=(16, 16)).astype(np.float32)
    ref = NO.bf16_round(A) @ NO.bf16_round(B)
    np.testing.assert_allclose(mfma_product(A, B), ref, rtol=1e-5, atol=1e-5)
    # and differs from the unrounded product by more than fp32 noise
    assert np.max(np.abs(mfma_product(A, B) - A.astype(np.float64) @ B)) > 1e-4


def _net(hidden, seed=3):
    from amdflow.cnn import WarmStartNet
    net = WarmStartNet(hidden=hidden, seed=0)
    p = NO.init_params(hidden, seed=seed)
    net.set_params(p)
    return net, p


@pytest.mark.parametrize("hidden,B,H,W", [(0, 1, 5, 7), (1, 2, 9, 13), (5, 3, 17, 11)])
def test_forward_matches_oracle(hidden, B, H, W):
    torch = _torch()
    rng = np.random.default_rng(hidden + 10)
    X = rng.normal(size=(B, H, W, 5)).astype(np.float32)
    net, p = _net(hidden)
    dX = torch.from_numpy(X.reshape(-1)).cuda()
    dY = torch.empty(B * H * W * 2, dtype=torch.float32, device="cuda")
    net.forward(dX.data_ptr(), B, H, W, dY.data_ptr())
    torch.cuda.synchronize()
    got = dY.cpu().numpy().reshape(B, H, W, 2)
    emu = NO.forward(p, X, hidden, bf16=True)[-1]
    exact = NO.forward(p, X, hidden, bf16=False)[-1]
    assert _rel(got, emu) <= 2e-3
    assert np.max(np.abs(got - emu)) <= 5e-3 * np.max(np.abs(emu))
    assert _rel(got, exact) <= 3e-2


@pytest.mark.parametrize("hidden,B,H,W", [(0, 1, 6, 5), (1, 2, 9, 13), (5, 2, 20, 17)])
def test_gradients_match_oracle(hidden, B, H, W):
    torch = _torch()
    rng = np.random.default_rng(hidden + 20)
    X = rng.normal(size=(B, H, W, 5)).astype(np.float32)
    T = rng.normal(size=(B, 2, H, W)).astype(np.float32)
    net, p = _net(hidden)
    dX = torch.from_numpy(X.reshape(-1)).cuda()
    dT = torch.from_numpy(T.reshape(-1)).cuda()
    loss, g = net.gradients(dX.data_ptr(), dT.data_ptr(), B, H, W)
    l_emu, g_emu = NO.loss_and_grads(p, X, T, hidden, bf16=True)
    l_ex, g_ex = NO.loss_and_grads(p, X, T, hidden, bf16=False)
    assert abs(loss - l_emu) <= 1e-4 * l_emu
    # per layer: weights and bias blocks
    off = 0
    for cin, cout in NO.layer_dims(hidden):
        nw = cout * 9 * cin
        sl_w, sl_b = slice(off, off + nw), slice(off + nw, off + nw + cout)
        assert _rel(g[sl_w], g_emu[sl_w]) <= 5e-3, (cin, cout)
        assert _rel(g[sl_b], g_emu[sl_b]) <= 5e-3, (cin, cout)
        off += nw + cout
    assert _rel(g, g_ex) <= 5e-2


def test_training_is_deterministic_and_learns():
    torch = _torch()
    rng = np.random.default_rng(4)
    B, H, W = 4, 12, 16
    X = rng.normal(size=(B, H, W, 5)).astype(np.float32)
    # a target the network can represent: a fixed linear map of the input
    T = np.stack([X[..., 0] - 0.5 * X[..., 4], 0.3 * X[..., 2]], 1).astype(np.float32)
    dX = torch.from_numpy(X.reshape(-1)).cuda()
    dT = torch.from_numpy(T.reshape(-1)).cuda()
    params = []
    losses = None
    for _ in range(2):
        net, _ = _net(1, seed=9)
        ls = [net.train_step(dX.data_ptr(), dT.data_ptr(), B, H, W, 1e-3) for _ in range(60)]
        params.append(net.get_params())
        losses = ls
        net.close()
    assert np.array_equal(params[0], params[1])
    assert losses[-1] < 0.5 * losses[0]


# ------------------------------------------------------ two-phase solve API


def _pair(w, h, cap, seed):
    I0, I1, fu, fv = synth.make_pair(w, h, cap, seed, bumps=True)
    return synth.quantize(I0), synth.quantize(I1)


def _oracle_pyramid(q0, q1, ns):
    a, b = O.image_normalization(q0.astype(np.float32), q1.astype(np.float32))
    s08 = float(np.float32(0.8))
    a, b = O.gaussian(a, s08), O.gaussian(b, s08)
    pyr = [(a, b)]
    for _ in range(1, ns):
        a, b = O.zoom_out(a), O.zoom_out(b)
        pyr.append((a, b))
    return pyr


def test_scale_images_match_oracle_pyramid():
    torch = _torch()
    q0, q1 = _pair(200, 150, 4.0, 31)
    h, w = q0.shape
    tv = TV_L1(w, h, nscales=5)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    pyr = _oracle_pyramid(q0, q1, tv.nscales)
    for s in range(tv.nscales):
        nx, ny = tv.scale_size(s)
        a, b = tv.scale_images(s)
        assert np.array_equal(_d2h(a, nx * ny), pyr[s][0].reshape(-1)), s
        assert np.array_equal(_d2h(b, nx * ny), pyr[s][1].reshape(-1)), s


def test_two_phase_zero_init_equals_plain_run():
    torch = _torch()
    q0, q1 = _pair(160, 120, 3.0, 5)
    h, w = q0.shape
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    dA = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    dB = torch.empty_like(dA)
    tv = TV_L1(w, h)
    tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), dA.data_ptr())
    itA = tv.iterations().copy()
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    tv.solve_prepared(dB.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dA.cpu().numpy(), dB.cpu().numpy())
    assert np.array_equal(itA, tv.iterations())


@pytest.mark.parametrize("warps,max_it", [(2, 100), (1, 7), (5, 300)])
def test_solve_from_initial_flow_matches_oracle(warps, max_it):
    torch = _torch()
    q0, q1 = _pair(192, 128, 4.0, 77)
    h, w = q0.shape
    tv = TV_L1(w, h)
    ns = tv.nscales
    nx, ny = tv.scale_size(ns - 1)
    rng = np.random.default_rng(warps)
    yy, xx = np.mgrid[0:ny, 0:nx]
    uinit = np.concatenate([(0.3 * np.sin(xx / 5.0) + 0.1 * rng.normal(size=(ny, nx))).ravel(),
                            (0.2 * np.cos(yy / 4.0)).ravel()]).astype(np.float32)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    dI = torch.from_numpy(uinit).cuda()
    dU = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    tv.solve_prepared(dU.data_ptr(), dI.data_ptr(), -1, warps, max_it)
    torch.cuda.synchronize()
    uo, ito = O.tvl1_init(q0, q1, uinit, O.default_params(warps=warps, max_iterations=max_it))
    assert np.array_equal(dU.cpu().numpy(), uo)
    it = tv.iterations()
    assert np.array_equal(it[:, :warps], ito)
    assert not it[:, warps:].any()
    # the coarsest scale's final flow, as the oracle's single-scale solve gives it
    dC = torch.empty(2 * nx * ny, dtype=torch.float32, device="cuda")
    tv.scale_flow(ns - 1, dC.data_ptr())
    pyr = _oracle_pyramid(q0, q1, ns)
    u1, u2 = uinit[:nx * ny].reshape(ny, nx).copy(), uinit[nx * ny:].reshape(ny, nx).copy()
    c1, c2, _ = O.dual_tvl1(pyr[-1][0], pyr[-1][1], u1, u2,
                            O.default_params(warps=warps, max_iterations=max_it))
    assert np.array_equal(dC.cpu().numpy(), np.concatenate([c1.ravel(), c2.ravel()]))


@pytest.mark.parametrize("first,warps,caps", [(1, [5, 1], [300, 30]), (2, [3, 1, 2], [100, 100, 7]),
                                              (-1, [1, 2, 3, 4, 5], [0, 9, 0, 50, 0])])
def test_solve_schedule_matches_oracle(first, warps, caps):
    """Per-scale warpings and iteration caps (of_tvl1_solve_prepared_schedule)
    from a given flow at `first` are bit-exact with the oracle's schedule
    solve, with the iteration counts of every (scale, warping)."""
    torch = _torch()
    q0, q1 = _pair(256, 160, 4.0, 79)
    h, w = q0.shape
    tv = TV_L1(w, h)
    ns = tv.nscales
    fs = ns - 1 if first < 0 else first
    nx, ny = tv.scale_size(fs)
    rng = np.random.default_rng(fs)
    uinit = (rng.normal(size=2 * nx * ny) * 0.5).astype(np.float32)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    dI = torch.from_numpy(uinit).cuda()
    dU = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    tv.solve_prepared_schedule(dU.data_ptr(), warps, caps, dI.data_ptr(), first)
    torch.cuda.synchronize()
    uo, ito = O.tvl1_schedule(q0, q1, warps, caps, uinit, first=first)
    assert np.array_equal(dU.cpu().numpy(), uo)
    assert np.array_equal(tv.iterations(), ito)


@pytest.mark.parametrize("first", [0, 1, 2])
def test_solve_from_a_finer_scale_matches_oracle(first):
    """Scales coarser than `first` are skipped; the solve of the rest from a
    given flow at `first` is bit-exact with the oracle, and solving the same
    prepared pair twice gives the same result."""
    torch = _torch()
    q0, q1 = _pair(256, 160, 4.0, 78)
    h, w = q0.shape
    tv = TV_L1(w, h)
    nx, ny = tv.scale_size(first)
    rng = np.random.default_rng(first)
    uinit = (rng.normal(size=2 * nx * ny) * 0.5).astype(np.float32)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    dI = torch.from_numpy(uinit).cuda()
    dU = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    prm = O.default_params(warps=2, max_iterations=100)
    uo, ito = O.tvl1_init(q0, q1, uinit, prm, first=first)
    for _ in range(2):
        tv.solve_prepared(dU.data_ptr(), dI.data_ptr(), first, 2, 100)
        torch.cuda.synchronize()
        assert np.array_equal(dU.cpu().numpy(), uo)
        it = tv.iterations()
        assert np.array_equal(it[:, :2], ito) and not it[first + 1:].any()


def test_lift_and_downsample_match_oracle():
    """of_tvl1_lift_flow = zoom_in x 1/zf per scale; of_tvl1_flow_to_scale =
    zoom_out per scale x zf^s (exact: zf^s is a power of two)."""
    torch = _torch()
    w, h = 200, 150
    tv = TV_L1(w, h)
    ns = tv.nscales
    sizes = [tv.scale_size(s) for s in range(ns)]
    rng = np.random.default_rng(9)
    nxc, nyc = sizes[-1]
    uc = rng.normal(size=2 * nxc * nyc).astype(np.float32)
    a, b = uc[:nxc * nyc].reshape(nyc, nxc), uc[nxc * nyc:].reshape(nyc, nxc)
    for s in range(ns - 1, 0, -1):
        nx, ny = sizes[s - 1]
        a = (O.zoom_in(a, nx, ny) * np.float32(2)).astype(np.float32)
        b = (O.zoom_in(b, nx, ny) * np.float32(2)).astype(np.float32)
        dC = torch.from_numpy(uc).cuda()
        dL = torch.empty(2 * nx * ny, dtype=torch.float32, device="cuda")
        tv.lift_flow(ns - 1, dC.data_ptr(), s - 1, dL.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(dL.cpu().numpy(), np.concatenate([a.ravel(), b.ravel()])), s
    U = rng.normal(size=2 * w * h).astype(np.float32) * 3
    dU = torch.from_numpy(U).cuda()
    a, b = U[:w * h].reshape(h, w), U[w * h:].reshape(h, w)
    for s in range(1, ns):
        a, b = O.zoom_out(a), O.zoom_out(b)
        nx, ny = sizes[s]
        dS = torch.empty(2 * nx * ny, dtype=torch.float32, device="cuda")
        tv.flow_to_scale(dU.data_ptr(), s, dS.data_ptr())
        f = np.float32(0.5 ** s)
        assert np.array_equal(dS.cpu().numpy(), np.concatenate([(a * f).ravel(), (b * f).ravel()])), s


def test_energy_matches_oracle():
    torch = _torch()
    q0, q1 = _pair(240, 135, 6.0, 12)
    h, w = q0.shape
    tv = TV_L1(w, h, nscales=3)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    pyr = _oracle_pyramid(q0, q1, tv.nscales)
    rng = np.random.default_rng(2)
    for s in range(tv.nscales):
        nx, ny = tv.scale_size(s)
        # flows pointing inside and far outside the image
        u = (rng.normal(size=2 * nx * ny) * (3.0 if s else 40.0)).astype(np.float32)
        du = torch.from_numpy(u).cuda()
        e, dsum, tsum = tv.energy(s, du.data_ptr())
        eo, do, to = O.tvl1_energy(pyr[s][0], pyr[s][1], u, 0.15)
        assert abs(dsum - do) <= 1e-12 * do and abs(tsum - to) <= 1e-12 * to
        assert abs(e - eo) <= 1e-12 * eo


# ------------------------------------------------------------- warm start


def _warm_setup(w, h, seed):
    torch = _torch()
    q0, q1 = _pair(w, h, 4.0, seed)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    return q0, q1, d0, d1


def test_warm_run_keeps_zero_prediction_and_matches_oracle():
    """All-zero network: the prediction is exactly zero flow, whose energy
    ties with zero flow's, so it is kept and the limited solve runs."""
    torch = _torch()
    from amdflow.cnn import WarmStartNet
    q0, q1, d0, d1 = _warm_setup(160, 120, 8)
    h, w = q0.shape
    net = WarmStartNet()
    net.set_params(np.zeros(net.num_params, np.float32))
    tv = TV_L1(w, h)
    dU = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    for first in (-1, 1):
        fell, ep, ez = net.warm_run(tv, d0.data_ptr(), d1.data_ptr(), dU.data_ptr(), first, 2, 100)
        assert not fell and ep == ez
        uo, ito = O.tvl1_init(q0, q1, None, O.default_params(warps=2, max_iterations=100), first=first)
        assert np.array_equal(dU.cpu().numpy(), uo)
        assert np.array_equal(tv.iterations()[:, :2], ito)


def test_warm_run_falls_back_to_plain_solve():
    """A network whose output is a 30 px constant flow: its energy is worse
    than zero flow's, so the plain solve (handle parameters) runs."""
    torch = _torch()
    from amdflow.cnn import WarmStartNet
    q0, q1, d0, d1 = _warm_setup(160, 120, 9)
    h, w = q0.shape
    net = WarmStartNet()
    p = np.zeros(net.num_params, np.float32)
    p[-2:] = 30.0   # head bias
    net.set_params(p)
    tv = TV_L1(w, h)
    dU = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    fell, ep, ez = net.warm_run(tv, d0.data_ptr(), d1.data_ptr(), dU.data_ptr(), 1, 2, 100)
    assert fell and ep > ez
    uo, ito = O.tvl1(q0, q1)
    assert np.array_equal(dU.cpu().numpy(), uo)
    assert np.array_equal(tv.iterations(), ito)


def test_warm_run_with_predicted_flow_matches_oracle():
    """A random network's prediction (computed on the GPU) fed to the
    oracle's solve from an initial flow gives the same flow, when kept."""
    torch = _torch()
    from amdflow.cnn import WarmStartNet
    q0, q1, d0, d1 = _warm_setup(192, 128, 10)
    h, w = q0.shape
    net = WarmStartNet(seed=5)
    p = net.get_params()
    p[-2:] = 0.05
    net.set_params(p)
    tv = TV_L1(w, h)
    ns = tv.nscales
    nx, ny = tv.scale_size(ns - 1)
    fx, fy = tv.scale_size(1)
    tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
    a, b = tv.scale_images(ns - 1)
    dP = torch.empty(2 * nx * ny, dtype=torch.float32, device="cuda")
    dL = torch.empty(2 * fx * fy, dtype=torch.float32, device="cuda")
    net.predict(a, b, ny, nx, dP.data_ptr())
    tv.lift_flow(ns - 1, dP.data_ptr(), 1, dL.data_ptr())
    torch.cuda.synchronize()
    lifted = dL.cpu().numpy()
    dU = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    fell, ep, ez = net.warm_run(tv, d0.data_ptr(), d1.data_ptr(), dU.data_ptr(), 1, 2, 100)
    if fell:
        uo, _ = O.tvl1(q0, q1)
    else:
        uo, _ = O.tvl1_init(q0, q1, lifted, O.default_params(warps=2, max_iterations=100), first=1)
    assert np.array_equal(dU.cpu().numpy(), uo)


WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "sycl-optic-flow_amd", "weights", "warmstart_c32_h5.npy")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained warm-start weights not present")
def test_trained_network_beats_zero_init_at_same_limits():
    """At 1920x1080, starting at 480x270 with at most 2 warpings and 100
    iterations: from the trained network's prediction the solve stays close
    to the plain solve and to the true motion; from zero flow it does not."""
    from amdflow.cnn import WarmStartNet
    from amdflow.warmstart import evaluate
    net = WarmStartNet.pretrained(WEIGHTS)
    res = evaluate(net, npairs=4, reps=1)
    e, g = res["mean_epe_vs_plain"], res["mean_epe_vs_ground_truth"]
    assert res["fallbacks"] == 0, res
    assert e["warm"] < 0.5 * e["zero_init_same_limits"], res
    assert g["warm"] < g["zero"], res


def test_set_init_flow_matches_solve_from_scale():
    """of_tvl1_set_init_flow (host arrays) seeds the next run exactly as
    solve_prepared seeds scale `level`, is used once, and can be cleared."""
    torch = _torch()
    q0, q1 = _pair(320, 240, 4.0, 31)
    h, w = q0.shape
    tv = TV_L1(w, h, nscales=4)
    d0 = torch.from_numpy(q0.reshape(-1)).cuda()
    d1 = torch.from_numpy(q1.reshape(-1)).cuda()
    plain = torch.empty(2 * w * h, dtype=torch.float32, device="cuda")
    tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), plain.data_ptr())
    for level in (1, 2):
        nx, ny = tv.scale_size(level)
        seed = torch.empty(2 * nx * ny, dtype=torch.float32, device="cuda")
        tv.flow_to_scale(plain.data_ptr(), level, seed.data_ptr())
        torch.cuda.synchronize()
        ref = torch.empty_like(plain)
        tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr())
        tv.solve_prepared(ref.data_ptr(), seed.data_ptr(), level)
        s = seed.cpu().numpy()
        tv.set_init_flow(level, s[:nx * ny], s[nx * ny:])
        got = torch.empty_like(plain)
        tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), got.data_ptr())
        assert torch.equal(got, ref), level
        again = torch.empty_like(plain)                      # used once
        tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), again.data_ptr())
        assert torch.equal(again, plain), level
    nx, ny = tv.scale_size(1)
    tv.set_init_flow(1, np.zeros(nx * ny, np.float32), np.zeros(nx * ny, np.float32))
    tv.set_init_flow(1, None, None)                          # cleared
    cleared = torch.empty_like(plain)
    tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), cleared.data_ptr())
    assert torch.equal(cleared, plain)
    tv.close()
