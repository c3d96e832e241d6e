"""Training and evaluation of the learned TV-L1 warm start.

Data: seeded synthetic 1920x1080 pairs generated on the GPU (of_synth_pair,
sinusoid + bump motion up to 20 px, noise sigma 1), each solved by the plain
TV-L1 solver (5 scales, 5 warpings, at most 300 iterations, zero initial
flow).  The network input is the coarsest scale's (I0, I1, Ix, Iy, It); the
label is the plain solver's final flow resampled to the coarsest scale
(of_tvl1_flow_to_scale).  Training: Adam from learning rate 1e-3 (half-cosine
decay), mean-squared flow error, on the GPU through the library's own
forward / backward kernels.

Use: the prediction, carried to scale `first_scale` the way the solver
carries flow between scales, seeds a solve of scales first_scale .. 0 with
at most `warps` warpings and `max_iterations` iterations per warping; the
coarser scales are not solved.  (Seeding the coarsest scale instead gains
nothing: the coarse scales of TV-L1 pull any initial flow towards their own
solution, which is close to zero at 120x68.)  Evaluation compares, by mean
end-point error (EPE), against the plain solve and the true motion: the
warm-started solve, the same limited solve from zero flow (the method
without the network), and the same limited solve seeded with the plain
flow itself resampled to the coarsest scale (what a perfect network would
give).

    PYTHONPATH=sycl-optic-flow_amd python -m amdflow.warmstart --pairs 20000 \
        --epochs 8 --out sycl-optic-flow_amd/weights/warmstart_c32_h5.npy
    PYTHONPATH=sycl-optic-flow_amd python -m amdflow.warmstart --bound
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import TV_L1, tuning
from .cnn import FIRST_SCALE, WarmStartNet
from .synth import DevicePairs

WIDTH, HEIGHT, CAP = 1920, 1080, 20.0
TRAIN_SEED0 = 1_000_000
EVAL_SEED0 = 9_000_000


def _solver(width: int, height: int, **kw) -> TV_L1:
    tv = TV_L1(width, height, nscales=5, **kw)
    entry = tuning.lookup(width, height, tv.nscales, int(tv.params.warps))
    if entry is not None:
        tuning.apply(tv, [tuple(s) for s in entry["shapes"]])
    return tv


def make_dataset(npairs: int, seed0: int = TRAIN_SEED0, width: int = WIDTH, height: int = HEIGHT,
                 cap: float = CAP, log=print, every: int = 1000):
    """Returns (X [n, h*w*5], T [n, 2*h*w], (h, w)) as CUDA tensors."""
    import torch
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    gen = DevicePairs(width, height, cap, bumps=True)
    tv = _solver(width, height)
    ns = tv.nscales
    wc, hc = tv.scale_size(ns - 1)
    X = torch.empty((npairs, hc * wc * 5), dtype=torch.float32, device="cuda")
    T = torch.empty((npairs, 2 * hc * wc), dtype=torch.float32, device="cuda")
    dU = torch.empty(2 * width * height, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(npairs):
        d0, d1, _ = gen.generate(seed0 + i, sp)
        tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), dU.data_ptr(), sp)
        a, b = tv.scale_images(ns - 1)
        WarmStartNet.features(a, b, hc, wc, X[i].data_ptr(), sp)
        tv.flow_to_scale(dU.data_ptr(), ns - 1, T[i].data_ptr(), sp)
        if log and (i + 1) % every == 0:
            log(f"dataset: {i + 1}/{npairs} pairs, {time.perf_counter() - t0:.1f} s")
    stream.synchronize()
    tv.close()
    return X, T, (hc, wc)


def train(net: WarmStartNet, X, T, hw, epochs: int, batch: int = 32, lr: float = 1e-3,
          seed: int = 0, log=print):
    """Adam from learning rate `lr`, decayed along a half cosine to lr/100
    over the run."""
    import math

    import torch
    hc, wc = hw
    n = X.shape[0]
    g = torch.Generator().manual_seed(seed)
    history = []
    t0 = time.perf_counter()
    steps = 0
    total = max(1, epochs * (n // batch))
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).cuda()
        losses = []
        for k in range(0, n - batch + 1, batch):
            idx = perm[k:k + batch]
            xb = X.index_select(0, idx).contiguous()
            tb = T.index_select(0, idx).contiguous()
            lr_t = lr * (0.01 + 0.99 * 0.5 * (1.0 + math.cos(math.pi * steps / total)))
            losses.append(net.train_step(xb.data_ptr(), tb.data_ptr(), batch, hc, wc, lr_t))
            steps += 1
        history.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean loss {history[-1]:.5f}, "
                f"{time.perf_counter() - t0:.1f} s, {steps} steps")
    return history


def epe(a, b, n: int) -> float:
    """Mean end-point error between two planar flows of n pixels (tensors)."""
    d = (a - b).view(2, n)
    return float(d.square().sum(0).sqrt().mean())


def evaluate(net: WarmStartNet, npairs: int = 16, seed0: int = EVAL_SEED0,
             first_scale: int = FIRST_SCALE, warps: int = 2, max_iterations: int = 100,
             width: int = WIDTH, height: int = HEIGHT, cap: float = CAP, reps: int = 2,
             keep_first: bool = False):
    """EPE against the converged plain solve and the true motion, and pairs/s,
    of the plain solve, the warm-started solve and the same limited solve
    from zero flow, over `npairs` held-out pairs.  keep_first adds
    "first_pair": host copies of the first pair's frames, its plain and
    warm-started flows and the plain solve's iteration count."""
    import torch
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    n = width * height
    gen = DevicePairs(width, height, cap, bumps=True)
    pairs = []
    for i in range(npairs):
        a, b, f = gen.generate(seed0 + i, sp)
        stream.synchronize()
        pairs.append((a.clone(), b.clone(), f.clone()))
    torch.cuda.synchronize()
    tv = _solver(width, height)
    ns = tv.nscales
    fs = ns - 1 if first_scale < 0 else first_scale
    cw, ch = tv.scale_size(ns - 1)
    fw, fh = tv.scale_size(fs)
    kinds = ("plain", "warm", "zero", "perfect")
    outs = {k: [torch.empty(2 * n, dtype=torch.float32, device="cuda") for _ in range(npairs)]
            for k in kinds}
    dC = torch.empty(2 * cw * ch, dtype=torch.float32, device="cuda")
    dL = torch.empty(2 * fw * fh, dtype=torch.float32, device="cuda")
    fell, e_pred, e_zero, its = [], [], [], {k: 0 for k in kinds}
    first_its = {}

    def run(kind, i):
        d0, d1, _ = pairs[i]
        dU = outs[kind][i].data_ptr()
        if kind == "plain":
            tv.run_device_u8(d0.data_ptr(), d1.data_ptr(), dU, sp)
        elif kind == "warm":
            return net.warm_run(tv, d0.data_ptr(), d1.data_ptr(), dU, first_scale, warps,
                                max_iterations, sp)
        elif kind == "zero":
            tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr(), sp)
            tv.solve_prepared(dU, 0, fs, warps, max_iterations, sp)
        else:   # the plain flow itself, resampled to the coarsest scale, as the seed
            tv.flow_to_scale(outs["plain"][i].data_ptr(), ns - 1, dC.data_ptr(), sp)
            tv.prepare_device_u8(d0.data_ptr(), d1.data_ptr(), sp)
            tv.lift_flow(ns - 1, dC.data_ptr(), fs, dL.data_ptr(), sp)
            tv.solve_prepared(dU, dL.data_ptr(), fs, warps, max_iterations, sp)
        return None

    timing = {}
    for kind in kinds:
        for i in range(npairs):                  # warm-up and results
            r = run(kind, i)
            its[kind] += int(tv.iterations().sum())
            if i == 0:
                first_its[kind] = int(tv.iterations().sum())
            if kind == "warm":
                fell.append(r[0]); e_pred.append(r[1]); e_zero.append(r[2])
        stream.synchronize()
        if kind == "perfect":
            continue   # not a method: it needs the plain solve first
        t0 = time.perf_counter()
        for _ in range(reps):
            for i in range(npairs):
                run(kind, i)
        stream.synchronize()
        timing[kind] = reps * npairs / (time.perf_counter() - t0)
    res = {
        "pairs": npairs, "first_scale": fs, "first_scale_size": [fw, fh], "warps": warps,
        "max_iterations": max_iterations,
        "pairs_per_s": {k: round(v, 2) for k, v in timing.items()},
        "speedup_warm_vs_plain": round(timing["warm"] / timing["plain"], 3),
        "mean_epe_vs_plain": {
            "warm": float(np.mean([epe(outs["warm"][i], outs["plain"][i], n) for i in range(npairs)])),
            "zero_init_same_limits": float(np.mean([epe(outs["zero"][i], outs["plain"][i], n)
                                                    for i in range(npairs)])),
            "perfect_seed_same_limits": float(np.mean([epe(outs["perfect"][i], outs["plain"][i], n)
                                                       for i in range(npairs)])),
        },
        "mean_epe_vs_ground_truth": {
            k: float(np.mean([epe(outs[k][i], pairs[i][2], n) for i in range(npairs)]))
            for k in kinds},
        "iterations_per_pair": {k: its[k] / npairs for k in its},
        "fallbacks": int(sum(fell)),
        "mean_energy_ratio_pred_vs_zero": float(np.mean(np.asarray(e_pred) / np.asarray(e_zero))),
    }
    if keep_first:
        res["first_pair"] = {
            "I0": pairs[0][0].cpu().numpy().reshape(height, width),
            "I1": pairs[0][1].cpu().numpy().reshape(height, width),
            "plain": outs["plain"][0].cpu().numpy(), "warm": outs["warm"][0].cpu().numpy(),
            "iterations_plain": first_its["plain"]}
    tv.close()
    return res


def schedule_bound(npairs: int = 8, first_scales=(1, 2, 3), finest_warps=(1, 2, 3, 5),
                   coarse_warps=(1, 2), caps=(30, 100, 300), seed0: int = EVAL_SEED0,
                   width: int = WIDTH, height: int = HEIGHT, cap: float = CAP, reps: int = 2,
                   log=None):
    """What any seed can gain: each held-out pair is seeded at scale s with
    the plain solve's own final flow resampled to that scale (a perfect
    prediction, far better than a network working at the coarsest scale),
    then scales s .. 0 are solved with `finest_warps` warpings at scale 0,
    `coarse_warps` at scales 1 .. s and an iteration cap per warping.  Each
    schedule's time (frames resident; preparation included, the seed's
    cost not) and mean EPE against the plain solve; returns the plain time,
    every schedule and the fastest one within 0.05 px."""
    import itertools

    import torch
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    n = width * height
    gen = DevicePairs(width, height, cap, bumps=True)
    pairs = []
    for i in range(npairs):
        a, b, _ = gen.generate(seed0 + i, sp)
        stream.synchronize()
        pairs.append((a.clone(), b.clone()))
    tv = _solver(width, height)
    plain = [torch.empty(2 * n, dtype=torch.float32, device="cuda") for _ in range(npairs)]

    def run_plain():
        for i, (a, b) in enumerate(pairs):
            tv.run_device_u8(a.data_ptr(), b.data_ptr(), plain[i].data_ptr(), sp)

    run_plain()
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run_plain()
    stream.synchronize()
    t_plain = (time.perf_counter() - t0) / (reps * npairs)
    seeds = {}
    for s in first_scales:
        nx, ny = tv.scale_size(s)
        seeds[s] = [torch.empty(2 * nx * ny, dtype=torch.float32, device="cuda")
                    for _ in range(npairs)]
        for i in range(npairs):
            tv.flow_to_scale(plain[i].data_ptr(), s, seeds[s][i].data_ptr(), sp)
    out = torch.empty(2 * n, dtype=torch.float32, device="cuda")
    rows = []
    for s in first_scales:
        for w0, wc, c in itertools.product(finest_warps, coarse_warps, caps):
            warps = [w0] + [wc] * s
            limits = [c] * (s + 1)

            def run(errs=None):
                for i, (a, b) in enumerate(pairs):
                    tv.prepare_device_u8(a.data_ptr(), b.data_ptr(), sp)
                    tv.solve_prepared_schedule(out.data_ptr(), warps, limits,
                                               seeds[s][i].data_ptr(), s, sp)
                    if errs is not None:
                        stream.synchronize()
                        errs.append(epe(out, plain[i], n))

            errs = []
            run(errs)
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            stream.synchronize()
            t = (time.perf_counter() - t0) / (reps * npairs)
            rows.append({"first_scale": s, "warps": warps, "cap": c, "ms": round(t * 1e3, 3),
                         "speedup": round(t_plain / t, 3), "mean_epe_vs_plain": float(np.mean(errs))})
            if log:
                log(json.dumps(rows[-1]))
    tv.close()
    ok = [r for r in rows if r["mean_epe_vs_plain"] <= 0.05]
    return {"pairs": npairs, "plain_ms": round(t_plain * 1e3, 3), "schedules": rows,
            "fastest_within_0_05": max(ok, key=lambda r: r["speedup"]) if ok else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bound", action="store_true",
                    help="only the perfect-seed bound over warping schedules (schedule_bound)")
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-pairs", type=int, default=16)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    log = lambda m: print(m, flush=True)   # noqa: E731
    if args.bound:
        print(json.dumps(schedule_bound(log=log)), flush=True)
        return
    X, T, hw = make_dataset(args.pairs, log=log)
    net = WarmStartNet(seed=args.seed)
    log(f"network: {net.num_params} parameters; coarsest scale {hw[1]}x{hw[0]}")
    hist = train(net, X, T, hw, args.epochs, args.batch, args.lr, args.seed, log=log)
    if args.out:
        net.save(args.out)
        log(f"saved {args.out}")
    del X, T
    res = evaluate(net, args.eval_pairs)
    res["train_loss"] = hist
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
