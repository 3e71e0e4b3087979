"""Surrogate pipeline: data from the GPU simulator, training, warm start.

1. ``generate``: batches of seeded 256x256 fBm terrains run as one ensemble
   (one launch per kernel for the whole batch) from rest for ``cycles``
   cycles; the sample is (initial height, zero flow) -> the mean of
   (discharge, momentumx, momentumy) over the last ``avg`` cycles.
2. ``train``: the U-Net of :mod:`surrogate` with Adam, every operation a
   HIP kernel (BF16 MFMA convolutions, FP32 accumulation).
3. ``warm_start``: tile-wise inference with a halo, written into a terrain
   as its initial discharge/momentum, then the plain particle method runs K
   verification cycles and the warm start is rejected if the discharge field
   moved by more than the stated L1 bound (the particle method is the judge
   and the fallback).

Run ``python -m simplehydrology_amd.train_surrogate --help``.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import Ensemble, Params, World
from . import surrogate as sg

HERE = os.path.dirname(os.path.abspath(__file__))


def l1_rel(a, b):
    """sum |a-b| / sum |b| (torch tensors)"""
    return float((a - b).abs().sum() / b.abs().sum().clamp_min(1e-30))


def flow_tensor(world_or_ens, terrain, n, device):
    import torch
    buf = torch.empty(n, n, 4, dtype=torch.float32, device=device)
    if terrain is None:
        world_or_ens.export(buf.data_ptr())
    else:
        world_or_ens.export(terrain, buf.data_ptr())
    return buf


def generate(n_terrains, size=256, cycles=80, avg=20, batch=256, seed0=0, device="cuda:0",
             log=print):
    """Simulator-made training pairs.  Returns (x bf16 [N,S,S,8], y f32 [N,S,S,3])."""
    import torch
    drops = size * size // 16
    xs = torch.empty(n_terrains, size, size, 8, dtype=torch.bfloat16, device=device)
    ys = torch.empty(n_terrains, size, size, 3, dtype=torch.float32, device=device)
    steps = 0
    t0 = time.perf_counter()
    for b0 in range(0, n_terrains, batch):
        nb = min(batch, n_terrains - b0)
        specs = [dict(nx=size, ny=size, drops=drops, seed=seed0 + b0 + i) for i in range(nb)]
        e = Ensemble(specs)
        for i, s in enumerate(specs):
            e.generate_fbm(i, s["seed"], 8)
        buf = torch.empty(size, size, 4, dtype=torch.float32, device=device)
        for i in range(nb):
            e.export(i, buf.data_ptr())
            z = torch.zeros_like(buf[..., 0])
            xs[b0 + i] = sg.make_input(buf[..., 0], z, z, z)
        e.erode(cycles - avg, graph=True)
        acc = torch.zeros(nb, size, size, 3, dtype=torch.float32, device=device)
        for _ in range(avg):
            e.erode(1, graph=False)
            for i in range(nb):
                e.export(i, buf.data_ptr())
                acc[i] += buf[..., 1:4]
        acc /= avg
        ys[b0:b0 + nb] = sg.encode_flow(acc[..., 0], acc[..., 1], acc[..., 2])
        steps += int(e.stats()["steps"])
        e.close()
        log(f"generated {b0 + nb}/{n_terrains} terrains, {time.perf_counter() - t0:.1f} s")
    return xs, ys, {"seconds": time.perf_counter() - t0, "particle_steps": steps,
                    "cycles": cycles, "avg": avg, "size": size}


def train(net, xs, ys, epochs=4, batch=8, lr=2e-3, log=print, val=None):
    import torch
    n = xs.shape[0]
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    t0 = time.perf_counter()
    hist = []
    g = torch.Generator(device="cpu").manual_seed(0)
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        run = 0.0
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch].to(xs.device)
            x = xs.index_select(0, idx).contiguous()
            y = ys.index_select(0, idx)
            out = net.forward(x)
            loss, gy = sg.mse_loss_and_grad(out, y)
            net.backward(gy)
            # cosine decay
            cur = lr * 0.5 * (1.0 + np.cos(np.pi * k / max(1, total)))
            net.step(cur)
            k += 1
            run += float(loss.item()) if s % 20 == 0 else 0.0
        torch.cuda.synchronize()
        entry = {"epoch": ep + 1, "seconds": time.perf_counter() - t0}
        if val is not None:
            entry["val_loss"] = evaluate(net, *val)
        hist.append(entry)
        log(json.dumps(entry))
    return hist


def evaluate(net, xs, ys, batch=8):
    import torch
    tot, cnt = 0.0, 0
    for b0 in range(0, xs.shape[0], batch):
        out = net.forward(xs[b0:b0 + batch].contiguous())
        tot += float(((out[..., :3] - ys[b0:b0 + batch]) ** 2).sum())
        cnt += out[..., :3].numel()
    torch.cuda.synchronize()
    return tot / cnt


def discharge_error(net, xs, ys, batch=8):
    """Mean relative L1 error of the decoded discharge field."""
    errs = []
    for b0 in range(0, xs.shape[0], batch):
        out = net.forward(xs[b0:b0 + batch].contiguous())
        d, _, _ = sg.decode_flow(out[..., :3])
        dt, _, _ = sg.decode_flow(ys[b0:b0 + batch])
        for i in range(d.shape[0]):
            errs.append(l1_rel(d[i], dt[i]))
    return float(np.mean(errs))


def conv_timing(net, batch=8, size=256, reps=5):
    """FLOP/s of one training step (forward + backward) from HIP events."""
    import torch
    x = torch.zeros(batch, size, size, 8, dtype=torch.bfloat16, device=net.device)
    y = torch.zeros(batch, size, size, 3, dtype=torch.float32, device=net.device)
    out = net.forward(x)
    loss, gy = sg.mse_loss_and_grad(out, y)
    net.backward(gy)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        out = net.forward(x)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    e0.record()
    for _ in range(reps):
        out = net.forward(x)
        loss, gy = sg.mse_loss_and_grad(out, y)
        net.backward(gy)
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / reps
    f = net.flops(batch, size, size)
    return {"batch": batch, "forward_ms": fwd_ms, "train_step_ms": step_ms,
            "forward_tflops": f / fwd_ms / 1e9, "train_tflops": 3 * f / step_ms / 1e9,
            "gflop_per_sample_forward": f / batch / 1e9}


def steady_probe(size=256, terrains=16, cycles=200, device="cuda:0", log=print):
    """Per-cycle relative L1 change of discharge and of erf(0.4 d), averaged
    over a batch of terrains run cold from rest."""
    import torch
    drops = size * size // 16
    e = Ensemble([dict(nx=size, ny=size, drops=drops, seed=500 + i) for i in range(terrains)])
    for i in range(terrains):
        e.generate_fbm(i, 500 + i, 8)
    prev = [None] * terrains
    out = []
    buf = torch.empty(size, size, 4, dtype=torch.float32, device=device)
    for c in range(cycles):
        e.erode(1, graph=False)
        dd, de = [], []
        for i in range(terrains):
            e.export(i, buf.data_ptr())
            d = buf[..., 1].clone()
            if prev[i] is not None:
                dd.append(l1_rel(d, prev[i]))
                de.append(l1_rel(torch.erf(0.4 * d), torch.erf(0.4 * prev[i])))
            prev[i] = d
        if dd:
            out.append((c, float(np.mean(dd)), float(np.mean(de))))
    e.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="surrogate data, training and warm start")
    ap.add_argument("--probe-steady", action="store_true")
    ap.add_argument("--terrains", type=int, default=512)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--cycles", type=int, default=80)
    ap.add_argument("--avg", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    import torch

    def log(s):
        print(s, flush=True)

    if args.probe_steady:
        for c, a, b in steady_probe(log=log):
            if c % 5 == 0 or c < 10:
                log(f"cycle {c:4d}  dL1(d) {a:.5f}  dL1(erf) {b:.5f}")
        return
    xs, ys, ginfo = generate(args.terrains + args.val, cycles=args.cycles, avg=args.avg, log=log)
    xv, yv = xs[args.terrains:], ys[args.terrains:]
    xs, ys = xs[:args.terrains], ys[:args.terrains]
    net = sg.UNet(seed=0)
    log(json.dumps({"params": net.n_params(), "gen": ginfo}))
    log(json.dumps(conv_timing(net, batch=args.batch)))
    hist = train(net, xs, ys, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log,
                 val=(xv, yv))
    res = {"gen": ginfo, "history": hist, "val_discharge_l1": discharge_error(net, xv, yv)}
    log(json.dumps(res))
    if args.out:
        np.savez(args.out, **net.state_dict())
    return res


if __name__ == "__main__":
    main()
