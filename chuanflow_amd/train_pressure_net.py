"""Trains the pressure-preconditioner network on the GPU and writes its
parameters (the committed weights file is produced by the defaults).

    python -m chuanflow_amd.train_pressure_net [--cases 12] [--size 64x32x32]
        [--steps 600] [--lr 3e-4] [--out chuanflow_amd/weights/mpnn_pressure.npy]

Data: pressure systems dumped from the GPU SIMPLE loop on seeded obstacle
channels (three per channel, after SIMPLE iterations 1, 11 and 21).

    python -m chuanflow_amd.train_pressure_net --stream [--channels 20]
        [--size 128x128x128] [--per-channel 100] [--steps-per-system 2]

streams the systems instead: every pressure solve of the PISO transient from
rest on each channel (unit box, 128^3 jittered hex by default; 2 correctors
per step) leaves its pressure system in device memory, the network takes Adam
steps on each, and then held-out channels report flexible-CG iterations with
the tuned constants and with the network.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import learned as LN
from . import native as N


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=12)
    ap.add_argument("--size", default="64x32x32")
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=None, help="default 3e-4; 1e-4 with --stream")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--data-seed", type=int, default=1)
    ap.add_argument("--out", default=LN.DEFAULT_WEIGHTS)
    ap.add_argument("--stream", action="store_true")
    ap.add_argument("--channels", type=int, default=20)
    ap.add_argument("--per-channel", type=int, default=100)
    ap.add_argument("--steps-per-system", type=int, default=2)
    ap.add_argument("--heldout", type=int, default=2)
    ap.add_argument("--compare", default=None, help="weights file also evaluated on the held-out systems")
    a = ap.parse_args(argv)
    size = tuple(int(v) for v in a.size.split("x"))
    ctx = N.Context(0)
    if a.stream:
        a.lr = 1e-4 if a.lr is None else a.lr
        return stream(ctx, a, size)
    a.lr = 3e-4 if a.lr is None else a.lr
    t0 = time.perf_counter()
    systems = LN.training_set(ctx, cases=a.cases, size=size, seed=a.data_seed, per_case=3, iterations=30,
                              verbose=True)
    t1 = time.perf_counter()
    net, hist = LN.train(ctx, systems, steps=a.steps, lr=a.lr, seed=a.seed)
    ctx.sync()
    t2 = time.perf_counter()
    LN.save(net, a.out)
    print(json.dumps({"systems": len(systems), "size": size, "steps": a.steps, "lr": a.lr,
                      "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
                      "loss_first20": float(np.mean(hist[:20])), "loss_last20": float(np.mean(hist[-20:])),
                      "out": a.out}), flush=True)


def stream(ctx, a, size):
    def log(msg):
        print(f"[train] {msg}", file=sys.stderr, flush=True)
    length = (1.0, 1.0, 1.0)
    t0 = time.perf_counter()
    net = N.Mpnn(ctx, seed=a.seed)
    hist, used, rejected = LN.stream_train(ctx, net, channels=a.channels, size=size, length=length,
                                           per_channel=a.per_channel, steps_per_system=a.steps_per_system,
                                           lr=a.lr, seed=a.data_seed, log=log)
    ctx.sync()
    t1 = time.perf_counter()
    LN.save(net, a.out)
    log(f"{used} systems, {len(hist)} steps in {t1 - t0:.1f}s; held-out evaluation")
    nets = {"network": net}
    if a.compare:
        nets["compare"] = LN.load(ctx, a.compare)
    rows = LN.heldout_iterations(ctx, nets, channels=a.heldout, size=size, length=length, seed=a.data_seed + 76)
    finite = np.isfinite(hist)
    print(json.dumps({"mode": "stream", "systems": used, "size": size, "length": length,
                      "per_channel": a.per_channel, "steps_per_system": a.steps_per_system, "lr": a.lr,
                      "seconds": round(t1 - t0, 1), "nonfinite_steps": int((~finite).sum()),
                      "channels_rejected": rejected,
                      "loss_first100": float(np.mean(hist[finite][:100])),
                      "loss_last100": float(np.mean(hist[finite][-100:])),
                      "heldout": rows, "out": a.out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
