"""Trains the pressure-preconditioner network on the GPU and writes its
parameters.  The committed weights come from two stages:

    python -m chuanflow_amd.train_pressure_net
        -> chuanflow_amd/weights/mpnn_pressure_stage1.npy
    python -m chuanflow_amd.train_pressure_net --stream \
        --init chuanflow_amd/weights/mpnn_pressure_stage1.npy
        -> chuanflow_amd/weights/mpnn_pressure.npy

Stage 1 (defaults: --cases 12 --size 64x32x32 --steps 600 --lr 3e-4): 600 Adam
steps on pressure systems dumped from the GPU SIMPLE loop on seeded obstacle
channels (three per channel, after SIMPLE iterations 1, 11 and 21).

Stage 2 (--stream; defaults --channels 20 --size 128x128x128 --per-channel 100
--steps-per-system 2 --lr 1e-4): the systems are streamed instead.  Every pressure
solve of the PISO transient from rest on each channel (unit box, 128^3 jittered
hex; 2 correctors per step) leaves its pressure system in device memory, and
the network takes Adam steps on each: 2,000 systems in all.  Every 10 systems
the update is kept only if the network solves the current system in no more
flexible-CG iterations than the tuned constants and the last kept parameters.
Held-out channels then report iterations with the tuned constants and with
the network.
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
    ap.add_argument("--size", default=None, help="default 64x32x32; 128x128x128 with --stream")
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=None, help="default 3e-4; 1e-4 with --stream")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--data-seed", type=int, default=1)
    ap.add_argument("--out", default=None, help="default weights/mpnn_pressure_stage1.npy; with --stream "
                    "weights/mpnn_pressure.npy")
    ap.add_argument("--stream", action="store_true")
    ap.add_argument("--channels", type=int, default=20)
    ap.add_argument("--per-channel", type=int, default=100)
    ap.add_argument("--steps-per-system", type=int, default=2)
    ap.add_argument("--heldout", type=int, default=2)
    ap.add_argument("--compare", default=None, help="weights file also evaluated on the held-out systems")
    ap.add_argument("--init", default=None, help="start the streamed training from this weights file")
    a = ap.parse_args(argv)
    size = tuple(int(v) for v in (a.size or ("128x128x128" if a.stream else "64x32x32")).split("x"))
    ctx = N.Context(0)
    if a.stream:
        a.lr = 1e-4 if a.lr is None else a.lr
        a.out = LN.DEFAULT_WEIGHTS if a.out is None else a.out
        return stream(ctx, a, size)
    a.lr = 3e-4 if a.lr is None else a.lr
    a.out = LN.STAGE1_WEIGHTS if a.out is None else a.out
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
    if a.init:
        net.set_params(np.load(a.init, allow_pickle=False).astype(np.float32))
    hist, used, rejected, accepted = LN.stream_train(ctx, net, channels=a.channels, size=size, length=length,
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
                      "updates_accepted": accepted, "updates_rejected": rejected, "init": a.init,
                      "loss_first100": float(np.mean(hist[finite][:100])),
                      "loss_last100": float(np.mean(hist[finite][-100:])),
                      "heldout": rows, "out": a.out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
