"""Trains the pressure-preconditioner network on the GPU and writes its
parameters (the committed weights file is produced by the defaults).

    python -m chuanflow_amd.train_pressure_net [--cases 12] [--size 64x32x32]
        [--steps 600] [--lr 3e-4] [--out chuanflow_amd/weights/mpnn_pressure.npy]

Data: pressure systems dumped from the GPU SIMPLE loop on seeded obstacle
channels (three per channel, after SIMPLE iterations 1, 11 and 21).
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
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--data-seed", type=int, default=1)
    ap.add_argument("--out", default=LN.DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    size = tuple(int(v) for v in a.size.split("x"))
    ctx = N.Context(0)
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


if __name__ == "__main__":
    sys.exit(main())
