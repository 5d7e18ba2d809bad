"""Training data and training of the learned pressure preconditioner.

The pressure systems come from the GPU SIMPLE loop itself: laminar channel
flows past a seeded cylindrical obstacle (radius U[0.05, 0.15] of a unit-length
channel, inlet speed U[0.005, 0.05] m/s, nu = 1e-3) on jittered hex meshes;
every few outer iterations the assembled pressure matrix and right-hand side
are copied out (cfx_mesh_set_pressure_dump).  The network is then trained
on the GPU (cfx_mpnn_train) against the A-norm error left by one
preconditioned step (optimal step length) on smooth error samples, and its
parameters saved as a .npy array.
"""
from __future__ import annotations

import os

import numpy as np

from . import native as N

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "mpnn_pressure.npy")
STAGE1_WEIGHTS = os.path.join(HERE, "weights", "mpnn_pressure_stage1.npy")


def channel_bcs(patch_start, patch_size, n_internal, n_boundary, Uin):
    """Per boundary face: patch 0 inlet (U fixed (Uin,0,0), p zeroGradient),
    patch 1 outlet (U zeroGradient, p fixed 0), every other patch a wall."""
    pid = np.empty(n_boundary, np.int64)
    for p, (s, n) in enumerate(zip(patch_start, patch_size)):
        pid[s - n_internal:s - n_internal + n] = p
    Ut = np.where(pid == 1, 2, 0).astype(np.int32)
    Uv = np.zeros((n_boundary, 3))
    Uv[pid == 0, 0] = Uin
    pt = np.where(pid == 1, 0, 2).astype(np.int32)
    pv = np.zeros((n_boundary, 1))
    return (Ut, Uv), (pt, pv)


def slot_geometry(arrays, geom, rowPtr, col):
    """Per CSR slot: 1 - cos(Sf, d) of the face between the row and column
    cells (0 on the diagonal), from the host mesh."""
    own, nei = np.asarray(arrays["owner"]), np.asarray(arrays["neighbour"])
    nI = len(nei)
    Sf, d = geom["Sf"][:nI], geom["dvec"][:nI]
    g = 1.0 - (Sf * d).sum(axis=1) / (np.linalg.norm(Sf, axis=1) * np.linalg.norm(d, axis=1))
    geo = np.zeros(len(col))
    n = len(rowPtr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowPtr))
    keys = rows * n + col  # ascending: rows ascending, columns sorted within a row
    for a, b in ((own[:nI], nei), (nei, own[:nI])):
        geo[np.searchsorted(keys, a.astype(np.int64) * n + b)] = g
    return geo


def channel_mesh(nx, ny, nz, seed, jitter=0.15, radius=None, length=(1.0, 0.5, 0.5)):
    rng = np.random.default_rng(seed)
    if radius is None:
        radius = rng.uniform(0.05, 0.15)
    cx = rng.uniform(0.25, 0.4) * length[0]
    cy = length[1] * 0.5 + rng.uniform(-0.05, 0.05)
    hm = N.HostMesh.obstacle_channel(nx, ny, nz, (0, 0, 0), length, jitter, seed, cx, cy, radius)
    hm.shuffle(seed + 7)
    hm.renumber_rcm()
    return hm, radius


def simple_pressure_systems(ctx, hm, Uin, iterations=40, every=10, max_dumps=4, nu=1e-3, verbose=False):
    """Runs the GPU SIMPLE loop on a channel mesh and returns its dumped
    pressure systems as host (rowStart, cols, vals, rhs, geo) tuples."""
    arrays, geom = hm.arrays(), hm.geometry()
    dm = N.Mesh(ctx, arrays, geom)
    n, nnz = dm.nCells, dm.sizes()["nnz"]
    Ubc, pbc = channel_bcs(arrays["patchStart"], arrays["patchSize"], len(arrays["neighbour"]), dm.nB, Uin)
    dU, dp, dphi = ctx.zeros((3, n)), ctx.zeros(n), ctx.zeros(dm.nFaces)
    lc, lk = dm.linearise(3, *Ubc)
    dm.correct_boundary_flux(dU, lc, lk, dphi)
    dval, drhs = ctx.empty((max_dumps, nnz)), ctx.empty((max_dumps, n))
    dm.set_pressure_dump(dval, drhs, every, max_dumps)
    # relaxation U 0.9 / p 0.03 ramping from 0.02: the reference's own U 0.7 / p 0.3
    # diverges on these channels within a few iterations
    ctl = N.SimpleControls(maxIterations=iterations, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-6,
                           linearMaxIter=3000, nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02,
                           pPrecision=N.FP32_IR)
    res = dm.simple(nu, N.Schemes(0, 1, 1, 0.9), ctl, Ubc, pbc, dU, dp, dphi)
    count = dm.pressure_dump_count()
    if verbose:
        print(f"[learned] SIMPLE {res.iterations} iterations, continuity {res.finalContinuityError:.3e}, "
              f"diverged {res.diverged}, {count} pressure systems", flush=True)
    rp, cl = dm.pattern()
    geo = slot_geometry(arrays, geom, rp.astype(np.int64), cl.astype(np.int64))
    if res.diverged:
        return []  # the dumped systems of a diverging run are not pressure systems of a flow
    vals, rhs = dval.get(), drhs.get()
    out = []
    for k in range(count):
        if np.all(np.isfinite(vals[k])) and np.all(np.isfinite(rhs[k])):
            out.append((rp.astype(np.int64), cl.astype(np.int64), vals[k].copy(), rhs[k].copy(), geo))
    return out


def training_set(ctx, cases=8, size=(48, 24, 24), seed=1, per_case=3, iterations=30, verbose=False):
    """Pressure systems from `cases` seeded obstacle channels."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(cases):
        hm, _ = channel_mesh(*size, seed=seed * 1000 + c)
        Uin = rng.uniform(0.005, 0.05)
        out += simple_pressure_systems(ctx, hm, Uin, iterations=iterations, every=max(1, iterations // per_case),
                                       max_dumps=per_case, verbose=verbose)
    return out


def channel_run_device(ctx, hm, Uin, iterations, every=1, max_dumps=None, nu=1e-3, pressure=None, driver="simple",
                       h=None):
    """Runs the GPU SIMPLE loop (`iterations` outer iterations) or the PISO
    transient from rest (`iterations` pressure solves: iterations / 2 steps of
    2 correctors, dt = 0.25 h / Uin) on a channel mesh, copying every
    `every`-th pressure system into device memory (values [max][nnz], rhs
    [max][n]).  Returns a dict with the device mesh, the dump arrays and count,
    the pattern, the per-slot geometry and the driver result (`diverged`)."""
    arrays, geom = hm.arrays(), hm.geometry()
    dm = N.Mesh(ctx, arrays, geom)
    n, nnz = dm.nCells, dm.sizes()["nnz"]
    Ubc, pbc = channel_bcs(arrays["patchStart"], arrays["patchSize"], len(arrays["neighbour"]), dm.nB, Uin)
    dU, dp, dphi = ctx.zeros((3, n)), ctx.zeros(n), ctx.zeros(dm.nFaces)
    lc, lk = dm.linearise(3, *Ubc)
    dm.correct_boundary_flux(dU, lc, lk, dphi)
    max_dumps = max_dumps or (iterations + every - 1) // every
    dval, drhs = ctx.empty((max_dumps, nnz)), ctx.empty((max_dumps, n))
    dm.set_pressure_dump(dval, drhs, every, max_dumps)
    if pressure == N.FP32_MG:
        dm.set_mg(**TUNED_MG)
    prec = N.FP32_IR if pressure is None else pressure
    if driver == "simple":
        ctl = N.SimpleControls(maxIterations=iterations, continuityTol=0, pcorrTol=0, momentumTol=0,
                               linearTol=1e-6, linearMaxIter=3000, nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03,
                               pRelaxStart=0.02, pPrecision=prec)
        res = dm.simple(nu, N.Schemes(0, 1, 1, 0.9), ctl, Ubc, pbc, dU, dp, dphi)
        diverged = bool(res.diverged)
    else:
        ctl = N.PisoControls(dt=0.25 * h / Uin, correctors=2, continuityTol=0, linearTol=1e-6, linearMaxIter=3000,
                             nonOrthCorrectors=1, pPrecision=prec)
        diverged = False
        for _ in range((iterations + 1) // 2):
            res = dm.piso_step(nu, N.Schemes(0, 1, 1, 0.9), ctl, Ubc, pbc, dU, dp, dphi)
            if not np.isfinite(res.continuityError):
                diverged = True
                break
    count = dm.pressure_dump_count()
    dm.set_pressure_dump(None, None, 0, 0)
    rp, cl = dm.pattern()
    rp, cl = rp.astype(np.int64), cl.astype(np.int64)
    return dict(dm=dm, dval=dval, drhs=drhs, count=count, n=n, nnz=nnz, rp=rp, cl=cl,
                geo=slot_geometry(arrays, geom, rp, cl), res=res, diverged=diverged)


def _fcg_iterations(G, b, net, dgeo, maxIter):
    G.set_mpnn(net, dgeo if net is not None else None)
    _, rep = G.solve(b, None, 1e-8, maxIter, N.SOLVER_CG, N.PRECOND_LEARNED if net is not None else N.PRECOND_MG,
                     N.FP32_IR)
    G.set_mpnn(None)
    return rep.iterations if rep.converged and rep.precondFallbacks == 0 else None


def stream_train(ctx, net, channels=20, size=(128, 128, 128), length=(1.0, 1.0, 1.0), per_channel=100,
                 steps_per_system=2, lr=1e-4, seed=11, samples=2, smooth=1, log=None, driver="piso",
                 validate_every=10):
    """Trains `net` on pressure systems streamed from the GPU flow solver: per
    seeded obstacle channel (radius U[0.05, 0.15], inlet speed U[0.005, 0.05])
    the transient from rest (PISO, or SIMPLE iterations) leaves `per_channel`
    pressure systems in device memory, and the network takes
    `steps_per_system` Adam steps on each before the next channel runs.  No
    system is kept beyond its channel.

    The error samples alternate between smooth ones (`smooth` cycles applied
    to a random vector: the error flexible CG leaves) and raw random vectors
    (which keep the smoother weights where high frequencies are damped: a
    smoother tuned on smooth error alone drifts to weights that amplify them).
    One cycle is enough: after several cycles of the default three-sweep cycle
    a sample is close to an eigenvector of M A, on which one preconditioned
    step with the optimal length leaves almost nothing whatever the weights,
    and the objective stops telling weights apart.
    Every `validate_every` systems the updated network must solve the
    current system (K-cycle, FP64 residual 1e-8) in no more iterations than
    the tuned constants and than the last accepted parameters; otherwise
    those parameters are restored.  Returns (loss history, systems used,
    updates rejected, updates accepted)."""
    rng = np.random.default_rng(seed)
    hist, used, rejected, accepted = [], 0, 0, 0
    keep = net.params()
    for c in range(channels):
        hm, radius = channel_mesh(*size, seed=seed * 1000 + c, length=length)
        Uin = rng.uniform(0.005, 0.05)
        run = channel_run_device(ctx, hm, Uin, per_channel, pressure=N.FP32_MG, driver=driver,
                                 h=length[0] / size[0])
        if run["diverged"]:
            if log:
                log(f"channel {c}: {driver} diverged, skipped")
            continue
        G = N.Csr(ctx, run["rp"], run["cl"], np.zeros(run["nnz"]))
        G.set_mg(**TUNED_MG)
        dgeo = ctx.array(run["geo"])
        before = len(hist)
        for k in range(run["count"]):
            G.set_values_device(run["dval"].offset(k * run["nnz"]))
            h = net.train([G], [dgeo], [run["drhs"].offset(k * run["n"])], steps=steps_per_system,
                          samples=samples, smooth=smooth if k % 2 == 0 else 0, lr=lr,
                          seed=seed * 100003 + used)
            hist.extend(h.tolist())
            used += 1
            if (k + 1) % validate_every == 0 or k + 1 == run["count"]:
                b = run["drhs"].get()[k]
                base = _fcg_iterations(G, b, None, None, 1000)
                got = _fcg_iterations(G, b, net, dgeo, 1000 if base is None else base)
                cand = net.params()
                net.set_params(keep)
                prev = _fcg_iterations(G, b, net, dgeo, 1000 if base is None else base)
                ok = got is not None and base is not None and got <= base and (prev is None or got <= prev)
                if ok:
                    keep = cand
                    net.set_params(cand)
                    accepted += 1
                else:
                    rejected += 1
        if log:
            seg = np.asarray(hist[before:])
            log(f"channel {c}: radius {radius:.3f}, Uin {Uin:.4f}, {run['count']} systems, "
                f"loss {seg[:10].mean():.4f} -> {seg[-10:].mean():.4f}; updates accepted {accepted}, "
                f"rejected {rejected} so far")
        del G, run
    return np.asarray(hist), used, rejected, accepted


def heldout_iterations(ctx, nets, channels=2, size=(128, 128, 128), length=(1.0, 1.0, 1.0), seed=77,
                       iterations=61, every=30, tol=1e-8, driver="piso"):
    """Flexible-CG iterations to tol (FP32 + FP64 refinement) on held-out
    channel pressure systems with the tuned constants and with each network
    of `nets` (name -> Mpnn); one dict per system."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(channels):
        hm, radius = channel_mesh(*size, seed=seed * 1000 + c, length=length)
        Uin = rng.uniform(0.005, 0.05)
        run = channel_run_device(ctx, hm, Uin, iterations, every=every, pressure=N.FP32_MG, driver=driver,
                                 h=length[0] / size[0])
        if run["diverged"]:
            continue
        G = N.Csr(ctx, run["rp"], run["cl"], np.zeros(run["nnz"]))
        dgeo = ctx.array(run["geo"])
        for k in range(run["count"]):
            G.set_values_device(run["dval"].offset(k * run["nnz"]))
            b = run["drhs"].get()[k]
            row = {"channel": c, "system": k, "cells": run["n"]}
            for name, net in [("tuned", None)] + list(nets.items()):
                G.set_mg(**TUNED_MG)
                G.set_mpnn(net, dgeo if net is not None else None)
                _, rep = G.solve(b, None, tol, 2000, N.SOLVER_CG,
                                 N.PRECOND_LEARNED if net is not None else N.PRECOND_MG, N.FP32_IR)
                row[name] = rep.iterations
                row[name + "_converged"] = bool(rep.converged)
            G.set_mpnn(None)
            out.append(row)
    return out


def to_device(ctx, systems, mg=None):
    """Csr objects (with multigrid constants) and device geometry / rhs."""
    csrs, geos, rhss = [], [], []
    for rp, cl, v, rhs, geo in systems:
        G = N.Csr(ctx, rp, cl, v)
        if mg:
            G.set_mg(**mg)
        csrs.append(G)
        geos.append(ctx.array(geo))
        rhss.append(ctx.array(rhs))
    return csrs, geos, rhss


# the multigrid's default constants (tuned on a 128^3 jittered pressure system), written out
TUNED_MG = dict(omegaPre=0.8, omegaPost=0.8, scale=1.8, restrictScale=1.8, coarseOmega=0.9, coarseScale=3.0,
                smoothFine=4, smoothCoarse=2, smoothDeep=2)


def train(ctx, systems, steps=300, lr=1e-3, seed=1, samples=2, smooth=1, fp16=True, mg=TUNED_MG):
    """Trains a fresh network on the systems; returns (net, loss history)."""
    csrs, geos, rhss = to_device(ctx, systems, mg)
    net = N.Mpnn(ctx, seed=seed, fp16=fp16)
    hist = net.train(csrs, geos, rhss, steps=steps, samples=samples, smooth=smooth, lr=lr, seed=seed)
    return net, hist


def load(ctx, path=DEFAULT_WEIGHTS, fp16=True):
    net = N.Mpnn(ctx, fp16=fp16)
    net.set_params(np.load(path, allow_pickle=False).astype(np.float32))
    return net


def save(net, path=DEFAULT_WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, net.params())
