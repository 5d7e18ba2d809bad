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


TUNED_MG = dict(scale=1.8, coarseScale=1.8)


def train(ctx, systems, steps=300, lr=1e-3, seed=1, samples=2, smooth=4, fp16=True, mg=TUNED_MG):
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
