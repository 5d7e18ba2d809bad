"""Batched detail regions: many independent channel blocks (the 3D detail
regions of the 1D-3D coupling) whose pressure systems are solved together.

Each block is a jittered hex channel with seeded cell count (log-uniform),
aspect ratio, inlet rate and inlet profile (flat or parabolic).  The block's
DetailBlock runs a few GPU SIMPLE iterations at its rate and its last
pressure system (CSR values and right-hand side) is kept; the systems are
then solved to the reference's stop test one by one (FP64 Jacobi-PCG, the
reference's serial loop over blocks) and all at once (cfx_batch_solve).
"""
from __future__ import annotations

import time

import numpy as np

from . import native as N


def block_specs(count=64, cells=(20_000, 300_000), aspect=(4.0, 20.0), speed=(0.005, 0.05), seed=5,
                width=0.05):
    """(nx, ny, nz, length, width, inlet speed, shape) per block, seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1])))
        ar = rng.uniform(*aspect)
        ny = max(4, int(round((n / ar) ** (1.0 / 3.0))))
        nx = max(8, int(round(ar * ny)))
        out.append(dict(nx=nx, ny=ny, nz=ny, length=ar * width, width=width, speed=rng.uniform(*speed),
                        shape="parabolic" if rng.uniform() < 0.5 else "flat", seed=int(rng.integers(1 << 31))))
    return out


# The coupled cases' relaxation (relaxU 0.9, relaxP 0.03, pRelaxStart 0.02) diverges on these finer
# blocks within 10-40 SIMPLE iterations, in the oracle as on the GPU (the pressure loop's gain grows
# with resolution); relaxP 0.01 from a 0.005 floor converges on every block tried, the largest included.
SIMPLE_CTL = dict(continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-6, linearMaxIter=3000,
                  nonOrthCorrectors=1, relaxU=0.9, relaxP=0.01, pRelaxStart=0.005, pPrecision=N.FP32_IR)


def build_blocks(ctx, specs, simple_iterations=3, jitter=0.1, nu=1e-3, rho=1000.0, log=None, own_context=False):
    """Blocks with their last SIMPLE pressure system: list of dicts with the
    Csr (on `ctx`), device rhs, dp and sizes.  With own_context each block's
    DetailBlock runs on a context (stream) of its own and is kept ("detail",
    "q"), for concurrent coupling passes."""
    blocks = []
    for k, s in enumerate(specs):
        t0 = time.perf_counter()
        hm = N.HostMesh.channel(s["nx"], s["ny"], s["nz"], (0, 0, 0), (s["length"], s["width"], s["width"]), jitter,
                                s["seed"])
        bctx = N.Context(0) if own_context else ctx
        blk = N.Detail(bctx, hm, nu=nu, rho=rho, shape=s["shape"])
        dm = blk.mesh()
        nnz = dm.sizes()["nnz"]
        n = blk.nCells
        dval, drhs = bctx.empty((simple_iterations, nnz)), bctx.empty((simple_iterations, n))
        dm.set_pressure_dump(dval, drhs, 1, simple_iterations)
        q = s["speed"] * s["width"] * s["width"]
        dp, m = blk.loss_at(q, N.SimpleControls(maxIterations=simple_iterations, **SIMPLE_CTL))
        last = dm.pressure_dump_count() - 1
        dm.set_pressure_dump(None, None, 0, 0)
        if last < 0:
            raise RuntimeError(f"block {k}: SIMPLE produced no pressure system")
        rp, cl = dm.pattern()
        vals = dval.get()[last].copy()
        rhs = drhs.get()[last].copy()
        G = N.Csr(ctx, rp.astype(np.int64), cl.astype(np.int64), vals)
        blocks.append(dict(csr=G, b=ctx.array(rhs), n=n, nnz=nnz, dp=dp, measure=m, spec=s, q=q,
                           detail=blk if own_context else None, ctx=bctx if own_context else None,
                           host=(rp.astype(np.int64), cl.astype(np.int64), vals, rhs)))
        del blk, dval, drhs
        if log:
            log(f"block {k}: {s['nx']}x{s['ny']}x{s['nz']} = {n} cells, {s['shape']}, dp {dp:.4g} Pa "
                f"({time.perf_counter() - t0:.1f}s)")
    return blocks


def solve_one_by_one(ctx, blocks, tol=1e-8, maxIter=20000):
    xs, reps = [], []
    for b in blocks:
        x = ctx.zeros(b["n"])
        reps.append(b["csr"].solve_device(b["b"], x, tol, maxIter, N.SOLVER_CG, N.PRECOND_JACOBI, N.FP64))
        xs.append(x)
    return xs, reps


def solve_batched(ctx, blocks, tol=1e-8, maxIter=20000, schedule=None, x0=None):
    """x0: host initial guesses (default zero)."""
    xs = [ctx.zeros(b["n"]) if x0 is None else ctx.array(x0[k]) for k, b in enumerate(blocks)]
    reps, fin, st = N.batch_solve(ctx, [b["csr"] for b in blocks], [b["b"] for b in blocks], xs, tol, maxIter,
                                  schedule)
    return xs, reps, fin, st


def coupling_pass(blocks, simple_iterations=3, threads=1, pressure=N.FP32_IR):
    """One pass of the coupling loop over the blocks: lossAt at each block's rate, warm-started
    from its fields; pressure is the SIMPLE pressure solve (FP64 / FP32_IR Jacobi-PCG, FP32_MG
    multigrid FCG); returns (wall ms, dp list)."""
    ctl = N.SimpleControls(maxIterations=simple_iterations, **dict(SIMPLE_CTL, pPrecision=pressure))
    t0 = time.perf_counter()
    dps, _ = N.detail_loss_at_many([b["detail"] for b in blocks], [b["q"] for b in blocks], ctl, threads)
    return (time.perf_counter() - t0) * 1e3, dps
