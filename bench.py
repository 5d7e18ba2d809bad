#!/usr/bin/env python3
"""p-solve time to 1e-8 FP64 residual on the 128^3 non-orthogonal pressure
system, and SpMV bandwidth, on one MI355X.

Workload (one "step" = one pressure solve from x0 = 0):
  128^3 hex box [0,1]^3, interior vertices jittered uniformly by +-15% of the
  cell size (MT19937-64, seed 128), faces and cells shuffled (seed 7) and then
  RCM-renumbered; Laplacian coefficients (Sf.d)/|d|^2 (gamma = 1), all
  boundaries zeroGradient (pure Neumann); rhs N(0,1) (seed 2024) projected to
  zero mean.  Solved to ||r||inf / ||b||inf < 1e-8 (the reference stop test,
  decided on true FP64 residuals) by
    - plain FP64 Jacobi-PCG (the reference's solveCG on the GPU),
    - FP32-inner Jacobi-PCG with FP64 reliable updates,
    - flexible CG with the aggregation multigrid cycle (FP32, FP64 refinement),
      tuned constants,
    - the same with the message-passing network's per-row finest-level weights
      (network inference included in the time): `value`.
  The explicit non-orthogonal corrector is not part of the timed system: its
  boundary-face terms do not sum to zero, so with it the pure Neumann system
  has no solution and no solver reaches 1e-8 (the reference's included).

Other workloads: --workload c3 (PISO lid-driven cavity on 6M Kuhn tets),
c5 (64 batched detail regions), spmv (SpMV and a copy of known size, for the
HBM traffic counters).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "p-solve time to 1e-8 FP64 residual: learned+mixed PCG vs plain PCG; SpMV GB/s"
HBM_PEAK_GBPS = 8000.0
# HBM bytes per FP64 SpMV launch on the 128^3 system from rocprofv3 PMC (profiles/spmv_128cube_hbm_traffic.md):
# FETCH_SIZE 84,851.1 KiB / 0.50005 (the copy-kernel calibration) + WRITE_SIZE 16,384 KiB; the staging reads
# int16 column offsets, so the traffic is below the canonical bytes, which count 4-byte columns
SPMV_FP64_TRAFFIC_128 = 190534903
# HBM bytes per k_fcg_A launch on the 128^3 system (on the multigrid's sliced copy of the values), the same way:
# FETCH_SIZE 86,891.2 KiB / 0.50005 + WRITE_SIZE 8,322.2 KiB (profiles/spmv_128cube_hbm_traffic.md)
FCG_A_TRAFFIC_128 = 186457796
# HBM bytes per finest-level k_mg_up launch on the 128^3 system (64-row slices, int16 columns), the same way:
# FETCH_SIZE 55,747.9 KiB / 0.50005 + WRITE_SIZE 8,192.0 KiB (profiles/spmv_128cube_hbm_traffic.md)
MG_UP_TRAFFIC_128 = 122549474
SEED_MESH, SEED_SHUFFLE, SEED_RHS = 128, 7, 2024


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def canonical_fcg_bytes(n, nnz):
    # flexible CG's q = A p: FP64 value + int32 column per nonzero, int32 row pointer, FP32 p read once,
    # FP32 r read (p.r partial), FP32 q written
    return 12 * nnz + 4 * (n + 1) + 12 * n


def fine_level_bytes(n, nnz, sweeps=1):
    # per FCG iteration on the finest level: k_fcg_A, then `sweeps` pre-smoothing passes (FP32 column-scaled
    # values + int32 columns, row pointer, r and w read; x and r written) and as many post-smoothing passes
    # (FP32 values + columns, row pointer, x, b, w read; z written)
    down = 8 * nnz + 4 * (n + 1) + 16 * n
    up = 8 * nnz + 4 * (n + 1) + 16 * n
    return canonical_fcg_bytes(n, nnz) + sweeps * (down + up)


def canonical_up_bytes(n, nnz):
    # one finest-level post-smoothing sweep z = x + W (b - A x): FP32 value + int32 column per nonzero, int32
    # row pointer, x, b and W read once, z written once
    return 8 * nnz + 4 * (n + 1) + 16 * n


def canonical_spmv_bytes(n, nnz, s):
    # value + int32 column per nonzero, int32 row pointer, x read once, y written once
    return (s + 4) * nnz + 4 * (n + 1) + 2 * s * n


def build_system(N, ctx, n):
    t0 = time.perf_counter()
    hm = N.HostMesh.box(n, n, n, (0, 0, 0), (1, 1, 1), 0.15, SEED_MESH, False)
    hm.shuffle(SEED_SHUFFLE)
    bw_shuffled = hm.bandwidth()
    hm.renumber_rcm()
    arrays, geom = hm.arrays(), hm.geometry()
    t1 = time.perf_counter()
    log(f"mesh {n}^3: {arrays['nCells']} cells, bandwidth {bw_shuffled} -> {hm.bandwidth()} (RCM), "
        f"{t1 - t0:.1f}s host")
    dm = N.Mesh(ctx, arrays, geom)
    nB = dm.nB
    lin_c = ctx.array(np.ones((nB, 1)))
    lin_k = ctx.zeros((nB, 1))
    gamma = ctx.array(np.ones(dm.nFaces))
    A = N.Matrix(dm, 1)
    A.laplacian(gamma, lin_c, lin_k, -1)
    rng = np.random.default_rng(SEED_RHS)
    b = rng.standard_normal(dm.nCells)
    b -= b.mean()
    d_b = ctx.array(b)
    A.add_rhs(0, d_b)  # the matrix rhs is the system rhs (boundary terms vanish for zeroGradient)
    ctx.sync()
    log(f"device mesh + assembly {time.perf_counter() - t1:.1f}s; sizes {dm.sizes()}")
    return hm, arrays, geom, dm, A, b, bw_shuffled


def time_solve(N, ctx, A, d_x, precision, steps, warmup, tol, maxIter, precond=None, warm=None):
    """Solves from x = 0, `warmup` untimed then `steps` timed: [(wall ms, report)]; the warm-up walls are
    appended to `warm` when given."""
    reps = []
    precond = N.PRECOND_JACOBI if precond is None else precond
    for i in range(warmup + steps):
        ctx.check(N.lib().cfx_zero(ctx.h, d_x.ptr, d_x.nbytes))
        ctx.sync()
        t0 = time.perf_counter()
        rep = A.solve(d_x, N.SOLVER_CG, tol, maxIter, 0, None, None, None, -1, True, precond, precision)
        ctx.sync()
        wall = (time.perf_counter() - t0) * 1e3
        if i >= warmup:
            reps.append((wall, rep.as_dict()))
        elif warm is not None:
            warm.append(wall)
        log(f"  {'warmup' if i < warmup else 'step'} {i}: {wall:.2f} ms wall, {rep.ms:.2f} ms device, "
            f"{rep.iterations} it, outer {rep.outerIterations}, conv {rep.converged}")
    return reps


def cpu_info():
    try:
        model = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")),
                     platform.processor())
    except OSError:
        model = platform.processor()
    return model, os.cpu_count()


def cpu_baseline_and_check(arrays, b, x64_gpu, xmix_gpu, gpu_iters, sample_iters, A_dev, xmg_gpu=None, xln_gpu=None):
    """Checker + CPU baseline: the oracle restates the reference; the
    reference build itself (oracle/_ref) is timed when present."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402  (checker / CPU baseline only)

    out = {}
    t0 = time.perf_counter()
    m = O.Mesh(arrays["points"], arrays["faceStart"], arrays["faceVerts"], arrays["owner"], arrays["neighbour"],
               arrays["patchStart"], arrays["patchSize"], arrays["nCells"]).compute_geometry()
    lc, lk = O.linearise(m, 1, np.full(m.nB, 2, np.int32), np.zeros((m.nB, 1)))
    L = O.Ldu(m)
    O.assemble_laplacian(m, L, np.ones(m.nFaces), lc, lk, -1)
    csr = O.csr_from_ldu(m, L)
    d, r, u, lo = A_dev.get()
    out["assembly_bitwise_equal"] = bool(np.array_equal(d[0], L.diag[0]) and np.array_equal(u, L.upper[:m.nInternal])
                                         and np.array_equal(lo, L.lower[:m.nInternal]))
    bn = np.abs(b).max()
    out["fp64_true_residual"] = float(np.abs(b - O.spmv(csr, x64_gpu)).max() / bn)
    out["mixed_true_residual"] = float(np.abs(b - O.spmv(csr, xmix_gpu)).max() / bn)
    out["mixed_vs_fp64_rel_err"] = float(np.abs((xmix_gpu - xmix_gpu.mean()) - (x64_gpu - x64_gpu.mean())).max()
                                         / np.abs(x64_gpu - x64_gpu.mean()).max())
    if xln_gpu is not None:
        out["learned_true_residual"] = float(np.abs(b - O.spmv(csr, xln_gpu)).max() / bn)
        out["learned_vs_fp64_rel_err"] = float(np.abs((xln_gpu - xln_gpu.mean()) - (x64_gpu - x64_gpu.mean())).max()
                                               / np.abs(x64_gpu - x64_gpu.mean()).max())
    if xmg_gpu is not None:
        out["mg_true_residual"] = float(np.abs(b - O.spmv(csr, xmg_gpu)).max() / bn)
        out["mg_vs_fp64_rel_err"] = float(np.abs((xmg_gpu - xmg_gpu.mean()) - (x64_gpu - x64_gpu.mean())).max()
                                          / np.abs(x64_gpu - x64_gpu.mean()).max())
    log(f"checker: assembly bit-exact {out['assembly_bitwise_equal']}, true residuals "
        f"fp64 {out['fp64_true_residual']:.2e} mixed {out['mixed_true_residual']:.2e} ({time.perf_counter()-t0:.1f}s)")

    model, ncpu = cpu_info()
    kind = "port"
    if O.have_ref():
        # the reference's own solveCG (compiled from its sources): time a bounded
        # number of iterations, subtracting the CSR build (measured with 0 iterations)
        R = O.ref()
        rs, cl, vv = csr.rowStart, csr.cols, csr.vals
        x = np.zeros(m.nCells)
        res = np.zeros(3)

        def ref_run(k):
            x[:] = 0
            t = time.perf_counter()
            R.cfr_solve_cg(O.C.c_int64(m.nCells), O._ptr(rs, O.C.c_int64), O._ptr(cl, O.C.c_int64), O._ptr(vv),
                           O._ptr(b), O._ptr(x), O.C.c_double(1e-8), O.C.c_int(k), O.C.c_int(0), O._ptr(res))
            return time.perf_counter() - t
        t_build = ref_run(0)
        t_full = ref_run(sample_iters)
        per_it = (t_full - t_build) / sample_iters
        kind = "reference"
    else:
        t = time.perf_counter()
        O.solve_cg(csr, b, None, 1e-8, sample_iters)
        per_it = (time.perf_counter() - t) / sample_iters
    est_ms = per_it * gpu_iters * 1e3
    log(f"cpu baseline ({kind}): {per_it*1e3:.1f} ms/iteration x {gpu_iters} iterations = {est_ms/1e3:.1f} s")
    out["cpu_baseline"] = {"value": est_ms, "unit": "ms", "cores": 1, "kind": kind,
                           "sample": f"{sample_iters} Jacobi-PCG iterations of the same 128^3 system "
                                     f"(FP64, 1 thread, {model}), extrapolated to the {gpu_iters} iterations "
                                     f"the FP64 solve takes; {ncpu} host cores present"}
    return out


def bench_c5(args):
    """64 detail-region pressure systems (about 6M cells): one by one vs batched."""
    from chuanflow_amd import detail_batch as DB
    from chuanflow_amd import native as N
    ctx = N.Context(0)
    ctx.set_graphs(bool(args.graphs))
    t0 = time.perf_counter()
    specs = DB.block_specs(count=args.blocks, seed=5)
    blocks = DB.build_blocks(ctx, specs, simple_iterations=3, log=log, own_context=True)
    cells = sum(b["n"] for b in blocks)
    log(f"{len(blocks)} blocks, {cells} cells, built in {time.perf_counter() - t0:.1f}s")
    xs = [ctx.zeros(b["n"]) for b in blocks]
    tt = time.perf_counter()
    best, table = N.batch_autotune(ctx, [b["csr"] for b in blocks], [b["b"] for b in blocks], xs, iters=20)
    tune_s = time.perf_counter() - tt
    log(f"autotuned {best.as_dict()} in {tune_s:.2f}s")
    one, bat, bdef, b3 = [], [], [], []
    persist = N.BatchSchedule(4, 2048, 1, 16, persistent=1)  # one launch for the whole batch
    for i in range(args.warmup + args.steps):
        ctx.sync()
        t1 = time.perf_counter()
        x1, r1 = DB.solve_one_by_one(ctx, blocks, args.tol)
        ctx.sync()
        t2 = time.perf_counter()
        xb, rb, fin, st = DB.solve_batched(ctx, blocks, args.tol, schedule=best)
        ctx.sync()
        t3 = time.perf_counter()
        xd, rd, find, std = DB.solve_batched(ctx, blocks, args.tol, schedule=N.BatchSchedule())
        ctx.sync()
        t4 = time.perf_counter()
        x3, r3, fin3, st3 = DB.solve_batched(ctx, blocks, args.tol, schedule=persist)
        ctx.sync()
        t5 = time.perf_counter()
        if i >= args.warmup:
            one.append((t2 - t1) * 1e3)
            bat.append((t3 - t2) * 1e3)
            bdef.append((t4 - t3) * 1e3)
            b3.append((t5 - t4) * 1e3)
        log(f"step {i}: one by one {(t2 - t1) * 1e3:.1f} ms (device {sum(r.ms for r in r1):.1f}), batched tuned "
            f"{(t3 - t2) * 1e3:.1f} ms (device {st.ms:.1f}), batched default {(t4 - t3) * 1e3:.1f} ms "
            f"(device {std.ms:.1f}), persistent {(t5 - t4) * 1e3:.1f} ms (device {st3.ms:.1f})")
    # the coupling loop's pass over its blocks (lossAt: 3 warm-started GPU SIMPLE iterations each), one
    # block after another and on concurrent streams (one context per block, host threads)
    passes = {}
    for pname, prec in (("jacobi", N.FP32_IR), ("multigrid", N.FP32_MG)) if args.c5_passes else ():
        for threads in (1, 4, 8):
            ms, dps = DB.coupling_pass(blocks, 3, threads, prec)
            passes[f"{pname}_{threads}"] = ms
            log(f"coupling pass, {pname} pressure, {threads} thread(s): {ms:.1f} ms, dp[0] {dps[0]:.6g} Pa")
    # CPU baseline: the reference build's solveCG on every block's own system (one core), a bounded number of
    # iterations per block (the CSR build, timed with 0 iterations, subtracted), extrapolated to the iterations
    # that block takes
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402  (CPU baseline only)
    cpu = None
    if O.have_ref() and not args.no_cpu:
        R = O.ref()
        total_ms, sample = 0.0, args.c5_cpu_sample
        t_all = time.perf_counter()
        for blk, rep in zip(blocks, rb):
            rs, cl, vv, rhs = blk["host"]
            x, res = np.zeros(blk["n"]), np.zeros(3)

            def ref_run(k):
                x[:] = 0
                t = time.perf_counter()
                R.cfr_solve_cg(O.C.c_int64(blk["n"]), O._ptr(rs, O.C.c_int64), O._ptr(cl, O.C.c_int64), O._ptr(vv),
                               O._ptr(rhs), O._ptr(x), O.C.c_double(1e-8), O.C.c_int(k), O.C.c_int(0), O._ptr(res))
                return time.perf_counter() - t
            per_it = (ref_run(sample) - ref_run(0)) / sample
            total_ms += per_it * rep.iterations * 1e3
        model, _ = cpu_info()
        log(f"cpu baseline (reference solveCG, one core): {total_ms / 1e3:.1f} s for the 64 blocks "
            f"({time.perf_counter() - t_all:.1f} s of sampling)")
        cpu = {"value": round(total_ms, 1), "unit": "ms", "cores": 1, "kind": "reference",
               "sample": f"{sample} Jacobi-PCG iterations of every block's system with the reference build's "
                         f"solveCG (FP64, 1 thread, {model}), each extrapolated to the iterations that block takes"}
    bitwise = all(np.array_equal(a.get(), b.get()) and np.array_equal(a.get(), c.get()) for a, b, c in zip(x1, xd, x3))
    conv = all(r.converged for r in rb) and all(r.converged for r in r1)
    tb, to = float(np.mean(bat)), float(np.mean(one))
    per_block = [{"cells": b["n"], "shape": b["spec"]["shape"], "aspect": round(b["spec"]["length"] / b["spec"]["width"], 2),
                  "iterations": r.iterations, "alone_ms": round(ra.ms, 3), "batched_finish_ms": round(f, 3),
                  "persistent_finish_ms": round(fp, 3), "dp_Pa": round(b["dp"], 6)}
                 for b, r, ra, f, fp in zip(blocks, rb, r1, fin, fin3)]
    result = {
        "metric": "detail-region pressure systems solved to 1e-8: batched vs one by one",
        "value": round(tb, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(tb, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(blocks)} jittered-hex channel blocks, cells log-uniform [20k,300k], aspect "
                               f"U[4,20], inlet U[0.005,0.05] m/s, flat/parabolic; last of 3 GPU SIMPLE pressure "
                               f"systems each; FP64 Jacobi-PCG to 1e-8", "cells": cells, "blocks": len(blocks),
                   "parallelism": "single GPU"},
        "one_by_one_ms": round(to, 3), "batched_default_ms": round(float(np.mean(bdef)), 3),
        "batched_persistent_ms": round(float(np.mean(b3)), 3),
        "batched_forms": "default: A, B, C launches per iteration over all systems (captured hipGraph), host "
                         "polling every 16 iterations; persistent: one launch (captured hipGraph) for the whole "
                         "batch, workgroups claiming chunks of any unfinished system; both bitwise equal to the "
                         "solves alone",
        "speedup_batched_vs_one_by_one": round(to / tb, 3), "autotune_s": round(tune_s, 3),
        "schedule": best.as_dict(), "candidates": [{"schedule": c, "ms": round(m, 4)} for c, m in table],
        "max_iterations": st.maxIterations, "all_converged": conv, "default_schedule_bitwise_equal": bitwise,
        "coupling_pass_ms": {k: round(v, 1) for k, v in passes.items()},
        "coupling_pass": "lossAt of every block (3 warm-started GPU SIMPLE iterations each, relaxU 0.9, relaxP 0.01; "
                         "momentum BiCGStab, pressure to 1e-6 by Jacobi-PCG FP32+FP64 refinement or multigrid FCG), "
                         "serial vs host threads on one stream per block; passes continue one from another",
        "per_block": per_block,
    }
    if cpu:
        result["cpu_baseline"] = cpu
    return result


def cavity_bcs(patch_start, patch_size, n_internal, n_boundary, lid=1.0):
    """Lid-driven cavity: the y-max patch (4th of the box) moves at (lid, 0, 0),
    the other walls are at rest; p zeroGradient everywhere."""
    pid = np.empty(n_boundary, np.int64)
    for p, (s, n) in enumerate(zip(patch_start, patch_size)):
        pid[s - n_internal:s - n_internal + n] = p
    Ut = np.zeros(n_boundary, np.int32)
    Uv = np.zeros((n_boundary, 3))
    Uv[pid == 3, 0] = lid
    return (Ut, Uv), (np.full(n_boundary, 2, np.int32), np.zeros((n_boundary, 1)))


@contextlib.contextmanager
def stdout_to_stderr():
    """Send what native code writes to stdout (the reference build's progress lines) to stderr, so stdout
    carries only the JSON line."""
    sys.stdout.flush()
    libc = ctypes.CDLL(None)
    libc.fflush(None)
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def channel_bcs(patch_start, patch_size, n_internal, n_boundary, u_in):
    """Box channel along x: inlet (1st patch) fixed U = (u_in, 0, 0), outlet (2nd) zeroGradient U and
    fixed p = 0, the other patches no-slip walls with zeroGradient p; per boundary face and per patch
    (the reference driver's patch-level form)."""
    pid = np.empty(n_boundary, np.int64)
    for p, (s, n) in enumerate(zip(patch_start, patch_size)):
        pid[s - n_internal:s - n_internal + n] = p
    Ut = np.where(pid == 1, 2, 0).astype(np.int32)
    Uv = np.zeros((n_boundary, 3))
    Uv[pid == 0, 0] = u_in
    pt = np.where(pid == 1, 0, 2).astype(np.int32)
    pv = np.zeros((n_boundary, 1))
    npatch = len(patch_start)
    Upt = np.array([0, 2] + [0] * (npatch - 2), np.int32)
    Upv = np.zeros((npatch, 3))
    Upv[0, 0] = u_in
    ppt = np.array([2, 0] + [2] * (npatch - 2), np.int32)
    return (Ut, Uv), (pt, pv), (Upt, Upv, ppt, np.zeros(npatch))


def bench_c1(args):
    """Steady laminar channel, 40x20x20 hex (16,000 cells, 1 x 0.5 x 0.5 m, BlockMesh face order), nu 1e-3,
    inlet 0.01 m/s: 200 SIMPLE iterations from rest per step.  Relaxation U 0.9 / p 0.03 from a 0.02
    floor: the literal 0.7 / 0.3 diverges, in the reference build as on the GPU, at the same iteration.
    The fields of a 20-iteration prefix are checked against the oracle; the reference build's SIMPLE on
    the same case, timed over a 10-iteration sample, is the CPU baseline."""
    from chuanflow_amd import native as N
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402  (checker and CPU baseline only)
    ctx = N.Context(0)
    ctx.set_graphs(bool(args.graphs))
    hm = N.HostMesh.box(40, 20, 20, (0, 0, 0), (1.0, 0.5, 0.5), 0.0, 0, False)
    arrays, geom = hm.arrays(), hm.geometry()
    dm = N.Mesh(ctx, arrays, geom)
    nI = len(arrays["neighbour"])
    Ubc, pbc, patch_bcs = channel_bcs(arrays["patchStart"], arrays["patchSize"], nI, dm.nB, 0.01)
    om = O.Mesh(arrays["points"], arrays["faceStart"], arrays["faceVerts"], arrays["owner"], arrays["neighbour"],
                arrays["patchStart"], arrays["patchSize"], arrays["nCells"]).compute_geometry()

    def rest():
        U, p, phi = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
        lc, lk = O.linearise(om, 3, *Ubc)
        O.correct_boundary_flux(om, U, lc, lk, phi)
        return U, p, phi

    kw = dict(continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-8, linearMaxIter=2000,
              nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02)
    sch = N.Schemes(0, 1, 1, 0.9)
    runs = {}
    variants = [v for v in (("fp64", N.FP64), ("fp32ir", N.FP32_IR), ("multigrid", N.FP32_MG))
                if v[0] in args.c1_variants.split(",")]
    if not variants or variants[0][0] != "fp64":
        raise SystemExit("--c1-variants must start with fp64 (the reference's pressure solve, the metric)")
    for name, pp in variants:
        ctl = N.SimpleControls(maxIterations=args.c1_iterations, **kw, pPrecision=pp)
        times, devs = [], []
        for step in range(args.warmup + args.steps):
            U, p, phi = rest()
            dU, dp, dphi = ctx.array(np.ascontiguousarray(U.T)), ctx.array(p), ctx.array(phi)
            ctx.sync()
            t0 = time.perf_counter()
            r = dm.simple(1e-3, sch, ctl, Ubc, pbc, dU, dp, dphi)
            ctx.sync()
            if step >= args.warmup:
                times.append((time.perf_counter() - t0) * 1e3)
                devs.append(r.ms)
        runs[name] = dict(ms=float(np.mean(times)), device_ms=float(np.mean(devs)), iterations=r.iterations,
                          final_continuity=r.finalContinuityError, diverged=bool(r.diverged))
        log(f"C1 {args.c1_iterations} SIMPLE iterations, pressure {name}: {runs[name]['ms']:.1f} ms "
            f"(device {runs[name]['device_ms']:.1f}), continuity {r.finalContinuityError:.3e}")
    # parity: a 20-iteration prefix, GPU (reference pressure solve) against the oracle
    nchk = 20
    U, p, phi = rest()
    dU, dp, dphi = ctx.array(np.ascontiguousarray(U.T)), ctx.array(p), ctx.array(phi)
    dm.simple(1e-3, sch, N.SimpleControls(maxIterations=nchk, **kw), Ubc, pbc, dU, dp, dphi)
    O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(maxIterations=nchk, **kw), Ubc, pbc)
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    parity = {"iterations": nchk, "U": rel(dU.get().T, U), "p": rel(dp.get(), p), "phi": rel(dphi.get(), phi)}
    log(f"parity after {nchk} iterations: {parity}")
    # CPU baseline: the reference build's SIMPLE (one core), a sample of iterations, scaled to the full run
    sample = args.c1_cpu_sample
    U, p, phi = rest()
    ctl_vec = [sample, 0, 0, 0, 1e-8, 2000, 2, 0.9, 0.03, 0.02, 0, 0, 0]
    if O.have_ref():
        Upt, Upv, ppt, ppv = patch_bcs
        rm = O.RefMesh(om)
        with stdout_to_stderr():
            t0 = time.perf_counter()
            rm.simple(Upt, Upv, ppt, ppv, 1e-3, (0, 1, 1), ctl_vec, U, p, phi)
            cpu_s = time.perf_counter() - t0
        kind = "reference"
    else:
        t0 = time.perf_counter()
        O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(maxIterations=sample, **kw),
                 Ubc, pbc)
        cpu_s = time.perf_counter() - t0
        kind = "port"
    cpu_ms = cpu_s * 1e3 * args.c1_iterations / sample
    model, _ = cpu_info()
    log(f"CPU ({kind}, 1 core, {model}): {sample} iterations in {cpu_s:.2f} s -> {cpu_ms / 1e3:.1f} s for "
        f"{args.c1_iterations}")
    g = runs["fp64"]
    result = {
        "metric": f"C1 steady channel: {args.c1_iterations} SIMPLE iterations, wall time",
        "value": round(g["ms"], 2), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(g["ms"], 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "steady laminar channel 40x20x20 hex (16,000 cells, 1 x 0.5 x 0.5 m, BlockMesh "
                               "face order), nu 1e-3, inlet 0.01 m/s, SIMPLE from rest, relaxation U 0.9 / p "
                               "0.03 (floor 0.02), Jacobi-PCG pressure to 1e-8, 2 non-orthogonal correctors",
                   "cells": dm.nCells, "iterations": args.c1_iterations, "parallelism": "single GPU"},
        "device_ms": round(g["device_ms"], 2), "ms_per_iteration": round(g["ms"] / args.c1_iterations, 3),
        "final_continuity": g["final_continuity"],
        "pressure_fp32ir_ms": round(runs["fp32ir"]["ms"], 2) if "fp32ir" in runs else None,
        "pressure_multigrid_ms": round(runs["multigrid"]["ms"], 2) if "multigrid" in runs else None,
        "parity_vs_oracle": parity,
        "cpu_baseline": {"value": round(cpu_ms, 1), "unit": "ms", "cores": 1, "kind": kind,
                         "sample": f"{sample} SIMPLE iterations of the same case from rest, scaled to "
                                   f"{args.c1_iterations}", "cpu": model},
    }
    return result


def kinetic_energy(V, U3):
    return 0.5 * float((V * (U3 ** 2).sum(axis=0)).sum())


def bench_c3(args):
    """Transient lid-driven cavity, Re 1000, on Kuhn-split jittered tets: PISO per-step time and
    kinetic energy; the kinetic-energy drift against the FP64 CPU oracle on a small instance."""
    from chuanflow_amd import learned as LN
    from chuanflow_amd import native as N
    ctx = N.Context(0)
    ctx.set_graphs(bool(args.graphs))
    prec = {"fp64": N.FP64, "fp32ir": N.FP32_IR, "mg": N.FP32_MG}[args.c3_pressure]

    def setup(n, learned):
        hm = N.HostMesh.box(n, n, n, (0, 0, 0), (1, 1, 1), 0.10, SEED_MESH, True)
        hm.renumber_rcm()
        arrays, geom = hm.arrays(), hm.geometry()
        dm = N.Mesh(ctx, arrays, geom)
        Ubc, pbc = cavity_bcs(arrays["patchStart"], arrays["patchSize"], len(arrays["neighbour"]), dm.nB)
        dU, dp, dphi = ctx.zeros((3, dm.nCells)), ctx.zeros(dm.nCells), ctx.zeros(dm.nFaces)
        # the cavity's constants: measured at 61 ms per step with these (the 128^3 defaults are tuned for cold
        # solves to 1e-8 with four sweeps); the coarse sweeps on every K-cycle level
        dm.set_mg(omegaPre=2.0 / 3.0, omegaPost=2.0 / 3.0, scale=1.8, restrictScale=1.0, coarseOmega=2.0 / 3.0,
                  coarseScale=1.8, kcycleDepth=args.kcycle_depth, fusion=args.mg_fusion,
                  smoothFine=args.c3_smooth_fine, smoothCoarse=args.c3_smooth_coarse,
                  smoothDeep=args.c3_smooth_coarse)
        if learned:
            dm.set_mpnn(LN.load(ctx))
        return hm, arrays, geom, dm, Ubc, pbc, dU, dp, dphi

    kw = dict(dt=2e-3, correctors=2, continuityTol=1e-10, linearTol=args.tol, linearMaxIter=20000,
              nonOrthCorrectors=1)
    # laplacian "uncorrected" by default: on Kuhn tets the reference's explicit non-orthogonal correction is a
    # diverging fixed-point iteration (continuity error x4 every 5 steps, worse with more correctors, at any dt)
    corrected = 1 if args.c3_laplacian == "corrected" else 0
    sch = N.Schemes(0, corrected, 1, 0.9)
    t0 = time.perf_counter()
    learned = prec == N.FP32_MG and args.c3_network
    hm, arrays, geom, dm, Ubc, pbc, dU, dp, dphi = setup(args.c3_n, learned)
    V = np.asarray(geom["V"])
    log(f"cavity: {dm.nCells} tets, mesh {time.perf_counter() - t0:.1f}s, pressure solve {args.c3_pressure}")
    # the closed cavity's pressure is a singular Neumann system; --c3-project makes both the GPU and the oracle
    # project its right-hand side to zero sum (the reference PISO does not)
    uprec = N.FP32_IR if args.c3_momentum == "fp32ir" else N.FP64
    ctl = N.PisoControls(**kw, pPrecision=prec, projectNeumann=args.c3_project, uPrecision=uprec)
    times, ke, dev = [], [], []
    for step in range(args.c3_steps):
        ctx.sync()
        t1 = time.perf_counter()
        r = dm.piso_step(1e-3, sch, ctl, Ubc, pbc, dU, dp, dphi)
        ctx.sync()
        times.append((time.perf_counter() - t1) * 1e3)
        dev.append(r.ms)
        if step % 10 == 9 or step == args.c3_steps - 1:
            ke.append((step + 1, kinetic_energy(V, dU.get())))
            log(f"step {step + 1}: {times[-1]:.1f} ms, KE {ke[-1][1]:.6e}, continuity {r.continuityError:.2e}")
    # kinetic-energy drift against the FP64 CPU oracle: the same case on a small mesh, a step prefix
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # noqa: E402  (checker and CPU baseline only)
    hs, a_s, g_s, ds, Ubs, pbs, sU, sp_, sphi = setup(args.c3_check_n, learned)
    om = O.Mesh(a_s["points"], a_s["faceStart"], a_s["faceVerts"], a_s["owner"], a_s["neighbour"], a_s["patchStart"],
                a_s["patchSize"], a_s["nCells"]).compute_geometry()
    Uo, po, phio = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    drift = []
    tc = time.perf_counter()
    Vs = np.asarray(g_s["V"])
    for step in range(args.c3_check_steps):
        ds.piso_step(1e-3, sch, ctl, Ubs, pbs, sU, sp_, sphi)
        O.piso_step(om, Uo, po, phio, 1e-3, O.Schemes(0, corrected, 1, 0.9),
                    O.PisoControls(**kw, projectNeumann=args.c3_project), Ubs, pbs)
        kg, ko = kinetic_energy(Vs, sU.get()), kinetic_energy(Vs, Uo.T)
        drift.append({"step": step + 1, "ke_gpu": kg, "ke_oracle": ko, "rel": abs(kg - ko) / ko})
    cpu_ms = (time.perf_counter() - tc) * 1e3 / max(1, args.c3_check_steps)
    mean_ms = float(np.mean(times[1:] if len(times) > 1 else times))
    # CPU baseline: one step of the reference build's PISO (Jacobi-PCG pressure, FP64, one core) on a smaller
    # instance of the same cavity, scaled by the cell count (a step of the 6M-tet case takes minutes on one core)
    cpu = None
    if O.have_ref() and args.c3_cpu_n > 0:
        nc = args.c3_cpu_n
        hc = N.HostMesh.box(nc, nc, nc, (0, 0, 0), (1, 1, 1), 0.10, SEED_MESH, True)
        hc.renumber_rcm()
        ac = hc.arrays()
        mc = O.Mesh(ac["points"], ac["faceStart"], ac["faceVerts"], ac["owner"], ac["neighbour"], ac["patchStart"],
                    ac["patchSize"], ac["nCells"]).compute_geometry()
        rm = O.RefMesh(mc)
        npatch = len(ac["patchStart"])
        Upv = np.zeros((npatch, 3))
        Upv[3, 0] = 1.0  # the lid, as cavity_bcs
        Uc, pc, phic = np.zeros((mc.nCells, 3)), np.zeros(mc.nCells), np.zeros(mc.nFaces)
        ctl_vec = [2e-3, 2, 1e-10, args.tol, 20000, 1, 0, 0, 0]
        with stdout_to_stderr():
            t0c = time.perf_counter()
            rm.piso_step(np.zeros(npatch, np.int32), Upv, np.full(npatch, 2, np.int32), np.zeros(npatch), 1e-3,
                         (0, corrected, 1), ctl_vec, Uc, pc, phic)
            sec = time.perf_counter() - t0c
        scaled = sec * 1e3 * dm.nCells / mc.nCells
        model, _ = cpu_info()
        log(f"cpu baseline (reference PISO, one core): {sec:.1f} s for one step on {mc.nCells} tets -> "
            f"{scaled / 1e3:.1f} s per step scaled to {dm.nCells}")
        cpu = {"value": round(scaled, 1), "unit": "ms", "cores": 1, "kind": "reference",
               "sample": f"one PISO step (the first, from rest) of the same cavity on {nc}^3 hex -> {mc.nCells} "
                         f"tets with the reference build (Jacobi-PCG pressure and BiCGStab momentum, FP64, 1 thread, "
                         f"{model}), scaled by the cell count to {dm.nCells}"}
    result = {
        "metric": "PISO step time, lid-driven cavity Re 1000 on Kuhn tets; kinetic-energy drift vs FP64 CPU oracle",
        "value": round(mean_ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.c3_steps, "warmup": 1,
        "ms_per_step": round(mean_ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f64 assembly; momentum {args.c3_momentum}; pressure {args.c3_pressure}{' + network weights' if args.c3_network else ' (constant weights)'}", "data": "synthetic",
        "config": {"workload": f"{args.c3_n}^3 hex unit box, +-10% jitter (seed {SEED_MESH}), 6 Kuhn tets per hex, "
                               f"RCM; lid U=(1,0,0), nu 1e-3, PISO 2 correctors, dt 2e-3, laplacian "
                               f"{args.c3_laplacian}, rhs projection {args.c3_project}", "cells": dm.nCells,
                   "parallelism": "single GPU"},
        "step_ms_first": round(times[0], 3), "step_ms_mean": round(mean_ms, 3),
        "device_ms_mean": round(float(np.mean(dev[1:] if len(dev) > 1 else dev)), 3), "kinetic_energy": ke,
        "ke_check": {"cells": ds.nCells, "steps": args.c3_check_steps, "max_rel_drift": max(d["rel"] for d in drift),
                     "per_step": drift, "oracle_step_ms": round(cpu_ms, 1)},
    }
    if cpu:
        result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)


def bench_spmv(args):
    """SpMV alone on the 128^3 pressure system (FP64 and FP32 values, the flexible CG's fine-level
    k_fcg_A and the multigrid's finest-level post-smoothing sweep k_mg_up; args.spmv_reps launches each) and a
    device-to-device copy of a known byte count: the workload the HBM traffic counters are read on
    (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE; the copy's counters give the scale to bytes)."""
    from chuanflow_amd import native as N
    ctx = N.Context(0)
    ctx.set_graphs(bool(args.graphs))
    hm, arrays, geom, dm, A, b, bw = build_system(N, ctx, args.n)
    n, nnz = dm.nCells, dm.sizes()["nnz"]
    xs = ctx.array(np.random.default_rng(1).standard_normal(n))
    ys = ctx.empty(n)
    xs32 = ctx.array(np.random.default_rng(1).standard_normal(n).astype(np.float32))
    ys32 = ctx.empty(n, np.float32)
    ctx.sync()  # the uploads stay out of the timed loops
    A.time_spmv(0, xs, ys, args.spmv_reps, N.FP64)  # the device's first work since the assembly: warm-up
    ms64 = A.time_spmv(0, xs, ys, args.spmv_reps, N.FP64)
    ms32 = A.time_spmv(0, xs32, ys32, args.spmv_reps, N.FP32_IR)
    # the finest-level post-smoothing sweep (the call builds the multigrid hierarchy on the pattern; no solve),
    # then k_fcg_A on the hierarchy's sliced copy of the values, as the learned solve runs both
    msu = A.time_smooth(args.spmv_reps)
    msf = A.time_fcg_spmv(args.spmv_reps)
    copy_bytes = 256 << 20
    src, dst = ctx.zeros(copy_bytes // 8), ctx.empty(copy_bytes // 8)
    for _ in range(args.spmv_reps):
        dst.copy_from(src)
    ctx.sync()
    b64, b32 = canonical_spmv_bytes(n, nnz, 8), canonical_spmv_bytes(n, nnz, 4)
    print(json.dumps({"workload": "spmv", "cells": n, "nnz": nnz, "reps": args.spmv_reps,
                      "spmv_fp64_us": round(ms64 * 1e3, 2), "spmv_fp64_bytes": b64,
                      "spmv_fp64_GBps": round(b64 / (ms64 * 1e-3) / 1e9, 1),
                      "spmv_fp32_us": round(ms32 * 1e3, 2), "spmv_fp32_bytes": b32,
                      "spmv_fp32_GBps": round(b32 / (ms32 * 1e-3) / 1e9, 1),
                      "fcg_a_us": round(msf * 1e3, 2), "fcg_a_bytes": canonical_fcg_bytes(n, nnz),
                      "fcg_a_GBps": round(canonical_fcg_bytes(n, nnz) / (msf * 1e-3) / 1e9, 1),
                      "mg_up_us": round(msu * 1e3, 2), "mg_up_bytes": canonical_up_bytes(n, nnz),
                      "mg_up_GBps": round(canonical_up_bytes(n, nnz) / (msu * 1e-3) / 1e9, 1),
                      "copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="cells per edge (128 = the 2.1M-cell workload)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-sample-iters", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--mg-omega", type=float, default=None, help="finest-level smoother weight (pre = post); "
                    "default: the library's tuned constants")
    ap.add_argument("--mg-scale", type=float, default=None, help="finest-level prolongation = restriction scale")
    ap.add_argument("--mg-coarse-omega", type=float, default=None)
    ap.add_argument("--mg-coarse-scale", type=float, default=None)
    ap.add_argument("--kcycle-depth", type=int, default=2)
    ap.add_argument("--mg-fusion", type=int, default=125)
    ap.add_argument("--graphs", type=int, default=0, choices=[0, 1],
                    help="flexible CG: replay a captured iteration (1) or launch kernel by kernel (0)")
    ap.add_argument("--smooth-fine", type=int, default=4, help="multigrid smoothing sweeps on the finest level")
    ap.add_argument("--smooth-coarse", type=int, default=2, help="multigrid smoothing sweeps on the K-cycle levels")
    ap.add_argument("--inner-tol", type=float, default=0.0,
                    help="reliable-update factor of the FP32-refined solves (0: each solver's default, 1e-2 for "
                         "Jacobi-PCG and 1e-4 for flexible CG)")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c5", "spmv"],
                    help="c2: the 128^3 pressure system (the headline metric); c1: SIMPLE channel 40x20x20; "
                         "c3: PISO cavity on tets; "
                         "c5: batched detail regions; spmv: SpMV and a copy of known size (traffic counters)")
    ap.add_argument("--c1-iterations", type=int, default=200)
    ap.add_argument("--c1-cpu-sample", type=int, default=10)
    ap.add_argument("--c1-variants", default="fp64,fp32ir,multigrid",
                    help="pressure solves timed: fp64 (reference Jacobi-PCG, the metric), fp32ir, multigrid")
    ap.add_argument("--c3-n", type=int, default=100)
    ap.add_argument("--c3-cpu-n", type=int, default=40,
                    help="hex cells per edge of the cavity the reference build's CPU step is timed on (0: skip)")
    ap.add_argument("--c3-network", action="store_true",
                    help="multigrid pressure solve with the network's finest-level weights (default: constant "
                         "weights, faster on these tets: 92.7 against 100.1 ms per step over 40 steps)")
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--c3-pressure", default="mg", choices=["fp64", "fp32ir", "mg"])
    # the cavity's pressure solves start from the last step's pressure and take a few iterations each: one
    # smoothing sweep per side costs less there (61.1 against 66.0 ms per step over 60 steps with 3/2)
    ap.add_argument("--c3-smooth-fine", type=int, default=1)
    ap.add_argument("--c3-smooth-coarse", type=int, default=1)
    ap.add_argument("--c3-check-n", type=int, default=12)
    ap.add_argument("--c3-check-steps", type=int, default=20)
    ap.add_argument("--c3-laplacian", default="uncorrected", choices=["uncorrected", "corrected"])
    ap.add_argument("--c3-project", type=int, default=0, choices=[0, 1])
    ap.add_argument("--c3-momentum", default="fp32ir", choices=["fp32ir", "fp64"],
                    help="momentum BiCGStab: FP32 with FP64 refinement, or FP64")
    ap.add_argument("--blocks", type=int, default=64)
    ap.add_argument("--c5-passes", type=int, default=1, choices=[0, 1],
                    help="c5: also time the coupling pass over the blocks")
    ap.add_argument("--c5-cpu-sample", type=int, default=20,
                    help="c5: reference solveCG iterations sampled per block for the CPU baseline")
    ap.add_argument("--legs", default="c1,c5",
                    help="c2: also run short C1 (fp64 pressure, one timed run) and C5 (no coupling pass) legs and "
                         "report them as sub-objects of the line (empty: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is described, "
              "not measured, in this build", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c1":
        print(json.dumps(bench_c1(args)), flush=True)
        return
    if args.workload == "c5":
        print(json.dumps(bench_c5(args)), flush=True)
        return
    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "spmv":
        return bench_spmv(args)
    from chuanflow_amd import native as N
    ctx = N.Context(0)
    ctx.set_graphs(bool(args.graphs))
    arch, cus = ctx.device_info()
    log(f"device {arch}, {cus} CUs")
    hm, arrays, geom, dm, A, b, bw_shuffled = build_system(N, ctx, args.n)
    A.set_inner_tol(args.inner_tol)
    n, nnz = dm.nCells, dm.sizes()["nnz"]
    maxIter = 20000

    d_x = ctx.zeros(n)
    log("FP64 Jacobi-PCG")
    r64 = time_solve(N, ctx, A, d_x, N.FP64, args.steps, args.warmup, args.tol, maxIter)
    x64 = d_x.get()
    log("FP32-inner Jacobi-PCG + FP64 refinement")
    rmix = time_solve(N, ctx, A, d_x, N.FP32_IR, args.steps, args.warmup, args.tol, maxIter)
    xmix = d_x.get()
    log("multigrid-preconditioned flexible CG (FP32 + FP64 refinement), tuned constants")
    mg_kw = {}
    if args.mg_omega is not None:
        mg_kw.update(omegaPre=args.mg_omega, omegaPost=args.mg_omega)
    if args.mg_scale is not None:
        mg_kw.update(scale=args.mg_scale, restrictScale=args.mg_scale)
    if args.mg_coarse_omega is not None:
        mg_kw.update(coarseOmega=args.mg_coarse_omega)
    if args.mg_coarse_scale is not None:
        mg_kw.update(coarseScale=args.mg_coarse_scale)
    A.set_mg(kcycleDepth=args.kcycle_depth, fusion=args.mg_fusion, smoothFine=args.smooth_fine,
             smoothCoarse=args.smooth_coarse, **mg_kw)
    warm_mg = []
    rmg = time_solve(N, ctx, A, d_x, N.FP32_IR, args.steps, args.warmup, args.tol, maxIter, N.PRECOND_MG, warm_mg)
    xmg = d_x.get()
    log("learned: network weights on the finest level, flexible CG (FP32 + FP64 refinement)")
    from chuanflow_amd import learned as LN
    net = LN.load(ctx)
    A.set_mpnn(net)
    rln = time_solve(N, ctx, A, d_x, N.FP32_IR, args.steps, args.warmup, args.tol, maxIter, N.PRECOND_LEARNED)
    xln = d_x.get()

    # the learned solve's fine-level kernel (k_fcg_A: FP64 values on FP32 vectors, once per FCG iteration),
    # timed inside one more learned solve (not in `value`): that solve launches its kernels one by one with
    # HIP events around every k_fcg_A on the library stream, so each launch sees the caches the solve leaves
    ctx.set_fcg_probe(True)
    ctx.check(N.lib().cfx_zero(ctx.h, d_x.ptr, d_x.nbytes))
    A.solve(d_x, N.SOLVER_CG, args.tol, maxIter, 0, None, None, None, -1, True, N.PRECOND_LEARNED, N.FP32_IR)
    ctx.sync()
    ctx.set_fcg_probe(False)
    ms_fcg, n_fcg = A.fcg_probe()
    ms_up, n_up = A.smooth_probe()
    ms_fcg_loop = A.time_fcg_spmv(args.spmv_reps)
    ms_up_loop = A.time_smooth(args.spmv_reps)
    b_up = canonical_up_bytes(n, nnz)
    gbs_up = b_up / (ms_up * 1e-3) / 1e9
    log(f"k_mg_up (finest post-smoothing) in the solve {ms_up*1e3:.1f} us ({n_up} launches) = {gbs_up:.0f} GB/s "
        f"(canonical bytes); back to back {ms_up_loop*1e3:.1f} us")
    b_fcg = canonical_fcg_bytes(n, nnz)
    gbs_fcg = b_fcg / (ms_fcg * 1e-3) / 1e9
    log(f"k_fcg_A in the solve {ms_fcg*1e3:.1f} us ({n_fcg} launches) = {gbs_fcg:.0f} GB/s (canonical bytes); "
        f"back to back {ms_fcg_loop*1e3:.1f} us")
    # SpMV roofline: average launch time over reps on the library stream (HIP events)
    xs = ctx.array(np.random.default_rng(1).standard_normal(n))
    ys = ctx.empty(n)
    ms64 = A.time_spmv(0, xs, ys, args.spmv_reps, N.FP64)
    xs32 = ctx.array(np.random.default_rng(1).standard_normal(n).astype(np.float32))
    ys32 = ctx.empty(n, np.float32)
    ms32 = A.time_spmv(0, xs32, ys32, args.spmv_reps, N.FP32_IR)
    b64 = canonical_spmv_bytes(n, nnz, 8)
    b32 = canonical_spmv_bytes(n, nnz, 4)
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    log(f"SpMV FP64 {ms64*1e3:.1f} us = {gbs64:.0f} GB/s; FP32 {ms32*1e3:.1f} us = {gbs32:.0f} GB/s (canonical bytes)")
    # what plain streams reach on this device (1 GiB, best of a few grids): a read-only pass and a copy
    read_gbs, copy_gbs = ctx.stream_bandwidth(1 << 30, 20)
    # the same read stream over the SpMV's own byte count: that working set (217 MB at 128^3) fits the 256 MB
    # last-level cache, which serves part of back-to-back passes over it
    read_ws_gbs, _ = ctx.stream_bandwidth(b64 // (1 << 20) * (1 << 20), 50)
    log(f"plain streams: read {read_gbs:.0f} GB/s (1 GiB), {read_ws_gbs:.0f} GB/s ({b64 / 1e6:.0f} MB), "
        f"copy {copy_gbs:.0f} GB/s (read + written)")

    t64 = float(np.mean([w for w, _ in r64]))
    tmix = float(np.mean([w for w, _ in rmix]))
    it64 = r64[-1][1]["iterations"]
    tmg = float(np.mean([w for w, _ in rmg]))
    tln = float(np.mean([w for w, _ in rln]))
    result = {
        "metric": METRIC,
        "value": round(tln, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tln, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner + f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"pressure-Poisson {args.n}^3 jittered hex (+-15%, seed {SEED_MESH}), faces/cells shuffled "
                        f"(seed {SEED_SHUFFLE}) then RCM, Neumann, N(0,1) zero-mean rhs (seed {SEED_RHS}), "
                        f"tol {args.tol:g} inf-norm relative, x0 = 0",
            "cells": n, "nnz": nnz, "rcm_bandwidth": hm.bandwidth(), "shuffled_bandwidth": bw_shuffled,
            "parallelism": "single GPU",
        },
        "plain_fp64_pcg_ms": round(t64, 3),
        "plain_fp64_pcg_iterations": it64,
        "mixed_ms": round(tmix, 3),
        "mixed_inner_iterations": rmix[-1][1]["iterations"],
        "mixed_outer_iterations": rmix[-1][1]["outerIterations"],
        "mixed_true_residual": rmix[-1][1]["trueResidual"],
        "speedup_mixed_vs_plain": round(t64 / tmix, 3),
        "mg_fcg_ms": round(tmg, 3),
        "mg_fcg_iterations": rmg[-1][1]["iterations"],
        "mg_fcg_outer_iterations": rmg[-1][1]["outerIterations"],
        "mg_fcg_true_residual": rmg[-1][1]["trueResidual"],
        "speedup_mg_vs_plain": round(t64 / tmg, 3),
        "learned_ms": round(tln, 3),
        "learned_iterations": rln[-1][1]["iterations"],
        "learned_outer_iterations": rln[-1][1]["outerIterations"],
        "learned_true_residual": rln[-1][1]["trueResidual"],
        "learned_precond_fallbacks": rln[-1][1]["precondFallbacks"],
        "speedup_learned_vs_plain": round(t64 / tln, 3),
        # what the network buys and costs, measured: FCG device time (the solver's own events, network not
        # included) saved against the multigrid with constants, and the time outside the solver (network
        # inference and the finest-level weight setup) over the constant-weight solve's
        "learned_fcg_device_ms": round(float(np.mean([r["ms"] for _, r in rln])), 3),
        "mg_fcg_device_ms": round(float(np.mean([r["ms"] for _, r in rmg])), 3),
        "learned_saved_fcg_ms": round(float(np.mean([r["ms"] for _, r in rmg]) - np.mean([r["ms"] for _, r in rln])),
                                      3),
        "learned_inference_ms": round(float(np.mean([w - r["ms"] for w, r in rln]) -
                                            np.mean([w - r["ms"] for w, r in rmg])), 3),
        "learned_preconditioner": "3-layer MPNN (width 32, FP16 MFMA) giving the finest-level smoother weight and "
                                  "transfer scale of an aggregation K-cycle inside flexible CG; weights "
                                  "chuanflow_amd/weights/mpnn_pressure.npy trained on the GPU on pressure "
                                  "systems of obstacle channels (33 SIMPLE systems at 64x32x32, then 2,000 PISO "
                                  "systems streamed at 128^3 with validated updates)",
        "spmv_fp64_us": round(ms64 * 1e3, 2),
        "spmv_fp64_GBps": round(gbs64, 1),
        "spmv_fp32_us": round(ms32 * 1e3, 2),
        "spmv_fp32_GBps": round(gbs32, 1),
        "device_stream_read_GBps": round(read_gbs, 1),
        "device_stream_copy_GBps": round(copy_gbs, 1),
        "spmv_fp64_vs_stream_read": round(gbs64 / read_gbs, 3),
        "device_stream_read_spmv_bytes_GBps": round(read_ws_gbs, 1),
        "spmv_fp64_vs_stream_read_same_bytes": round(gbs64 / read_ws_gbs, 3),
        # the solve's most-launched fine-level kernel: the post-smoothing sweep runs smoothFine times per FCG
        # iteration (the pre-smoothing as many, k_fcg_A once)
        "roofline": {"bound": "hbm", "achieved": round(gbs_up, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbs_up / HBM_PEAK_GBPS, 4),
                     "traffic": MG_UP_TRAFFIC_128 if args.n == 128 else None,
                     "kernel": "k_mg_up (the learned solve's finest-level post-smoothing sweep z = x + W (b - A x) on "
                               "the values in 64-row slices: FP32 values, FP32 vectors; smoothFine launches per FCG "
                               "iteration)",
                     "bytes_per_launch": b_up, "us_per_launch": round(ms_up * 1e3, 2),
                     "launches_timed": n_up, "us_back_to_back": round(ms_up_loop * 1e3, 2)},
        "fcg_a_roofline": {"achieved": round(gbs_fcg, 1), "frac": round(gbs_fcg / HBM_PEAK_GBPS, 4),
                           "traffic": FCG_A_TRAFFIC_128 if args.n == 128 else None,
                           "kernel": "k_fcg_A (the learned solve's fine-level q = A p: FP64 values, FP32 vectors, "
                                     "p.q and p.r partials; once per FCG iteration)",
                           "bytes_per_launch": b_fcg, "us_per_launch": round(ms_fcg * 1e3, 2),
                           "launches_timed": n_fcg, "us_back_to_back": round(ms_fcg_loop * 1e3, 2)},
        "spmv_fp64_roofline": {"achieved": round(gbs64, 1), "frac": round(gbs64 / HBM_PEAK_GBPS, 4),
                               "traffic": SPMV_FP64_TRAFFIC_128 if args.n == 128 else None,
                               "kernel": "k_spmv<double> (CSR-stream SpMV, FP64; the plain PCG's product)",
                               "bytes_per_launch": b64},
        # one-time cost of a new pattern, not in `value`: the first multigrid solve builds the aggregation
        # hierarchy on the host and captures the FCG iteration as a graph (later solves on the pattern reuse both)
        "setup_ms": round(warm_mg[0] - tmg, 1) if warm_mg else None,
        "first_solve_ms": round(warm_mg[0], 1) if warm_mg else None,
        # the fine-level matrix passes of every learned FCG iteration (k_fcg_A, the pre- and post-smoothing
        # sweeps) at their canonical bytes, over the whole learned solve's wall time: a floor on what it streams
        "learned_fine_level_GBps": round(
            (rln[-1][1]["iterations"] + rln[-1][1]["outerIterations"]) * fine_level_bytes(n, nnz, args.smooth_fine)
            / (tln * 1e-3) / 1e9,
            1),
    }
    if not args.no_cpu:
        chk = cpu_baseline_and_check(arrays, b, x64, xmix, it64, args.cpu_sample_iters, A, xmg, xln)
        result["cpu_baseline"] = chk.pop("cpu_baseline")
        result["checks"] = chk
    result["legs"] = run_legs(args)
    print(json.dumps(result), flush=True)


def run_legs(args):
    """Short runs of the other configurations, reported beside the headline: C1 (200 SIMPLE iterations, the
    reference's Jacobi-PCG pressure solve, one warm-up and one timed run) and C5 (64 batched detail regions,
    without the coupling-pass timing), each with its oracle check and its own CPU baseline sample."""
    import copy
    legs = {}
    names = [x for x in args.legs.split(",") if x]
    if "c1" in names:
        a = copy.copy(args)
        a.c1_variants, a.steps, a.warmup, a.c1_cpu_sample = "fp64", 1, 1, 3
        log("leg C1")
        r = bench_c1(a)
        legs["c1"] = {k: r.get(k) for k in ("metric", "value", "unit", "ms_per_iteration", "device_ms",
                                            "final_continuity", "parity_vs_oracle", "cpu_baseline", "config")}
    if "c5" in names:
        a = copy.copy(args)
        a.c5_passes, a.steps, a.warmup, a.c5_cpu_sample = 0, 1, 1, 5
        log("leg C5")
        r = bench_c5(a)
        legs["c5"] = {k: r.get(k) for k in ("metric", "value", "unit", "one_by_one_ms", "batched_default_ms",
                                            "batched_persistent_ms", "speedup_batched_vs_one_by_one", "schedule",
                                            "max_iterations", "all_converged", "default_schedule_bitwise_equal",
                                            "cpu_baseline", "config")}
        fin = [b["batched_finish_ms"] for b in r["per_block"]]
        legs["c5"]["block_finish_ms"] = {"first": min(fin), "last": max(fin)}
    return legs


if __name__ == "__main__":
    main()
