"""Krylov solvers on the GPU vs the oracle (solveCG / solveBiCGStab and the
fvMatrix pass loop), the FP32-inner / FP64-refined PCG, and the
reference's failure behaviour (no exceptions for non-convergence).

Tolerances: the SpMV and every vector update are the reference's
arithmetic; only the dot products are summed in a different (tree) order,
so iterates drift by rounding.  Iteration counts agree within +-2 at
tol 1e-8, and solutions agree to 1e-10 relative (max norm) when both solve
to 1e-13."""
import numpy as np
import pytest

import oracle as O
from channel import channel_bcs, neumann_bcs
from helpers import host_and_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pmesh(ctx):
    from chuanflow_amd import native as N
    print("\n[case] 16x12x10 jittered hex, shuffled + RCM", flush=True)
    hm, a, om = host_and_oracle(16, 12, 10, (0, 0, 0), (1, 0.75, 0.6), jitter=0.15, seed=21, shuffle=3, rcm=True)
    return N, ctx, om, N.Mesh.from_host(ctx, hm)


def _pressure_system(om, neumann, rng):
    if neumann:
        t, v = neumann_bcs(om.nB)
    else:
        _, _, t, v, *_ = channel_bcs(om)
    lc, lk = O.linearise(om, 1, t, v)
    A = O.Ldu(om)
    gf = rng.uniform(0.5, 2.0, om.nFaces)
    O.assemble_laplacian(om, A, gf, lc, lk, -1)
    csr = O.csr_from_ldu(om, A)
    b = rng.standard_normal(om.nCells)
    if neumann:
        b -= b.mean()
    return csr, b


@pytest.mark.parametrize("neumann", [False, True])
def test_cg_matches_oracle(pmesh, neumann):
    N, ctx, om, _ = pmesh
    rng = np.random.default_rng(5)
    csr, b = _pressure_system(om, neumann, rng)
    G = N.Csr(ctx, csr.rowStart, csr.cols, csr.vals)
    # SpMV bit-exact
    x = rng.standard_normal(om.nCells)
    assert np.array_equal(G.spmv(ctx.array(x)).get(), O.spmv(csr, x))
    xo, ro, ho = O.solve_cg(csr, b, None, 1e-8, 2000, hist=True)
    xg, rg = G.solve(b, None, 1e-8, 2000)
    assert rg.converged and ro.converged
    assert abs(rg.iterations - ro.iterations) <= 2
    hg = G.residual_history(rg.iterations + 1)
    k = min(20, ro.iterations)
    np.testing.assert_allclose(hg[:k], ho[:k], rtol=1e-6)
    assert hg[0] == ho[0]  # initial residual: bit-exact SpMV and max
    xo, ro = O.solve_cg(csr, b, None, 1e-13, 5000)
    xg, rg = G.solve(b, None, 1e-13, 5000)
    if neumann:  # compare modulo the null space (constants)
        xo, xg = xo - xo.mean(), xg - xg.mean()
    err = np.abs(xg - xo).max() / np.abs(xo).max()
    print(f"CG at tol 1e-13 vs oracle: rel diff {err:.2e}", flush=True)
    assert err <= 1e-10


def test_cg_warm_start_and_initial_convergence(pmesh):
    N, ctx, om, _ = pmesh
    rng = np.random.default_rng(6)
    csr, b = _pressure_system(om, False, rng)
    G = N.Csr(ctx, csr.rowStart, csr.cols, csr.vals)
    xo, _ = O.solve_cg(csr, b, None, 1e-12, 5000)
    x, rep = G.solve(b, xo, 1e-8, 100)  # already converged: 0 iterations
    assert rep.converged and rep.iterations == 0 and np.array_equal(x, xo)


def test_bicgstab_matches_oracle(pmesh):
    N, ctx, om, dm = pmesh
    rng = np.random.default_rng(7)
    Ut, Uv, *_ = channel_bcs(om)
    lc, lk = O.linearise(om, 3, Ut, Uv)
    A = O.Ldu(om, 3)
    O.assemble_ddt(om, A, rng.standard_normal((om.nCells, 3)), 5e-3)
    O.assemble_div(om, A, rng.standard_normal(om.nFaces) * 1e-3, lc, lk)
    O.assemble_laplacian(om, A, O.harmonic_gamma_uniform(om, 1e-3), lc, lk, -1)
    csr = O.csr_from_ldu(om, A, 0)
    b = A.rhs[0]
    G = N.Csr(ctx, csr.rowStart, csr.cols, csr.vals)
    xo, ro = O.solve_bicgstab(csr, b, None, 1e-8, 2000)
    xg, rg = G.solve(b, None, 1e-8, 2000, solver=N.SOLVER_BICGSTAB)
    assert ro.converged and rg.converged and abs(rg.iterations - ro.iterations) <= 2
    xo, ro = O.solve_bicgstab(csr, b, None, 1e-13, 5000)
    xg, rg = G.solve(b, None, 1e-13, 5000, solver=N.SOLVER_BICGSTAB)
    err = np.abs(xg - xo).max() / np.abs(xo).max()
    print(f"BiCGStab at tol 1e-13 vs oracle: rel diff {err:.2e}", flush=True)
    assert err <= 1e-10


def test_bicgstab_fp32_with_fp64_refinement(pmesh):
    """BiCGStab in FP32 with FP64 iterative refinement on a momentum system
    (ddt + upwind convection + laplacian, non-symmetric): the true FP64
    residual meets the tolerance, and the solution matches the oracle's FP64
    BiCGStab to the tolerance's scale."""
    N, ctx, om, dm = pmesh
    rng = np.random.default_rng(17)
    Ut, Uv, *_ = channel_bcs(om)
    lc, lk = O.linearise(om, 3, Ut, Uv)
    A = O.Ldu(om, 3)
    O.assemble_ddt(om, A, rng.standard_normal((om.nCells, 3)), 5e-3)
    O.assemble_div(om, A, rng.standard_normal(om.nFaces) * 1e-2, lc, lk)
    O.assemble_laplacian(om, A, O.harmonic_gamma_uniform(om, 1e-3), lc, lk, -1)
    csr = O.csr_from_ldu(om, A, 0)
    b = A.rhs[0]
    G = N.Csr(ctx, csr.rowStart, csr.cols, csr.vals)
    for tol in (1e-8, 1e-12):
        xg, rg = G.solve(b, None, tol, 2000, solver=N.SOLVER_BICGSTAB, precision=N.FP32_IR)
        true = np.abs(b - O.spmv(csr, xg)).max() / np.abs(b).max()
        print(f"tol {tol:g}: {rg.iterations} FP32 iterations in {rg.outerIterations} corrections, "
              f"true residual {true:.2e}, fallbacks {rg.precondFallbacks}", flush=True)
        assert rg.converged and true < tol
        assert rg.outerIterations >= 1
        xo, ro = O.solve_bicgstab(csr, b, None, 1e-14, 5000)
        assert np.abs(xg - xo).max() <= 100 * tol * np.abs(xo).max()


def test_fvmatrix_pass_loop_matches_oracle(pmesh):
    """fvMatrix::solve with nonOrth=2: 3 passes, rhs restored and corrected
    from the current solution each pass (fvMatrix.cpp:86-121)."""
    N, ctx, om, dm = pmesh
    rng = np.random.default_rng(8)
    _, _, pt, pv, *_ = channel_bcs(om)
    lc, lk = O.linearise(om, 1, pt, pv)
    gf = rng.uniform(0.5, 2.0, om.nFaces)
    src = rng.standard_normal(om.nCells)
    A = O.Ldu(om)
    O.assemble_laplacian(om, A, gf, lc, lk, -1)
    A.rhs[0] += src
    csr = O.csr_from_ldu(om, A)
    x = np.zeros(om.nCells)
    for _ in range(3):
        rhs = A.rhs[0].copy()
        O.nonorth_correct(om, gf, x, lc, lk, rhs=rhs)
        x, ro = O.solve_cg(csr, rhs, x, 1e-13, 5000)
    M = N.Matrix(dm, 1)
    d_gf, d_lc, d_lk = ctx.array(gf), ctx.array(lc), ctx.array(lk)
    M.laplacian(d_gf, d_lc, d_lk, -1)
    M.add_rhs(0, ctx.array(src))
    dx = ctx.zeros(om.nCells)
    rep = M.solve(dx, N.SOLVER_CG, 1e-13, 5000, 2, d_gf, d_lc, d_lk, -1, True)
    err = np.abs(dx.get() - x).max() / np.abs(x).max()
    print(f"fvMatrix 3 passes at tol 1e-13 vs oracle: rel diff {err:.2e}", flush=True)
    assert rep.converged and err <= 1e-10
    # rhs_ after the solve holds the last pass' corrections
    _, r, _, _ = M.get()
    np.testing.assert_allclose(r[0], rhs, rtol=0, atol=1e-9 * np.abs(rhs).max())


@pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not built")
def test_fvmatrix_relaxation_matches_reference_build(pmesh):
    """Schemes::relaxation (fvMatrix.cpp:110-118): after the final of three
    passes x = xOld + relaxation (x - xOld), xOld the solution before that
    pass.  Against the reference's own fvMatrix::solve (compiled in
    oracle/_ref) with relaxation 0.7 on the jittered mesh, and, for one pass
    from a nonzero guess, exactly the reference's expression applied to the
    unrelaxed solve."""
    N, ctx, om, dm = pmesh
    rng = np.random.default_rng(18)
    _, _, pt, pv, _, _, ppt, ppv = channel_bcs(om)  # per boundary face (product), per patch (reference)
    lc, lk = O.linearise(om, 1, pt, pv)
    gf = rng.uniform(0.5, 2.0, om.nFaces)
    src = rng.standard_normal(om.nCells)
    ref = O.RefMesh(om).laplacian_face(gf, ppt, ppv, 1, 1, -1, src, 1e-13, 5000, 2, np.zeros(om.nCells),
                                       np.zeros(om.nCells), relaxation=0.7)
    M = N.Matrix(dm, 1)
    d_gf, d_lc, d_lk = ctx.array(gf), ctx.array(lc), ctx.array(lk)
    M.laplacian(d_gf, d_lc, d_lk, -1)
    M.add_rhs(0, ctx.array(src))
    M.set_relaxation(0.7)
    dx = ctx.zeros(om.nCells)
    rep = M.solve(dx, N.SOLVER_CG, 1e-13, 5000, 2, d_gf, d_lc, d_lk, -1, True)
    err = np.abs(dx.get() - ref["x"]).max() / np.abs(ref["x"]).max()
    print(f"relaxed 3-pass solve vs reference build: rel diff {err:.2e}", flush=True)
    assert rep.converged and ref["converged"] and err <= 1e-10
    # one pass from x0: the relaxed result is the reference's expression on the unrelaxed one, bit for bit
    x0 = rng.standard_normal(om.nCells)
    M.set_relaxation(1.0)
    d1 = ctx.array(x0)
    M.solve(d1, N.SOLVER_CG, 1e-12, 5000, 0)
    M.set_relaxation(0.7)
    d2 = ctx.array(x0)
    M.solve(d2, N.SOLVER_CG, 1e-12, 5000, 0)
    x1 = d1.get()
    assert np.array_equal(d2.get(), x0 + 0.7 * (x1 - x0))
    with pytest.raises(RuntimeError, match="relaxation"):
        M.set_relaxation(1.5)


@pytest.mark.parametrize("neumann", [False, True])
def test_mixed_precision_reaches_fp64_tolerance(pmesh, neumann):
    N, ctx, om, _ = pmesh
    rng = np.random.default_rng(9)
    csr, b = _pressure_system(om, neumann, rng)
    G = N.Csr(ctx, csr.rowStart, csr.cols, csr.vals)
    xg, rg = G.solve(b, None, 1e-8, 5000, precision=N.FP32_IR)
    assert rg.converged and rg.outerIterations >= 2
    r = b - O.spmv(csr, xg)  # true FP64 residual on the CPU
    assert np.abs(r).max() / np.abs(b).max() < 1e-8
    assert abs(rg.trueResidual - np.abs(r).max() / np.abs(b).max()) < 1e-10
    xo, _ = O.solve_cg(csr, b, None, 1e-13, 5000)
    if neumann:
        xo, xg = xo - xo.mean(), xg - xg.mean()
    err = np.abs(xg - xo).max() / np.abs(xo).max()
    print(f"mixed-precision error vs FP64 oracle: {err:.2e}")
    assert err < 1e-5


def test_failure_modes_report_not_throw(ctx):
    from chuanflow_amd import native as N
    # indefinite matrix: pAp <= 0 -> breakdown, not converged (Matrix.cpp:191-194)
    rs = np.array([0, 2, 4], np.int64)
    cols = np.array([0, 1, 0, 1], np.int64)
    vals = np.array([1.0, 0.0, 0.0, -1.0])
    G = N.Csr(ctx, rs, cols, vals)
    x, rep = G.solve(np.array([0.0, 1.0]), None, 1e-8, 10)
    assert not rep.converged and rep.breakdown == 1
    # maxIter too small: not converged, no error
    n = 50
    rs = np.arange(0, 3 * n - 1, 3)
    A = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    rs = np.concatenate([[0], np.cumsum((A != 0).sum(axis=1))]).astype(np.int64)
    cols = np.nonzero(A)[1].astype(np.int64)
    G = N.Csr(ctx, rs, cols, A[A != 0])
    x, rep = G.solve(A @ np.ones(n), None, 1e-10, 1)
    assert rep.iterations == 1 and not rep.converged
    # NaN right-hand side: the residual norm is NaN, never converges
    b = A @ np.ones(n)
    b[3] = np.nan
    x, rep = G.solve(b, None, 1e-10, 5)
    assert not rep.converged and np.isnan(rep.initialResidual)
    # tridiagonal Poisson: x = 1 (solverBackendApp.cpp:89-109)
    x, rep = G.solve(A @ np.ones(n), None, 1e-12, 500)
    assert rep.converged and np.abs(x - 1).max() < 1e-8


def test_stream_bandwidth_yardstick(ctx):
    """cfx_stream_bandwidth: plain read and copy streams land between a
    fraction of the HBM peak and the peak itself (8 TB/s), and a read-only
    stream is not slower than a copy's reads and writes together by more than
    timing noise."""
    print("\n[case] stream bandwidth, 256 MiB", flush=True)
    rd, cp = ctx.stream_bandwidth(256 << 20, 5)
    print(f"  read {rd:.0f} GB/s, copy {cp:.0f} GB/s", flush=True)
    assert 1000 < rd < 8000 * 1.02 and 1000 < cp < 8000 * 1.02
    assert rd > 0.8 * cp
