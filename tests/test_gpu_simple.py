"""The SIMPLE and PISO drivers on the GPU vs the oracle (which is
bit-identical to the reference build, see test_oracle.py).

Only the Krylov dot products are summed in another order, so with tight
linear tolerances the fields track the oracle to ~1e-9; on the C1 channel
(40x20x20, 200 iterations, the reference's converging controls U 0.9 /
p 0.03 / pRelaxStart 0.02, linear tol 1e-8) they must agree to 1e-6
relative."""
import numpy as np
import pytest

import oracle as O
from channel import channel_bcs
from helpers import aos, host_and_oracle, oracle_mesh_from_arrays, soa

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup_case(ctx, nx, ny, nz, mx, jitter, seed, Uin, rcm=True):
    from chuanflow_amd import native as N
    hm, a, om = host_and_oracle(nx, ny, nz, (0, 0, 0), mx, jitter=jitter, seed=seed, rcm=rcm)
    dm = N.Mesh.from_host(ctx, hm)
    Ut, Uv, pt, pv, *_ = channel_bcs(om, Uin)
    return N, om, dm, (Ut, Uv), (pt, pv)


def initial_fields(om, Ubc):
    U = np.zeros((om.nCells, 3))
    p = np.zeros(om.nCells)
    phi = np.zeros(om.nFaces)
    lc, lk = O.linearise(om, 3, *Ubc)
    O.correct_boundary_flux(om, U, lc, lk, phi)
    return U, p, phi


def test_linearise_bitwise(ctx):
    N, om, dm, Ubc, pbc = setup_case(ctx, 6, 5, 4, (1, 0.5, 0.5), 0.15, 2, 0.02)
    rng = np.random.default_rng(3)
    t = rng.integers(0, 4, om.nB).astype(np.int32)
    v = rng.standard_normal((om.nB, 3))
    for nc in (1, 3):
        lc, lk = dm.linearise(nc, t, v[:, :nc])
        rc, rk = O.linearise(om, nc, t, v[:, :nc].copy())
        assert np.array_equal(lc.get(), rc) and np.array_equal(lk.get(), rk), nc


@pytest.mark.parametrize("relax", [(0.9, 0.03, 0.02), (0.7, 0.3, 0.0)])
def test_simple_tracks_oracle_tight(ctx, relax):
    N, om, dm, Ubc, pbc = setup_case(ctx, 10, 8, 6, (1, 0.5, 0.5), 0.15, 5, 0.01)
    print(f"\n[case] SIMPLE 10x8x6 jittered, relax {relax}", flush=True)
    U, p, phi = initial_fields(om, Ubc)
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    iters = 10
    kw = dict(maxIterations=iters, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-12, linearMaxIter=5000,
              nonOrthCorrectors=2, relaxU=relax[0], relaxP=relax[1], pRelaxStart=relax[2])
    r, hg = dm.simple(1e-3, N.Schemes(0, 1, 1, 0.9), N.SimpleControls(**kw), Ubc, pbc, dU, dp, dphi, hist=True)
    ro, ho = O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(**kw), Ubc, pbc, hist=True)
    assert r.iterations == ro.iterations == iters and not r.diverged
    eU, ep, ephi = rel(aos(dU.get()), U), rel(dp.get(), p), rel(dphi.get(), phi)
    print(f"SIMPLE {relax}: rel diff U {eU:.2e} p {ep:.2e} phi {ephi:.2e}", flush=True)
    assert eU < 1e-10 and ep < 1e-10 and ephi < 1e-10
    np.testing.assert_allclose(hg, ho, rtol=1e-6, atol=1e-14)


def test_simple_with_pressure_reference_adjusts_outflow(ctx):
    """A channel whose pressure is zeroGradient on every patch (no pressure
    reference): SIMPLE scales the outlet flux so that the net boundary flux
    vanishes (adjustPhi, simpleSolver.cpp:211-247; the patch sums on the
    device in a fixed tree order) and solves the singular pressure system;
    10 iterations at tight solver tolerances agree with the oracle.  The
    laplacian is "uncorrected": with the corrected one the explicit
    non-orthogonal boundary terms make the pure Neumann pressure system
    inconsistent on a jittered mesh, and the reference diverges (the oracle
    with it)."""
    N, om, dm, Ubc, pbc = setup_case(ctx, 10, 8, 6, (1, 0.5, 0.5), 0.15, 5, 0.01)
    print("\n[case] SIMPLE 10x8x6 jittered, pressure zeroGradient everywhere", flush=True)
    pbc = (np.full_like(pbc[0], 2), np.zeros_like(pbc[1]))
    U, p, phi = initial_fields(om, Ubc)
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    kw = dict(maxIterations=10, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-12, linearMaxIter=5000,
              nonOrthCorrectors=0, relaxU=0.7, relaxP=0.3, pRelaxStart=0.0)
    r = dm.simple(1e-3, N.Schemes(0, 0, 1, 0.9), N.SimpleControls(**kw), Ubc, pbc, dU, dp, dphi)
    ro = O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 0, 1, 0.9), O.SimpleControls(**kw), Ubc, pbc)
    assert not r.diverged and not ro.diverged and np.all(np.isfinite(dU.get()))
    pg = dp.get()
    eU, ep, ephi = rel(aos(dU.get()), U), rel(pg - pg.mean(), p - p.mean()), rel(dphi.get(), phi)
    print(f"adjusted-flux SIMPLE: rel diff U {eU:.2e} p {ep:.2e} phi {ephi:.2e}", flush=True)
    assert eU < 1e-10 and ep < 1e-10 and ephi < 1e-10
    # the outlet flux balances the inlet's after the adjustment
    nI = om.nInternal
    assert abs(dphi.get()[nI:].sum()) <= 1e-10 * np.abs(dphi.get()[nI:]).sum()


def c1_case(ctx):
    return setup_case(ctx, 40, 20, 20, (1, 0.5, 0.5), 0.0, 0, 0.01, rcm=False)


@pytest.mark.parametrize("pprec", [0, 1])
def test_c1_channel_200_iterations_within_1e6_of_oracle(ctx, pprec):
    N, om, dm, Ubc, pbc = c1_case(ctx)
    print(f"\n[case] C1 40x20x20 channel, 200 SIMPLE iterations, pressure precision {pprec}", flush=True)
    U, p, phi = initial_fields(om, Ubc)
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    kw = dict(maxIterations=200, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-8, linearMaxIter=2000,
              nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02)
    r = dm.simple(1e-3, N.Schemes(0, 1, 1, 0.9), N.SimpleControls(**kw, pPrecision=pprec), Ubc, pbc, dU, dp, dphi)
    ro = O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(**kw), Ubc, pbc)
    print(f"GPU {r.ms:.0f} ms for {r.iterations} iterations; final continuity {r.finalContinuityError:.3e} "
          f"(oracle {ro.finalContinuityError:.3e})")
    assert r.iterations == ro.iterations == 200 and not r.diverged
    for a, b, name in ((aos(dU.get()), U, "U"), (dp.get(), p, "p"), (dphi.get(), phi, "phi")):
        e = rel(a, b)
        print(f"  {name}: max rel diff {e:.2e}")
        assert e < 1e-6, name


def test_c1_channel_fp32_momentum_and_pressure_within_1e6_of_oracle(ctx):
    """C1 with both linear solves in FP32 and FP64 refinement (momentum
    BiCGStab and pressure PCG decide convergence on true FP64 residuals):
    after 200 SIMPLE iterations the fields stay within 1e-6 of the FP64
    oracle."""
    N, om, dm, Ubc, pbc = c1_case(ctx)
    print("\n[case] C1 40x20x20 channel, 200 SIMPLE iterations, FP32 momentum and pressure", flush=True)
    U, p, phi = initial_fields(om, Ubc)
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    kw = dict(maxIterations=200, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-8, linearMaxIter=2000,
              nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02)
    r = dm.simple(1e-3, N.Schemes(0, 1, 1, 0.9), N.SimpleControls(**kw, pPrecision=1, uPrecision=1), Ubc, pbc,
                  dU, dp, dphi)
    ro = O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(**kw), Ubc, pbc)
    print(f"GPU {r.ms:.0f} ms for {r.iterations} iterations")
    assert r.iterations == ro.iterations == 200 and not r.diverged
    for a, b, name in ((aos(dU.get()), U, "U"), (dp.get(), p, "p"), (dphi.get(), phi, "phi")):
        e = rel(a, b)
        print(f"  {name}: max rel diff {e:.2e}")
        assert e < 1e-6, name


def test_c1_literal_controls_reproduce_reference_divergence(ctx):
    """U 0.7 / p 0.3 without pRelaxStart diverges on the reference (iterations
    4-8); the GPU run must diverge as well, within two iterations of it."""
    N, om, dm, Ubc, pbc = c1_case(ctx)
    U, p, phi = initial_fields(om, Ubc)
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    kw = dict(maxIterations=30, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-8, linearMaxIter=2000,
              nonOrthCorrectors=2, relaxU=0.7, relaxP=0.3, pRelaxStart=0.0)
    r = dm.simple(1e-3, N.Schemes(0, 1, 1, 0.9), N.SimpleControls(**kw), Ubc, pbc, dU, dp, dphi)
    ro = O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(**kw), Ubc, pbc)
    print(f"GPU diverged at {r.iterations}, oracle at {ro.iterations}")
    assert ro.diverged and r.diverged and abs(r.iterations - ro.iterations) <= 2


def test_piso_tracks_oracle(ctx):
    N, om, dm, Ubc, pbc = setup_case(ctx, 10, 8, 6, (1, 0.5, 0.5), 0.12, 9, 0.02)
    rng = np.random.default_rng(4)
    U = rng.standard_normal((om.nCells, 3)) * 0.01
    p = np.zeros(om.nCells)
    phi = np.zeros(om.nFaces)
    lc, lk = O.linearise(om, 3, *Ubc)
    O.update_phi_internal(om, U, phi)
    O.correct_boundary_flux(om, U, lc, lk, phi)
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    kw = dict(dt=0.01, correctors=2, continuityTol=1e-10, linearTol=1e-12, linearMaxIter=5000, nonOrthCorrectors=2)
    for step in range(3):
        r = dm.piso_step(1e-3, N.Schemes(0, 1, 1, 0.9), N.PisoControls(**kw), Ubc, pbc, dU, dp, dphi)
        ro = O.piso_step(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.PisoControls(**kw), Ubc, pbc)
        assert r.correctorsUsed == ro.correctorsUsed
        eU, ep, ephi = rel(aos(dU.get()), U), rel(dp.get(), p), rel(dphi.get(), phi)
        print(f"PISO step {step}: rel diff U {eU:.2e} p {ep:.2e} phi {ephi:.2e}", flush=True)
        assert eU < 1e-10 and ep < 1e-10 and ephi < 1e-10, step


def test_piso_inviscid_body_force_closed_form(ctx):
    """pisoTests.cpp:97-119: on the 0.2 x 0.1 x 0.05 channel (2 x 10 x 1
    cells, walls fixedValue 0 at y, zeroGradient elsewhere, p gauge 0 at
    xmin) with nu = 1e-8 and a body force f = 0.1, one step of dt = 0.002
    gives u = f dt everywhere (1e-9), leaves p at the gauge (1e-9), and the
    predicted flux is already divergence free (one corrector, < 1e-14)."""
    from chuanflow_amd import native as N
    hm, a, om = host_and_oracle(2, 10, 1, (0, 0, 0), (0.2, 0.1, 0.05))
    dm = N.Mesh.from_host(ctx, hm)
    pid = om.patch_of_boundary_face()
    Ut = np.where((pid == 2) | (pid == 3), 0, 2).astype(np.int32)
    pt = np.where(pid == 0, 0, 2).astype(np.int32)
    Ubc, pbc = (Ut, np.zeros((om.nB, 3))), (pt, np.zeros((om.nB, 1)))
    dU, dp, dphi = ctx.zeros((3, om.nCells)), ctx.zeros(om.nCells), ctx.zeros(om.nFaces)
    f, dt = 0.1, 0.002
    ctl = N.PisoControls(dt=dt, correctors=2, continuityTol=1e-10, linearTol=1e-8, linearMaxIter=2000,
                         nonOrthCorrectors=2, bodyForce=(f, 0, 0))
    r = dm.piso_step(1e-8, N.Schemes(0, 1, 1, 0.9), ctl, Ubc, pbc, dU, dp, dphi)
    U = aos(dU.get())
    assert np.abs(U[:, 0] - f * dt).max() < 1e-9
    assert np.abs(dp.get()).max() < 1e-9
    assert r.continuityError < 1e-14 and r.correctorsUsed == 1
    # and the oracle agrees
    Uo, po, phio = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    ro = O.piso_step(om, Uo, po, phio, 1e-8, O.Schemes(0, 1, 1, 0.9),
                     O.PisoControls(dt=dt, correctors=2, continuityTol=1e-10, linearTol=1e-8, linearMaxIter=2000,
                                    nonOrthCorrectors=2, bodyForce=(f, 0, 0)), Ubc, pbc)
    assert ro.correctorsUsed == 1 and rel(U, Uo) < 1e-12


def test_piso_closed_cavity_on_tets_with_neumann_projection(ctx):
    """A closed lid-driven cavity on Kuhn tets (+-10% jitter): the pressure is a
    singular Neumann system; with its right-hand side projected to zero sum
    (not in the reference, whose PISO diverges here) GPU and oracle agree."""
    from chuanflow_amd import native as N
    import bench
    print("\n[case] cavity 6x6x6 hex -> 1296 tets, PISO with Neumann projection", flush=True)
    hm = N.HostMesh.box(6, 6, 6, (0, 0, 0), (1, 1, 1), 0.10, 3, True)
    hm.renumber_rcm()
    a, g = hm.arrays(), hm.geometry()
    om = oracle_mesh_from_arrays(a)
    dm = N.Mesh(ctx, a, g)
    Ubc, pbc = bench.cavity_bcs(a["patchStart"], a["patchSize"], len(a["neighbour"]), om.nB)
    U, p, phi = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    dU, dp, dphi = ctx.zeros((3, om.nCells)), ctx.zeros(om.nCells), ctx.zeros(om.nFaces)
    kw = dict(dt=2e-3, correctors=2, continuityTol=1e-10, linearTol=1e-12, linearMaxIter=5000, nonOrthCorrectors=1)
    for step in range(3):
        r = dm.piso_step(1e-3, N.Schemes(0, 1, 1, 0.9), N.PisoControls(**kw, projectNeumann=1), Ubc, pbc, dU, dp,
                         dphi)
        O.piso_step(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.PisoControls(**kw, projectNeumann=1), Ubc, pbc)
        Ug = aos(dU.get())
        assert np.all(np.isfinite(Ug)) and rel(Ug, U) < 1e-8, step
        pg = dp.get()
        assert rel(pg - pg.mean(), p - p.mean()) < 1e-8, step


def test_piso_closed_cavity_on_tets_uncorrected_laplacian(ctx):
    """The cavity as the tet benchmark runs it: laplacian "uncorrected", no
    right-hand-side projection (the reference algorithm unchanged); 5 steps on
    the GPU agree with the oracle and continuity closes to the solver
    tolerance after every step."""
    from chuanflow_amd import native as N
    import bench
    print("\n[case] cavity 6x6x6 hex -> 1296 tets, PISO, uncorrected laplacian", flush=True)
    hm = N.HostMesh.box(6, 6, 6, (0, 0, 0), (1, 1, 1), 0.10, 3, True)
    hm.renumber_rcm()
    a, g = hm.arrays(), hm.geometry()
    om = oracle_mesh_from_arrays(a)
    dm = N.Mesh(ctx, a, g)
    Ubc, pbc = bench.cavity_bcs(a["patchStart"], a["patchSize"], len(a["neighbour"]), om.nB)
    U, p, phi = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    dU, dp, dphi = ctx.zeros((3, om.nCells)), ctx.zeros(om.nCells), ctx.zeros(om.nFaces)
    kw = dict(dt=2e-3, correctors=2, continuityTol=1e-10, linearTol=1e-12, linearMaxIter=5000, nonOrthCorrectors=1)
    for step in range(5):
        r = dm.piso_step(1e-3, N.Schemes(0, 0, 1, 0.9), N.PisoControls(**kw), Ubc, pbc, dU, dp, dphi)
        ro = O.piso_step(om, U, p, phi, 1e-3, O.Schemes(0, 0, 1, 0.9), O.PisoControls(**kw), Ubc, pbc)
        Ug = aos(dU.get())
        assert np.all(np.isfinite(Ug)) and rel(Ug, U) < 1e-8, step
        pg = dp.get()
        assert rel(pg - pg.mean(), p - p.mean()) < 1e-8, step
        assert r.continuityError < 1e-9 and ro.continuityError < 1e-9, (step, r.continuityError, ro.continuityError)


def test_simple_on_gmsh_tet_channel_matches_oracle(ctx, tmp_path):
    """The imported-mesh path end to end: a Gmsh tet channel (the tests'
    writer, 3,072 jittered Kuhn tets) read by the library's reader,
    RCM-renumbered, uploaded, and 15 SIMPLE iterations on the GPU against the
    oracle's SIMPLE on the same arrays (inlet/outlet/walls from the file's
    physical surfaces; laplacian uncorrected, as tets need; the coupled
    cases' relaxation, which converges here where 0.7 / 0.3 diverges in the
    oracle too)."""
    from chuanflow_amd import native as N
    from helpers import write_gmsh_tet_channel
    print("\n[case] SIMPLE on a Gmsh tet channel 16x4x8 (3072 tets), 15 iterations", flush=True)
    path = str(tmp_path / "channel.msh")
    write_gmsh_tet_channel(path, 16, 4, 8, L=(1.0, 0.25, 0.25), jitter=0.1, seed=11)
    hm = N.HostMesh.read_gmsh(path, {"inlet": "inlet", "outlet": "outlet"})
    hm.renumber_rcm()
    a = hm.arrays()
    om = oracle_mesh_from_arrays(a)
    dm = N.Mesh.from_host(ctx, hm)
    names = [n for n, _ in hm.patches()]
    pid = np.repeat(np.arange(len(names)), np.asarray(a["patchSize"]))
    nB = om.nFaces - om.nInternal
    assert len(pid) == nB
    inlet, outlet = names.index("inlet"), names.index("outlet")
    Ut = np.where(pid == outlet, 2, 0).astype(np.int32)
    Uv = np.zeros((nB, 3))
    Uv[pid == inlet, 0] = 0.02
    pt = np.where(pid == outlet, 0, 2).astype(np.int32)
    pv = np.zeros((nB, 1))
    U, p, phi = initial_fields(om, (Ut, Uv))
    dU, dp, dphi = ctx.array(soa(U)), ctx.array(p), ctx.array(phi)
    kw = dict(maxIterations=15, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-12, linearMaxIter=5000,
              nonOrthCorrectors=1, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02)
    r = dm.simple(1e-3, N.Schemes(0, 0, 1, 0.9), N.SimpleControls(**kw), (Ut, Uv), (pt, pv), dU, dp, dphi)
    ro = O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 0, 1, 0.9), O.SimpleControls(**kw), (Ut, Uv), (pt, pv))
    assert r.iterations == ro.iterations == 15 and not r.diverged
    for x, y, name in ((aos(dU.get()), U, "U"), (dp.get(), p, "p"), (dphi.get(), phi, "phi")):
        e = rel(x, y)
        print(f"  {name}: max rel diff {e:.2e}", flush=True)
        assert e < 1e-10, name
    assert np.abs(U).max() > 0.01  # the flow is driven through the imported patches


@pytest.mark.parametrize("momentum", ["fp64", "fp32ir"])
def test_piso_tet_cavity_multigrid_pressure_tracks_oracle(ctx, momentum):
    """The tet cavity with the benchmark's solves: multigrid flexible CG for
    the pressure (constant weights; the second corrector keeps the
    hierarchy of the first, whose matrix it repeats) and FP64 or FP32-refined
    momentum, 6 steps against the oracle's FP64 Jacobi solves at the same
    tolerance."""
    from chuanflow_amd import native as N
    import bench
    print(f"\n[case] cavity 8x8x8 hex -> 3072 tets, PISO, multigrid pressure, momentum {momentum}", flush=True)
    hm = N.HostMesh.box(8, 8, 8, (0, 0, 0), (1, 1, 1), 0.10, 5, True)
    hm.renumber_rcm()
    a, g = hm.arrays(), hm.geometry()
    om = oracle_mesh_from_arrays(a)
    dm = N.Mesh(ctx, a, g)
    dm.set_mg(scale=1.8, coarseScale=1.8, kcycleDepth=2)
    Ubc, pbc = bench.cavity_bcs(a["patchStart"], a["patchSize"], len(a["neighbour"]), om.nB)
    U, p, phi = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    dU, dp, dphi = ctx.zeros((3, om.nCells)), ctx.zeros(om.nCells), ctx.zeros(om.nFaces)
    kw = dict(dt=2e-3, correctors=2, continuityTol=1e-10, linearTol=1e-10, linearMaxIter=5000, nonOrthCorrectors=1)
    uprec = N.FP64 if momentum == "fp64" else N.FP32_IR
    ctl = N.PisoControls(**kw, pPrecision=N.FP32_MG, uPrecision=uprec)
    for step in range(6):
        r = dm.piso_step(1e-3, N.Schemes(0, 0, 1, 0.9), ctl, Ubc, pbc, dU, dp, dphi)
        ro = O.piso_step(om, U, p, phi, 1e-3, O.Schemes(0, 0, 1, 0.9), O.PisoControls(**kw), Ubc, pbc)
        Ug = aos(dU.get())
        pg = dp.get()
        eu, ep = rel(Ug, U), rel(pg - pg.mean(), p - p.mean())
        print(f"  step {step + 1}: U {eu:.2e}, p {ep:.2e}, continuity {r.continuityError:.2e}", flush=True)
        assert np.all(np.isfinite(Ug)) and eu < 1e-7 and ep < 1e-6, step
        assert r.continuityError < 1e-8, step
