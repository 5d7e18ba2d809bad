"""HIP kernels vs the oracle, through the C ABI: the FV operators and the
assembly are bit-exact (per-cell gathers in ascending face order reproduce
the reference's serial scatter sums), and so is the CSR-stream SpMV."""
import numpy as np
import pytest

import oracle as O
from channel import channel_bcs
from helpers import aos, host_and_oracle, soa

pytestmark = pytest.mark.gpu

CASES = {
    "hex_jitter_rcm": dict(nx=12, ny=8, nz=6, mx=(1, 0.5, 0.5), jitter=0.15, seed=4, shuffle=5, rcm=True),
    "hex_plain": dict(nx=10, ny=6, nz=5, mx=(1, 0.5, 0.5), jitter=0.0, seed=0),
    "tet_jitter": dict(nx=5, ny=4, nz=3, mx=(1, 0.5, 0.5), jitter=0.1, seed=8, tets=True, rcm=True),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, ctx):
    from chuanflow_amd import native as N
    kw = dict(CASES[request.param])
    nx, ny, nz, mx = kw.pop("nx"), kw.pop("ny"), kw.pop("nz"), kw.pop("mx")
    print(f"\n[case] {request.param}", flush=True)
    hm, a, om = host_and_oracle(nx, ny, nz, (0, 0, 0), mx, **kw)
    dm = N.Mesh.from_host(ctx, hm)
    Ut, Uv, pt, pv, *_ = channel_bcs(om, 0.02)
    Ulc, Ulk = O.linearise(om, 3, Ut, Uv)
    plc, plk = O.linearise(om, 1, pt, pv)
    rng = np.random.default_rng(17)
    return dict(N=N, ctx=ctx, om=om, dm=dm, Ulc=Ulc, Ulk=Ulk, plc=plc, plk=plk, rng=rng,
                d=dict(Ulc=ctx.array(Ulc), Ulk=ctx.array(Ulk), plc=ctx.array(plc), plk=ctx.array(plk)))


def test_grad_bitwise(case):
    c, om, dm, ctx = case, case["om"], case["dm"], case["ctx"]
    x = c["rng"].standard_normal(om.nCells)
    dx = ctx.array(x)
    for corrected, quirk in ((True, True), (False, True), (False, False)):
        g = dm.grad(dx, c["d"]["plc"], c["d"]["plk"], 1, 0, corrected, quirk).get()
        ref = O.grad(om, x, c["plc"], c["plk"], 1, 0, corrected, quirk)
        assert np.array_equal(aos(g), ref), (corrected, quirk)
    U = c["rng"].standard_normal((3, om.nCells))
    dU = ctx.array(U)
    for comp in range(3):
        g = dm.grad(dU.offset(comp * om.nCells), c["d"]["Ulc"], c["d"]["Ulk"], 3, comp, True, False).get()
        ref = O.grad(om, U[comp], c["Ulc"], c["Ulk"], 3, comp, True, False)
        assert np.array_equal(aos(g), ref), comp


def test_interpolation_and_gamma_bitwise(case):
    om, dm, ctx = case["om"], case["dm"], case["ctx"]
    cell = case["rng"].uniform(0.1, 2.0, om.nCells)
    assert np.array_equal(dm.interpolate_to_faces(ctx.array(cell)).get(), O.interpolate_to_faces(om, cell))
    assert np.array_equal(dm.harmonic_gamma_uniform(1.3e-3).get(), O.harmonic_gamma_uniform(om, 1.3e-3))
    w = ctx.empty(max(1, om.nInternal))
    ctx.check(case["N"].lib().cfx_owner_weights(dm.h, w.ptr))
    assert np.array_equal(w.get()[:om.nInternal], O.owner_weights(om))


def _oracle_momentum(om, c, phi, gf, old, dt, scheme):
    A = O.Ldu(om, 3)
    if dt:
        O.assemble_ddt(om, A, old, dt)
    O.assemble_div(om, A, phi, c["Ulc"], c["Ulk"], scheme=scheme, blend=0.7)
    O.assemble_laplacian(om, A, gf, c["Ulc"], c["Ulk"], -1)
    return A


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_momentum_assembly_bitwise(case, scheme):
    c, om, dm, ctx, N = case, case["om"], case["dm"], case["ctx"], case["N"]
    phi = c["rng"].standard_normal(om.nFaces)
    gf = O.harmonic_gamma_uniform(om, 1e-3)
    old = c["rng"].standard_normal((om.nCells, 3))
    for dt in (0.0, 2e-3):
        A = _oracle_momentum(om, c, phi, gf, old, dt, scheme)
        M = N.Matrix(dm, 3)
        if dt:
            M.ddt(ctx.array(soa(old)), dt)
        M.div(ctx.array(phi), c["d"]["Ulc"], c["d"]["Ulk"], scheme=scheme, blend=0.7)
        M.laplacian(ctx.array(gf), c["d"]["Ulc"], c["d"]["Ulk"], -1)
        d, r, u, lo = M.get()
        assert np.array_equal(d, A.diag) and np.array_equal(r, A.rhs)
        assert np.array_equal(u, A.upper[:om.nInternal]) and np.array_equal(lo, A.lower[:om.nInternal])
        # CSR values through SpMV: bit-identical products
        x = c["rng"].standard_normal((3, om.nCells))
        M.solve(ctx.array(np.zeros((3, om.nCells))), solver=N.SOLVER_BICGSTAB, maxIter=0, nonOrth=0)
        for comp in range(3):
            y = M.spmv(comp, ctx.array(x[comp])).get()
            assert np.array_equal(y, O.spmv(O.csr_from_ldu(om, A, comp), x[comp])), comp


def test_pressure_assembly_and_corrector_bitwise(case):
    c, om, dm, ctx, N = case, case["om"], case["dm"], case["ctx"], case["N"]
    rAUf = c["rng"].uniform(0.05, 1.0, om.nFaces)
    A = O.Ldu(om, 1)
    O.assemble_laplacian(om, A, rAUf, c["plc"], c["plk"], -1)
    M = N.Matrix(dm, 1)
    d_gf = ctx.array(rAUf)
    M.laplacian(d_gf, c["d"]["plc"], c["d"]["plk"], -1)
    d, r, u, lo = M.get()
    assert np.array_equal(d, A.diag) and np.array_equal(r, A.rhs) and np.array_equal(u, A.upper[:om.nInternal])
    x = c["rng"].standard_normal(om.nCells)
    rhs0 = c["rng"].standard_normal(om.nCells)
    d_rhs = ctx.array(rhs0)
    dm.nonorth_correct(d_gf, ctx.array(x), c["d"]["plc"], c["d"]["plk"], d_rhs)
    ref = O.nonorth_correct(om, rAUf, x, c["plc"], c["plk"], rhs=rhs0.copy())
    assert np.array_equal(d_rhs.get(), ref)


def test_flux_operators_bitwise(case):
    c, om, dm, ctx = case, case["om"], case["dm"], case["ctx"]
    U = c["rng"].standard_normal((om.nCells, 3))
    p = c["rng"].standard_normal(om.nCells)
    rAU = c["rng"].uniform(0.1, 1.0, om.nCells)
    phi = np.zeros(om.nFaces)
    O.update_phi_internal(om, U, phi)
    O.correct_boundary_flux(om, U, c["Ulc"], c["Ulk"], phi)
    d_phi = ctx.zeros(om.nFaces)
    d_U = ctx.array(soa(U))
    dm.update_phi_internal(d_U, d_phi)
    dm.correct_boundary_flux(d_U, c["d"]["Ulc"], c["d"]["Ulk"], d_phi)
    assert np.array_equal(d_phi.get(), phi)
    assert dm.continuity_error(d_phi) == O.continuity_error(om, phi)
    gP = O.grad(om, p, c["plc"], c["plk"])
    rAUf = O.interpolate_to_faces(om, rAU)
    for corrected in (True, False):
        ph = phi.copy()
        O.correct_fluxes(om, p, gP, rAUf, c["plc"], c["plk"], corrected, ph)
        d2 = ctx.array(phi)
        dm.correct_fluxes(ctx.array(p), ctx.array(soa(gP)), ctx.array(rAUf), c["d"]["plc"], c["d"]["plk"],
                          corrected, d2)
        assert np.array_equal(d2.get(), ph), corrected
        assert np.array_equal(dm.div_phi(d2).get(), O.div_phi(om, ph))
        assert dm.continuity_error(d2) == O.continuity_error(om, ph)


def test_explicit_operator_matches_reference_definition(case):
    """explicitOperator: h = rhs - (A x - D x) (fvMatrix.cpp:131-146)."""
    c, om, dm, ctx, N = case, case["om"], case["dm"], case["ctx"], case["N"]
    phi = c["rng"].standard_normal(om.nFaces)
    gf = O.harmonic_gamma_uniform(om, 1e-3)
    A = _oracle_momentum(om, c, phi, gf, None, 0.0, 0)
    M = N.Matrix(dm, 3)
    M.div(ctx.array(phi), c["d"]["Ulc"], c["d"]["Ulk"])
    M.laplacian(ctx.array(gf), c["d"]["Ulc"], c["d"]["Ulk"], -1)
    M.solve(ctx.array(np.zeros((3, om.nCells))), solver=N.SOLVER_BICGSTAB, maxIter=0, nonOrth=0)
    x = c["rng"].standard_normal(om.nCells)
    for comp in range(3):
        h = M.explicit_operator(ctx.array(x), comp).get()
        ref = A.rhs[comp] - (O.spmv(O.csr_from_ldu(om, A, comp), x) - A.diag[comp] * x)
        assert np.array_equal(h, ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32ir"])
def test_vector_bicgstab_components_equal_separate_solves(case, precision):
    """The BiCGStab of a vector equation (FP64, or FP32 with FP64 refinement)
    with its components in shared launches gives the fields and report of the
    reference's component-after-component loop (fvMatrix.cpp:93-122) bit for
    bit, with and without the non-orthogonal corrector passes, and both track
    the oracle's solves (the FP32 solve reaches tol 1e-12 in 2-3 refinement
    sweeps here)."""
    c, om, dm, ctx, N = case, case["om"], case["dm"], case["ctx"], case["N"]
    prec = N.FP64 if precision == "fp64" else N.FP32_IR
    phi = c["rng"].standard_normal(om.nFaces) * 1e-3
    gf = O.harmonic_gamma_uniform(om, 1e-3)
    old = c["rng"].standard_normal((om.nCells, 3))
    A = _oracle_momentum(om, c, phi, gf, old, 2e-3, 0)
    x0 = c["rng"].standard_normal((3, om.nCells)) * 1e-2
    out = {}
    for batched in (True, False):
        ctx.set_component_batching(batched)
        M = N.Matrix(dm, 3)
        M.ddt(ctx.array(soa(old)), 2e-3)
        M.div(ctx.array(phi), c["d"]["Ulc"], c["d"]["Ulk"], scheme=0, blend=0.7)
        M.laplacian(ctx.array(gf), c["d"]["Ulc"], c["d"]["Ulk"], -1)
        # None: no corrector (plain solve); 0 and 2: corrector applied before each of 1 and 3 passes
        for nonOrth in (None, 0, 2):
            x = ctx.array(x0)
            corr = {} if nonOrth is None else dict(gammaFace=ctx.array(gf), linc=c["d"]["Ulc"], link=c["d"]["Ulk"])
            rep = M.solve(x, solver=N.SOLVER_BICGSTAB, tol=1e-12, maxIter=400, nonOrth=nonOrth or 0, precision=prec,
                          **corr)
            out[batched, nonOrth] = (x.get(), rep.iterations, rep.finalResidual, rep.converged,
                                     rep.outerIterations, rep.precondFallbacks)
    ctx.set_component_batching(True)
    for nonOrth in (None, 0, 2):
        a, b = out[True, nonOrth], out[False, nonOrth]
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:], nonOrth
        assert a[3] == 1
    # the oracle's BiCGStab per component on the same system
    for comp in range(3):
        xo, _ = O.solve_bicgstab(O.csr_from_ldu(om, A, comp), A.rhs[comp].copy(), x0[comp].copy(), 1e-12, 400)
        assert np.abs(out[True, None][0][comp] - xo).max() <= 1e-9 * max(np.abs(xo).max(), 1e-300), comp
    if precision == "fp32ir":
        print(f"  FP32 refinement: outer {out[True, None][4]}, FP64 fallback {out[True, None][5]}", flush=True)


def test_mixed_vector_bicgstab_hard_systems_equal_separate_solves(case):
    """Vector diffusion systems whose face coefficients span 1e-3, 1e-6 and
    1e-9 (log-uniform, seeded): the FP32-refined component solve and the
    one-by-one solves agree bit for bit, reports included.  The first two
    converge in FP32 refinement sweeps.  At 1e-9 the FP32 solve can fail (on
    the tet mesh its iterates blow up): the failed correction is then not kept
    and the solve finishes in FP64, so x stays finite."""
    c, om, dm, ctx, N = case, case["om"], case["dm"], case["ctx"], case["N"]
    rng = np.random.default_rng(23)
    for span in (1e-3, 1e-6, 1e-9):
        gf = np.exp(rng.uniform(np.log(span), 0.0, om.nFaces)) * 1e-3
        x0 = rng.standard_normal((3, om.nCells)) * 1e-2
        src = rng.standard_normal((3, om.nCells))
        out = {}
        for batched in (True, False):
            ctx.set_component_batching(batched)
            M = N.Matrix(dm, 3)
            M.laplacian(ctx.array(gf), c["d"]["Ulc"], c["d"]["Ulk"], -1)
            for comp in range(3):
                M.add_rhs(comp, ctx.array(src[comp]))
            x = ctx.array(x0)
            rep = M.solve(x, solver=N.SOLVER_BICGSTAB, tol=1e-10, maxIter=3000, nonOrth=0, precision=N.FP32_IR)
            out[batched] = (x.get(), rep.iterations, rep.finalResidual, rep.converged, rep.outerIterations,
                            rep.precondFallbacks)
        ctx.set_component_batching(True)
        a, b = out[True], out[False]
        print(f"  span {span:g}: iterations {a[1]}, outer {a[4]}, fallback {a[5]}, converged {a[3]}", flush=True)
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:], span
        assert np.all(np.isfinite(a[0])), span
        if span > 1e-9:
            assert a[3] == 1 and a[5] == 0, span
