"""The oracle pinned: against the reference's golden vectors (tests/golden),
against the reference build itself where it is present (oracle/_ref), and
against closed forms and invariants that catch dropped terms, wrong signs,
wrong indices and transposed operands."""
import os

import numpy as np
import pytest

import oracle as O
from channel import channel_bcs, neumann_bcs
from helpers import oracle_mesh_from_arrays

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


def golden_mesh():
    g = gold("geom_jitter")
    return oracle_mesh_from_arrays(g), g


# ------------------------------------------------------------------ RNG
def test_mt19937_64_known_answer():
    # the C++ standard fixes the 10000th output of a default-seeded mt19937_64
    assert O.mt64_nth(5489, 10000) == 9981545732273789042


# ----------------------------------------------------- golden vectors
def test_blockmesh_matches_reference_golden():
    g = gold("blockmesh_3x4x5")
    m = O.blockmesh(3, 4, 5, (0, 0, 0), (1, 2, 3))
    for k in ("points", "faceVerts", "owner", "neighbour", "patchStart", "patchSize"):
        assert np.array_equal(getattr(m, k), g[k]), k


def test_geometry_matches_reference_golden_bitwise():
    m, g = golden_mesh()
    for k in ("Sf", "Cf", "magSf", "C", "V", "dvec", "invDist"):
        assert np.array_equal(m.geom[k], g["g_" + k]), k
    assert np.array_equal(O.owner_weights(m), g["ow"])


def test_jitter_generator_reproduces_golden_points():
    g = gold("geom_jitter")
    m = O.blockmesh(6, 5, 4, (0, 0, 0), (1, 0.5, 0.5), jitter=0.15, seed=11)
    assert np.array_equal(m.points, g["points"])


def test_grad_matches_reference_golden_bitwise():
    m, _ = golden_mesh()
    g = gold("grad_jitter")
    lc, lk = O.linearise(m, 1, *channel_bcs(m)[2:4])
    assert np.array_equal(O.grad(m, g["data"], lc, lk, corrected=True), g["grad_corr"])
    assert np.array_equal(O.grad(m, g["data"], lc, lk, corrected=False, quirk=True), g["grad_lin"])


def test_laplacian_corrector_cg_matches_reference_golden():
    m, _ = golden_mesh()
    g = gold("lap_jitter")
    _, _, pt, pv, *_ = channel_bcs(m)
    lc, lk = O.linearise(m, 1, pt, pv)
    A = O.Ldu(m)
    O.assemble_laplacian(m, A, g["gammaFace"], lc, lk, -1)
    A.rhs[0] += g["extra"]
    assert np.array_equal(A.diag[0], g["diag"])
    csr = O.csr_from_ldu(m, A)
    x = np.zeros(m.nCells)
    for _ in range(3):  # nonOrth = 2 -> 3 passes, rhs restored each pass
        rhs = A.rhs[0].copy()
        O.nonorth_correct(m, g["gammaFace"], x, lc, lk, rhs=rhs)
        x, rep = O.solve_cg(csr, rhs, x, 1e-10, 500)
    assert np.array_equal(rhs, g["rhs"])          # rhs_ after the last pass
    assert np.array_equal(x, g["x"])
    assert rep.iterations == int(g["iterations"])
    # off-diagonal probe: h(u) - h(0) = -(A u - D u)
    offU = -(O.spmv(csr, g["u"]) - A.diag[0] * g["u"])
    np.testing.assert_allclose(offU, g["offU"], rtol=0, atol=1e-13 * np.abs(g["offU"]).max())


def test_cg_and_bicgstab_match_reference_on_tridiagonal():
    g = gold("cg_tridiag")
    A = O.Csr(len(g["rowStart"]) - 1, len(g["cols"]))
    A.rowStart[:], A.cols[:], A.vals[:] = g["rowStart"], g["cols"], g["vals"]
    x, rep = O.solve_cg(A, g["b"], None, 1e-10, 500)
    assert np.array_equal(x, g["cg_x"]) and rep.iterations == int(g["cg_iters"])
    x, rep = O.solve_bicgstab(A, g["b"], None, 1e-10, 500)
    assert np.array_equal(x, g["bicg_x"]) and rep.iterations == int(g["bicg_iters"])
    # and x = 1 (solverBackendApp.cpp:89-109)
    np.testing.assert_allclose(g["cg_x"], 1.0, atol=1e-6)


def test_flux_operators_match_reference_golden_bitwise():
    m, _ = golden_mesh()
    g = gold("flux_jitter")
    Ut, Uv, pt, pv, *_ = channel_bcs(m)
    Ulc, Ulk = O.linearise(m, 3, Ut, Uv)
    plc, plk = O.linearise(m, 1, pt, pv)
    phi = np.zeros(m.nFaces)
    O.update_phi_internal(m, g["U"], phi)
    O.correct_boundary_flux(m, g["U"], Ulc, Ulk, phi)
    assert O.continuity_error(m, phi) == g["contErrBefore"]
    rAUf = O.interpolate_to_faces(m, g["rAU"])
    assert np.array_equal(rAUf, g["rAUf"])
    O.correct_fluxes(m, g["p"], g["gradP"], rAUf, plc, plk, True, phi)
    assert np.array_equal(phi, g["phi"])
    assert np.array_equal(O.div_phi(m, phi), g["divPhi"])
    assert O.continuity_error(m, phi) == g["contErrAfter"]


@pytest.mark.parametrize("tag,relax", [("a", (0.9, 0.03, 0.02)), ("b", (0.7, 0.3, 0.0))])
def test_simple_matches_reference_golden_bitwise(tag, relax):
    m, _ = golden_mesh()
    g = gold("simple_channel_" + tag)
    Ut, Uv, pt, pv, *_ = channel_bcs(m)
    Ulc, Ulk = O.linearise(m, 3, Ut, Uv)
    U = np.zeros((m.nCells, 3))
    p = np.zeros(m.nCells)
    phi = np.zeros(m.nFaces)
    O.correct_boundary_flux(m, U, Ulc, Ulk, phi)
    ctl = O.SimpleControls(maxIterations=10, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-8,
                           linearMaxIter=2000, nonOrthCorrectors=2, relaxU=relax[0], relaxP=relax[1],
                           pRelaxStart=relax[2])
    r = O.simple(m, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), ctl, (Ut, Uv), (pt, pv))
    assert r.iterations == int(g["result"][0])
    assert np.array_equal(U, g["U"]) and np.array_equal(p, g["p"]) and np.array_equal(phi, g["phi"])
    assert r.finalContinuityError == g["result"][1] and r.finalMomentumResidual == g["result"][2]


def test_piso_matches_reference_golden_bitwise():
    m, _ = golden_mesh()
    g = gold("piso_channel")
    Ut, Uv, pt, pv, *_ = channel_bcs(m)
    U, p, phi = g["U0"].copy(), np.zeros(m.nCells), g["phi0"].copy()
    ctl = O.PisoControls(dt=0.01, correctors=2, continuityTol=1e-10, linearTol=1e-8, linearMaxIter=2000,
                         nonOrthCorrectors=2)
    for k in range(3):
        r = O.piso_step(m, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), ctl, (Ut, Uv), (pt, pv))
        assert r.correctorsUsed == int(g["results"][k][0])
        assert r.continuityError == g["results"][k][2]
    assert np.array_equal(U, g["U"]) and np.array_equal(p, g["p"]) and np.array_equal(phi, g["phi"])


# ------------------------------------------- live reference (if built)
@pytest.mark.skipif(not O.have_ref(), reason="reference build not present")
def test_oracle_matches_reference_build_on_fresh_case():
    m = O.blockmesh(5, 4, 3, (0, 0, 0), (2, 1, 1), jitter=0.12, seed=99).compute_geometry()
    rm = O.RefMesh(m)
    g = rm.geometry()
    for k in g:
        assert np.array_equal(g[k], m.geom[k]), k
    Ut, Uv, pt, pv, Upt, Upv, ppt, ppv = channel_bcs(m, 0.05)
    Ulc, Ulk = O.linearise(m, 3, Ut, Uv)
    U, p, phi = np.zeros((m.nCells, 3)), np.zeros(m.nCells), np.zeros(m.nFaces)
    O.correct_boundary_flux(m, U, Ulc, Ulk, phi)
    Ur, pr, phir = U.copy(), p.copy(), phi.copy()
    # blend convection (the reference's default blendCoeff 0.9), one
    # non-orthogonal corrector: other branches of every operator
    ctl = O.SimpleControls(maxIterations=6, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-9,
                           linearMaxIter=2000, nonOrthCorrectors=1, relaxU=0.8, relaxP=0.2, pRelaxStart=0.0)
    O.simple(m, U, p, phi, 2e-3, O.Schemes(2, 1, 1, 0.9), ctl, (Ut, Uv), (pt, pv))
    rm.simple(Upt, Upv, ppt, ppv, 2e-3, (2, 1, 1), [6, 0, 0, 0, 1e-9, 2000, 1, 0.8, 0.2, 0, 0, 0, 0], Ur, pr, phir)
    assert np.array_equal(U, Ur) and np.array_equal(p, pr) and np.array_equal(phi, phir)


@pytest.mark.skipif(not O.have_ref(), reason="reference build not present")
def test_reference_relaxation_is_applied_after_the_final_pass_only():
    """Schemes::relaxation (fvMatrix.cpp:110-118) in the reference build: three
    passes of the oracle's corrected laplacian solve, then
    x = xOld + 0.7 (x - xOld) with xOld the solution before the final pass,
    equal bit for bit to the reference's relaxed fvMatrix::solve (the
    unrelaxed one equals the oracle's pass loop)."""
    m = O.blockmesh(7, 5, 4, (0, 0, 0), (1, 0.7, 0.6), jitter=0.15, seed=5).compute_geometry()
    rng = np.random.default_rng(3)
    _, _, pt, pv, _, _, ppt, ppv = channel_bcs(m)
    lc, lk = O.linearise(m, 1, pt, pv)
    gf = rng.uniform(0.5, 2.0, m.nFaces)
    src = rng.standard_normal(m.nCells)
    A = O.Ldu(m)
    O.assemble_laplacian(m, A, gf, lc, lk, -1)
    A.rhs[0] += src
    csr = O.csr_from_ldu(m, A)
    x = np.zeros(m.nCells)
    for _ in range(3):
        xOld = x.copy()
        rhs = A.rhs[0].copy()
        O.nonorth_correct(m, gf, x, lc, lk, rhs=rhs)
        x, _ = O.solve_cg(csr, rhs, x, 1e-12, 2000)
    rm = O.RefMesh(m)
    z = np.zeros(m.nCells)
    r1 = rm.laplacian_face(gf, ppt, ppv, 1, 1, -1, src, 1e-12, 2000, 2, z, z)
    r7 = rm.laplacian_face(gf, ppt, ppv, 1, 1, -1, src, 1e-12, 2000, 2, z, z, relaxation=0.7)
    assert np.array_equal(r1["x"], x)
    assert np.array_equal(r7["x"], xOld + 0.7 * (x - xOld)) and not np.array_equal(r7["x"], x)


# ----------------------------------------------------------- closed forms
def test_unit_cube_geometry_closed_forms():
    # kernelTests.cpp:58-83
    m = O.blockmesh(2, 2, 2).compute_geometry()
    assert m.nCells == 8 and m.nInternal == 12 and m.nB == 24
    assert abs(m.geom["V"].sum() - 1.0) < 1e-12 and np.allclose(m.geom["V"], 0.125)
    s = np.zeros((m.nCells, 3))
    np.add.at(s, m.owner, m.geom["Sf"])
    np.subtract.at(s, m.neighbour, m.geom["Sf"][:m.nInternal])
    assert np.abs(s).max() < 1e-12


def _cell_volumes_by_apex(m):
    """Independent volume: every face fan triangle (v0, vi, vi+1) of the
    cell's faces forms a tet with the cell's vertex mean as apex; each tet
    is oriented outward by the triangle centroid's side of the apex (not by
    the oracle's area-vector flip) and the signed volumes are summed."""
    cells = [[] for _ in range(m.nCells)]
    for f in range(m.nFaces):
        cells[m.owner[f]].append(f)
        if f < m.nInternal:
            cells[m.neighbour[f]].append(f)
    P = m.points
    vols = np.zeros(m.nCells)
    for c, faces in enumerate(cells):
        vid = np.unique(np.concatenate([m.faceVerts[m.faceStart[f]:m.faceStart[f + 1]] for f in faces]))
        apex = P[vid].mean(axis=0)
        v = 0.0
        for f in faces:
            vs = m.faceVerts[m.faceStart[f]:m.faceStart[f + 1]]
            fc = P[vs].mean(axis=0)
            n_face = sum(np.cross(P[vs[i]] - P[vs[0]], P[vs[i + 1]] - P[vs[0]]) for i in range(1, len(vs) - 1))
            sgn = 1.0 if np.dot(n_face, fc - apex) > 0 else -1.0
            for i in range(1, len(vs) - 1):
                a, b, cc = P[vs[0]], P[vs[i]], P[vs[i + 1]]
                v += sgn * np.dot(np.cross(b - a, cc - a), a - apex) / 6.0
        vols[c] = v
    return vols


def test_jittered_volumes_closed_forms():
    nx, ny, nz = 5, 4, 3
    m = O.blockmesh(nx, ny, nz, (0, 0, 0), (1, 1, 1), jitter=0.2, seed=5).compute_geometry()
    assert (m.geom["V"] > 0).all()
    assert abs(m.geom["V"].sum() - 1.0) < 1e-12
    s = np.zeros((m.nCells, 3))
    np.add.at(s, m.owner, m.geom["Sf"])
    np.subtract.at(s, m.neighbour, m.geom["Sf"][:m.nInternal])
    assert np.abs(s).max() < 1e-14
    np.testing.assert_allclose(_cell_volumes_by_apex(m), m.geom["V"], rtol=1e-12)


def test_orthogonal_laplacian_is_seven_point_stencil():
    nx, ny, nz = 4, 3, 5
    L = np.array([2.0, 1.5, 2.5])
    m = O.blockmesh(nx, ny, nz, (0, 0, 0), tuple(L)).compute_geometry()
    h = L / np.array([nx, ny, nz])
    lc, lk = O.linearise(m, 1, *neumann_bcs(m.nB))
    A = O.Ldu(m)
    O.assemble_laplacian(m, A, np.ones(m.nFaces), lc, lk, -1)
    d = m.geom["dvec"][:m.nInternal]
    axis = np.argmax(np.abs(d), axis=1)
    area = np.prod(h) / h[axis]
    np.testing.assert_allclose(-A.upper[:m.nInternal], area / h[axis], rtol=1e-13)
    np.testing.assert_allclose(A.upper[:m.nInternal], A.lower[:m.nInternal], rtol=0)
    csr = O.csr_from_ldu(m, A)
    D = csr.todense()
    assert np.allclose(D, D.T) and np.abs(D.sum(axis=1)).max() < 1e-12  # Neumann: row sums zero
    off = D - np.diag(np.diag(D))
    assert (off <= 0).all() and (np.diag(D) > 0).all()                 # M-matrix signs


def test_single_cell_dirichlet_closed_form():
    # kernelTests.cpp:102-113: one cell, T=0 at xmin, T=1 at xmax -> T=0.5
    m = O.blockmesh(1, 1, 1).compute_geometry()
    pid = m.patch_of_boundary_face()
    t = np.where((pid == 0) | (pid == 1), 0, 2).astype(np.int32)
    v = np.where(pid == 1, 1.0, 0.0)[:, None]
    lc, lk = O.linearise(m, 1, t, v)
    A = O.Ldu(m)
    O.assemble_laplacian(m, A, np.ones(m.nFaces), lc, lk, -1)
    x, rep = O.solve_cg(O.csr_from_ldu(m, A), A.rhs[0], None, 1e-10, 100)
    assert rep.converged and abs(x[0] - 0.5) < 1e-12
    assert A.diag[0, 0] == 4.0  # 2 faces * |Sf|/|d| = 2 * 1/0.5


def test_cg_matches_dense_solve_and_a_norm_error_decreases():
    m = O.blockmesh(6, 5, 4, jitter=0.15, seed=3).compute_geometry()
    _, _, pt, pv, *_ = channel_bcs(m)
    lc, lk = O.linearise(m, 1, pt, pv)  # fixed p at outlet -> SPD
    A = O.Ldu(m)
    O.assemble_laplacian(m, A, np.ones(m.nFaces), lc, lk, -1)
    csr = O.csr_from_ldu(m, A)
    D = csr.todense()
    rng = np.random.default_rng(0)
    b = rng.standard_normal(m.nCells)
    xs = np.linalg.solve(D, b)
    x, rep = O.solve_cg(csr, b, None, 1e-13, 2000)
    assert rep.converged
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-10 * np.abs(xs).max())
    errs = []
    for k in range(1, 40):
        xk, _ = O.solve_cg(csr, b, None, 1e-30, k)
        e = xk - xs
        errs.append(e @ D @ e)
    assert all(errs[i + 1] <= errs[i] * (1 + 1e-12) for i in range(len(errs) - 1))


def test_grad_exact_for_linear_field_on_hex():
    # gradientTests.cpp:131-145: linear field exact on hex
    m = O.blockmesh(5, 4, 6, (0, 0, 0), (1, 2, 1)).compute_geometry()
    a = np.array([0.3, -1.2, 2.5])
    C = m.geom["C"]
    T = C @ a + 0.7
    bf = m.geom["Cf"][m.nInternal:]
    t = np.zeros(m.nB, np.int32)  # fixedValue with the exact face values
    lc, lk = O.linearise(m, 1, t, (bf @ a + 0.7)[:, None])
    g = O.grad(m, T, lc, lk, corrected=True)
    assert np.abs(g - a).max() < 1e-9


def test_corrected_flux_is_divergence_free():
    # check_piso.py:110-111: after correctFluxes with the solved pressure the
    # flux is divergence free to the solver tolerance
    m = O.blockmesh(6, 5, 4, (0, 0, 0), (1, 0.5, 0.5)).compute_geometry()
    Ut, Uv, pt, pv, *_ = channel_bcs(m)
    Ulc, Ulk = O.linearise(m, 3, Ut, Uv)
    U, p, phi = np.zeros((m.nCells, 3)), np.zeros(m.nCells), np.zeros(m.nFaces)
    O.correct_boundary_flux(m, U, Ulc, Ulk, phi)
    ctl = O.SimpleControls(maxIterations=3, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-13,
                           linearMaxIter=5000, nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02)
    r, hist = O.simple(m, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), ctl, (Ut, Uv), (pt, pv), hist=True)
    assert (hist[:, 1] < 1e-10).all()  # errAfter
