"""The library's host mesh layer (no GPU): BlockMesh, jitter, Kuhn tets,
shuffle, RCM and polyMesh geometry, against the oracle and invariants."""
import numpy as np
import pytest

import oracle as O
from helpers import host_and_oracle, oracle_mesh_from_arrays

N = pytest.importorskip("chuanflow_amd.native")


@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_blockmesh_topology_and_points_match_oracle(jitter):
    hm = N.HostMesh.box(7, 5, 4, (0, 0, 0), (1, 0.5, 0.5), jitter, 11, False)
    a = hm.arrays()
    o = O.blockmesh(7, 5, 4, (0, 0, 0), (1, 0.5, 0.5), jitter=jitter, seed=11)
    for k in ("points", "faceVerts", "owner", "neighbour", "patchStart", "patchSize"):
        assert np.array_equal(a[k], getattr(o, k)), k


@pytest.mark.parametrize("kw", [dict(jitter=0.15), dict(jitter=0.1, tets=True),
                                dict(jitter=0.15, shuffle=7, rcm=True), dict(jitter=0.1, tets=True, rcm=True)])
def test_geometry_bitwise_equal_to_oracle(kw):
    hm, a, om = host_and_oracle(6, 5, 4, (0, 0, 0), (1, 0.5, 0.5), seed=3, **kw)
    g = hm.geometry()
    for k in ("Sf", "Cf", "magSf", "C", "V", "dvec", "invDist"):
        assert np.array_equal(g[k], om.geom[k]), k


def _closure(a, g):
    s = np.zeros((a["nCells"], 3))
    np.add.at(s, a["owner"], g["Sf"])
    nI = len(a["neighbour"])
    np.subtract.at(s, a["neighbour"], g["Sf"][:nI])
    return np.abs(s).max()


def test_kuhn_tets_invariants():
    nx, ny, nz = 4, 3, 5
    hm = N.HostMesh.box(nx, ny, nz, (0, 0, 0), (1, 2, 3), 0.1, 4, True)
    a, g = hm.arrays(), hm.geometry()
    assert a["nCells"] == 6 * nx * ny * nz
    assert np.all(np.diff(a["faceStart"]) == 3)
    assert (g["V"] > 0).all() and abs(g["V"].sum() - 6.0) < 1e-12
    assert _closure(a, g) < 1e-14
    assert np.all(a["owner"][:len(a["neighbour"])] < a["neighbour"])
    # every boundary face lies on its box side
    nI = len(a["neighbour"])
    Cf = g["Cf"]
    for p, (s, n) in enumerate(zip(a["patchStart"], a["patchSize"])):
        ax, hi = p // 2, p % 2
        want = (0.0, (1.0, 2.0, 3.0)[ax])[hi]
        assert np.allclose(Cf[s:s + n, ax], want, atol=1e-12)
    # boundary face count: 2 triangles per boundary quad
    assert len(a["owner"]) - nI == 4 * (nx * ny + ny * nz + nz * nx)


def test_rcm_is_a_permutation_reduces_bandwidth_and_sorts_faces():
    hm = N.HostMesh.box(12, 10, 8, jitter=0.1, seed=2)
    hm.shuffle(123)
    bw0 = hm.bandwidth()
    newId = hm.renumber_rcm()
    assert np.array_equal(np.sort(newId), np.arange(len(newId)))
    bw1 = hm.bandwidth()
    assert bw1 < bw0 and bw1 <= 2 * 12 * 10
    a = hm.arrays()
    nI = len(a["neighbour"])
    key = a["owner"][:nI] * (a["nCells"] + 1) + a["neighbour"]
    assert np.all(np.diff(key) > 0) and np.all(a["owner"][:nI] < a["neighbour"])
    for s, n in zip(a["patchStart"], a["patchSize"]):
        assert np.all(np.diff(a["owner"][s:s + n]) >= 0)


def test_shuffle_and_rcm_leave_the_solution_invariant():
    """The oracle solves the same Poisson problem on the original and on the
    shuffled + RCM-renumbered mesh; mapped back, the solutions agree to
    summation-order rounding."""
    hm0, a0, m0 = host_and_oracle(8, 6, 5, (0, 0, 0), (1, 0.5, 0.5), jitter=0.15, seed=9)
    hm1 = N.HostMesh.box(8, 6, 5, (0, 0, 0), (1, 0.5, 0.5), 0.15, 9, False)
    hm1.shuffle(77)
    a_sh = hm1.arrays()  # cell ids permuted by the shuffle
    newId = hm1.renumber_rcm()
    m1 = oracle_mesh_from_arrays(hm1.arrays())
    # the shuffle's cell permutation: identify cells by centroid
    from scipy.spatial import cKDTree
    t = cKDTree(m1.geom["C"])
    _, to1 = t.query(m0.geom["C"])
    assert np.array_equal(np.sort(to1), np.arange(m0.nCells))
    np.testing.assert_allclose(m1.geom["V"][to1], m0.geom["V"], rtol=1e-13)

    def solve(m, b):
        pid = m.patch_of_boundary_face()
        t = np.where(pid == 1, 0, 2).astype(np.int32)
        lc, lk = O.linearise(m, 1, t, np.zeros((m.nB, 1)))
        A = O.Ldu(m)
        O.assemble_laplacian(m, A, np.ones(m.nFaces), lc, lk, -1)
        x, rep = O.solve_cg(O.csr_from_ldu(m, A), b, None, 1e-13, 5000)
        assert rep.converged
        return x

    b0 = np.sin(7 * m0.geom["C"][:, 0]) * np.cos(3 * m0.geom["C"][:, 1])
    b1 = np.empty_like(b0)
    b1[to1] = b0
    x0 = solve(m0, b0)
    x1 = solve(m1, b1)
    np.testing.assert_allclose(x1[to1], x0, rtol=0, atol=1e-10 * np.abs(x0).max())


def test_from_arrays_roundtrip_and_bad_input_rejected():
    hm, a, _ = host_and_oracle(3, 3, 2)
    h2 = N.HostMesh.from_arrays(a["points"], a["faceStart"], a["faceVerts"], a["owner"], a["neighbour"],
                                a["patchStart"], a["patchSize"])
    g1, g2 = hm.geometry(), h2.geometry()
    for k in g1:
        assert np.array_equal(g1[k], g2[k])
    bad = a["faceVerts"].copy()
    bad[0] = 10 ** 6
    with pytest.raises(N.CfxError):
        N.HostMesh.from_arrays(a["points"], a["faceStart"], bad, a["owner"], a["neighbour"], a["patchStart"],
                               a["patchSize"])
    with pytest.raises(N.CfxError):  # a patch that covers an internal face
        N.HostMesh.from_arrays(a["points"], a["faceStart"], a["faceVerts"], a["owner"], a["neighbour"],
                               a["patchStart"] - 1, a["patchSize"])


def test_obstacle_channel_invariants():
    from chuanflow_amd import native as N
    mn, mx, (nx, ny, nz) = (0, 0, 0), (1.0, 0.5, 0.1), (40, 20, 4)
    cx, cy, R = 0.3, 0.25, 0.08
    box = N.HostMesh.box(nx, ny, nz, mn, mx, 0.1, 5)
    hm = N.HostMesh.obstacle_channel(nx, ny, nz, mn, mx, 0.1, 5, cx, cy, R)
    a, g = hm.arrays(), hm.geometry()
    gb = box.geometry()
    inside = (gb["C"][:, 0] - cx) ** 2 + (gb["C"][:, 1] - cy) ** 2 < R * R
    assert a["nCells"] == box.arrays()["nCells"] - inside.sum() and inside.sum() > 0
    # kept cells keep their volumes and centres (same order; a reversed cut
    # face changes the summation order of its cell only)
    np.testing.assert_allclose(g["V"], gb["V"][~inside], rtol=1e-13)
    np.testing.assert_allclose(g["C"], gb["C"][~inside], rtol=1e-13)
    # every cell is closed: the outward face areas sum to zero
    nI, nF, n = len(a["neighbour"]), len(a["owner"]), a["nCells"]
    S = np.zeros((n, 3))
    np.add.at(S, a["owner"], g["Sf"])
    np.add.at(S, a["neighbour"], -g["Sf"][:nI])
    assert np.abs(S).max() < 1e-14
    # the 7th patch faces the cylinder axis: normals point out of the fluid, into the obstacle
    ps, sz = a["patchStart"], a["patchSize"]
    assert len(ps) == 7 and ps[6] + sz[6] == nF and sz[6] > 0
    f = np.arange(ps[6], ps[6] + sz[6])
    toward = np.column_stack([cx - g["Cf"][f, 0], cy - g["Cf"][f, 1]])
    assert np.all((g["Sf"][f, :2] * toward).sum(axis=1) > 0)


def test_oracle_piso_cavity_on_tets_uncorrected_closes_continuity():
    """With the laplacian "uncorrected" the corrected face flux is exactly the
    flux of the pressure equation just solved, so after every PISO step of the
    closed Kuhn-tet cavity the continuity error is at the solver tolerance (no
    right-hand-side projection needed)."""
    import bench
    hm, a, om = host_and_oracle(6, 6, 6, jitter=0.10, seed=3, tets=True, rcm=True)
    Ubc, pbc = bench.cavity_bcs(a["patchStart"], a["patchSize"], len(a["neighbour"]), om.nB)
    U, p, phi = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    ctl = O.PisoControls(dt=2e-3, correctors=2, continuityTol=1e-10, linearTol=1e-12, linearMaxIter=5000,
                         nonOrthCorrectors=1)
    for step in range(8):
        r = O.piso_step(om, U, p, phi, 1e-3, O.Schemes(0, 0, 1, 0.9), ctl, Ubc, pbc)
        assert r.continuityError < 1e-9, (step, r.continuityError)
    assert 0 < np.abs(U).max() < 1.0
