"""Detail regions (the coupling's DetailBlock) against the detail oracle:
CPU pins of the oracle and, on the GPU, the block's profile, boundary
conditions and pressure drop against the oracle's restatement driven by
the oracle's SIMPLE."""
import numpy as np
import pytest

import detail_oracle as DO
import oracle as O
from helpers import oracle_mesh_from_arrays

TYPES = ["inlet", "outlet", "wall", "wall", "wall", "wall"]


def channel(nx, ny, nz, L=(1.0, 0.2, 0.2), jitter=0.0, seed=1):
    from chuanflow_amd import native as N
    hm = N.HostMesh.channel(nx, ny, nz, (0, 0, 0), L, jitter, seed)
    return hm, hm.arrays(), hm.geometry()


def test_oracle_profile_and_sections_closed_forms():
    hm, a, g = channel(8, 6, 5)
    pts = np.asarray(a["points"]).reshape(-1, 3)
    pin, pout, area, axis, fs = DO.flow_axis(g["Sf"], a["patchStart"], a["patchSize"], TYPES)
    assert (pin, pout, axis, fs) == (0, 1, 0, 1) and abs(area - 0.04) < 1e-15
    f, w = DO.inlet_profile(pts, g["Sf"], g["Cf"], g["magSf"], a["patchStart"], a["patchSize"], TYPES, axis, pin,
                            area, "parabolic")
    a_f = g["magSf"][f]
    assert abs((w * a_f).sum() - a_f.sum()) < 1e-14 * a_f.sum()
    # on a uniform grid the weights are prod(1 - eta^2) up to one factor
    eta_y = (g["Cf"][f, 1] - 0.1) / 0.1
    eta_z = (g["Cf"][f, 2] - 0.1) / 0.1
    shape = (1 - eta_y ** 2) * (1 - eta_z ** 2)
    np.testing.assert_allclose(w / shape, (w / shape)[0], rtol=1e-13)
    assert len(DO.inlet_profile(pts, g["Sf"], g["Cf"], g["magSf"], a["patchStart"], a["patchSize"], TYPES, axis,
                                pin, area, "flat")[0]) == 0
    secs = DO.sections(g["C"], pts, axis, fs)
    assert len(secs) == 8 and all(len(s) == 30 for s in secs)


def prism_channel(nx, ny, nz, L=(1.0, 0.2, 0.2), jitter=0.3, seed=2):
    """A channel whose interior vertices move only across the flow, by an
    amount that depends on the vertex's (j, k) alone: every cell is an exact
    prism along x with a quadrilateral cross-section, so the cells of one
    x-slab differ in volume while the slab itself stays the box
    [x_i, x_i+1] x [0, Ly] x [0, Lz] (its volume centroid is the box centre)."""
    _, a, _ = channel(nx, ny, nz, L)
    pts = np.asarray(a["points"], np.float64).reshape(-1, 3).copy()
    h = np.array(L) / np.array([nx, ny, nz])
    j = np.rint(pts[:, 1] / h[1]).astype(int)
    k = np.rint(pts[:, 2] / h[2]).astype(int)
    rng = np.random.default_rng(seed)
    dy, dz = rng.uniform(-jitter, jitter, (2, ny + 1, nz + 1))
    inner = (j > 0) & (j < ny) & (k > 0) & (k < nz)
    pts[inner, 1] += dy[j[inner], k[inner]] * h[1]
    pts[inner, 2] += dz[j[inner], k[inner]] * h[2]
    a = dict(a, points=pts.reshape(np.shape(a["points"])))
    m = oracle_mesh_from_arrays(a)
    return a, m, pts


def test_oracle_measure_closed_forms():
    """DetailBlock::measure (coupledNetwork.cpp:1154-1200) on fields whose
    section means are known in closed form.  p = a + g x + h x (y - Ly/2) at
    the cell centroids: the volume-weighted mean of a slab is a + g x_k
    (the (y - Ly/2) term averages to zero over a slab's volume because the
    slab is a box), so dp = rho g (x_first - x_last) exactly; an unweighted
    or area-less mean would leave a term of the jittered centroids behind.
    rho scales the drop linearly; the sections run in the flow's direction
    (reversed when the inlet is the xmax patch); a uniform velocity U = u x
    gives fluxIn = fluxOut = u * (section area)."""
    nx, ny, nz, L = 10, 6, 5, (1.0, 0.2, 0.2)
    a, m, pts = prism_channel(nx, ny, nz, L)
    g = m.geom
    C, V = g["C"], g["V"]
    assert V.std() > 1e-3 * V.mean()  # the cells of a slab really differ in volume
    A, G, H, u = 0.7, -3.0, 50.0, 0.02
    p = A + G * C[:, 0] + H * C[:, 0] * (C[:, 1] - 0.5 * L[1])
    Sf = g["Sf"]
    phi = Sf[:, 0] * u
    for types, first, last in ((TYPES, 0.05, 0.95), (["outlet", "inlet"] + TYPES[2:], 0.95, 0.05)):
        pin, pout, area, axis, fs = DO.flow_axis(Sf, a["patchStart"], a["patchSize"], types)
        assert axis == 0 and abs(area - L[1] * L[2]) < 1e-15
        secs = DO.sections(C, pts, axis, fs)
        assert len(secs) == nx and all(len(s) == ny * nz for s in secs)
        assert abs(C[secs[0], 0] - first).max() < 1e-14 and abs(C[secs[-1], 0] - last).max() < 1e-14
        for rho in (1.0, 998.2):
            dp, fin, fout = DO.measure(p, phi, V, secs, rho, a["patchStart"], a["patchSize"], pin, pout)
            exact = rho * G * (first - last)
            assert abs(dp - exact) <= 1e-12 * abs(exact), (dp, exact)
            # the planted mistake: an unweighted section mean misses the closed form by far more
            wrong = rho * (p[secs[0]].mean() - p[secs[-1]].mean())
            assert abs(wrong - exact) > 1e-6 * abs(exact)
            q = u * area
            assert abs(fin - fs * q) <= 1e-14 and abs(fout - fs * q) <= 1e-14


@pytest.mark.parametrize("shape,mean", [("power1over7", 7.0 / 8.0), ("parabolic", 2.0 / 3.0)])
def test_oracle_inlet_profile_closed_forms(shape, mean):
    """buildInletProfile (coupledNetwork.cpp:768-876): the weights integrate to
    the section area (so the inlet carries exactly the rate), are symmetric
    about the section centre, and on a fine uniform section their peak tends
    to 1/mean^2, where mean is the 1D profile's average over [-1, 1]:
    int_0^1 (1 - t)^(1/7) dt = 7/8 for the one-seventh law and
    int_0^1 (1 - t^2) dt = 2/3 for the parabola.  A wrong exponent (1/6 or
    1/8) moves the peak by 3-4 %.  The imposed inlet flux sum(U_f . Sf)
    equals the rate on a transversely jittered section too."""
    n = 256
    hm, a, g = channel(2, n, n)
    pts = np.asarray(a["points"]).reshape(-1, 3)
    pin, pout, area, axis, fs = DO.flow_axis(g["Sf"], a["patchStart"], a["patchSize"], TYPES)
    f, w = DO.inlet_profile(pts, g["Sf"], g["Cf"], g["magSf"], a["patchStart"], a["patchSize"], TYPES, axis, pin,
                            area, shape)
    af = g["magSf"][f]
    assert abs((w * af).sum() - af.sum()) <= 1e-12 * af.sum() and w.min() > 0
    tol = 2e-3 if shape == "power1over7" else 1e-4
    assert abs(w.max() * mean * mean - 1.0) < tol, w.max() * mean * mean
    # mirror symmetry: the weight at (y, z) equals the weight at (Ly - y, z)
    key = {(round(y, 12), round(z, 12)): wi for (y, z), wi in zip(g["Cf"][f, 1:], w)}
    for (y, z), wi in list(key.items())[:500]:
        assert abs(key[(round(0.2 - y, 12), z)] - wi) <= 1e-13 * wi

    a2, m2, pts2 = prism_channel(4, 7, 6)
    g2 = m2.geom
    pin, pout, area, axis, fs = DO.flow_axis(g2["Sf"], a2["patchStart"], a2["patchSize"], TYPES)
    f2, w2 = DO.inlet_profile(pts2, g2["Sf"], g2["Cf"], g2["magSf"], a2["patchStart"], a2["patchSize"], TYPES, axis,
                              pin, area, shape)
    q = 0.013 * area
    Ut, Uv, pt, pv = DO.boundary(q, m2.nInternal, a2["patchStart"], a2["patchSize"], TYPES, axis, fs, area, pin, pout,
                                 f2, w2)
    inflow = -(Uv[f2 - m2.nInternal] * g2["Sf"][f2]).sum()
    assert abs(inflow - q) <= 1e-14 * q
    assert np.all(Ut[f2 - m2.nInternal] == 0) and np.all(pt[np.arange(a2["patchStart"][pout],
                                                                    a2["patchStart"][pout] + a2["patchSize"][pout])
                                                          - m2.nInternal] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["flat", "parabolic", "power1over7"])
def test_detail_block_matches_oracle(ctx, shape):
    from chuanflow_amd import native as N
    print(f"\n[case] detail block 24x6x6 jittered, {shape}", flush=True)
    hm, a, g = channel(24, 6, 6, jitter=0.1, seed=4)
    pts = np.asarray(a["points"]).reshape(-1, 3)
    blk = N.Detail(ctx, hm, nu=1e-3, rho=1000.0, shape=shape)
    info = blk.info()
    pin, pout, area, axis, fs = DO.flow_axis(g["Sf"], a["patchStart"], a["patchSize"], TYPES)
    assert info["axis"] == axis and info["flowSign"] == fs and info["inletArea"] == area
    fo, wo = DO.inlet_profile(pts, g["Sf"], g["Cf"], g["magSf"], a["patchStart"], a["patchSize"], TYPES, axis, pin,
                              area, shape)
    f, w = blk.inlet()
    assert np.array_equal(f, fo) and np.allclose(w, wo, rtol=1e-14, atol=0)
    secs = DO.sections(g["C"], pts, axis, fs)
    assert info["sections"] == len(secs)
    q = 0.02 * area
    kw = dict(maxIterations=15, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-10, linearMaxIter=5000,
              nonOrthCorrectors=2, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02)
    dp, m = blk.loss_at(q, N.SimpleControls(**kw))
    # the oracle: same boundary conditions, the oracle's SIMPLE from zero fields, the same measurement
    om = oracle_mesh_from_arrays(a)
    Ut, Uv, pt, pv = DO.boundary(q, om.nInternal, a["patchStart"], a["patchSize"], TYPES, axis, fs, area, pin, pout,
                                 fo, wo)
    U, p, phi = np.zeros((om.nCells, 3)), np.zeros(om.nCells), np.zeros(om.nFaces)
    lc, lk = O.linearise(om, 3, Ut, Uv)
    O.correct_boundary_flux(om, U, lc, lk, phi)
    O.simple(om, U, p, phi, 1e-3, O.Schemes(0, 1, 1, 0.9), O.SimpleControls(**kw), (Ut, Uv), (pt, pv))
    dpo, fino, fouto = DO.measure(p, phi, g["V"], secs, 1000.0, a["patchStart"], a["patchSize"], pin, pout)
    print(f"dp gpu {dp:.9e} oracle {dpo:.9e}; flux in {m['fluxIn']:.6e} (q {q:.6e}) out {m['fluxOut']:.6e}",
          flush=True)
    assert dp > 0 and abs(dp - dpo) <= 1e-6 * abs(dpo)
    # on a jittered mesh SIMPLE moves the inlet flux slightly off q, as the oracle (= reference) does
    assert abs(m["fluxIn"] - fino) <= 1e-9 * q and abs(m["fluxOut"] - fouto) <= 1e-6 * q
    Ug, pg, phig = blk.fields()
    assert np.abs(pg - p).max() <= 1e-6 * np.abs(p).max()


@pytest.mark.gpu
@pytest.mark.parametrize("pressure", ["jacobi_fp32ir", "multigrid"])
def test_detail_blocks_concurrent_loss_at_equals_serial(pressure):
    """The coupling pass over several blocks run by host threads on one
    stream per block (cfx_detail_loss_at_many) gives, block by block, the
    pressure drop and fluxes of the blocks run one after another, with the
    reference's Jacobi-PCG pressure solve and with the multigrid FCG one."""
    from chuanflow_amd import native as N
    print(f"\n[case] 6 detail blocks, lossAt serial vs 4 threads (one context per block), {pressure}", flush=True)
    prec = N.FP32_IR if pressure == "jacobi_fp32ir" else N.FP32_MG
    ctl = N.SimpleControls(maxIterations=5, continuityTol=0, pcorrTol=0, momentumTol=0, linearTol=1e-8,
                           linearMaxIter=3000, nonOrthCorrectors=1, relaxU=0.9, relaxP=0.03, pRelaxStart=0.02,
                           pPrecision=prec)
    specs = [(12 + 2 * k, 5, 5, 0.001 * (k + 1), "parabolic" if k % 2 else "flat") for k in range(6)]

    def make():
        out = []
        for k, (nx, ny, nz, q, shape) in enumerate(specs):
            ctx = N.Context(0)
            hm = N.HostMesh.channel(nx, ny, nz, (0, 0, 0), (1.0, 0.2, 0.2), 0.1, 40 + k)
            out.append((ctx, N.Detail(ctx, hm, nu=1e-3, rho=1000.0, shape=shape), q))
        return out

    serial, conc = make(), make()
    ref = [blk.loss_at(q, ctl)[1] for _, blk, q in serial]
    dps, ms = N.detail_loss_at_many([b for _, b, _ in conc], [q for _, _, q in conc], ctl, threads=4)
    for r, m, dp in zip(ref, ms, dps):
        assert np.isfinite(dp) and dp == r["dpPa"] and m["fluxIn"] == r["fluxIn"] and m["fluxOut"] == r["fluxOut"]
    with pytest.raises(RuntimeError, match="one context"):
        shared = N.Context(0)
        hm = N.HostMesh.channel(12, 5, 5, (0, 0, 0), (1.0, 0.2, 0.2), 0.1, 3)
        N.detail_loss_at_many([N.Detail(shared, hm), N.Detail(shared, hm)], [0.001, 0.001], ctl, threads=2)
