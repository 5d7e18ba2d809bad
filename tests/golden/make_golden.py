"""Write the golden vectors under tests/golden/ from the REFERENCE build.

Run in a container where the reference tree is present and
``make -C oracle ref`` has produced ``oracle/_ref/libcfref.so``:

    python tests/golden/make_golden.py

Each fixture is a small .npz of inputs and the reference's outputs on them
(no reference source is stored).  Cases:
  blockmesh_3x4x5   BlockMesh::create topology and points
  geom_jitter       polyMesh geometry of a jittered 6x5x4 hex mesh
  grad_jitter       fvc::grad corrected/linear on it
  lap_jitter        laplacianFace + corrector + solveCG (aDiag, rhs_, offdiag probe, x)
  cg_tridiag        solveCG / solveBiCGStab on a 1D tridiagonal Poisson (x = 1)
  flux_jitter       updatePhiInternal/correctBoundaryFlux/correctFluxes/divPhi/continuityError
  simple_channel    solveIncompressibleSimple, 10 iterations, both relaxation settings
  piso_channel      3 solveTransientIncompressiblePiso steps
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle as O  # noqa: E402

from channel import channel_bcs  # noqa: E402


def save(name, **arrs):
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)
    print("wrote", name)


def mesh_arrays(m):
    return dict(points=m.points, faceStart=m.faceStart, faceVerts=m.faceVerts, owner=m.owner,
                neighbour=m.neighbour, patchStart=m.patchStart, patchSize=m.patchSize, nCells=m.nCells)


def main():
    if not O.have_ref():
        sys.exit("reference build missing: run `make -C oracle ref` first")
    b = O.ref_blockmesh(3, 4, 5, (0, 0, 0), (1, 2, 3))
    save("blockmesh_3x4x5", **mesh_arrays(b))

    m = O.blockmesh(6, 5, 4, (0, 0, 0), (1, 0.5, 0.5), jitter=0.15, seed=11)
    rm = O.RefMesh(m)
    g = rm.geometry()
    save("geom_jitter", **mesh_arrays(m), **{"g_" + k: v for k, v in g.items()},
         ow=rm.owner_weights())
    m.compute_geometry()

    rng = np.random.default_rng(2024)
    Ut, Uv, pt, pv, Upt, Upv, ppt, ppv = channel_bcs(m, 0.01)
    data = rng.standard_normal(m.nCells)
    save("grad_jitter", data=data, ptype=ppt, pval=ppv, grad_corr=rm.grad(data, ppt, ppv, True),
         grad_lin=rm.grad(data, ppt, ppv, False))

    gf = rng.uniform(0.5, 2.0, m.nFaces)
    extra = rng.standard_normal(m.nCells)
    u = rng.standard_normal(m.nCells)
    r = rm.laplacian_face(gf, ppt, ppv, True, True, -1, extra, 1e-10, 500, 2, np.zeros(m.nCells), u)
    save("lap_jitter", gammaFace=gf, extra=extra, u=u, ptype=ppt, pval=ppv, x=r["x"], diag=r["diag"],
         rhs=r["rhs"], offU=r["offU"], iterations=r["iterations"])

    n = 64
    rs = [0]
    cols, vals = [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                cols.append(j)
                vals.append(v)
        rs.append(len(cols))
    rs, cols, vals = np.array(rs, np.int64), np.array(cols, np.int64), np.array(vals)
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(rs[i], rs[i + 1]):
            A[i, cols[k]] = vals[k]
    bvec = A @ np.ones(n)
    out = {}
    for name, bicg in (("cg", 0), ("bicg", 1)):
        x = np.zeros(n)
        res = np.zeros(3)
        it = O.ref().cfr_solve_cg(O.C.c_int64(n), O._ptr(rs, O.C.c_int64), O._ptr(cols, O.C.c_int64), O._ptr(vals),
                                  O._ptr(bvec), O._ptr(x), O.C.c_double(1e-10), O.C.c_int(500), O.C.c_int(bicg),
                                  O._ptr(res))
        out[name + "_x"], out[name + "_iters"], out[name + "_res"] = x, it, res
    save("cg_tridiag", rowStart=rs, cols=cols, vals=vals, b=bvec, **out)

    U = rng.standard_normal((m.nCells, 3))
    p = rng.standard_normal(m.nCells)
    plc, plk = O.linearise(m, 1, pt, pv)
    gP = O.grad(m, p, plc, plk)
    rAU = rng.uniform(0.1, 1.0, m.nCells)
    fr = rm.flux_ops(U, p, gP, rAU, Upt, Upv, ppt, ppv, True)
    save("flux_jitter", U=U, p=p, gradP=gP, rAU=rAU, **fr)

    for tag, relax in (("a", (0.9, 0.03, 0.02)), ("b", (0.7, 0.3, 0.0))):
        ctlv = [10, 0, 0, 0, 1e-8, 2000, 2, relax[0], relax[1], relax[2], 0, 0, 0]
        Uo = np.zeros((m.nCells, 3))
        po = np.zeros(m.nCells)
        phio = np.zeros(m.nFaces)
        Ulc, Ulk = O.linearise(m, 3, Ut, Uv)
        O.correct_boundary_flux(m, Uo, Ulc, Ulk, phio)
        res = rm.simple(Upt, Upv, ppt, ppv, 1e-3, (0, 1, 1), ctlv, Uo, po, phio)
        save("simple_channel_" + tag, ctl=np.array(ctlv), U=Uo, p=po, phi=phio, result=res)

    pctlv = [0.01, 2, 1e-10, 1e-8, 2000, 2, 0, 0, 0]
    U0 = rng.standard_normal((m.nCells, 3)) * 0.01
    Ulc, Ulk = O.linearise(m, 3, Ut, Uv)
    phi0 = np.zeros(m.nFaces)
    O.update_phi_internal(m, U0, phi0)
    O.correct_boundary_flux(m, U0, Ulc, Ulk, phi0)
    Us, ps, phis = U0.copy(), np.zeros(m.nCells), phi0.copy()
    steps = []
    for _ in range(3):
        steps.append(rm.piso_step(Upt, Upv, ppt, ppv, 1e-3, (0, 1, 1), pctlv, Us, ps, phis))
    save("piso_channel", ctl=np.array(pctlv), U0=U0, phi0=phi0, U=Us, p=ps, phi=phis, results=np.array(steps))


if __name__ == "__main__":
    main()
