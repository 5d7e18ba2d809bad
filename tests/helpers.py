"""Shared test helpers: the same mesh as an oracle Mesh, a product HostMesh
and a device Mesh."""
import numpy as np

import oracle as O


def oracle_mesh_from_arrays(a) -> "O.Mesh":
    m = O.Mesh(np.ascontiguousarray(a["points"], np.float64), np.ascontiguousarray(a["faceStart"], np.int64),
               np.ascontiguousarray(a["faceVerts"], np.int64), np.ascontiguousarray(a["owner"], np.int64),
               np.ascontiguousarray(a["neighbour"], np.int64), np.ascontiguousarray(a["patchStart"], np.int64),
               np.ascontiguousarray(a["patchSize"], np.int64), int(a["nCells"]))
    return m.compute_geometry()


def host_and_oracle(nx, ny, nz, mn=(0, 0, 0), mx=(1, 1, 1), jitter=0.0, seed=0, tets=False, shuffle=None,
                    rcm=False):
    from chuanflow_amd import native as N
    hm = N.HostMesh.box(nx, ny, nz, mn, mx, jitter, seed, tets)
    if shuffle is not None:
        hm.shuffle(shuffle)
    if rcm:
        hm.renumber_rcm()
    a = hm.arrays()
    return hm, a, oracle_mesh_from_arrays(a)


def soa(v):
    """[n][3] -> [3][n] contiguous"""
    return np.ascontiguousarray(np.asarray(v).T)


def aos(v):
    return np.ascontiguousarray(np.asarray(v).reshape(3, -1).T)


def write_gmsh_tet_channel(path, nx, ny, nz, L=(1.0, 0.25, 0.25), jitter=0.1, seed=0):
    """An ASCII Gmsh 2.2 file of a box channel along x: an nx*ny*nz hex grid,
    interior nodes jittered by up to `jitter` of the spacing, each hex split
    into the 6 Kuhn tets about its (0,0,0)-(1,1,1) diagonal (a conforming
    split), tets positively oriented; boundary triangles in the physical
    surfaces "inlet" (x = 0), "outlet" (x = L) and "walls", the volume in
    "fluid"."""
    rng = np.random.default_rng(seed)
    h = np.array(L) / np.array([nx, ny, nz])
    idx = lambda i, j, k: (k * (ny + 1) + j) * (nx + 1) + i
    pts = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                p = np.array([i, j, k], float) * h
                inner = np.array([0 < i < nx, 0 < j < ny, 0 < k < nz])
                p += np.where(inner, rng.uniform(-jitter, jitter, 3) * h, 0.0)
                pts[idx(i, j, k)] = p
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    kuhn = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    tets, tris = [], []
    faces = {"x0": [(0, 3, 7), (0, 7, 4)], "x1": [(1, 2, 6), (1, 6, 5)], "y0": [(0, 1, 5), (0, 5, 4)],
             "y1": [(2, 3, 6), (3, 7, 6)], "z0": [(0, 1, 2), (0, 2, 3)], "z1": [(4, 5, 6), (4, 6, 7)]}
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                v = [idx(i + a, j + b, k + c) for a, b, c in corner]
                for t in kuhn:
                    a, b, c, d = (v[q] for q in t)
                    if np.linalg.det(np.array([pts[b] - pts[a], pts[c] - pts[a], pts[d] - pts[a]])) < 0:
                        b, c = c, b
                    tets.append((a, b, c, d))
                for name, on in (("x0", i == 0), ("x1", i == nx - 1), ("y0", j == 0), ("y1", j == ny - 1),
                                 ("z0", k == 0), ("z1", k == nz - 1)):
                    if on:
                        group = 1 if name == "x0" else 2 if name == "x1" else 3
                        tris += [(group, tuple(v[q] for q in tri)) for tri in faces[name]]
    out = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "4",
           '2 1 "inlet"', '2 2 "outlet"', '2 3 "walls"', '3 4 "fluid"', "$EndPhysicalNames",
           "$Nodes", str(len(pts))]
    out += [f"{q + 1} {x:.17g} {y:.17g} {z:.17g}" for q, (x, y, z) in enumerate(pts)]
    out += ["$EndNodes", "$Elements", str(len(tris) + len(tets))]
    e = 1
    for g, tri in tris:
        out.append(f"{e} 2 2 {g} {g} " + " ".join(str(q + 1) for q in tri))
        e += 1
    for t in tets:
        out.append(f"{e} 4 2 4 4 " + " ".join(str(q + 1) for q in t))
        e += 1
    out.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
