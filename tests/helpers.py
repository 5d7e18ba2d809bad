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
