"""Boundary conditions of the laminar channel case used across the tests.

Box patches in BlockMesh order (xmin xmax ymin ymax zmin zmax):
xmin inlet  U fixedValue (Uin, 0, 0), p zeroGradient;
xmax outlet U zeroGradient,           p fixedValue 0;
four walls  U fixedValue 0,            p zeroGradient.
Returned per boundary face (types, values) and per patch (for the
reference build's per-patch API).  Type codes: 0 fixedValue,
1 fixedGradient, 2 zeroGradient, 3 symmetry.
"""
import numpy as np


def patch_of_boundary_face(patchStart, patchSize, nInternal, nB):
    out = np.empty(nB, np.int64)
    for p, (s, n) in enumerate(zip(patchStart, patchSize)):
        out[s - nInternal:s - nInternal + n] = p
    return out


def channel_bcs(m, Uin=0.01):
    pid = patch_of_boundary_face(m.patchStart, m.patchSize, m.nInternal, m.nFaces - m.nInternal)
    Ut = np.where(pid == 1, 2, 0).astype(np.int32)
    Uv = np.zeros((len(pid), 3))
    Uv[pid == 0, 0] = Uin
    pt = np.where(pid == 1, 0, 2).astype(np.int32)
    pv = np.zeros((len(pid), 1))
    Upt = np.array([0, 2, 0, 0, 0, 0], np.int32)
    Upv = np.zeros((6, 3))
    Upv[0, 0] = Uin
    ppt = np.array([2, 0, 2, 2, 2, 2], np.int32)
    ppv = np.zeros(6)
    return Ut, Uv, pt, pv, Upt, Upv, ppt, ppv


def neumann_bcs(nB):
    """All-zeroGradient pressure (pure Neumann)."""
    return np.full(nB, 2, np.int32), np.zeros((nB, 1))
