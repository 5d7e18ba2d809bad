"""Detail-region oracle: inlet profile, boundary conditions and section
measurement of the reference's DetailBlock, restated in numpy (FP64); the
SIMPLE solve itself is the oracle's C restatement (oracle.simple).

TEST INFRASTRUCTURE ONLY.  Importable from tests/, from
``__graft_entry__.smoke()`` and from bench.py's cpu_baseline leg, as the
checker; never from the product package.

Follows src/coupling/src/coupledNetwork.cpp: patchAreaVector :30-39, the
flow axis from the inlet area vector :532-551, buildSections :644-671,
buildInletProfile :768-876, applyBoundary :878-926, measure :1154-1200.
The reference's coupling library does not build from a few of its own
files (4,981 lines with turbulence, phase and energy models), so these are
pinned by closed forms instead: flat inlets carry no weights, parabolic
weights are prod(1 - eta^2) up to one normalisation whose area integral is
the section area, the inlet flux equals the imposed rate, and a uniform
grid's sections are its cell slabs.
"""
from __future__ import annotations

import numpy as np


def patch_area(Sf, start, size):
    return Sf[start:start + size].sum(axis=0)


def flow_axis(Sf, patch_start, patch_size, types):
    pin = types.index("inlet")
    pout = types.index("outlet")
    a_in = patch_area(Sf, patch_start[pin], patch_size[pin])
    area = np.linalg.norm(a_in)
    nrm = a_in / area
    axis = 0
    if abs(nrm[1]) > abs(nrm[axis]):
        axis = 1
    if abs(nrm[2]) > abs(nrm[axis]):
        axis = 2
    return pin, pout, area, axis, (1 if nrm[axis] < 0 else -1)


def sections(C, points, axis, flow_sign):
    lo, hi = points[:, axis].min(), points[:, axis].max()
    band = 1e-9 * max(abs(hi - lo), 1e-30)
    n = len(C)
    order = sorted(range(n), key=lambda c: (C[c, axis], c))
    if flow_sign < 0:
        order = order[::-1]
    out, g = [], 0
    while g < n:
        anchor = C[order[g], axis]
        h = g + 1
        while h < n and abs(C[order[h], axis] - anchor) <= band:
            h += 1
        out.append(order[g:h])
        g = h
    return out


def inlet_profile(points, Sf, Cf, magSf, patch_start, patch_size, types, axis, pin, area, shape):
    if shape == "flat":
        return np.zeros(0, np.int64), np.zeros(0)
    bounds = np.zeros((3, 2), bool)
    for p, t in enumerate(types):
        if t != "wall":
            continue
        sf = patch_area(Sf, patch_start[p], patch_size[p])
        if not np.linalg.norm(sf) > 0:
            continue
        k = 0
        for c in (1, 2):
            if abs(sf[c]) > abs(sf[k]):
                k = c
        if k == axis:
            continue
        bounds[k, 0 if sf[k] < 0 else 1] = True
    axes = [k for k in range(3) if k != axis and bounds[k].all()]
    lo, hi = points.min(axis=0), points.max(axis=0)
    f = np.arange(patch_start[pin], patch_start[pin] + patch_size[pin])
    w = np.ones(len(f))
    for k in axes:
        mid, half = 0.5 * (hi[k] + lo[k]), 0.5 * (hi[k] - lo[k])
        eta = (Cf[f, k] - mid) / half
        w = w * ((1.0 - eta * eta) if shape == "parabolic" else (1.0 - np.abs(eta)) ** (1.0 / 7.0))
    a = magSf[f]
    norm = (w * a).sum() / a.sum()
    return f, w / norm


def boundary(q, n_internal, patch_start, patch_size, types, axis, flow_sign, area, pin, pout, faces, weights):
    nB = patch_start[-1] + patch_size[-1] - n_internal
    Ut = np.full(nB, 2, np.int32)
    pt = np.full(nB, 2, np.int32)
    Uv = np.zeros((nB, 3))
    pv = np.zeros((nB, 1))
    speed = flow_sign * q / area
    for p, t in enumerate(types):
        b = np.arange(patch_start[p], patch_start[p] + patch_size[p]) - n_internal
        if p == pin:
            Ut[b] = 0
            Uv[b, axis] = speed
        elif p == pout:
            pt[b] = 0
        elif t == "wall":
            Ut[b] = 0
    if len(faces):
        Uv[faces - n_internal, axis] = speed * weights
    return Ut, Uv, pt, pv


def measure(p, phi, V, secs, rho, patch_start, patch_size, pin, pout):
    def mean(cells):
        c = np.asarray(cells)
        return rho * (p[c] * V[c]).sum() / V[c].sum()
    dp = mean(secs[0]) - mean(secs[-1])
    fin = phi[patch_start[pin]:patch_start[pin] + patch_size[pin]].sum()
    fout = phi[patch_start[pout]:patch_start[pout] + patch_size[pout]].sum()
    return dp, -fin, fout
