"""Multigrid / flexible-CG oracle: a CPU restatement (numpy + scipy, FP64)
of the aggregation hierarchy, the Galerkin coarse operators, the V-cycle
with per-row finest-level weights, and flexible CG.

TEST INFRASTRUCTURE ONLY.  Importable from tests/, from
``__graft_entry__.smoke()`` and from bench.py's cpu_baseline leg, as the
checker; never from the product package.

The multigrid preconditioner has no counterpart in the reference, whose
solvers are Jacobi-PCG and Jacobi-BiCGStab (src/common/src/Matrix.cpp:153-319);
its stop test (||r||inf / max(||b||inf, 1e-300) < tol, Matrix.cpp:176-186) is
kept.  Pinned by: aggregation invariants (a partition into connected
aggregates, every phase-1 root owning its whole neighbourhood); Galerkin
values = scipy's P^T A P; the V-cycle against its closed-form error
propagation (I - Wpost A)(I - S P Ac^-1 P^T A)(I - Wpre A) with an exact
coarse solve; the coarsest-level sweeps against their polynomial closed form;
the Jacobi limit a = 1, s = 0, c = 0 giving z = D^-1 r; and flexible CG against
plain PCG for a fixed SPD preconditioner (identical iterates in exact
arithmetic).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp


def aggregate(rp, col):
    """Greedy aggregation of a symmetric pattern (three phases):
    (1) a row whose whole neighbourhood is unaggregated becomes a root and
        takes that neighbourhood;
    (2) each still-free row joins the phase-1 aggregate of its first
        (lowest slot) aggregated neighbour;
    (3) each remaining row starts an aggregate with its free neighbours.
    Returns (agg, nAggregates)."""
    rp = np.asarray(rp)
    col = np.asarray(col)
    n = len(rp) - 1
    agg = np.full(n, -1, np.int64)
    na = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        nb = col[rp[i]:rp[i + 1]]
        if np.any(agg[nb] >= 0):
            continue
        agg[nb] = na
        agg[i] = na
        na += 1
    a2 = agg.copy()
    for i in range(n):
        if agg[i] >= 0:
            continue
        for j in col[rp[i]:rp[i + 1]]:
            if agg[j] >= 0:
                a2[i] = agg[j]
                break
    agg = a2
    for i in range(n):
        if agg[i] >= 0:
            continue
        agg[i] = na
        for j in col[rp[i]:rp[i + 1]]:
            if agg[j] < 0:
                agg[j] = na
        na += 1
    return agg, na


def prolongation(agg, na):
    n = len(agg)
    return sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, na))


@dataclass
class Params:
    """Finest-level a (pre), s (scale), c (post) constants, coarse-level
    constants (pre = post weight), coarsest sweeps, hierarchy caps."""
    omegaPre: float = 0.800000011920929    # the defaults of the device cycle (FP32 constants)
    scale: float = 1.7999999523162842
    omegaPost: float = 0.800000011920929
    restrictScale: float = 1.7999999523162842
    coarseOmega: float = 0.8999999761581421
    coarseScale: float = 3.0
    coarseSweeps: int = 30
    coarsestOmega: float = 2.0 / 3.0
    maxLevels: int = 10
    coarseMax: int = 128
    kcycleDepth: int = 2
    smoothFine: int = 4     # smoothing sweeps before and after the coarse correction on the finest level,
    smoothCoarse: int = 2   # on level 1 of the K-cycle levels,
    smoothDeep: int = 2     # on the K-cycle levels below level 1 (1 below them)


DENSE_MAX = 128


@dataclass
class Hierarchy:
    A: list = field(default_factory=list)      # scipy CSR per level
    agg: list = field(default_factory=list)    # agg[l]: level-(l+1) row of each level-l row
    P: list = field(default_factory=list)


def hierarchy(A, maxLevels=10, coarseMax=128):
    """Aggregation hierarchy.  Each level is aggregated with its rows numbered
    in the order the aggregates of the level below were made; then every coarse
    level above the coarsest is renumbered aggregate by aggregate (rows sorted
    by their aggregate, ties by their number), from the top down, so that each
    level's rows follow the final numbers of the level above.  The renumbering
    permutes coarse rows only: the operators are the same up to that
    permutation."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    levels, aggs = [A], []
    while A.shape[0] > coarseMax and len(levels) < maxLevels:
        agg, na = aggregate(A.indptr, A.indices)
        if na >= A.shape[0]:
            break
        Ac = sp.csr_matrix(prolongation(agg, na).T @ A @ prolongation(agg, na))
        Ac.sort_indices()
        aggs.append(agg)
        levels.append(Ac)
        A = Ac
    last = len(levels) - 1
    for l in range(last - 1, 0, -1):
        order = np.argsort(aggs[l], kind="stable")
        pi = np.empty_like(order)
        pi[order] = np.arange(len(order))
        aggs[l - 1] = pi[aggs[l - 1]]
        aggs[l] = aggs[l][order]
    H = Hierarchy(A=[levels[0]])
    for l in range(last):
        P = prolongation(aggs[l], levels[l + 1].shape[0])
        Ac = sp.csr_matrix(P.T @ H.A[l] @ P)
        Ac.sort_indices()
        H.agg.append(aggs[l])
        H.P.append(P)
        H.A.append(Ac)
    return H


def dense_coarse_solve(Ac, b):
    """Exact coarsest solve; a pure Neumann operator (all row sums below
    1e-9 max|diag|) is solved as A + gamma 1 1^T, gamma = mean(diag) / n,
    i.e. the pseudo-inverse on zero-mean right-hand sides."""
    M = Ac.toarray()
    n = M.shape[0]
    d = np.diag(M)
    if np.abs(M.sum(axis=1)).max() <= 1e-9 * np.abs(d).max():
        M = M + d.sum() / n / n
    return np.linalg.solve(M, b)


def coarsest(Ac, b, sweeps, omega, exact=False):
    if exact:
        return dense_coarse_solve(Ac, b)
    dinv = 1.0 / Ac.diagonal()
    x = np.zeros_like(b)
    for _ in range(sweeps):
        x = x + omega * dinv * (b - Ac @ x)
    return x


def solve_level(H, b, p: Params, level, exact_coarse=False):
    """Coarse correction on level >= 1: exact / Jacobi on the coarsest
    level, the K-cycle (two flexible-CG steps preconditioned by the level's
    cycle) on levels 1..kcycleDepth, one cycle below."""
    A = H.A[level]
    if level == len(H.A) - 1:
        return coarsest(A, b, p.coarseSweeps, p.coarsestOmega, exact_coarse or A.shape[0] <= DENSE_MAX)
    M = lambda v: vcycle(H, v, p, exact_coarse=exact_coarse, level=level)
    if level > p.kcycleDepth:
        return M(b)
    c1 = M(b)
    v1 = A @ c1
    rho1, a1 = c1 @ v1, c1 @ b
    r2 = b - (a1 / rho1 if rho1 > 0 else 0.0) * v1
    c2 = M(r2)
    v2 = A @ c2
    g, be, a2 = c2 @ v1, c2 @ v2, c2 @ r2
    k1, k2 = 1.0, 0.0
    if rho1 > 0:
        rho2 = be - g * g / rho1
        if rho2 > 0:
            k1, k2 = a1 / rho1 - g * a2 / (rho1 * rho2), a2 / rho2
        else:
            k1 = a1 / rho1
    return k1 * c1 + k2 * c2


def vcycle(H, r, p: Params, pre=None, scale=None, post=None, exact_coarse=False, level=0, restrict=None):
    """z = M r: x = Wpre b, then (sweeps - 1) times x += Wpre (b - A x);
    r = b - A x; bc = P^T T r; e = coarse correction (solve_level);
    x += S P e; then sweeps times x += Wpost (b - A x); z = x.  sweeps is
    smoothFine on the finest level, smoothCoarse on level 1 and smoothDeep on
    the further K-cycle levels, 1 below them.  pre/scale/post/
    restrict are per-row finest-level a, s, c, t (None -> constant)."""
    A = H.A[level]
    b = np.asarray(r, np.float64)
    last = len(H.A) - 1
    dinv = 1.0 / A.diagonal()
    if level == 0:
        a = p.omegaPre if pre is None else np.asarray(pre, np.float64)
        s = p.scale if scale is None else np.asarray(scale, np.float64)
        c = p.omegaPost if post is None else np.asarray(post, np.float64)
        t = p.restrictScale if restrict is None else np.asarray(restrict, np.float64)
    else:
        a = c = p.coarseOmega
        s = p.coarseScale
        t = 1.0
    sweeps = p.smoothFine if level == 0 else (
        1 if level > p.kcycleDepth else (p.smoothCoarse if level == 1 else p.smoothDeep))
    x = a * dinv * b
    for _ in range(sweeps - 1):
        x = x + a * dinv * (b - A @ x)
    if level < last:
        res = b - A @ x
        agg = H.agg[level]
        bc = np.bincount(agg, weights=t * res, minlength=H.A[level + 1].shape[0])
        ec = solve_level(H, bc, p, level + 1, exact_coarse)
        x = x + s * ec[agg]
    for _ in range(sweeps):
        x = x + c * dinv * (b - A @ x)
    return x


def fcg(A, b, precond, tol=1e-8, maxIter=1000, x0=None):
    """Flexible CG, FCG(1): alpha = p.r / p.q, beta = -z.q / p.q, with the
    reference's stop test on the recurrence residual.  Returns (x, iters,
    converged)."""
    A = sp.csr_matrix(A)
    x = np.zeros(len(b)) if x0 is None else np.array(x0, np.float64)
    r = b - A @ x
    bnorm = max(np.abs(b).max(), 1e-300)
    if np.abs(r).max() / bnorm < tol:
        return x, 0, True
    z = precond(r)
    p = z.copy()
    for k in range(1, maxIter + 1):
        q = A @ p
        pq = p @ q
        alpha = (p @ r) / pq
        x += alpha * p
        r -= alpha * q
        if np.abs(r).max() / bnorm < tol:
            return x, k, True
        z = precond(r)
        beta = -(z @ q) / pq
        p = z + beta * p
    return x, maxIter, False


def pcg(A, b, precond, tol=1e-8, maxIter=1000):
    """Standard PCG (beta = r.z / r_old.z_old) for the FCG comparison."""
    A = sp.csr_matrix(A)
    x = np.zeros(len(b))
    r = b.copy()
    bnorm = max(np.abs(b).max(), 1e-300)
    z = precond(r)
    p = z.copy()
    rz = r @ z
    for k in range(1, maxIter + 1):
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if np.abs(r).max() / bnorm < tol:
            return x, k, True
        z = precond(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, maxIter, False
