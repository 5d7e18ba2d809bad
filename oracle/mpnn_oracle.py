"""Message-passing network and multigrid training objective: a CPU FP64
restatement (numpy + scipy) of the network's forward and backward pass and
of the multigrid training objective (the A-norm error left by one
preconditioned step with the optimal step length) with its adjoint.

TEST INFRASTRUCTURE ONLY.  Importable from tests/, from
``__graft_entry__.smoke()`` and from bench.py's cpu_baseline leg, as the
checker; never from the product package.

No counterpart in the reference (its preconditioner is Jacobi,
src/common/src/Matrix.cpp:160-165).  Pinned by central differences of its
own forward pass and objective (FP64, relative 1e-6), by the closed-form
features of an interior 7-point Laplacian row, and by the untrained network
giving the tuned constants a = 0.8, t = 1.6.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

import mg_oracle as MG

H, FIN, L, NOUT = 32, 8, 3, 2
K = 3 * H
AMAX = 1.600000023841858  # 1.6 and 3.2 as the FP32 constants of the kernels
TMAX = 3.200000047683716
OFF_WIN, OFF_BIN = 0, FIN * H
OFF_L0 = OFF_BIN + H
LAYER = K * H + H
OFF_WOUT = OFF_L0 + L * LAYER
OFF_BOUT = OFF_WOUT + H * NOUT
NPARAMS = OFF_BOUT + NOUT


def unpack(p):
    p = np.asarray(p, np.float64)
    d = {"Win": p[OFF_WIN:OFF_BIN].reshape(FIN, H), "bin": p[OFF_BIN:OFF_L0]}
    for l in range(L):
        o = OFF_L0 + l * LAYER
        d[f"W{l}"] = p[o:o + K * H].reshape(K, H)
        d[f"b{l}"] = p[o + K * H:o + LAYER]
    d["Wout"] = p[OFF_WOUT:OFF_BOUT].reshape(H, NOUT)
    d["bout"] = p[OFF_BOUT:NPARAMS]
    return d


def pack(d):
    parts = [d["Win"].ravel(), d["bin"]]
    for l in range(L):
        parts += [d[f"W{l}"].ravel(), d[f"b{l}"]]
    parts += [d["Wout"].ravel(), d["bout"]]
    return np.concatenate(parts)


def random_params(seed, out_scale=0.5):
    rng = np.random.default_rng(seed)
    d = {"Win": rng.uniform(-0.7, 0.7, (FIN, H)), "bin": rng.uniform(-0.1, 0.1, H)}
    for l in range(L):
        d[f"W{l}"] = rng.uniform(-0.25, 0.25, (K, H))
        d[f"b{l}"] = rng.uniform(-0.1, 0.1, H)
    d["Wout"] = rng.uniform(-out_scale, out_scale, (H, NOUT))
    d["bout"] = rng.uniform(-0.2, 0.2, NOUT)
    return pack(d)


def features(A, geo=None, r0=None):
    """Node features x0 [n, 8] and edge features ew, eg (one per stored
    entry of A, 0 on the diagonal)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    d = np.abs(A.diagonal())
    d = np.where(d > 1e-300, d, 1e-300)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    cols = A.indices
    off = rows != cols
    ew = np.zeros(A.nnz)
    ew[off] = np.abs(A.data[off]) / np.sqrt(d[rows[off]] * d[cols[off]])
    eg = np.zeros(A.nnz) if geo is None else np.where(off, np.asarray(geo, np.float64), 0.0)
    x = np.zeros((n, FIN))
    deg = np.bincount(rows[off], minlength=n)
    x[:, 0] = np.bincount(rows[off], weights=np.abs(A.data[off]), minlength=n) / d
    x[:, 1] = np.bincount(rows, weights=A.data, minlength=n) / d
    x[:, 2] = deg / 8.0
    wmax = np.zeros(n)
    wmin = np.full(n, np.inf)
    np.maximum.at(wmax, rows[off], ew[off])
    np.minimum.at(wmin, rows[off], ew[off])
    x[:, 3] = wmax
    x[:, 4] = np.where(deg > 0, wmin, 0.0)
    x[:, 5] = np.where(deg > 0, np.bincount(rows[off], weights=eg[off], minlength=n) / np.maximum(deg, 1), 0.0)
    if r0 is not None and np.abs(r0).max() > 0:
        x[:, 6] = np.abs(r0) / np.abs(r0).max()
    x[:, 7] = 1.0
    return x, ew, eg


def edge_matrices(A, ew, eg):
    A = sp.csr_matrix(A)
    shape = A.shape
    We = sp.csr_matrix((ew, A.indices, A.indptr), shape=shape)
    Ge = sp.csr_matrix((eg, A.indices, A.indptr), shape=shape)
    return We, Ge


def sigmoid(v):
    return 1.0 / (1.0 + np.exp(-v))


def forward(params, A, x0, ew, eg):
    """Returns (a, t, cache)."""
    P = unpack(params)
    We, Ge = edge_matrices(A, ew, eg)
    u0 = x0 @ P["Win"] + P["bin"]
    h = np.maximum(u0, 0.0)
    cache = {"x0": x0, "u0": u0, "We": We, "Ge": Ge, "X": [], "U": []}
    for l in range(L):
        X = np.hstack([h, We @ h, Ge @ h])
        U = X @ P[f"W{l}"] + P[f"b{l}"]
        cache["X"].append(X)
        cache["U"].append(U)
        h = h + np.maximum(U, 0.0)
    o = h @ P["Wout"] + P["bout"]
    cache["h"] = h
    cache["o"] = o
    return AMAX * sigmoid(o[:, 0]), TMAX * sigmoid(o[:, 1]), cache


def backward(params, cache, ga, gt):
    """dJ/dparams for J with dJ/da = ga, dJ/dt = gt."""
    P = unpack(params)
    o = cache["o"]
    s0, s1 = sigmoid(o[:, 0]), sigmoid(o[:, 1])
    go = np.column_stack([ga * AMAX * s0 * (1 - s0), gt * TMAX * s1 * (1 - s1)])
    G = {k: np.zeros_like(v) for k, v in P.items()}
    G["Wout"] = cache["h"].T @ go
    G["bout"] = go.sum(axis=0)
    dh = go @ P["Wout"].T
    We, Ge = cache["We"], cache["Ge"]
    for l in reversed(range(L)):
        X, U = cache["X"][l], cache["U"][l]
        du = dh * (U > 0)
        G[f"W{l}"] = X.T @ du
        G[f"b{l}"] = du.sum(axis=0)
        dX = du @ P[f"W{l}"].T
        dh = dh + dX[:, :H] + We.T @ dX[:, H:2 * H] + Ge.T @ dX[:, 2 * H:]
    du0 = dh * (cache["u0"] > 0)
    G["Win"] = cache["x0"].T @ du0
    G["bin"] = du0.sum(axis=0)
    return pack(G)


def objective(H_, e, a, t, p: MG.Params):
    """L = 1 - (z.A e)^2 / ((e.A e)(z.A z)), z = M A e, for the cycle with
    tied finest-level weights (pre = post = a, restriction = prolongation =
    t), p.smoothFine smoothing sweeps on each side and the cycle of p below
    level 1 (the K-cycle on levels 1..kcycleDepth): the A-norm error left by
    one preconditioned step with the optimal step length,
    min_alpha |e - alpha z|_A^2 / |e|_A^2."""
    q = p
    A = H_.A[0]
    b = A @ e
    z = MG.vcycle(H_, b, q, pre=a, scale=t, post=a, restrict=t)
    alpha = (z @ b) / (z @ (A @ z))
    eps = e - alpha * z
    return (eps @ (A @ eps)) / (e @ b)


def objective_grad(H_, e, a, t, p: MG.Params):
    """(L, dL/da, dL/dt) by the adjoint of the finest level: A is symmetric
    and W = a D^-1 diagonal, so each sweep x' = x + W (b - A x) passes a
    gradient g back as g - A W g and adds g * D^-1 (b - A x) to dL/da; the
    coarse correction Q = P (coarse cycle) P^T is taken as symmetric and
    linear, which it is for the V-cycle (kcycleDepth = 0) and approximately
    is for the K-cycle (a nearly exact coarse solve), applied to the
    gradient as it is to the residual."""
    q = p
    k = max(1, q.smoothFine)
    A = H_.A[0]
    dinv = 1.0 / A.diagonal()
    agg = H_.agg[0]
    nc = H_.A[1].shape[0]
    Q = lambda v: MG.solve_level(H_, np.bincount(agg, weights=v, minlength=nc), q, 1)[agg]
    b = A @ e
    db = dinv * b
    x = a * db
    pre = []  # D^-1 (b - A x) of the pre-smoothing sweeps 2..k
    for _ in range(k - 1):
        pj = dinv * (b - A @ x)
        pre.append(pj)
        x = x + a * pj
    r1 = b - A @ x
    u = Q(t * r1)
    y = x + t * u
    post = []  # D^-1 (b - A y) of the post-smoothing sweeps
    for _ in range(k):
        pj = dinv * (b - A @ y)
        post.append(pj)
        y = y + a * pj
    z = y
    Az = A @ z
    c, pz, qz = e @ b, z @ b, z @ Az
    Lval = 1.0 - pz * pz / (c * qz)
    g = (2.0 * pz / (c * qz)) * (-b + (pz / qz) * Az)
    ga = np.zeros_like(a)
    for pj in reversed(post):
        ga = ga + g * pj
        g = g - A @ (dinv * (a * g))
    gt = g * u
    py = Q(t * g)
    gt = gt + r1 * py
    g = g - A @ (t * py)
    for pj in reversed(pre):
        ga = ga + g * pj
        g = g - A @ (dinv * (a * g))
    ga = ga + g * db
    return Lval, ga, gt
