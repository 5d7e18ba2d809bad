"""The network / objective oracle pinned by FP64 central differences and
closed forms (no counterpart in the reference)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mg_oracle as MG
import mpnn_oracle as NN
from test_oracle_mg import laplacian3d


def small_graph(seed=0):
    A = laplacian3d(5, 4, 3, seed=seed)
    rng = np.random.default_rng(seed + 100)
    # symmetric per-face geometry on the stored entries, 0 on the diagonal
    n = A.shape[0]
    R = rng.uniform(0.0, 0.3, (n, n))
    R = 0.5 * (R + R.T)
    np.fill_diagonal(R, 0.0)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    geo = R[rows, A.indices]
    r0 = rng.standard_normal(A.shape[0])
    return A, geo, r0


def test_interior_laplacian_row_features_closed_form():
    # unit 7-point stencil: interior rows have sum|a_ij|/a_ii = 1, zero row sum,
    # 6 neighbours, w_ij = 1/6
    n = 5
    A = laplacian3d(n, n, n, seed=0, dirichlet=False)
    A.data[:] = np.where(A.data < 0, -1.0, 0.0)
    A = A - sp.diags(np.asarray(A.sum(axis=1)).ravel())
    A = sp.csr_matrix(A)
    A.sort_indices()
    x, ew, _ = NN.features(A)
    i = 2 + n * (2 + n * 2)  # centre cell
    np.testing.assert_allclose(x[i], [1.0, 0.0, 6 / 8, 1 / 6, 1 / 6, 0.0, 0.0, 1.0], atol=1e-15)
    assert np.all(ew[A.indptr[i]:A.indptr[i + 1]][A.indices[A.indptr[i]:A.indptr[i + 1]] != i] == 1 / 6)


def test_untrained_network_gives_the_head_centre():
    """A zero output layer gives a = 0.8 (the multigrid's tuned smoother
    weight) and t = 1.6, the transfer scale the network was trained around."""
    A, geo, r0 = small_graph(1)
    p = NN.random_params(3)
    d = NN.unpack(p)
    d["Wout"][:] = 0
    d["bout"][:] = 0
    a, t, _ = NN.forward(NN.pack(d), A, *NN.features(A, geo, r0))
    np.testing.assert_allclose(a, 0.8, rtol=1e-7)
    np.testing.assert_allclose(t, 1.6, rtol=1e-7)


def test_network_backward_matches_central_differences_fp64():
    A, geo, r0 = small_graph(2)
    x0, ew, eg = NN.features(A, geo, r0)
    p = NN.random_params(4)
    rng = np.random.default_rng(5)
    al, be = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    J = lambda q: (lambda a, t: al @ a + be @ t)(*NN.forward(q, A, x0, ew, eg)[:2])
    _, _, cache = NN.forward(p, A, x0, ew, eg)
    g = NN.backward(p, cache, al, be)
    idx = rng.choice(NN.NPARAMS, 200, replace=False)
    idx = np.union1d(idx, [NN.OFF_BOUT, NN.OFF_BOUT + 1, NN.OFF_BIN, NN.OFF_L0 + NN.K * NN.H])
    h = 1e-6
    fd = np.array([(J(p + h * np.eye(1, NN.NPARAMS, k)[0]) - J(p - h * np.eye(1, NN.NPARAMS, k)[0])) / (2 * h)
                   for k in idx])
    assert np.linalg.norm(g[idx] - fd) <= 1e-6 * np.linalg.norm(fd)


@pytest.mark.parametrize("sweeps", [1, 3])
def test_objective_gradient_matches_central_differences_fp64(sweeps):
    A = laplacian3d(9, 8, 7, seed=6)
    Hh = MG.hierarchy(A, coarseMax=30)
    assert len(Hh.A) >= 3
    n = A.shape[0]
    rng = np.random.default_rng(7)
    a = rng.uniform(0.5, 0.8, n)
    t = rng.uniform(1.1, 1.5, n)
    e = rng.standard_normal(n)
    # the V-cycle below level 1: the coarse correction is linear, so the adjoint is the exact gradient
    p = MG.Params(coarseScale=1.6, smoothFine=sweeps, kcycleDepth=0)
    Lv, ga, gt = NN.objective_grad(Hh, e, a, t, p)
    # L = 1 - ratio: its rounding is absolute, of the order of the ratio's (1e-16), however small L is
    assert abs(Lv - NN.objective(Hh, e, a, t, p)) <= 1e-14
    h = 1e-6
    for g, which in ((ga, "a"), (gt, "t")):
        idx = rng.choice(n, 40, replace=False)
        fd = []
        for k in idx:
            d = np.zeros(n)
            d[k] = h
            if which == "a":
                fd.append((NN.objective(Hh, e, a + d, t, p) - NN.objective(Hh, e, a - d, t, p)) / (2 * h))
            else:
                fd.append((NN.objective(Hh, e, a, t + d, p) - NN.objective(Hh, e, a, t - d, p)) / (2 * h))
        fd = np.array(fd)
        assert np.linalg.norm(g[idx] - fd) <= 1e-6 * np.linalg.norm(fd), which
