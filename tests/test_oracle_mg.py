"""The multigrid / flexible-CG oracle pinned by invariants and closed forms
(the preconditioner has no counterpart in the reference)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mg_oracle as MG


def laplacian3d(nx, ny, nz, seed=0, dirichlet=True):
    """7-point FV Laplacian with random positive face conductances; a
    Dirichlet face on x = 0 makes it SPD."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for axis in range(3):
        a = np.take(idx, np.arange(idx.shape[2 - axis] - 1), axis=2 - axis).ravel()
        b = np.take(idx, np.arange(1, idx.shape[2 - axis]), axis=2 - axis).ravel()
        g = rng.uniform(0.5, 2.0, len(a))
        rows += [a, b]
        cols += [b, a]
        vals += [-g, -g]
        np.add.at(diag, a, g)
        np.add.at(diag, b, g)
    if dirichlet:
        diag[idx[:, :, 0].ravel()] += 2.0
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sort_indices()
    return A


def test_aggregation_is_a_partition_into_connected_aggregates():
    A = laplacian3d(9, 8, 7, seed=1)
    agg, na = MG.aggregate(A.indptr, A.indices)
    assert agg.min() == 0 and agg.max() == na - 1
    assert len(np.unique(agg)) == na
    G = sp.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
    for c in range(na):
        m = np.flatnonzero(agg == c)
        sub = G[m][:, m]
        ncomp, _ = sp.csgraph.connected_components(sub, directed=False)
        assert ncomp == 1, f"aggregate {c} is not connected"
    # coarsening: a 7-point stencil gives aggregates of up to 7 rows
    assert A.shape[0] / na > 3.0


def test_phase_one_roots_own_their_whole_neighbourhood():
    A = laplacian3d(6, 6, 6, seed=2)
    agg, _ = MG.aggregate(A.indptr, A.indices)
    # row 0 is always a root: it and all its neighbours share aggregate 0
    nb = A.indices[A.indptr[0]:A.indptr[1]]
    assert np.all(agg[nb] == 0)


def test_coarse_levels_are_numbered_aggregate_by_aggregate_and_equal_the_plain_hierarchy():
    """Every coarse level above the coarsest lists its rows aggregate by
    aggregate, in the order of the level above (so a block of consecutive rows
    restricts to a consecutive range of coarse rows); the renumbering is a
    permutation of each coarse level: with Q_l the permutation matrix, the
    operators are Q_l A_l Q_l^T of the hierarchy numbered as its aggregates
    were made (aggregation and Galerkin products repeated here without it)."""
    A = laplacian3d(24, 22, 20, seed=5)
    H = MG.hierarchy(A, coarseMax=8)
    assert len(H.A) >= 4
    for l in range(1, len(H.A) - 1):
        assert np.all(np.diff(H.agg[l]) >= 0), l  # aggregate ids non-decreasing along the rows
        assert np.array_equal(np.unique(H.agg[l]), np.arange(H.A[l + 1].shape[0]))
    # the plain hierarchy
    plain, aggs = [sp.csr_matrix(A)], []
    while plain[-1].shape[0] > 8:
        agg, na = MG.aggregate(plain[-1].indptr, plain[-1].indices)
        P = MG.prolongation(agg, na)
        plain.append(sp.csr_matrix(P.T @ plain[-1] @ P))
        aggs.append(agg)
    assert [m.shape[0] for m in plain] == [m.shape[0] for m in H.A]
    # recover each permutation from the aggregate maps: old row of every new row
    old_of_new = [np.arange(A.shape[0])]
    for l in range(1, len(H.A)):
        new_agg = H.agg[l - 1]  # new level-l row of each (final-numbered) level-(l-1) row
        old_agg = aggs[l - 1][old_of_new[l - 1]]  # old level-l row of the same rows
        m = np.full(H.A[l].shape[0], -1)
        m[new_agg] = old_agg
        assert np.all(m >= 0)
        old_of_new.append(m)
        Pm = H.A[l].toarray()
        Po = plain[l].toarray()[np.ix_(m, m)]
        np.testing.assert_allclose(Pm, Po, rtol=1e-13, atol=1e-13 * np.abs(Po).max())
    assert np.array_equal(old_of_new[-1], np.arange(H.A[-1].shape[0]))  # the coarsest keeps its numbering


def test_hierarchy_galerkin_is_ptap_and_preserves_row_sums():
    A = laplacian3d(14, 13, 12, seed=3)
    H = MG.hierarchy(A, coarseMax=50)
    assert len(H.A) >= 3
    for l in range(len(H.agg)):
        P = H.P[l]
        Ac = (P.T @ H.A[l] @ P).toarray()
        np.testing.assert_allclose(H.A[l + 1].toarray(), Ac, rtol=0, atol=1e-13 * np.abs(Ac).max())
        # P^T A P 1 = P^T A 1: coarse row sums are sums of fine row sums
        fine_rs = np.asarray(H.A[l].sum(axis=1)).ravel()
        coarse_rs = np.asarray(H.A[l + 1].sum(axis=1)).ravel()
        np.testing.assert_allclose(coarse_rs, np.bincount(H.agg[l], weights=fine_rs), atol=1e-12)


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_vcycle_two_level_matches_closed_form_error_propagation(sweeps):
    """E = (I - Wpost A)^k (I - S P Ac^-1 P^T T A) (I - Wpre A)^k with k
    smoothing sweeps on each side, M = (I - E) A^-1."""
    A = laplacian3d(6, 5, 4, seed=4)
    H = MG.hierarchy(A, maxLevels=2, coarseMax=10)
    assert len(H.A) == 2
    n = A.shape[0]
    rng = np.random.default_rng(5)
    pre, scale, post = rng.uniform(0.3, 1.0, n), rng.uniform(0.5, 2.5, n), rng.uniform(0.3, 1.0, n)
    restrict = rng.uniform(0.5, 1.5, n)
    p = MG.Params(smoothFine=sweeps)
    Ad = A.toarray()
    Dinv = np.diag(1.0 / np.diag(Ad))
    P = H.P[0].toarray()
    Ac = P.T @ Ad @ P
    I = np.eye(n)
    mp = np.linalg.matrix_power
    E = mp(I - np.diag(post) @ Dinv @ Ad, sweeps) @ (
        I - np.diag(scale) @ P @ np.linalg.solve(Ac, P.T @ np.diag(restrict) @ Ad)) @ mp(
        I - np.diag(pre) @ Dinv @ Ad, sweeps)
    M = (I - E) @ np.linalg.inv(Ad)
    r = rng.standard_normal(n)
    z = MG.vcycle(H, r, p, pre, scale, post, exact_coarse=True, restrict=restrict)
    np.testing.assert_allclose(z, M @ r, rtol=0, atol=1e-12 * np.abs(M @ r).max())


def test_tied_weights_give_a_symmetric_cycle():
    # a = c and s = t per row: M = M^T (with the Jacobi coarsest sweeps too)
    A = laplacian3d(7, 6, 5, seed=12)
    H = MG.hierarchy(A, coarseMax=20)
    assert len(H.A) >= 3
    n = A.shape[0]
    rng = np.random.default_rng(13)
    w, t = rng.uniform(0.4, 0.9, n), rng.uniform(0.8, 1.6, n)
    p = MG.Params(coarseScale=1.5, coarseSweeps=7, kcycleDepth=0)  # the K-cycle is nonlinear
    M = np.column_stack([MG.vcycle(H, e, p, w, t, w, restrict=t) for e in np.eye(n)])
    np.testing.assert_allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    # untied weights break the symmetry
    M2 = np.column_stack([MG.vcycle(H, e, p, w, t, 0.5 * w, restrict=t) for e in np.eye(n)])
    assert np.abs(M2 - M2.T).max() > 1e-3 * np.abs(M2).max()


def test_kcycle_equals_two_flexible_cg_steps_on_the_level():
    A = laplacian3d(16, 15, 14, seed=14)
    H = MG.hierarchy(A, coarseMax=10)
    assert len(H.A) >= 4
    p = MG.Params(scale=1.6, coarseScale=1.6, kcycleDepth=2)
    b = np.random.default_rng(15).standard_normal(H.A[1].shape[0])
    e = MG.solve_level(H, b, p, 1)
    x, _, _ = MG.fcg(H.A[1], b, lambda r: MG.vcycle(H, r, p, level=1), tol=0.0, maxIter=2)
    np.testing.assert_allclose(e, x, rtol=0, atol=1e-12 * np.abs(x).max())
    # below kcycleDepth the correction is one cycle
    p0 = MG.Params(scale=1.6, coarseScale=1.6, kcycleDepth=0)
    np.testing.assert_array_equal(MG.solve_level(H, b, p0, 1), MG.vcycle(H, b, p0, level=1))


def test_coarsest_sweeps_match_polynomial_closed_form():
    A = laplacian3d(4, 3, 3, seed=6)
    Ad = A.toarray()
    D = np.diag(1.0 / np.diag(Ad))
    b = np.random.default_rng(7).standard_normal(A.shape[0])
    w, k = 0.6, 5
    T = np.eye(len(b)) - w * D @ Ad
    expect = sum(np.linalg.matrix_power(T, j) for j in range(k)) @ (w * D @ b)
    np.testing.assert_allclose(MG.coarsest(A, b, k, w), expect, atol=1e-13)


def test_jacobi_limit_of_the_cycle():
    A = laplacian3d(10, 9, 8, seed=8)
    H = MG.hierarchy(A, coarseMax=30)
    r = np.random.default_rng(9).standard_normal(A.shape[0])
    n = A.shape[0]
    z = MG.vcycle(H, r, MG.Params(smoothFine=1), np.ones(n), np.zeros(n), np.zeros(n))
    np.testing.assert_array_equal(z, (1.0 / A.diagonal()) * r)


def test_fcg_equals_pcg_for_a_fixed_spd_preconditioner_and_mg_cuts_iterations():
    A = laplacian3d(16, 15, 14, seed=10)
    b = np.random.default_rng(11).standard_normal(A.shape[0])
    dinv = 1.0 / A.diagonal()
    jac = lambda r: dinv * r
    _, kf, cf = MG.fcg(A, b, jac, 1e-10, 2000)
    _, kp, cp = MG.pcg(A, b, jac, 1e-10, 2000)
    assert cf and cp and abs(kf - kp) <= 1
    H = MG.hierarchy(A, coarseMax=60)
    p = MG.Params(scale=1.8, coarseScale=1.8)
    x, km, cm = MG.fcg(A, b, lambda r: MG.vcycle(H, r, p), 1e-10, 500)
    assert cm and km * 4 < kf
    assert np.abs(b - A @ x).max() / np.abs(b).max() < 1e-10
