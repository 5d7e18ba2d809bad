"""The DIC oracle pinned by closed forms (no counterpart in the reference)."""
import numpy as np
import scipy.sparse as sp

import dic_oracle as DIC
import mg_oracle as MG
from test_oracle_mg import laplacian3d


def test_tridiagonal_dic_is_exact_cholesky():
    n = 40
    rng = np.random.default_rng(1)
    off = -rng.uniform(0.5, 1.5, n - 1)
    diag = np.r_[off, 0] * -1 + np.r_[0, off] * -1 + rng.uniform(0.1, 0.5, n)
    A = sp.diags([off, diag, off], [-1, 0, 1], format="csr")
    d = DIC.factor(A)
    r = rng.standard_normal(n)
    np.testing.assert_allclose(DIC.apply(A, d, r), np.linalg.solve(A.toarray(), r), rtol=1e-12)
    # PCG with an exact preconditioner converges in one iteration
    _, its, conv = MG.pcg(A, r, lambda v: DIC.apply(A, d, v), 1e-10, 10)
    assert conv and its == 1


def test_dic_inverse_is_symmetric_and_diagonal_case_is_exact():
    A = laplacian3d(5, 4, 3, seed=2)
    d = DIC.factor(A)
    M = np.column_stack([DIC.apply(A, d, e) for e in np.eye(A.shape[0])])
    np.testing.assert_allclose(M, M.T, atol=1e-13 * np.abs(M).max())
    D = sp.diags(np.arange(1.0, 7.0), format="csr")
    r = np.arange(6.0) - 2.5
    np.testing.assert_allclose(DIC.apply(D, DIC.factor(D), r), r / np.arange(1.0, 7.0), rtol=1e-15)


def test_dic_pcg_needs_fewer_iterations_than_jacobi():
    A = laplacian3d(14, 12, 10, seed=3)
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    d = DIC.factor(A)
    _, kd, cd = MG.pcg(A, b, lambda v: DIC.apply(A, d, v), 1e-8, 2000)
    _, kj, cj = MG.pcg(A, b, lambda v: v / A.diagonal(), 1e-8, 2000)
    assert cd and cj and kd < 0.75 * kj
