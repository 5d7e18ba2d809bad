"""Diagonal incomplete Cholesky (DIC) oracle: factor and substitutions in
FP64, summing each row's terms in ascending column order.

TEST INFRASTRUCTURE ONLY.  Importable from tests/, from
``__graft_entry__.smoke()`` and from bench.py's cpu_baseline leg, as the
checker; never from the product package.

No counterpart in the reference (its preconditioner is Jacobi,
src/common/src/Matrix.cpp:160-165); the definition is OpenFOAM's DIC,
rD_i = 1 / (a_ii - sum_{j<i} a_ij^2 rD_j).  Pinned by: on a tridiagonal
matrix IC(0) is the exact Cholesky factorisation, so M = A (M^-1 r = A^-1 r,
and PCG converges in one iteration); M^-1 is symmetric; on a diagonal
matrix M = A.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def factor(A):
    """D (the diagonal of the DIC factor, D_ii = 1 / rD_i)."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    d = np.zeros(n)
    for i in range(n):
        s, aii = 0.0, 0.0
        for k in range(A.indptr[i], A.indptr[i + 1]):
            j = A.indices[k]
            if j < i:
                s += A.data[k] * A.data[k] / d[j]
            elif j == i:
                aii = A.data[k]
        di = aii - s
        d[i] = di if di > 1e-14 * abs(aii) else aii
    return d


def apply(A, d, r):
    """z = M^-1 r, M = (D + L) D^-1 (D + U)."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    v = np.zeros(n)
    for i in range(n):
        s = r[i]
        for k in range(A.indptr[i], A.indptr[i + 1]):
            j = A.indices[k]
            if j >= i:
                break
            s -= A.data[k] * v[j]
        v[i] = s / d[i]
    z = v.copy()
    for i in range(n - 1, -1, -1):
        s = 0.0
        for k in range(A.indptr[i + 1] - 1, A.indptr[i] - 1, -1):
            j = A.indices[k]
            if j <= i:
                break
            s += A.data[k] * z[j]
        z[i] -= s / d[i]
    return z
