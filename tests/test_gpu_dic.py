"""DIC on the GPU against the DIC oracle: the FP64 factor bit for bit (same
ascending-column sums), the FP32 substitutions within 1e-4 of FP64, and
flexible CG with DIC to a true FP64 residual of 1e-8 in the oracle's DIC-PCG
iteration count (+-5%)."""
import numpy as np
import pytest
import scipy.sparse as sp

import dic_oracle as DIC
import mg_oracle as MG
from test_gpu_mg import fv_pressure_system
from test_oracle_mg import laplacian3d

pytestmark = pytest.mark.gpu


def make(ctx, A):
    from chuanflow_amd import native as N
    return N, N.Csr(ctx, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data)


@pytest.fixture(scope="module")
def sysfv(ctx):
    print("\n[case] DIC: 24^3 jittered FV pressure system (RCM)", flush=True)
    return fv_pressure_system(24, 24, 24, 8)


def test_factor_bitwise_and_apply(ctx, sysfv):
    A, _ = sysfv
    N, G = make(ctx, A)
    levels, d = G.dic_info()
    do = DIC.factor(A)
    assert np.array_equal(d, do)
    assert 1 < levels < A.shape[0]
    r = np.random.default_rng(1).standard_normal(A.shape[0]).astype(np.float32)
    z = ctx.empty(A.shape[0], np.float32)
    G.dic_apply(ctx.array(r), z)
    zo = DIC.apply(A, do, r.astype(np.float64))
    assert np.abs(z.get() - zo).max() / np.abs(zo).max() < 1e-4


def test_dic_fcg_converges_like_the_oracle(ctx, sysfv):
    A, b = sysfv
    N, G = make(ctx, A)
    x, rep = G.solve(b, None, 1e-8, 2000, N.SOLVER_CG, N.PRECOND_DIC, N.FP32_IR)
    true = np.abs(b - A @ x).max() / np.abs(b).max()
    d = DIC.factor(A)
    _, ko, co = MG.pcg(A, b, lambda v: DIC.apply(A, d, v), 1e-8, 2000)
    _, rj = G.solve(b, None, 1e-8, 5000, N.SOLVER_CG, N.PRECOND_JACOBI, N.FP64)
    print(f"DIC-FCG {rep.iterations} (oracle DIC-PCG {ko}), Jacobi-PCG {rj.iterations}", flush=True)
    assert rep.converged and true < 1e-8 and co
    assert abs(rep.iterations - ko) <= max(2, ko // 20)
    assert rep.iterations < rj.iterations


def test_tridiagonal_dic_solve_is_immediate(ctx):
    n = 200
    rng = np.random.default_rng(3)
    off = -rng.uniform(0.5, 1.5, n - 1)
    diag = -np.r_[off, 0] - np.r_[0, off] + rng.uniform(0.1, 0.5, n)
    A = sp.csr_matrix(sp.diags([off, diag, off], [-1, 0, 1]))
    A.sort_indices()
    b = rng.standard_normal(n)
    N, G = make(ctx, A)
    x, rep = G.solve(b, None, 1e-8, 50, N.SOLVER_CG, N.PRECOND_DIC, N.FP32_IR)
    assert rep.converged and rep.iterations <= 3
    assert np.abs(b - A @ x).max() / np.abs(b).max() < 1e-8
