"""The no-allocation Jacobi-PCG entry point (cfx_pcg_solve_noalloc) against
cfx_csr_solve_device with CG / Jacobi / FP64 on the same systems: two calls
in a row, the second on a smaller system (the shared scratch keeps the first
call's contents), each to the same iteration count within 2 and the same
solution within 1e-10 relative at tol 1e-12; a system above the scratch's row
limit is refused with a clear error before any device work."""
import numpy as np
import pytest

from test_oracle_mg import laplacian3d

pytestmark = pytest.mark.gpu


def _solve_both(ctx, A, b):
    from chuanflow_amd import native as N
    G = N.Csr(ctx, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data)
    db = ctx.array(b)
    x1 = ctx.zeros(A.shape[0])
    r1 = G.solve_device(db, x1, 1e-12, 5000, N.SOLVER_CG, N.PRECOND_JACOBI, N.FP64)
    rp = ctx.array(A.indptr.astype(np.int32))
    cl = ctx.array(A.indices.astype(np.int32))
    vv = ctx.array(A.data.astype(np.float64))
    x2 = ctx.zeros(A.shape[0])
    r2 = N.pcg_solve_noalloc(ctx, rp, cl, vv, db, x2, 1e-12, 5000)
    return x1.get(), r1, x2.get(), r2


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_solve_device_on_two_calls(ctx, order):
    from chuanflow_amd import native as N
    print("\n[case] no-allocation PCG: 24x24x24 then 10x12x14 FV Laplacians", flush=True)
    for dims, seed in (((24, 24, 24), 3), ((10, 12, 14), 4)):
        A = laplacian3d(*dims, seed=seed).tocsr()
        A.sort_indices()
        b = np.random.default_rng(seed).standard_normal(A.shape[0])
        x1, r1, x2, r2 = _solve_both(ctx, A, b)
        err = np.abs(x2 - x1).max() / np.abs(x1).max()
        res = np.abs(b - A @ x2).max() / np.abs(b).max()
        print(f"{dims}: its {r1.iterations} vs {r2.iterations}, rel diff {err:.2e}, residual {res:.2e}", flush=True)
        assert r1.converged and r2.converged and abs(r1.iterations - r2.iterations) <= 2
        assert err <= 1e-10 and res < 1e-11


def test_noalloc_refuses_systems_above_its_limit(ctx):
    from chuanflow_amd import native as N
    n = N.NOALLOC_MAX_ROWS + 1
    rp = ctx.zeros(2, np.int32)
    cl = ctx.zeros(1, np.int32)
    vv = ctx.zeros(1)
    b = ctx.zeros(n)
    x = ctx.zeros(n)
    with pytest.raises(RuntimeError, match="no-allocation solve"):
        N.pcg_solve_noalloc(ctx, rp, cl, vv, b, x)
