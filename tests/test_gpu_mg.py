"""The multigrid preconditioner and flexible CG on the GPU against the
multigrid oracle: hierarchy (aggregation bit-exact, Galerkin values to
1e-13), one V-cycle in FP32 against FP64 (rel 2e-4), FCG convergence to a
true FP64 residual of 1e-8, the Jacobi limit, and the Jacobi-PCG fallback
on a non-finite preconditioner."""
import numpy as np
import pytest
import scipy.sparse as sp

import mg_oracle as MG
import oracle as O
from channel import neumann_bcs
from helpers import host_and_oracle
from test_oracle_mg import laplacian3d

pytestmark = pytest.mark.gpu


def fv_pressure_system(nx, ny, nz, seed, pin=True):
    """Jittered, shuffled, RCM-renumbered FV Laplacian (gamma = 1, Neumann),
    by default with one pinned cell so the matrix is SPD, and a zero-mean
    N(0,1) right-hand side."""
    hm, a, om = host_and_oracle(nx, ny, nz, (0, 0, 0), (1, 1, 1), jitter=0.15, seed=seed, shuffle=seed + 1,
                                rcm=True)
    t, v = neumann_bcs(om.nB)
    lc, lk = O.linearise(om, 1, t, v)
    A = O.Ldu(om)
    O.assemble_laplacian(om, A, np.ones(om.nFaces), lc, lk, -1)
    csr = O.csr_from_ldu(om, A)
    M = sp.csr_matrix((csr.vals, csr.cols, csr.rowStart), shape=(om.nCells, om.nCells))
    if pin:
        M[0, 0] = M[0, 0] * 2.0  # pin the constant null space
    M.sort_indices()
    b = np.random.default_rng(seed).standard_normal(om.nCells)
    return M, b - b.mean()


@pytest.fixture(scope="module")
def systems(ctx):
    print("\n[case] multigrid systems: 40^3 random-conductance Laplacian, 32^3 jittered FV", flush=True)
    return {"lap40": (laplacian3d(40, 40, 40, seed=3), None), "fv32": fv_pressure_system(32, 32, 32, 5),
            "fv24n": fv_pressure_system(24, 24, 24, 6, pin=False)}


def make(ctx, A):
    from chuanflow_amd import native as N
    return N, N.Csr(ctx, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data)


@pytest.mark.parametrize("name", ["lap40", "fv32", "fv24n"])
def test_hierarchy_matches_oracle(ctx, systems, name):
    A, _ = systems[name]
    N, G = make(ctx, A)
    rows, nnz = G.mg_info()
    H = MG.hierarchy(A)
    assert rows == [m.shape[0] for m in H.A] and nnz == [m.nnz for m in H.A]
    for l in range(len(rows)):
        rp, cl, vv, agg = G.mg_level(l)
        assert np.array_equal(rp, H.A[l].indptr) and np.array_equal(cl, H.A[l].indices)
        np.testing.assert_allclose(vv, H.A[l].data, rtol=1e-13, atol=1e-13 * np.abs(H.A[l].data).max())
        if l + 1 < len(rows):
            assert np.array_equal(agg, H.agg[l])


@pytest.mark.parametrize("name,maxLevels,fine,coarse,deep", [("lap40", 10, 2, 1, 1), ("fv32", 10, 2, 1, 1),
                                                             ("fv32", 3, 2, 1, 1), ("fv24n", 10, 2, 1, 1),
                                                             ("lap40", 10, 1, 1, 1), ("lap40", 10, 3, 2, 2),
                                                             ("lap40", 10, 4, 2, 1), ("lap40", 10, 3, 1, 3)])
def test_vcycle_fp32_matches_oracle(ctx, systems, name, maxLevels, fine, coarse, deep):
    """Dense exact coarsest solve (<= 128 rows; pure Neumann in fv24n) and,
    with 3 levels, Jacobi sweeps on a 253-row coarsest level; one to four
    smoothing sweeps on the finest level, one to three on level 1 and on the
    K-cycle levels below it."""
    A, _ = systems[name]
    N, G = make(ctx, A)
    n = A.shape[0]
    rng = np.random.default_rng(1)
    pre, scale, post, restrict = (rng.uniform(0.4, 1.0, n).astype(np.float32),
                                  rng.uniform(0.5, 2.5, n).astype(np.float32),
                                  rng.uniform(0.4, 1.0, n).astype(np.float32),
                                  rng.uniform(0.5, 1.5, n).astype(np.float32))
    dpre, dscale, dpost, drestrict = ctx.array(pre), ctx.array(scale), ctx.array(post), ctx.array(restrict)
    G.set_mg(coarseScale=1.7, coarseSweeps=20, maxLevels=maxLevels, smoothFine=fine, smoothCoarse=coarse,
             smoothDeep=deep)
    G.set_mg_weights(dpre, dscale, dpost, drestrict)
    r = rng.standard_normal(n)
    r = (r - r.mean()).astype(np.float32)
    z = ctx.empty(n, np.float32)
    G.mg_apply(ctx.array(r), z)
    p = MG.Params(coarseScale=1.7, coarseSweeps=20, smoothFine=fine, smoothCoarse=coarse, smoothDeep=deep)
    H = MG.hierarchy(A, maxLevels=maxLevels)
    assert len(H.A) == len(G.mg_info()[0])
    ref = MG.vcycle(H, r.astype(np.float64), p, pre.astype(np.float64), scale.astype(np.float64),
                    post.astype(np.float64), restrict=restrict.astype(np.float64))
    err = np.abs(z.get() - ref).max() / np.abs(ref).max()
    # an FP32 cycle against FP64: on the pinned (nearly singular) fv32 system its rounding, amplified by the
    # exact coarsest solve, spans 4e-5 to 1.2e-4 of max|z| over weight seeds, for every launch fusion, with
    # two sweeps on the finest level (three on it and on the deep K-cycle levels reach 2.1e-4)
    assert err < 2e-4, err


@pytest.mark.parametrize("name", ["lap40", "fv32", "fv24n"])
def test_fcg_converges_to_true_fp64_residual(ctx, systems, name):
    A, b = systems[name]
    if b is None:
        b = np.random.default_rng(2).standard_normal(A.shape[0])
    N, G = make(ctx, A)
    G.set_mg(scale=1.8, coarseScale=1.8)
    x, rep = G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
    true = np.abs(b - A @ x).max() / np.abs(b).max()
    print(f"[{name}] fcg its {rep.iterations} outer {rep.outerIterations} true {true:.2e}", flush=True)
    assert rep.converged and rep.precondFallbacks == 0
    assert true < 1e-8 and abs(rep.trueResidual - true) <= 1e-3 * true + 1e-15
    _, rj = G.solve(b, None, 1e-8, 5000, N.SOLVER_CG, N.PRECOND_JACOBI, N.FP32_IR)
    assert rj.converged and rep.iterations * 5 < rj.iterations
    # the learned preconditioner without a network is refused (CFX_E_STATE), not silently replaced
    with pytest.raises(N.CfxError, match="network") as e:
        G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_LEARNED, N.FP32_IR)
    assert e.value.code == 3
    # same iterates as the FP64 oracle's FCG up to FP32 rounding: iteration counts agree closely
    H = MG.hierarchy(A)
    p = MG.Params(scale=1.8, coarseScale=1.8)
    _, ko, co = MG.fcg(A, b, lambda r: MG.vcycle(H, r, p), 1e-8, 500)
    assert co and abs(ko - rep.iterations) <= max(2, ko // 20)


def test_identity_network_limit_equals_jacobi_fcg(ctx, systems):
    A, b = systems["fv32"]
    N, G = make(ctx, A)
    n = A.shape[0]
    one, zero = ctx.array(np.ones(n, np.float32)), ctx.array(np.zeros(n, np.float32))
    # with one smoothing sweep, a = 1, s = 0, c = 0 makes the cycle z = D^-1 r
    G.set_mg(smoothFine=1)
    G.set_mg_weights(one, zero, zero)
    _, rl = G.solve(b, None, 1e-8, 5000, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
    N2, J = make(ctx, A)
    J.set_mg(maxLevels=1, omegaPre=1.0, omegaPost=0.0, smoothFine=1)
    _, rj = J.solve(b, None, 1e-8, 5000, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
    # FP64 Jacobi-PCG: FCG with a fixed SPD preconditioner has PCG's iterates in exact arithmetic
    _, rp = J.solve(b, None, 1e-8, 5000, N.SOLVER_CG, N.PRECOND_JACOBI, N.FP64)
    print(f"identity-limit FCG {rl.iterations}, Jacobi-FCG {rj.iterations}, FP64 Jacobi-PCG {rp.iterations}",
          flush=True)
    assert rl.converged and rj.converged and rp.converged
    assert rl.iterations == rj.iterations
    assert abs(rj.iterations - rp.iterations) <= max(3, rp.iterations // 50)


def test_nonfinite_weights_fall_back_to_jacobi_pcg(ctx, systems):
    A, b = systems["fv32"]
    N, G = make(ctx, A)
    n = A.shape[0]
    w = np.full(n, 0.7, np.float32)
    w[n // 3] = np.nan
    dw = ctx.array(w)
    G.set_mg_weights(dw, None, None)
    x, rep = G.solve(b, None, 1e-8, 5000, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
    true = np.abs(b - A @ x).max() / np.abs(b).max()
    assert rep.precondFallbacks == 1 and rep.converged and true < 1e-8


def test_new_values_on_same_pattern_rebuild_galerkin_values(ctx, systems):
    A, b = systems["lap40"]
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    N, G = make(ctx, A)
    G.set_mg(scale=1.8, coarseScale=1.8)
    A2 = A.copy()
    A2.data = A2.data * np.where(A2.indices == np.repeat(np.arange(A.shape[0]), np.diff(A.indptr)), 1.5, 1.0)
    G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
    G.set_values(A2.data)
    x, rep = G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
    assert rep.converged and np.abs(b - A2 @ x).max() / np.abs(b).max() < 1e-8
    _, _, vv, _ = G.mg_level(1)
    np.testing.assert_allclose(vv, MG.hierarchy(A2).A[1].data, rtol=1e-13, atol=1e-12)


def test_graph_replay_equals_eager_launches(ctx, systems):
    A, b = systems["fv32"]
    N, G = make(ctx, A)
    G.set_mg(scale=1.8, coarseScale=1.8)
    xs, reps = [], []
    for graphs in (True, False, True):
        ctx.set_graphs(graphs)
        x, rep = G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
        xs.append(x)
        reps.append(rep.iterations)
    ctx.set_graphs(False)  # the default
    assert reps[0] == reps[1] == reps[2]
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])


@pytest.mark.parametrize("case", ["lap40-k1", "fv24n-k0", "lap30-jacobi", "fv32-k2", "fv32-k2-sweeps"])
def test_fused_kernels_equal_separate_launches(ctx, systems, case):
    """Launch fusion changes no bit: the small coarsest levels' cycle in one
    workgroup (fusion bit 0; dense and Jacobi-swept coarsest level), the
    prolongation with the K-cycle combination folded into coarse
    post-smoothing (bit 1), that cycle repeated from LDS by every workgroup
    of the post-smoothing above it (bit 2), the K-cycle's r2 formed in the
    second visit's pre-smoothing (bit 3) and its combination formed in the
    prolongation (bit 4), the restriction from a coarse level formed by its
    pre-smoothing (bit 5, alone and with the tail's LDS cycle, which then
    stages that right-hand side), the finest level's smoothing on its values
    in 64-row slices (bit 6), alone and together, give the cycle of the
    separate kernels exactly, and the same flexible-CG iterates."""
    if case == "lap40-k1":
        A, b = systems["lap40"][0], None
        kw = dict(kcycleDepth=1)
    elif case == "fv24n-k0":
        A, b = systems["fv24n"]
        kw = dict(kcycleDepth=0)
    elif case == "lap30-jacobi":
        A, b = laplacian3d(30, 30, 30, seed=8), None
        kw = dict(kcycleDepth=0, coarseMax=1000, coarseSweeps=12)
    elif case == "fv32-k2":
        A, b = systems["fv32"]
        kw = dict(kcycleDepth=2)
    else:
        A, b = systems["fv32"]
        kw = dict(kcycleDepth=2, smoothFine=3, smoothCoarse=2)
    n = A.shape[0]
    if b is None:
        b = np.random.default_rng(9).standard_normal(n)
    N, G = make(ctx, A)
    G.set_mg(scale=1.8, coarseScale=1.8, **kw)
    rows, _ = G.mg_info()
    tail = G.mg_tail_level()
    print(f"[{case}] levels {rows}, tail from level {tail}", flush=True)
    if not case.startswith("fv32-k2"):
        assert 1 <= tail < len(rows) - 1 and tail > kw["kcycleDepth"]
    else:
        assert len(rows) >= 4  # a K-cycle level below level 1: its combination is deferred
    if case == "lap30-jacobi":
        assert rows[-1] > 128  # coarsest level swept, not inverted
    r = ctx.array(np.random.default_rng(3).standard_normal(n).astype(np.float32))
    zs, xs, its = [], [], []
    # the finest level's sliced layout (bit 6) applies to every case here
    assert A.nnz <= G.mg_slices() <= 1.25 * A.nnz
    fusions = (0, 1, 2, 3, 4, 8, 16, 29, 31, 32, 36, 61, 63, 64, 125, 127)
    for fusion in fusions:
        G.set_mg(scale=1.8, coarseScale=1.8, fusion=fusion, **kw)
        z = ctx.empty(n, np.float32)
        G.mg_apply(r, z)
        zs.append(z.get())
        x, rep = G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
        assert rep.converged
        xs.append(x)
        its.append(rep.iterations)
    for k in range(1, len(fusions)):
        assert np.array_equal(zs[0], zs[k]), fusions[k]
        assert its[0] == its[k] and np.array_equal(xs[0], xs[k]), fusions[k]


@pytest.mark.parametrize("case", ["far", "padded"])
def test_patterns_without_slices_keep_the_csr_stream_kernels(ctx, case):
    """A finest pattern whose column offsets leave int16 (couplings 40,000 rows
    apart), or whose 64-row slices would pad beyond a quarter of its nonzeros
    (every 64th row coupled to 40 rows), has no sliced layout: fusion bit 6
    then changes nothing, and the solve equals fusion 61's bit for bit."""
    A = laplacian3d(40, 40, 40, seed=21).tolil()
    n = A.shape[0]
    pairs = ([(i, i + 40000) for i in range(0, n - 40000, 997)] if case == "far" else
             [(i, i + d) for i in range(0, n - 400, 64) for d in range(2, 402, 10)])
    for i, j in pairs:
        A[i, j] = A[j, i] = A[i, j] - 0.05
        A[i, i] += 0.05
        A[j, j] += 0.05
    A = sp.csr_matrix(A)
    A.sort_indices()
    b = np.random.default_rng(22).standard_normal(n)
    N, G = make(ctx, A)
    assert G.mg_slices() == 0
    xs = []
    for fusion in (61, 125):
        G.set_mg(fusion=fusion)
        x, rep = G.solve(b, None, 1e-8, 500, N.SOLVER_CG, N.PRECOND_MG, N.FP32_IR)
        assert rep.converged and np.abs(b - A @ x).max() / np.abs(b).max() < 1e-8
        xs.append(x)
    assert np.array_equal(xs[0], xs[1])
