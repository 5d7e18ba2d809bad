"""The network's forward and backward kernels and the multigrid objective's
adjoint on the GPU against the FP64 CPU oracle.

Tolerances: FP32 storage (MFMA 32x32x2 f32) gives forward outputs within
1e-5 and parameter gradients within 1e-4 (normwise) of FP64 central
differences; FP16 storage (MFMA 32x32x8 f16, FP32 accumulation) within 2e-2
and 5e-2.  The objective's gradient in FP32 within 1e-3 (normwise)."""
import numpy as np
import pytest

import mg_oracle as MG
import mpnn_oracle as NN
from test_oracle_mg import laplacian3d

pytestmark = pytest.mark.gpu


def graph(seed):
    A = laplacian3d(9, 8, 7, seed=seed)
    n = A.shape[0]
    rng = np.random.default_rng(seed + 1)
    R = rng.uniform(0.0, 0.3, (n, n))
    R = 0.5 * (R + R.T)
    np.fill_diagonal(R, 0.0)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    geo = R[rows, A.indices]
    r0 = rng.standard_normal(n)
    return A, geo, r0


def setup(ctx, fp16, seed=11):
    from chuanflow_amd import native as N
    A, geo, r0 = graph(seed)
    G = N.Csr(ctx, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data)
    net = N.Mpnn(ctx, seed=3, fp16=fp16)
    p = NN.random_params(seed).astype(np.float32)
    net.set_params(p)
    return N, A, geo, r0, G, net, p.astype(np.float64)


def rel(x, y):
    return np.linalg.norm(np.asarray(x, np.float64) - y) / np.linalg.norm(y)


@pytest.mark.parametrize("fp16,tol", [(False, 1e-5), (True, 2e-2)])
def test_forward_matches_oracle(ctx, fp16, tol):
    print(f"\n[case] network forward, fp16={fp16}", flush=True)
    N, A, geo, r0, G, net, p = setup(ctx, fp16)
    a, t = net.forward(G, ctx.array(geo), ctx.array(r0))
    ao, to, _ = NN.forward(p, A, *NN.features(A, geo, r0))
    assert rel(a.get(), ao) < tol and rel(t.get(), to) < tol, (rel(a.get(), ao), rel(t.get(), to))
    # no geometry / no residual inputs
    a2, t2 = net.forward(G)
    ao2, to2, _ = NN.forward(p, A, *NN.features(A))
    assert rel(a2.get(), ao2) < tol and rel(t2.get(), to2) < tol


def test_backward_fp32_matches_central_differences(ctx):
    print("\n[case] network backward fp32 vs FP64 central differences", flush=True)
    N, A, geo, r0, G, net, p = setup(ctx, False)
    x0, ew, eg = NN.features(A, geo, r0)
    rng = np.random.default_rng(5)
    al, be = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    net.forward(G, ctx.array(geo), ctx.array(r0), train=True)
    g = net.backward(ctx.array(al.astype(np.float32)), ctx.array(be.astype(np.float32))).get()
    _, _, cache = NN.forward(p, A, x0, ew, eg)
    ga = NN.backward(p, cache, al, be)
    assert rel(g, ga) < 1e-4, rel(g, ga)
    J = lambda q: (lambda a, t: al @ a + be @ t)(*NN.forward(q, A, x0, ew, eg)[:2])
    idx = np.union1d(rng.choice(NN.NPARAMS, 300, replace=False),
                     [NN.OFF_WIN, NN.OFF_BIN, NN.OFF_L0, NN.OFF_L0 + 2 * NN.LAYER - 1, NN.OFF_WOUT, NN.OFF_BOUT + 1])
    h = 1e-6
    fd = np.array([(J(p + h * np.eye(1, NN.NPARAMS, k)[0]) - J(p - h * np.eye(1, NN.NPARAMS, k)[0])) / (2 * h)
                   for k in idx])
    assert rel(g[idx], fd) < 1e-4, rel(g[idx], fd)


def test_backward_fp16_close_to_oracle(ctx):
    print("\n[case] network backward fp16", flush=True)
    N, A, geo, r0, G, net, p = setup(ctx, True)
    rng = np.random.default_rng(6)
    al, be = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    net.forward(G, ctx.array(geo), ctx.array(r0), train=True)
    g = net.backward(ctx.array(al.astype(np.float32)), ctx.array(be.astype(np.float32))).get()
    _, _, cache = NN.forward(p, A, *NN.features(A, geo, r0))
    ga = NN.backward(p, cache, al, be)
    assert rel(g, ga) < 5e-2, rel(g, ga)


@pytest.mark.parametrize("sweeps,kcycle", [(1, 0), (3, 0), (3, 2)])
def test_objective_gradient_matches_oracle(ctx, sweeps, kcycle):
    """The device adjoint against the oracle's: the exact gradient with the
    V-cycle below level 1 (pinned by central differences on the CPU), and the
    same symmetric-coarse-correction gradient with the K-cycle."""
    print(f"\n[case] multigrid objective adjoint, {sweeps} smoothing sweeps, K-cycle depth {kcycle}", flush=True)
    from chuanflow_amd import native as N
    A = laplacian3d(20, 18, 16, seed=8)
    n = A.shape[0]
    G = N.Csr(ctx, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data)
    G.set_mg(coarseScale=1.6, smoothFine=sweeps, kcycleDepth=kcycle)
    rng = np.random.default_rng(9)
    a = rng.uniform(0.5, 0.8, n).astype(np.float32)
    t = rng.uniform(1.1, 1.5, n).astype(np.float32)
    e = rng.standard_normal(n).astype(np.float32)
    L, ga, gt = G.mg_loss_grad(ctx.array(e), ctx.array(a), ctx.array(t))
    Hh = MG.hierarchy(A)
    Lo, gao, gto = NN.objective_grad(Hh, e.astype(np.float64), a.astype(np.float64), t.astype(np.float64),
                                     MG.Params(coarseScale=1.6, smoothFine=sweeps, kcycleDepth=kcycle))
    print(f"L gpu {L:.6e} oracle {Lo:.6e}", flush=True)
    assert abs(L - Lo) <= 1e-4 * Lo
    assert rel(ga.get(), gao) < 1e-3 and rel(gt.get(), gto) < 1e-3, (rel(ga.get(), gao), rel(gt.get(), gto))
