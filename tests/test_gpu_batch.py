"""Batched solves of many detail-region pressure systems: with the default
schedule each system's iterates are bit for bit those of its solve alone
(same kernel bodies, same per-system chunks), in the three-launch form and
in the persistent form (one launch for the whole batch, any number of
workgroups and claims); other schedules (autotuned) agree to 1e-9 relative
at tol 1e-12; graph replay equals kernel-by-kernel launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def blocks(ctx):
    from chuanflow_amd import detail_batch as DB
    print("\n[case] 6 detail blocks, 2k-20k cells", flush=True)
    specs = DB.block_specs(count=6, cells=(2_000, 20_000), seed=11)
    return DB.build_blocks(ctx, specs, simple_iterations=2)


def test_blocks_have_positive_drops_and_spd_systems(blocks):
    for b in blocks:
        # SIMPLE moves the inlet flux of a jittered mesh by rounding-level
        # Rhie-Chow terms (the reference does the same); it stays near the rate
        assert b["dp"] > 0 and abs(b["measure"]["fluxIn"] - b["measure"]["imposedQ"]) <= 1e-2 * b["measure"]["imposedQ"]


def test_default_schedule_is_bitwise_the_single_solve(ctx, blocks):
    from chuanflow_amd import detail_batch as DB
    from chuanflow_amd import native as N
    x1, r1 = DB.solve_one_by_one(ctx, blocks)
    default = N.BatchSchedule(4, 2048, 1, 16)
    xb, rb, fin, st = DB.solve_batched(ctx, blocks, schedule=default)
    for a, b, ra, rbb in zip(x1, xb, r1, rb):
        assert ra.converged and rbb.converged and ra.iterations == rbb.iterations
        assert np.array_equal(a.get(), b.get())
    assert st.maxIterations == max(r.iterations for r in r1)
    assert np.all(fin > 0) and np.all(fin <= st.ms + 1e-9)
    # the three-launch form (graph replay and kernel-by-kernel launches) and the persistent
    # form (few workgroups and one claim each per step, no graph; many workgroups, four claims): the same bits
    for sched in (N.BatchSchedule(4, 2048, 1, 16, persistent=0), N.BatchSchedule(4, 2048, 0, 7, persistent=0),
                  N.BatchSchedule(4, 2048, 0, persistent=1, claim=1, wgs=7),
                  N.BatchSchedule(4, 2048, 1, persistent=1, claim=4, wgs=300)):
        print(f"schedule {sched.as_dict()}", flush=True)
        xn, rn, fn, sn = DB.solve_batched(ctx, blocks, schedule=sched)
        for a, b, ra, rbb in zip(xb, xn, rb, rn):
            assert ra.iterations == rbb.iterations and rbb.converged
            assert np.array_equal(a.get(), b.get())
        assert np.all(fn > 0) and np.all(fn <= sn.ms + 1e-9)


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_stops_at_the_cap_and_at_once(ctx, blocks, persistent):
    """The iteration cap (C(k) stops at k = maxIter) and an initial guess that
    already meets the tolerance (C(0) stops) retire systems -- inside the one
    launch in the persistent form -- exactly as the single solves do."""
    from chuanflow_amd import detail_batch as DB
    from chuanflow_amd import native as N
    sched = N.BatchSchedule(persistent=persistent)
    x1, r1 = DB.solve_one_by_one(ctx, blocks, tol=1e-8, maxIter=5)
    xb, rb, _, st = DB.solve_batched(ctx, blocks, tol=1e-8, maxIter=5, schedule=sched)
    for a, b, ra, rbb in zip(x1, xb, r1, rb):
        assert ra.iterations == rbb.iterations == 5 and not rbb.converged
        assert np.array_equal(a.get(), b.get())
    assert st.maxIterations == 5
    # start from converged solutions: no iteration, x untouched
    xs, _ = DB.solve_one_by_one(ctx, blocks, tol=1e-12)
    starts = [x.get() for x in xs]
    xc, rc, _, _ = DB.solve_batched(ctx, blocks, tol=1e-8, x0=starts, schedule=sched)
    for s0, x, r in zip(starts, xc, rc):
        assert r.converged and r.iterations == 0 and np.array_equal(s0, x.get())


def test_batched_blocks_match_oracle_cg(ctx, blocks):
    """Every block of the batched solve against the oracle's restatement of the
    reference's solveCG (Matrix.cpp:153-222) on the same CSR system: at
    tol 1e-13 the iteration counts agree within 2 and the solutions within
    1e-10 relative (summation order is the only difference)."""
    import oracle as O
    from chuanflow_amd import detail_batch as DB
    from chuanflow_amd import native as N
    xb, rb, _, _ = DB.solve_batched(ctx, blocks, tol=1e-13, schedule=N.BatchSchedule(4, 2048, 1, 16))
    for k, (b, x, r) in enumerate(zip(blocks, xb, rb)):
        rp, cl, vals, rhs = b["host"]
        A = O.Csr(b["n"], len(cl))
        A.rowStart[:], A.cols[:], A.vals[:] = rp, cl, vals
        xo, ro = O.solve_cg(A, rhs, None, 1e-13, 20000)
        xg = x.get()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"block {k}: {b['n']} cells, its gpu {r.iterations} oracle {ro.iterations}, rel err {err:.2e}",
              flush=True)
        assert r.converged and ro.converged and abs(r.iterations - ro.iterations) <= 2
        assert err <= 1e-10


def test_autotuned_schedule_agrees(ctx, blocks):
    from chuanflow_amd import detail_batch as DB
    from chuanflow_amd import native as N
    xs = [ctx.zeros(b["n"]) for b in blocks]
    best, table = N.batch_autotune(ctx, [b["csr"] for b in blocks], [b["b"] for b in blocks], xs, iters=10)
    print(f"autotuned {best.as_dict()} over {len(table)} candidates", flush=True)
    assert len(table) == 45 and min(ms for _, ms in table) > 0
    assert {c["persistent"] for c, _ in table} == {0, 1}
    x1, r1 = DB.solve_one_by_one(ctx, blocks, tol=1e-12)
    xb, rb, _, st = DB.solve_batched(ctx, blocks, tol=1e-12, schedule=best)
    for a, b, ra, rbb in zip(x1, xb, r1, rb):
        av, bv = a.get(), b.get()
        assert rbb.converged and abs(ra.iterations - rbb.iterations) <= max(3, ra.iterations // 50)
        assert np.abs(av - bv).max() <= 1e-9 * np.abs(av).max()
    # sched = null: the library autotunes and solves
    xa, ra2, _, st2 = DB.solve_batched(ctx, blocks, tol=1e-8)
    assert all(r.converged for r in ra2)
