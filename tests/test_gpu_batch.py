"""Batched mixed-form solves (icx_submit / icx_poll / icx_solve_batch): every job's
result equals its own icx_solve(MIXED) bit for bit, whatever else is in the batch and
whenever it was submitted."""
import numpy as np
import pytest

import icx

pytestmark = pytest.mark.gpu


def _jobs(ctx, seed, n):
    rng = np.random.default_rng(seed)
    return [icx.random_job(ctx, rng, size2d=(12, 60), size3d=(6, 16)) for _ in range(n)]


def _alone(jobs, tol=1e-8):
    out = []
    for j in jobs:
        u, info = j.solve(icx.SOLVE_MIXED, tol=tol, maxit=50000)
        assert info["converged"] and info["relres"] < tol
        out.append((u, info))
    return out


def test_batch_equals_alone(ctx):
    jobs = _jobs(ctx, 11, 6)
    ref = _alone(jobs)
    infos = ctx.solve_batch(jobs, tol=1e-8)
    for j, (u0, i0), i1 in zip(jobs, ref, infos):
        assert i1["iterations"] == i0["iterations"]
        assert i1["refinements"] == i0["refinements"]
        assert i1["relres"] == i0["relres"]
        assert np.array_equal(j.displacement(), u0)
    for j in jobs:
        j.close()


def test_staggered_submissions(ctx):
    """Jobs join a running batch at different times (continuous batching)."""
    jobs = _jobs(ctx, 12, 7)
    ref = _alone(jobs)
    pending = list(jobs)
    running, finished = set(), {}
    for _ in range(2):
        j = pending.pop(0)
        j.submit(tol=1e-8)
        running.add(j)
    polls = 0
    while running or pending:
        for j in ctx.poll(chunks=1):
            finished[id(j)] = (j.displacement(), j.last_info())
            running.discard(j)
        polls += 1
        if pending and polls % 2 == 0:
            j = pending.pop(0)
            j.submit(tol=1e-8)
            running.add(j)
        assert polls < 100000
    for j, (u0, i0) in zip(jobs, ref):
        u1, i1 = finished[id(j)]
        assert i1["converged"] and i1["iterations"] == i0["iterations"]
        assert np.array_equal(u1, u0)
    for j in jobs:
        j.close()


def test_batch_zero_load_and_iteration_limit(ctx):
    rng = np.random.default_rng(5)
    a = icx.random_job(ctx, rng, dims=3, size3d=(8, 8))
    a.set_fext(np.zeros(a.numeq))               # zero load: x = 0 at once
    b = icx.random_job(ctx, rng, dims=2, size2d=(40, 40))
    b.submit(tol=1e-12, maxit=5)                # stops at the iteration limit
    a.submit(tol=1e-8)
    seen = []
    while len(seen) < 2:
        seen += ctx.poll()
    ia, ib = a.last_info(), b.last_info()
    assert ia["converged"] and ia["iterations"] == 0
    assert not np.any(a.displacement())
    assert not ib["converged"] and ib["iterations"] == 5
    for j in (a, b):
        j.close()


def test_fused_batch_equals_alone(ctx):
    """FP64 fused-form solves (icx_submit_form / icx_solve_batch with SOLVE_FUSED) of
    quad4 and hexa8 jobs: each equals its own icx_solve(FUSED) bit for bit."""
    jobs = _jobs(ctx, 41, 6)
    ref = []
    for j in jobs:
        u, info = j.solve(icx.SOLVE_FUSED, tol=1e-9, maxit=50000)
        assert info["converged"]
        ref.append((u, info))
    infos = ctx.solve_batch(jobs, form=icx.SOLVE_FUSED, tol=1e-9)
    for j, (u0, i0), i1 in zip(jobs, ref, infos):
        for k in ("iterations", "matvecs", "relres"):
            assert i1[k] == i0[k], k
        assert np.array_equal(j.displacement(), u0)
    for j in jobs:
        j.close()


def test_mixed_and_fused_solves_in_one_batch(ctx):
    """Both forms in one batch, submitted at different times, with an iteration-limit
    stop and a zero load among the fused ones: every job equals its solo solve."""
    rng = np.random.default_rng(43)
    jobs = [icx.random_job(ctx, rng, size2d=(20, 90), size3d=(6, 20)) for _ in range(8)]
    forms = [icx.SOLVE_FUSED if q % 2 == 0 else icx.SOLVE_MIXED for q in range(len(jobs))]
    jobs[2].set_fext(np.zeros(jobs[2].numeq))
    limits = [50000] * len(jobs)
    limits[4] = 7
    ref = [j.solve(f, tol=1e-8, maxit=m) for j, f, m in zip(jobs, forms, limits)]
    pending = list(range(len(jobs)))
    running, finished = set(), {}
    polls = 0
    while pending or running:
        if pending and polls % 2 == 0:
            q = pending.pop(0)
            jobs[q].submit(tol=1e-8, maxit=limits[q], form=forms[q])
            running.add(q)
        for j in ctx.poll(chunks=1):
            q = jobs.index(j)
            finished[q] = (j.displacement(), j.last_info())
            running.discard(q)
        polls += 1
        assert polls < 100000
    for q, (u0, i0) in enumerate(ref):
        u1, i1 = finished[q]
        for k in ("iterations", "converged", "relres", "refinements"):
            assert i1[k] == i0[k], (q, k)
        assert np.array_equal(u1, u0), q
    assert not finished[4][1]["converged"] and finished[4][1]["iterations"] == 7
    assert finished[2][1]["iterations"] == 0 and not np.any(finished[2][0])
    with pytest.raises(icx.IcxError):
        jobs[0].submit(tol=1e-8, form=icx.SOLVE_PLAIN)
    for j in jobs:
        j.close()


def test_design_step_halves_equal_whole_step(ctx):
    """icx_opt_step == icx_opt_step_begin + submitted solve + icx_opt_step_end."""
    x0 = np.full(30 * 10, 0.5)
    A = icx.cantilever_2d(ctx, 30, 10)
    B = icx.cantilever_2d(ctx, 30, 10)
    for j in (A, B):
        j.opt_init(x0, driver=0, radius=1.5, form=icx.SOLVE_MIXED, tol=1e-9, maxit=20000)
    for _ in range(3):
        ra = A.opt_step()
        B.opt_step_begin()
        B.submit(tol=1e-9, maxit=20000)
        done = []
        while B not in done:
            done += ctx.poll()
        rb = B.opt_step_end(B.last_info())
        for k in ("objective", "constraint", "volume", "cg_iterations", "update_steps"):
            assert ra[k] == rb[k], k
    assert np.array_equal(A.opt_design(), B.opt_design())
    for j in (A, B):
        j.close()


def test_many_large_jobs_equal_alone(ctx):
    """Enough large hexa8 jobs that the blocks of one job span several rounds of a
    launch (a stop found by kernel A must reach every block of the job identically)."""
    rng = np.random.default_rng(21)
    jobs = [icx.random_job(ctx, rng, dims=3, size3d=(56, 72)) for _ in range(8)]
    jobs += [icx.random_job(ctx, rng, dims=2, size2d=(80, 160)) for _ in range(4)]
    ref = _alone(jobs)
    for _ in range(2):
        infos = ctx.solve_batch(jobs, tol=1e-8)
        for j, (u0, i0), i1 in zip(jobs, ref, infos):
            assert i1["iterations"] == i0["iterations"]
            assert np.array_equal(j.displacement(), u0)
    for j in jobs:
        j.close()


def test_single_reduction_batch_equals_alone(ctx):
    """With the single-reduction iteration switched on, hexa8 jobs run it in the batch
    (one launch of the iteration kernel for all of them per iteration) and quad4 jobs
    keep kernels A and B; every job still equals its own solve bit for bit."""
    ctx.set_option(icx.OPT_SINGLE_REDUCTION, 1)
    try:
        rng = np.random.default_rng(31)
        jobs = [icx.random_job(ctx, rng, dims=3, size3d=(6, 20)) for _ in range(4)]
        jobs += [icx.random_job(ctx, rng, dims=2, size2d=(12, 60)) for _ in range(2)]
        ref = _alone(jobs)
        infos = ctx.solve_batch(jobs, tol=1e-8)
        for j, (u0, i0), i1 in zip(jobs, ref, infos):
            assert i1["iterations"] == i0["iterations"]
            assert i1["relres"] == i0["relres"]
            assert np.array_equal(j.displacement(), u0)
    finally:
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 0)
    for j in jobs:
        j.close()


def test_large_hexa8_job_per_node_kernel_b(ctx):
    """A hexa8 job of more than 2^20 nodes takes kernel B with the per-node Jacobi factor
    (chunks of 4 nodes, 12 consecutive entries of the node-interleaved vectors, the nn mod
    4 tail nodes in one thread).  128 x 64 x 126 elements: 1 064 895 nodes (tail 3).  The
    mixed solve meets the true residual through the FP64 operator, agrees with the FP64
    plain form, and equals itself batched with a small job bit for bit."""
    big = icx.cantilever_3d(ctx, 128, 64, 126)
    assert big.numnp >= 1 << 20 and big.numnp % 4 == 3
    big.set_design(icx.random_density(big.nelx, 5))
    f = big.fext()
    u, info = big.solve(icx.SOLVE_MIXED, tol=1e-9, maxit=50000)
    assert info["converged"]
    assert np.linalg.norm(f - big.matvec(u)) / np.linalg.norm(f) < 1e-9
    up, _ = big.solve(icx.SOLVE_PLAIN, tol=1e-11, maxit=50000)
    assert np.linalg.norm(u - up) / np.linalg.norm(up) < 1e-6
    small = icx.cantilever_3d(ctx, 12, 6, 6)
    small.set_design(icx.random_density(small.nelx, 6))
    ref = _alone([big, small])
    infos = ctx.solve_batch([big, small], tol=1e-8)
    for j, (u0, i0), i1 in zip([big, small], ref, infos):
        assert i1["iterations"] == i0["iterations"]
        assert np.array_equal(j.displacement(), u0)
    big.close()
    small.close()
