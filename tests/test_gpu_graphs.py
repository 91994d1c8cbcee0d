"""Captured iteration chunks (HIP graphs) against eager launches: every replay of a
job's graph must give the same bits as the eager solve (the kernels are deterministic),
with eager work and synchronisations between the replays."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [icx.SOLVE_PLAIN, icx.SOLVE_FUSED, icx.SOLVE_MIXED, "sr"])
def test_graph_replays_equal_eager(ctx, form):
    """"sr": the mixed form with the single-reduction iteration."""
    job = icx.cantilever_3d(ctx, 14, 6, 5)
    rho = O.random_density(job.nelx, 21)
    job.set_design(rho)
    sr = form == "sr"
    if sr:
        form = icx.SOLVE_MIXED
    ctx.set_option(icx.OPT_SINGLE_REDUCTION, int(sr))
    ctx.set_option(icx.OPT_GRAPHS, 0)
    try:
        u_ref, info_ref = job.solve(form, tol=1e-9, maxit=20000)
        ctx.set_option(icx.OPT_GRAPHS, 1)
        xv = np.random.default_rng(0).standard_normal(job.numeq)
        for rep in range(3):
            u, info = job.solve(form, tol=1e-9, maxit=20000)
            assert np.array_equal(u, u_ref), (form, rep)
            assert info["iterations"] == info_ref["iterations"]
            job.matvec(xv)          # eager work between replays
            ctx.synchronize()
    finally:
        ctx.set_option(icx.OPT_GRAPHS, 1)
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 0)


def test_graph_helmholtz_and_design_change(ctx):
    job = icx.cantilever_3d(ctx, 12, 6, 6)
    job.helmholtz_setup(2.0)
    x = O.random_density(job.nelx, 3)
    ctx.set_option(icx.OPT_GRAPHS, 0)
    try:
        h_ref, _ = job.helmholtz(x, tol=1e-12)
        job.set_design(x)
        u_ref, _ = job.solve(icx.SOLVE_MIXED, tol=1e-9)
    finally:
        ctx.set_option(icx.OPT_GRAPHS, 1)
    for _ in range(2):
        h, _ = job.helmholtz(x, tol=1e-12)
        assert np.array_equal(h, h_ref)
        job.set_design(x * 0.5 + 0.25)      # a different design between replays
        job.solve(icx.SOLVE_MIXED, tol=1e-9)
        job.set_design(x)
        u, _ = job.solve(icx.SOLVE_MIXED, tol=1e-9)
        assert np.array_equal(u, u_ref)


@pytest.mark.parametrize("form", [icx.SOLVE_PLAIN, icx.SOLVE_FUSED, icx.SOLVE_MIXED, "sr"])
@pytest.mark.parametrize("maxit", [20000, 37])
def test_pipelined_polling_equals_chunk_by_chunk(ctx, form, maxit):
    """With pipelined polling a chunk is queued after the one whose state the host
    reads; a chunk queued after the stop must change nothing.  Converged solves and
    solves stopped by the iteration limit (37 is not a multiple of any chunk length)
    give the same bits and the same counts with and without it ("sr": the mixed form
    with the single-reduction iteration)."""
    job = icx.cantilever_3d(ctx, 14, 6, 5)
    job.set_design(O.random_density(job.nelx, 23))
    sr = form == "sr"
    if sr:
        form = icx.SOLVE_MIXED
    ctx.set_option(icx.OPT_SINGLE_REDUCTION, int(sr))
    ctx.set_option(icx.OPT_PIPELINE, 0)
    try:
        u_ref, info_ref = job.solve(form, tol=1e-9, maxit=maxit)
        ctx.set_option(icx.OPT_PIPELINE, 1)
        assert info_ref["converged"] == (maxit > 1000)
        for _ in range(2):
            u, info = job.solve(form, tol=1e-9, maxit=maxit)
            assert np.array_equal(u, u_ref)
            assert _counts(info) == _counts(info_ref)
    finally:
        ctx.set_option(icx.OPT_PIPELINE, 1)
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 0)


def _counts(info):
    return {k: v for k, v in info.items() if k != "seconds"}


def test_pipelined_polling_helmholtz_and_heat(ctx):
    job = icx.cantilever_3d(ctx, 12, 6, 6)
    job.helmholtz_setup(2.0)
    x = O.random_density(job.nelx, 5)
    ctx.set_option(icx.OPT_PIPELINE, 0)
    try:
        h_ref, hi_ref = job.helmholtz(x, tol=1e-12)
    finally:
        ctx.set_option(icx.OPT_PIPELINE, 1)
    h, hi = job.helmholtz(x, tol=1e-12)
    assert np.array_equal(h, h_ref) and hi == hi_ref

    heat = icx.Job(ctx, 8, (10, 6, 5), (10.0, 6.0, 5.0), icx.HEAT)
    heat.add_dirichlet([1, 0, 0, 0, 0, 0], [0] * 6, [0, 10, 0, 6, 0, 0])
    heat.make_dof()
    heat.add_neumann([0.1, 0, 0], [0, 10, 0, 6, 5, 5])
    heat.set_conductivity((100.0, 80.0, 60.0, 5.0, 3.0, 2.0), (0.1, 0.2, 0.3, 0.01, 0.02, 0.03))
    heat.set_design(O.random_density(heat.nelx, 6, lo=0.05, hi=1.0))
    ctx.set_option(icx.OPT_PIPELINE, 0)
    try:
        t_ref, ti_ref = heat.solve(icx.SOLVE_PLAIN, tol=1e-10)
    finally:
        ctx.set_option(icx.OPT_PIPELINE, 1)
    t, ti = heat.solve(icx.SOLVE_PLAIN, tol=1e-10)
    assert ti_ref["converged"]
    assert np.array_equal(t, t_ref) and _counts(ti) == _counts(ti_ref)
