"""Captured iteration chunks (HIP graphs) against eager launches: every replay of a
job's graph must give the same bits as the eager solve (the kernels are deterministic),
with eager work and synchronisations between the replays."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [icx.SOLVE_PLAIN, icx.SOLVE_FUSED, icx.SOLVE_MIXED])
def test_graph_replays_equal_eager(ctx, form):
    job = icx.cantilever_3d(ctx, 14, 6, 5)
    rho = O.random_density(job.nelx, 21)
    job.set_design(rho)
    ctx.set_option(icx.OPT_GRAPHS, 0)
    try:
        u_ref, info_ref = job.solve(form, tol=1e-9, maxit=20000)
    finally:
        ctx.set_option(icx.OPT_GRAPHS, 1)
    xv = np.random.default_rng(0).standard_normal(job.numeq)
    for rep in range(3):
        u, info = job.solve(form, tol=1e-9, maxit=20000)
        assert np.array_equal(u, u_ref), (form, rep)
        assert info["iterations"] == info_ref["iterations"]
        job.matvec(xv)          # eager work between replays
        ctx.synchronize()


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
