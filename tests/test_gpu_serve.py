"""Serving loop: jobs arriving over time, admitted by footprint, design iterations of
all admitted jobs batched together; each job's result equals the same job run alone."""
import numpy as np
import pytest

import icx
from icx import serve

pytestmark = pytest.mark.gpu

SPECS = [(2, (30, 10, 1)), (3, (10, 5, 5)), (2, (24, 8, 1)), (3, (8, 4, 4))]


def _alone(ctx, ndim, div, iters, tol):
    j = serve.make_job(ctx, ndim, div)
    j.opt_init(np.full(j.nelx, 0.4), driver=0, radius=1.5, projection=True,
               form=icx.SOLVE_MIXED, tol=tol, maxit=50000)
    for _ in range(iters):
        j.opt_step()
    x = j.opt_design()
    j.close()
    return x


def test_served_designs_equal_alone(ctx):
    iters, tol = 3, 1e-8
    out = serve.run(ctx, rate=200.0, iters=iters, specs=SPECS, tol=tol, keep_designs=True)
    assert out["design_iterations"] == iters * len(SPECS)
    assert out["iteration_ms_p99"] >= out["iteration_ms_p50"] > 0
    for i, (ndim, div) in enumerate(SPECS):
        assert np.array_equal(out["designs"][i], _alone(ctx, ndim, div, iters, tol)), i


def test_admission_by_footprint(ctx):
    """A budget that holds one job at a time: jobs are admitted one after another and
    the footprint estimate covers what the job really holds."""
    est = [serve.footprint_estimate(*s) for s in SPECS]
    out = serve.run(ctx, rate=1000.0, iters=1, specs=SPECS, budget=max(est) * 1.01)
    assert out["peak_admitted_jobs"] == 1
    assert out["max_used_within_budget"]
    for real, estimate in out["footprint_checks"]:
        assert real <= estimate, (real, estimate)
