"""GCMMA on the GPU against the oracle: the solver's entry points
(src/optimization/opt-solver/GCMMASolver.cpp:66-129) on the same data, and the design
loop of opt_macro_gcmma.cpp:83-216 (cone filter, GCMMA with inner loops that re-solve
trial designs) step by step."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.1, 0.9, n)
    df = -rng.uniform(0.0, 2.0, n) + 0.2 * rng.standard_normal(n)
    dg = rng.uniform(0.5, 1.5, n)
    return x, df, dg


@pytest.mark.parametrize("seed", [0, 1])
def test_gcmma_updates_match_the_oracle(ctx, seed):
    """Three outer updates, each followed by two inner updates with under-estimated
    approximations (raa grows), on a job's element count; the conservativeness check
    on both sides."""
    job = icx.cantilever_2d(ctx, 30, 10)
    n = job.nelx
    ref = O.GCMMA(n)
    lo, hi = np.full(n, 1e-4), np.ones(n)
    x, df, dg = _data(n, seed)
    f0, g0 = 3.0, 0.02 * n
    for k in range(3):
        xm_d, it_d = job.gcmma_outer(x, f0, df, g0, dg)
        xm_r, it_r = ref.outer(x, f0, df, g0, dg, lo, hi)
        assert it_d == it_r
        np.testing.assert_allclose(xm_d, xm_r, rtol=0, atol=1e-11)
        st = ref.state()
        for fn, gn, expect in ((st["f0app"] + 1e-4, st["fapp"][0] - 1.0, False),
                               (st["f0app"] - 1e-4, st["fapp"][0] - 1e-4, True)):
            assert job.gcmma_concheck(fn, gn) == ref.concheck(fn, gn) == expect
        for _ in range(2):
            st = ref.state()
            fn, gn = st["f0app"] + 0.05, st["fapp"][0] + 0.01
            xm_d, it_d = job.gcmma_inner(xm_d, fn, gn, x, f0, df, g0, dg)
            xm_r, it_r = ref.inner(xm_r, fn, gn, x, f0, df, g0, dg, lo, hi)
            assert it_d == it_r
            np.testing.assert_allclose(xm_d, xm_r, rtol=0, atol=1e-10)
        x = xm_r
        df = df * 0.9 + 0.05
        f0 *= 0.95


def test_gcmma_inner_before_outer_is_an_error(ctx):
    job = icx.cantilever_2d(ctx, 8, 4)
    n = job.nelx
    with pytest.raises(icx.IcxError):
        job.gcmma_inner(np.full(n, 0.5), 1.0, 0.0, np.full(n, 0.5), 1.0, -np.ones(n), 0.0, np.ones(n))
    with pytest.raises(icx.IcxError):
        job.gcmma_concheck(1.0, 0.0)


@pytest.mark.parametrize("form", [icx.SOLVE_PLAIN, icx.SOLVE_FUSED])
def test_gcmma_driver_matches_the_oracle(ctx, form):
    """Driver 2 against the oracle's loop (direct solves): objective and constraint per
    step, the number of inner loops per step and the design."""
    nx, ny, steps = 40, 14, 18     # inner loops at steps 1 and 16-18 on this problem
    job = icx.cantilever_2d(ctx, nx, ny)
    pr = O.cantilever_2d(nx, ny)
    x0 = np.full(pr.nelx, 0.3)
    ref = pr.run_driver(2, x0, steps, 2.0, projection=True, direct=True)
    job.opt_init(x0, driver=2, filter=icx.FILTER_CONE, radius=2.0, form=form, tol=1e-13, maxit=50000)
    recs = [job.opt_step() for _ in range(steps)]
    obj = np.array([r["objective"] for r in recs])
    con = np.array([r["constraint"] for r in recs])
    inner = np.array([r["inner_loops"] for r in recs])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    assert rel.max() < 1e-7, rel
    assert np.abs(con - ref["vol"]).max() < 1e-7 * pr.nelx
    assert np.array_equal(inner, ref["inner"]), (inner, ref["inner"])
    assert inner.sum() > 0            # the run exercises the inner loop
    assert all(r["trial_cg_iterations"] > 0 for r in recs)
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-6
    # the loop lowers the compliance and keeps the volume of the first design
    assert obj[-1] < obj[0]
    assert con[0] == 0.0 and con[-1] < 1e-3 * pr.nelx


def test_cxx_gcmma_driver_executable_matches_oracle():
    """host/opt_macro_gcmma (the reference driver's main loop over the C ABI) on the
    20 x 10 quad4 example, direct solves, against the oracle's driver loop."""
    import os
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "icarat-contact_amd", "_build", "opt_macro_gcmma")
    toml = os.path.join(root, "icarat-contact_amd", "host", "examples", "mbb_fdm_q4.toml")
    steps = 8
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([exe, toml, "Eigen_LDLT", str(steps)], cwd=d, capture_output=True,
                             text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        rows = np.loadtxt(os.path.join(d, "function.csv"), delimiter=",", ndmin=2)
    pr = O.Problem(4, (20, 10, 1), (20.0, 10.0, 1.0), O.PLANE_STRAIN, young=(0.001, 200.0))
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 10, 0, 10])
    pr.make_dof()
    pr.add_neumann([0.0, -0.01, 0.0], [20, 20, 0, 10, 0, 0])
    ref = pr.run_driver(2, np.full(pr.nelx, 0.4), steps, 2.0, projection=True, direct=True)
    assert rows.shape == (steps, 4)
    rel = np.abs(rows[:, 1] - ref["obj"]) / np.abs(ref["obj"])
    assert rel.max() <= 1e-8, rel
    assert np.abs(rows[:, 2] - ref["vol"]).max() <= 1e-8 * pr.nelx
    assert np.array_equal(rows[:, 3].astype(int), ref["inner"])


def test_gcmma_driver_through_the_split_step(ctx):
    """Driver 2 through icx_opt_step_begin / _end with the caller's own solve between
    them (the serving loop's pattern); the trial solves run inside _end."""
    nx, ny, steps = 40, 14, 18
    job = icx.cantilever_2d(ctx, nx, ny)
    pr = O.cantilever_2d(nx, ny)
    x0 = np.full(pr.nelx, 0.3)
    ref = pr.run_driver(2, x0, steps, 2.0, projection=True, direct=True)
    job.opt_init(x0, driver=2, filter=icx.FILTER_CONE, radius=2.0, form=icx.SOLVE_PLAIN, tol=1e-13, maxit=50000)
    recs = []
    for _ in range(steps):
        job.opt_step_begin()
        _, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=50000, want_u=False)
        recs.append(job.opt_step_end(info))
    obj = np.array([r["objective"] for r in recs])
    assert (np.abs(obj - ref["obj"]) / ref["obj"]).max() < 1e-7
    assert np.array_equal(np.array([r["inner_loops"] for r in recs]), ref["inner"])


def test_gcmma_driver_on_a_distorted_mesh(ctx):
    """Driver 2 on an unstructured quad4 mesh (per-element volumes in V0, the constraint
    and dg, cone filter on the mesh's neighbour lists) against the oracle."""
    from meshgen import box_mesh, distort
    ne, div, L, steps = 4, (30, 10), (30.0, 10.0), 12
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 8, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, icx.PLANE_STRESS)
    pr = O.Problem.from_mesh(ne, xd, conn, O.PLANE_STRESS)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, 0])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, -0.05, 0.0], [L[0], L[0], 0.0, 2.0, 0, 0])
    x0 = np.full(pr.nelx, 0.4)
    ref = pr.run_driver(2, x0, steps, 2.0, projection=True, direct=True)
    job.opt_init(x0, driver=2, filter=icx.FILTER_CONE, radius=2.0, form=icx.SOLVE_PLAIN, tol=1e-13, maxit=50000)
    recs = [job.opt_step() for _ in range(steps)]
    obj = np.array([r["objective"] for r in recs])
    con = np.array([r["constraint"] for r in recs])
    assert (np.abs(obj - ref["obj"]) / ref["obj"]).max() < 1e-7
    assert np.abs(con - ref["vol"]).max() <= 1e-8 * L[0] * L[1]
    assert np.array_equal(np.array([r["inner_loops"] for r in recs]), ref["inner"])
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-6
