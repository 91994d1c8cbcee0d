"""The stress constraint on the GPU (constraint/opt-stress-constraint/sensitivity.cpp:8-155)
against the oracle: g and its adjoint derivative on solved designs, and the C++ driver
(opt_stress_constraint.cpp:11-166) step by step against the oracle's loop."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("state", [O.PLANE_STRESS, O.PLANE_STRAIN])
@pytest.mark.parametrize("qp,pn,slim", [(0.7, 8.0, 1.0), (0.5, 16.0, 0.3)])
def test_stress_constraint_matches_the_oracle(ctx, state, qp, pn, slim):
    nx, ny = 30, 12
    job = icx.cantilever_2d(ctx, nx, ny, state=state)
    pr = O.cantilever_2d(nx, ny, state=state)
    rho = np.random.default_rng(4).uniform(0.2, 1.0, pr.nelx)
    u_ref = pr.direct(pr.assemble(rho))
    g_ref, dg_ref, _ = pr.stress_constraint(pr.node_values(u_ref), qp, pn, slim)
    job.set_design(rho)
    u, _ = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=50000)
    g, dg, its = job.stress_constraint(slim, pn, qp, tol=1e-13, maxit=50000)
    assert its > 0
    assert g == pytest.approx(g_ref, rel=1e-9)
    assert np.abs(dg - dg_ref).max() <= 1e-8 * np.abs(dg_ref).max()
    # g alone (no adjoint solve) and the job's displacement left as it was
    g2, none, _ = job.stress_constraint(slim, pn, qp, sens=False)
    assert none is None and g2 == g
    assert np.array_equal(job.displacement(), u)


def test_stress_constraint_rejects_3d_and_mesh_jobs(ctx):
    job = icx.cantilever_3d(ctx, 6, 3, 3)
    job.set_design(np.full(job.nelx, 0.5))
    job.solve(icx.SOLVE_PLAIN, tol=1e-10, maxit=10000)
    with pytest.raises(icx.IcxError):
        job.stress_constraint(1.0, 8.0, 0.7)


def test_cxx_stress_driver_matches_the_oracle():
    exe = os.path.join(ROOT, "icarat-contact_amd", "_build", "opt_stress_constraint")
    toml = os.path.join(ROOT, "icarat-contact_amd", "host", "examples", "stress_q4.toml")
    steps = 10
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([exe, toml, "Eigen_CG", str(steps)], cwd=d, capture_output=True,
                             text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        rows = np.loadtxt(os.path.join(d, "function.csv"), delimiter=",", ndmin=2)
    pr = O.Problem(4, (40, 20, 1), (40.0, 20.0, 0.25), O.PLANE_STRESS, young=(0.001, 100.0))
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 20, 0, 0])
    pr.make_dof()
    pr.add_neumann([0.0, -0.1, 0.0], [40, 40, 8, 12, 0, 0])
    ref = pr.run_driver(3, np.full(pr.nelx, 0.5), steps, 2.0, projection=True, direct=False,
                        cg_tol=1e-10, stress=(1.0, 8.0, 0.7))
    assert rows.shape == (steps, 3)
    np.testing.assert_allclose(rows[:, 1], ref["obj"], rtol=1e-7)
    np.testing.assert_allclose(rows[:, 2], ref["vol"], rtol=1e-6, atol=1e-8)
