"""The design loops on the GPU against the oracle's driver loops
(opt-macro-oc/opt_macro_oc.cpp:84-172 and opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:81-163).
Both sides start from the same design and run the same number of design steps; the
trajectories are compared step by step."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_driver_a_oc_cone(ctx):
    nx, ny, steps = 60, 20, 15
    job = icx.cantilever_2d(ctx, nx, ny)
    pr = O.cantilever_2d(nx, ny)
    x0 = np.full(pr.nelx, 0.5)
    ref = pr.run_driver(0, x0, steps, 1.5, projection=True, direct=True)
    job.opt_init(x0, driver=0, filter=icx.FILTER_CONE, radius=1.5, form=icx.SOLVE_PLAIN,
                 tol=1e-13, maxit=50000)
    obj = np.array([job.opt_step()["objective"] for _ in range(steps)])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    # measured: 2.0e-12 and 4.3e-11 (the CG at 1e-13 against the oracle's direct solve)
    assert rel.max() < 1e-10, rel
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-9


def test_driver_b_mma_helmholtz(ctx):
    nx, ny, steps = 40, 14, 10
    job = icx.cantilever_2d(ctx, nx, ny)
    pr = O.cantilever_2d(nx, ny)
    x0 = np.full(pr.nelx, 0.4)
    ref = pr.run_driver(1, x0, steps, 2.0, projection=True, direct=True, helm_direct=True)
    job.opt_init(x0, driver=1, filter=icx.FILTER_HELMHOLTZ, radius=2.0, form=icx.SOLVE_PLAIN,
                 tol=1e-13, maxit=50000, helm_tol=1e-13)
    recs = [job.opt_step() for _ in range(steps)]
    obj = np.array([r["objective"] for r in recs])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    # measured: 3.3e-13, 2.6e-14 and 2.9e-12
    assert rel.max() < 1e-11, rel
    con = np.array([r["constraint"] for r in recs])
    assert np.abs(con - ref["vol"]).max() < 1e-12
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-10


def test_cxx_driver_b_executable_matches_oracle():
    exe = os.path.join(ROOT, "icarat-contact_amd", "_build", "opt_mma_helmholtz")
    toml = os.path.join(ROOT, "icarat-contact_amd", "host", "examples", "mbb_fdm_q4.toml")
    steps = 6
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([exe, toml, "Eigen_LDLT", str(steps)], cwd=d, capture_output=True,
                             text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        rows = np.loadtxt(os.path.join(d, "function.csv"), delimiter=",", ndmin=2)
        # the output step: one binary VTK file per step (20 x 10 quad4: 231 points)
        for k in range(1, steps + 1):
            head = open(os.path.join(d, "res%d.vtk" % k), "rb").read(120).split(b"\n")
            assert head[2] == b"BINARY" and head[4] == b"POINTS 231 float"
    pr = O.Problem(4, (20, 10, 1), (20.0, 10.0, 1.0), O.PLANE_STRAIN, young=(0.001, 200.0))
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 10, 0, 10])
    pr.make_dof()
    pr.add_neumann([0.0, -0.01, 0.0], [20, 20, 0, 10, 0, 0])
    # both sides solve the Helmholtz filter with CG at the reference's 1e-10; the
    # filter error (1e-10 relative per solve, three solves per step) carried through
    # the MMA steps bounds the agreement of the trajectories
    ref = pr.run_driver(1, np.full(pr.nelx, 0.4), steps, 2.0, projection=True, direct=True,
                        helm_tol=1e-10)
    assert rows.shape[0] == steps
    rel = np.abs(rows[:, 1] - ref["obj"]) / np.abs(ref["obj"])
    assert rel.max() <= 1e-7, rel


def test_eigen_ldlt_substitute_on_a_design_with_voids(ctx):
    """Driver B's "Eigen_LDLT" (SimplicialLDLT, solvers.cpp:178-200) runs as the FP64 plain
    form at rel-res 1e-14.  On the C1 cantilever's design after 60 OC iterations (862
    elements at xmin = 1e-4, i.e. E = Emin + 1e-12 ~ 1e-9: contrast 1e9) that solve meets
    the direct solution (the oracle's banded Cholesky) within 1e-10 in u and c."""
    nx, ny = 120, 40
    pr = O.cantilever_2d(nx, ny)
    ref = pr.run_driver(0, np.full(pr.nelx, 0.5), 60, 1.5, projection=True, direct=True)
    rho = ref["design"]
    assert (rho <= 1.0001e-4).sum() > 500 and rho.max() == 1.0   # voids and solid
    b = pr.assemble(rho)
    ud = pr.direct(b)
    job = icx.cantilever_2d(ctx, nx, ny)
    job.set_design(rho)
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-14, maxit=200000)
    assert info["converged"], info
    cd = ud @ pr.fext
    print("LDLT substitute: %d iterations, |du|/|u| %.3e, dc %.3e"
          % (info["iterations"], np.linalg.norm(u - ud) / np.linalg.norm(ud), abs(job.compliance() - cd) / cd))
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-10
    assert abs(job.compliance() - cd) / cd <= 1e-10
    job.close()


def test_reference_shaped_driver_b_equals_the_model_driver():
    """Driver B written with the reference's constructors (MeshIcarat / BCIcarat /
    LinearAnalysis<E, N>(fem, element, node, force, dis) / FilterHelmholtz<E, N>(fem,
    element, node, R) / MMASolver(n, m) over LinearElastic and Node records) follows the
    model-based C++ driver B step for step on the same job file (its addPlot output keeps
    the reference's six significant digits)."""
    build = os.path.join(ROOT, "icarat-contact_amd", "_build")
    toml = os.path.join(ROOT, "icarat-contact_amd", "host", "examples", "mbb_fdm_q4.toml")
    steps = 6
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        r1 = subprocess.run([os.path.join(build, "opt_mma_helmholtz_ref"), toml, str(steps)], cwd=d1,
                            capture_output=True, text=True, timeout=300)
        assert r1.returncode == 0, r1.stderr[-2000:]
        r2 = subprocess.run([os.path.join(build, "opt_mma_helmholtz"), toml, "Eigen_LDLT", str(steps)], cwd=d2,
                            capture_output=True, text=True, timeout=300)
        assert r2.returncode == 0, r2.stderr[-2000:]
        ref_shaped = np.loadtxt(os.path.join(d1, "function.csv"), ndmin=2)
        model = np.loadtxt(os.path.join(d2, "function.csv"), delimiter=",", ndmin=2)
        assert os.path.exists(os.path.join(d1, "res%d.vtk" % steps))
        assert os.path.exists(os.path.join(d1, "objective.csv"))
    assert ref_shaped.shape == (steps, 2) and np.array_equal(ref_shaped[:, 0], np.arange(1, steps + 1))
    assert np.all(np.abs(ref_shaped[:, 1] - model[:, 1]) <= 5e-6 * np.abs(model[:, 1]))
