"""Serving loop: jobs arriving over time, admitted by footprint, design iterations of
all admitted jobs batched together; each job's result equals the same job run alone."""
import numpy as np
import pytest

import icx
from icx import serve

pytestmark = pytest.mark.gpu

SPECS = [(2, (30, 10, 1)), (3, (10, 5, 5)), (2, (24, 8, 1)), (3, (8, 4, 4))]


def _alone(ctx, ndim, div, iters, tol, form=None):
    j = serve.make_job(ctx, ndim, div)
    j.opt_init(np.full(j.nelx, 0.4), driver=0, radius=1.5, projection=True,
               form=serve.solve_form(ndim) if form is None else form, tol=tol, maxit=50000)
    for _ in range(iters):
        j.opt_step()
    x = j.opt_design()
    j.close()
    return x


def test_served_designs_equal_alone(ctx):
    """quad4 jobs are served in the FP64 fused form and hexa8 jobs in the mixed form,
    in one batch; every design equals its solo run bit for bit."""
    iters, tol = 3, 1e-8
    out = serve.run(ctx, rate=200.0, iters=iters, specs=SPECS, tol=tol, keep_designs=True)
    assert out["design_iterations"] == iters * len(SPECS)
    assert out["iteration_ms_p99"] >= out["iteration_ms_p50"] > 0
    assert out["unconverged_solves"] == 0
    forms = {r[0]: r[9] for r in out["iteration_records"]}
    assert forms == {i: serve.solve_form(nd) for i, (nd, _) in enumerate(SPECS)}
    assert set(forms.values()) == {icx.SOLVE_FUSED, icx.SOLVE_MIXED}
    for i, (ndim, div) in enumerate(SPECS):
        assert np.array_equal(out["designs"][i], _alone(ctx, ndim, div, iters, tol)), i


def test_served_with_one_form_for_all(ctx):
    """forms=SOLVE_MIXED serves every job in the mixed form (quad4 included)."""
    iters, tol = 2, 1e-8
    out = serve.run(ctx, rate=200.0, iters=iters, specs=SPECS, tol=tol, keep_designs=True, forms=icx.SOLVE_MIXED)
    assert {r[9] for r in out["iteration_records"]} == {icx.SOLVE_MIXED}
    for i, (ndim, div) in enumerate(SPECS):
        assert np.array_equal(out["designs"][i], _alone(ctx, ndim, div, iters, tol, icx.SOLVE_MIXED)), i


def test_admission_by_footprint(ctx):
    """A budget that holds one job at a time: jobs are admitted one after another and
    the footprint estimate covers what the job really holds."""
    est = [serve.footprint_estimate(*s) for s in SPECS]
    out = serve.run(ctx, rate=1000.0, iters=1, specs=SPECS, budget=max(est) * 1.01)
    assert out["peak_admitted_jobs"] == 1
    assert out["max_used_within_budget"]
    for real, estimate in out["footprint_checks"]:
        assert real <= estimate, (real, estimate)


TOML_Q4 = """# 40 x 12 quad4 cantilever, left edge clamped, traction strip at the free end
[mesh]
dimension = 2
type = "quad4"
division = [40, 12, 1]
length = [40.0, 12.0, 1.0]

[boundary_condition]
dirichlet = [
  [[1, 1, 0, 0, 0, 0], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [[0.0, 0.0], [0.0, 12.0], [0.0, 0.0]]],
]
neumann = [
  [[0.0, -0.01, 0.0], [[40.0, 40.0], [4.0, 8.0], [0.0, 0.0]]],
]

[material]
designs_0 = 0.45
penalty = 3.0
poisson = 0.3
state = "plane_stress"
type = "Elastic"
young = [0.001, 1.0]
"""


def test_served_toml_job_equals_the_cxx_driver(ctx, tmp_path):
    """A job given as the text of a reference job file is served (icx_job_create_toml,
    the front end the C++ drivers use) and follows the C++ driver A on the same file step
    by step (same FP64 fused-form solves at the reference's 1e-10, cone r = 3, projection);
    a malformed request is refused and the stream goes on."""
    import os
    import subprocess
    steps = 4
    bad = TOML_Q4.replace('type = "quad4"', 'type = "tria3"')
    out = serve.run(ctx, rate=500.0, iters=steps, specs=[TOML_Q4, bad, (2, (30, 10, 1))], tol=1e-10, radius=3.0,
                    keep_designs=True)
    assert [r[0] for r in out["rejected"]] == [1] and "tria3" in out["rejected"][0][1]
    assert sorted(out["designs"]) == [0, 2]
    f = tmp_path / "job.toml"
    f.write_text(TOML_Q4)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "icarat-contact_amd", "_build", "opt_macro_oc")
    res = subprocess.run([exe, str(f), "HIP_PCG_FUSED", str(steps), "3.0"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    rows = np.loadtxt(tmp_path / "function.csv", delimiter=",", ndmin=2)
    served = np.array(out["objectives"][0])
    assert rows.shape[0] == steps and len(served) == steps
    assert np.abs(served - rows[:, 1]).max() <= 1e-9 * np.abs(rows[:, 1]).max(), (served, rows[:, 1])


def test_resident_dof_cap(ctx):
    """With the resident DOFs capped the jobs queue at admission, one batch never holds
    more than the cap, and every design still equals its solo run."""
    iters, tol = 2, 1e-8
    cap = 2 * 31 * 11 + 3 * 11 * 6 * 6   # DOFs of the first two jobs: all four never fit
    out = serve.run(ctx, rate=1000.0, iters=iters, specs=SPECS, tol=tol, keep_designs=True, max_resident_dofs=cap)
    assert out["peak_resident_dofs"] <= cap and out["peak_admitted_jobs"] < len(SPECS)
    for i, (ndim, div) in enumerate(SPECS):
        assert np.array_equal(out["designs"][i], _alone(ctx, ndim, div, iters, tol)), i


def test_size_classes_equal_alone(ctx):
    """Two size classes (a context, stream and batch each, served by two host threads,
    kernels overlapping on the device): every design still equals its solo run, and the
    classes hold the jobs of their DOF ranges."""
    iters, tol = 3, 1e-8
    ndofs = [nd * (d0 + 1) * (d1 + 1) * ((d2 + 1) if nd == 3 else 1) for nd, (d0, d1, d2) in SPECS]
    out = serve.run(ctx, rate=500.0, iters=iters, specs=SPECS, tol=tol, keep_designs=True, classes=[700])
    assert out["size_classes_dofs"] == [700.0]
    assert sorted(ndofs) == [450, 675, 682, 1188]  # three jobs below the bound, one above
    assert out["design_iterations"] == iters * len(SPECS) and out["unconverged_solves"] == 0
    for i, (ndim, div) in enumerate(SPECS):
        assert np.array_equal(out["designs"][i], _alone(ctx, ndim, div, iters, tol)), i


def test_size_classes_share_the_budget(ctx):
    """A budget that holds one job: with two classes still one job at a time."""
    est = [serve.footprint_estimate(*s) for s in SPECS]
    out = serve.run(ctx, rate=1000.0, iters=1, specs=SPECS, budget=max(est) * 1.01, classes=[700])
    assert out["peak_admitted_jobs"] == 1 and out["max_used_within_budget"]
    assert out["design_iterations"] == len(SPECS)
