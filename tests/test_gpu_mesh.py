"""Unstructured-mesh jobs (icx_job_create_mesh, MeshMDPA meshes) against the oracle on
the same explicit meshes: the FP64 operator and its diagonal to rounding, the plain
solve to 1e-10 of the direct solution, compliance, sensitivities and volumes; a box
mesh given explicitly equals the structured job."""
import numpy as np
import pytest

import icx
import oracle as O
from meshgen import annulus_quad, box_mesh, distort, write_mdpa

pytestmark = pytest.mark.gpu

CASES = [
    (4, O.PLANE_STRESS, (12, 6), (12.0, 6.0)),
    (4, O.PLANE_STRAIN, (9, 7), (9.0, 7.0)),
    (8, O.SOLID3D, (7, 4, 5), (7.0, 4.0, 5.0)),
]


def _pair(ctx, ne, state, x, conn, L):
    job = icx.Job.from_mesh(ctx, ne, x, conn, state)
    pr = O.Problem.from_mesh(ne, x, conn, state)
    nd = 2 if ne == 4 else 3
    clamp = [0.0, 0.0, 0.0, L[1], 0.0, L[2] if nd == 3 else 0.0]
    flags = [1] * nd + [0] * (6 - nd)
    for o in (job, pr):
        o.add_dirichlet(flags, [0] * 6, clamp)
    assert job.make_dof() == pr.make_dof()
    load = [L[0], L[0], 0.0, L[1], 0.0, L[2] if nd == 3 else 0.0]
    val = [0.0, -0.1, 0.0] if nd == 2 else [0.0, 0.0, -0.1]
    for o in (job, pr):
        o.add_neumann(val, load)
    assert np.abs(job.fext() - pr.fext).max() <= 1e-15 * np.abs(pr.fext).max()
    rho = O.random_density(pr.nelx, 9, lo=0.05, hi=1.0)
    job.set_design(rho)
    b = pr.assemble(rho)
    return job, pr, rho, b


@pytest.mark.parametrize("ne,state,div,L", CASES)
def test_mesh_operator_solve_sensitivity(ctx, ne, state, div, L):
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 3, L)
    job, pr, rho, b = _pair(ctx, ne, state, xd, conn, L)
    v = np.random.default_rng(1).standard_normal(pr.numeq)
    y, yo = job.matvec(v), pr.spmv(v)
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    K = pr.csr()
    assert np.abs(job.diagonal() - K.diagonal()).max() <= 1e-14 * np.abs(K.diagonal()).max()
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=50000)
    ud = pr.direct(b)
    assert info["converged"]
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-10
    c = job.compliance()
    assert abs(c - ud @ pr.fext) / abs(ud @ pr.fext) < 1e-10
    dc, dv = job.sensitivity()
    vol = pr.volumes()
    dco, dvo = pr.sensitivity(rho, pr.element_energy(pr.node_values(ud)), vol)
    assert np.abs(dc - dco).max() <= 1e-8 * np.abs(dco).max()
    assert np.abs(dv - vol).max() <= 1e-14 * vol.max()
    assert np.abs(job.volumes() - vol).max() <= 1e-14 * vol.max()
    # element von Mises stress of the Gauss-point mean strain
    mo = pr.mises(pr.node_values(ud))
    assert np.abs(job.mises() - mo).max() <= 1e-8 * np.abs(mo).max()
    # K u over all DOFs: the applied load on the free DOFs, and zero in total (K 1 = 0:
    # the supports carry the load)
    R = job.reactions().reshape(-1, 3)
    assert np.abs(R.sum(axis=0)).max() <= 1e-12 * np.abs(R).max()
    free = [(n, c, pr.dof[n * pr.dofnp + c]) for n in range(pr.numnp) for c in range(pr.dofnp)
            if pr.dof[n * pr.dofnp + c] < pr.numeq]
    Rf = np.array([R[n, c] for n, c, _ in free])
    F = np.array([pr.fext[d] for _, _, d in free])
    assert np.abs(Rf - F).max() <= 1e-9 * np.abs(F).max()


@pytest.mark.parametrize("ne,state,div,L", CASES)
def test_mesh_mixed_form_meets_true_residual(ctx, ne, state, div, L):
    """The mixed form on a distorted mesh (FP32 Gauss-point element pass, FP32 inner
    Jacobi-PCG, FP64 residual replacement): the true residual through the oracle's CSR
    matrix meets the tolerance, and u matches the direct solution."""
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 3, L)
    job, pr, rho, b = _pair(ctx, ne, state, xd, conn, L)
    ud = pr.direct(b)
    K = pr.csr()
    for tol in (1e-8, 1e-11):
        u, info = job.solve(icx.SOLVE_MIXED, tol=tol, maxit=50000)
        assert info["converged"] and info["refinements"] >= 1, (tol, info)
        assert np.linalg.norm(b - K @ u) / np.linalg.norm(b) < tol, tol
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-8
    # the plain form on the same job afterwards still follows the direct solution
    u2, _ = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=50000)
    assert np.linalg.norm(u2 - ud) / np.linalg.norm(ud) < 1e-10


@pytest.mark.parametrize("ne,state,div,L", CASES)
def test_mesh_fused_form_matches_plain_form_and_oracle(ctx, ne, state, div, L):
    """The FP64 fused form on a distorted mesh (direction and deferred x step in place,
    FP64 element pass, gather with the p.Ap partials, FP64 kernel B): the same PCG as the
    plain form up to rounding, so the iteration counts agree within one and both follow
    the direct solution; the residual through the oracle's CSR meets the tolerance."""
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 5, L)
    job, pr, rho, b = _pair(ctx, ne, state, xd, conn, L)
    ud = pr.direct(b)
    K = pr.csr()
    for tol in (1e-10, 1e-13):
        up, pi = job.solve(icx.SOLVE_PLAIN, tol=tol, maxit=50000)
        uf, fi = job.solve(icx.SOLVE_FUSED, tol=tol, maxit=50000)
        assert fi["converged"] and pi["converged"], (fi, pi)
        assert abs(fi["iterations"] - pi["iterations"]) <= 1, (fi, pi)
        # the true residual meets the tolerance down to FP64 CG's attainable accuracy
        # (about 1e-12 here, the same for both forms)
        tr_f = np.linalg.norm(b - K @ uf) / np.linalg.norm(b)
        tr_p = np.linalg.norm(b - K @ up) / np.linalg.norm(b)
        assert tr_f < max(1.01 * tol, 10 * tr_p), (tol, tr_f, tr_p)
        assert np.linalg.norm(uf - up) / np.linalg.norm(up) < 1e-8
    assert np.linalg.norm(uf - ud) / np.linalg.norm(ud) < 1e-10
    c_d = float(b @ ud)
    assert abs(job.compliance() - c_d) <= 1e-10 * abs(c_d)


def test_mesh_driver_a_fused_form(ctx):
    """Driver A on a distorted quad4 mesh with the FP64 fused form at 1e-13: objectives
    and design follow the oracle's loop with direct solves."""
    ne, div, L, steps = 4, (24, 8), (24.0, 8.0), 8
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 21, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, icx.PLANE_STRESS)
    pr = O.Problem.from_mesh(ne, xd, conn, O.PLANE_STRESS)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, 0])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, -0.05, 0.0], [L[0], L[0], 0.0, 2.0, 0, 0])
    x0 = np.full(pr.nelx, 0.5)
    ref = pr.run_driver(0, x0, steps, 1.5, projection=True, direct=True)
    job.opt_init(x0, driver=0, filter=icx.FILTER_CONE, radius=1.5, form=icx.SOLVE_FUSED,
                 tol=1e-13, maxit=50000)
    obj = np.array([job.opt_step()["objective"] for _ in range(steps)])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    assert rel.max() < 1e-8, rel
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-7


def test_mesh_driver_a_mixed_form(ctx):
    """Driver A on a distorted hexa8 mesh with the mixed form at 1e-11: objectives and
    design follow the oracle's loop with direct solves."""
    ne, div, L, steps = 8, (8, 4, 4), (8.0, 4.0, 4.0), 6
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 14, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, icx.SOLID3D)
    pr = O.Problem.from_mesh(ne, xd, conn, O.SOLID3D)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, L[2]])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, 0.0, -0.1], [L[0], L[0], 0.0, L[1], 0.0, L[2]])
    x0 = np.full(pr.nelx, 0.5)
    ref = pr.run_driver(0, x0, steps, 1.5, projection=True, direct=True)
    job.opt_init(x0, driver=0, filter=icx.FILTER_CONE, radius=1.5, form=icx.SOLVE_MIXED,
                 tol=1e-11, maxit=50000)
    obj = np.array([job.opt_step()["objective"] for _ in range(steps)])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    assert rel.max() < 1e-8, rel
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-6


@pytest.mark.parametrize("ne,div,L", [(4, (10, 5), (10.0, 5.0)), (8, (6, 3, 4), (6.0, 3.0, 4.0))])
def test_box_mesh_equals_structured_job(ctx, ne, div, L):
    """The structured box given as an explicit mesh: same operator, same solve."""
    state = O.PLANE_STRESS if ne == 4 else O.SOLID3D
    x, conn = box_mesh(ne, div, L)
    a = icx.Job(ctx, ne, div, L, state)
    b = icx.Job.from_mesh(ctx, ne, x, conn, state)
    nd = 2 if ne == 4 else 3
    for j in (a, b):
        j.add_dirichlet([1] * nd + [0] * (6 - nd), [0] * 6, [0, 0, 0, L[1], 0, L[2] if nd == 3 else 0])
        j.make_dof()
        j.add_neumann([0.0, -1.0, 0.0] if nd == 2 else [0.0, 0.0, -1.0],
                      [L[0], L[0], 0, L[1], 0, L[2] if nd == 3 else 0])
        j.set_design(O.random_density(j.nelx, 4))
    assert np.array_equal(a.fext(), b.fext())
    v = np.random.default_rng(2).standard_normal(a.numeq)
    ya, yb = a.matvec(v), b.matvec(v)
    assert np.abs(ya - yb).max() <= 1e-13 * np.abs(ya).max()
    ua, _ = a.solve(icx.SOLVE_PLAIN, tol=1e-12, maxit=50000)
    ub, _ = b.solve(icx.SOLVE_PLAIN, tol=1e-12, maxit=50000)
    assert np.linalg.norm(ua - ub) / np.linalg.norm(ua) < 1e-9
    assert np.abs(a.volumes() - b.volumes()).max() <= 1e-14


def test_mdpa_annulus_with_curved_traction(ctx, tmp_path):
    """A quarter annulus read from an MDPA file, clamped on its x = 0 edge, loaded on the
    y = 0 edge: tractions integrated on the element sides the box selects (judgeLine)."""
    x, conn = annulus_quad(6, 12, 1.0, 2.0)
    path = tmp_path / "ring.mdpa"
    write_mdpa(path, x, conn, 4)
    xr, cr = icx.read_mdpa(path, icx.QUAD4)
    job = icx.Job.from_mesh(ctx, icx.QUAD4, xr, cr, icx.PLANE_STRAIN)
    pr = O.Problem.from_mesh(4, xr, cr, O.PLANE_STRAIN)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [-1e-9, 1e-9, 0.0, 3.0, 0, 0])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, 0.3, 0.0], [1.0, 2.0, -1e-9, 1e-9, 0, 0])
    assert np.abs(job.fext() - pr.fext).max() <= 1e-15
    rho = O.random_density(pr.nelx, 2, lo=0.1, hi=1.0)
    job.set_design(rho)
    b = pr.assemble(rho)
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=50000)
    ud = pr.direct(b)
    assert info["converged"] and np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-10


def test_mesh_job_rejects_box_only_calls(ctx):
    x, conn = box_mesh(4, (4, 2), (4.0, 2.0))
    job = icx.Job.from_mesh(ctx, icx.QUAD4, x, conn, icx.PLANE_STRESS)
    job.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 2, 0, 0])
    job.make_dof()
    job.set_design(np.full(job.nelx, 0.5))
    v = np.ones(job.numeq)
    for call in (lambda: job.matvec(v, prec=1), lambda: job.submit(tol=1e-8)):
        with pytest.raises(icx.IcxError) as e:
            call()
        assert e.value.code == 1
    bad = conn.reshape(-1, 4)[:, ::-1].reshape(-1).copy()  # clockwise: det J < 0
    with pytest.raises(icx.IcxError) as e:
        icx.Job.from_mesh(ctx, icx.QUAD4, x, bad, icx.PLANE_STRESS)
    assert e.value.code == 6


@pytest.mark.parametrize("ne,state,div,L,r", [(4, O.PLANE_STRESS, (14, 7), (14.0, 7.0), 1.5),
                                              (4, O.PLANE_STRAIN, (10, 6), (10.0, 6.0), 3.0),
                                              (8, O.SOLID3D, (6, 4, 4), (6.0, 4.0, 4.0), 1.6)])
def test_mesh_cone_filter_matches_brute_force(ctx, ne, state, div, L, r):
    """Neighbours by centre distance on a distorted mesh (cell grid on the device side,
    the O(n^2) scan of Filter::preNeighbour on the oracle side): forward and adjoint."""
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 6, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, state)
    f = O.Problem.from_mesh(ne, xd, conn, state).cone_filter(r, brute_force=True)
    job.cone_setup(r)
    v = np.random.default_rng(4).uniform(0.0, 1.0, job.nelx)
    fwd = f.apply(v)
    assert np.abs(job.cone(v) - fwd).max() <= 1e-14
    s = np.random.default_rng(5).standard_normal(job.nelx)
    assert np.abs(job.cone(s, adjoint=True) - f.apply(s, sens=True)).max() <= 1e-14 * np.abs(s).max()


def test_mesh_driver_a_matches_oracle(ctx):
    """Driver A (cone filter, projection, OC) on a distorted quad4 mesh: the device design
    loop against the oracle's driver loop, step by step."""
    ne, div, L, steps = 4, (30, 10), (30.0, 10.0), 10
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 8, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, icx.PLANE_STRESS)
    pr = O.Problem.from_mesh(ne, xd, conn, O.PLANE_STRESS)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, 0])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, -0.05, 0.0], [L[0], L[0], 0.0, 2.0, 0, 0])
    x0 = np.full(pr.nelx, 0.5)
    ref = pr.run_driver(0, x0, steps, 1.5, projection=True, direct=True)
    job.opt_init(x0, driver=0, filter=icx.FILTER_CONE, radius=1.5, form=icx.SOLVE_PLAIN,
                 tol=1e-13, maxit=50000)
    recs = [job.opt_step() for _ in range(steps)]
    obj = np.array([r["objective"] for r in recs])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    assert rel.max() < 1e-8, rel
    vol = np.array([r["volume"] for r in recs])
    assert np.abs(vol - ref["vol"]).max() <= 1e-8 * np.abs(ref["vol"]).max()
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-7


ROOT = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))


def test_cxx_driver_a_reads_mdpa_and_matches_oracle(tmp_path):
    """The C++ driver A (opt_macro_oc <toml> <mdpa>, as the reference's opt-macro-oc)
    on a distorted quad4 mesh: its function.csv against the oracle's driver loop, and
    a binary VTK file per step with the mesh's points and cells."""
    import os
    import subprocess
    ne, div, L, steps = 4, (24, 8), (24.0, 8.0), 6
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 12, L)
    write_mdpa(tmp_path / "plate.mdpa", xd, conn, ne)
    (tmp_path / "job.toml").write_text("""[mesh]
dimension = 2
type = ["quad4", "tria3"]

[boundary_condition]
dirichlet = [
  [[1, 1, 0, 0, 0, 0], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [[0.0, 0.0], [0.0, 8.0], [0.0, 0.0]]],
]
neumann = [
  [[0.0, -0.01, 0.0], [[24.0, 24.0], [0.0, 2.0], [0.0, 0.0]]],
]

[material]
designs_0 = 0.4
penalty = 3.0
poisson = 0.3
state = "plane_strain"
type = "Elastic"
young = [0.001, 100.0]
""")
    exe = os.path.join(ROOT, "icarat-contact_amd", "_build", "opt_macro_oc")
    out = subprocess.run([exe, "job.toml", "plate.mdpa", "Eigen_CG", str(steps), "1.5"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = np.loadtxt(tmp_path / "function.csv", delimiter=",", ndmin=2)
    assert rows.shape[0] == steps
    head = open(tmp_path / "res1.vtk", "rb").read(120).split(b"\n")
    assert head[2] == b"BINARY" and head[4] == ("POINTS %d float" % (len(xd) // 3)).encode()
    pr = O.Problem.from_mesh(ne, xd, conn, O.PLANE_STRAIN, young=(0.001, 100.0))
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 8, 0, 0])
    pr.make_dof()
    pr.add_neumann([0.0, -0.01, 0.0], [24, 24, 0, 2, 0, 0])
    ref = pr.run_driver(0, np.full(pr.nelx, 0.4), steps, 1.5, projection=True, direct=True)
    # the driver solves with CG to the reference's 1e-10, the oracle directly
    rel = np.abs(rows[:, 1] - ref["obj"]) / np.abs(ref["obj"])
    assert rel.max() <= 1e-7, rel
    # with a tria3 block in the file, the GPU path refuses the mesh
    with open(tmp_path / "plate.mdpa", "a") as f:
        f.write("\nBegin Elements SmallDisplacementElement2D3N\n1 0 1 2 3\nEnd Elements\n")
    out = subprocess.run([exe, "job.toml", "plate.mdpa", "Eigen_CG", "1"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode != 0 and "tria3" in out.stderr


@pytest.mark.parametrize("ne,state,div,L,R", [(4, O.PLANE_STRESS, (16, 8), (16.0, 8.0), 2.0),
                                              (4, O.PLANE_STRAIN, (9, 5), (9.0, 5.0), 3.5),
                                              (8, O.SOLID3D, (7, 4, 5), (7.0, 4.0, 5.0), 1.6)])
def test_mesh_helmholtz_filter_matches_oracle(ctx, ne, state, div, L, R):
    """FilterHelmholtz on a distorted mesh (element matrices integrated per element on the
    device, never stored): forward and sensitivity modes against the oracle's assembled
    CSR system solved directly, constants unchanged."""
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.22, 5, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, state)
    H = O.Problem.from_mesh(ne, xd, conn, state).helmholtz(R)
    job.helmholtz_setup(R)
    v = np.random.default_rng(8).uniform(0.0, 1.0, job.nelx)
    for sens in (False, True):
        ref, _ = H.apply(v, sens=sens, direct=True)
        out, it = job.helmholtz(v, sens=sens, tol=1e-13)
        assert it > 0
        assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max(), sens
    out, _ = job.helmholtz(np.full(job.nelx, 0.7), tol=1e-13)
    assert np.abs(out - 0.7).max() < 1e-11
    # a second radius on the same job rebuilds the diagonal
    job.helmholtz_setup(1.2 * R)
    ref, _ = O.Problem.from_mesh(ne, xd, conn, state).helmholtz(1.2 * R).apply(v, direct=True)
    out, _ = job.helmholtz(v, tol=1e-13)
    assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max()


def test_mesh_driver_b_matches_oracle(ctx):
    """Driver B (Helmholtz filter, projection, MMA) on a distorted quad4 mesh: the device
    design loop against the oracle's driver loop, step by step."""
    ne, div, L, steps = 4, (24, 10), (24.0, 10.0), 8
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 9, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, icx.PLANE_STRESS)
    pr = O.Problem.from_mesh(ne, xd, conn, O.PLANE_STRESS)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, 0])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, -0.05, 0.0], [L[0], L[0], 0.0, 2.0, 0, 0])
    x0 = np.full(pr.nelx, 0.4)
    ref = pr.run_driver(1, x0, steps, 2.0, projection=True, direct=True, helm_direct=True)
    job.opt_init(x0, driver=1, filter=icx.FILTER_HELMHOLTZ, radius=2.0, form=icx.SOLVE_PLAIN,
                 tol=1e-13, maxit=50000, helm_tol=1e-13)
    recs = [job.opt_step() for _ in range(steps)]
    obj = np.array([r["objective"] for r in recs])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    assert rel.max() < 1e-7, rel
    con = np.array([r["constraint"] for r in recs])
    assert np.abs(con - ref["vol"]).max() < 1e-8
    assert np.abs(job.opt_design() - ref["design"]).max() < 1e-6


def test_mesh_oc_with_helmholtz_filter(ctx):
    """OC with the Helmholtz filter on a hexa8 mesh (the bisection evaluates the volume of
    the filtered, projected candidate: one filter solve per bisection step) against the
    oracle's OC on the same filter."""
    ne, div, L, steps = 8, (8, 4, 4), (8.0, 4.0, 4.0), 4
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 10, L)
    job = icx.Job.from_mesh(ctx, ne, xd, conn, icx.SOLID3D)
    pr = O.Problem.from_mesh(ne, xd, conn, O.SOLID3D)
    for o in (job, pr):
        o.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, L[2]])
    assert job.make_dof() == pr.make_dof()
    for o in (job, pr):
        o.add_neumann([0.0, 0.0, -0.1], [L[0], L[0], 0.0, L[1], 0.0, L[2]])
    x0 = np.full(pr.nelx, 0.5)
    job.opt_init(x0, driver=0, filter=icx.FILTER_HELMHOLTZ, radius=1.5, form=icx.SOLVE_PLAIN,
                 tol=1e-13, maxit=50000, helm_tol=1e-13)
    recs = [job.opt_step() for _ in range(steps)]
    # the oracle: the same loop written out with its own filter, solve and OC
    H = pr.helmholtz(1.5)
    vol = pr.volumes()
    xk = x0.copy()
    for k in range(1, steps + 1):
        beta = min(2.0 ** (k // 40), 16.0)
        fil, _ = H.apply(xk, direct=True)
        thr = O.threshold(fil, beta)
        u = pr.direct(pr.assemble(thr))
        obj = u @ pr.fext
        assert abs(recs[k - 1]["objective"] - obj) <= 1e-8 * obj, k
        if k == 1:
            V0 = (thr * vol).sum()
        dfds, dgds = pr.sensitivity(thr, pr.element_energy(pr.node_values(u)), vol)
        dth = O.threshold_derivative(fil, beta)
        dfds, _ = H.apply(dfds * dth, sens=True, direct=True)
        dgds, _ = H.apply(dgds * dth, sens=True, direct=True)

        def gk(xc):
            f, _ = H.apply(xc, direct=True)
            return (O.threshold(f, beta) * vol).sum()
        n = len(xk)
        xk, _ = O.oc_update(xk, dfds, V0, dgds, np.full(n, 1e-4), np.full(n, 1.0), gk, move=0.15)
    assert np.abs(job.opt_design() - xk).max() < 1e-6
