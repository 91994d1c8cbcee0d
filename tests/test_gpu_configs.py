"""The benchmark workloads at their own sizes, on the GPU, against the oracle.

C1  Q4 120 x 40 plane-stress cantilever, the OC driver (opt-macro-oc/opt_macro_oc.cpp:
    84-172): cone r = 1.5, projection, x0 = 0.5, 50 iterations, FP64 plain form at 1e-13
    against the oracle's driver loop with direct solves.
C2  hexa8 96 x 48 x 48, densities U[0.01, 1] (seed 20261203): the plain form at 1e-13
    against the oracle's Eigen-CG restatement on its assembled CSR matrix (u, c, dc), and
    the mixed form at 1e-8 with its true residual taken through the oracle's CSR product.
C3  hexa8 192 x 96 x 96: the 20-iteration OC design loop with the Helmholtz filter
    (R = 3) in the mixed form, then on the loop's physical design the mixed form against
    the FP64 plain form at 1e-12, the energy identity sum_e E_e u_e^T K0 u_e = u . f, CPU
    spot checks of the FP64 operator and the element energies on sampled nodes and
    elements (oracle element matrix), and mass conservation of the Helmholtz filter.
C4  64 jobs (seed 20261204): the batch equals every job alone bit for bit, and every
    job's true residual through the oracle's CSR product meets 1e-8.
"""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu

SEED_C2 = 20261203
SEED_C4 = 20261204


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------------------------------------- C1
@pytest.mark.timeout(600)
def test_c1_oc_driver_50_iterations(ctx):
    nx, ny, steps = 120, 40, 50
    job = icx.cantilever_2d(ctx, nx, ny)
    pr = O.cantilever_2d(nx, ny)
    assert job.numeq == pr.numeq == 9840
    assert np.array_equal(job.dof_map(), pr.dof)
    x0 = np.full(pr.nelx, 0.5)
    ref = pr.run_driver(0, x0, steps, 1.5, projection=True, direct=True)
    job.opt_init(x0, driver=0, filter=icx.FILTER_CONE, radius=1.5, form=icx.SOLVE_PLAIN,
                 tol=1e-13, maxit=100000)
    recs = [job.opt_step() for _ in range(steps)]
    obj = np.array([r["objective"] for r in recs])
    vol = np.array([r["volume"] for r in recs])
    rel = np.abs(obj - ref["obj"]) / ref["obj"]
    print("C1: max rel objective diff %.3e over %d steps" % (rel.max(), steps))
    assert rel.max() <= 1e-10, rel
    assert np.abs(vol - ref["vol"]).max() <= 1e-10 * pr.nelx
    assert np.abs(job.opt_design() - ref["design"]).max() <= 1e-8
    job.close()


# ------------------------------------------------------------------------------- C2
@pytest.fixture(scope="module")
def c2(ctx):
    nx, ny, nz = 96, 48, 48
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    pr = O.cantilever_3d(nx, ny, nz)
    rho = O.random_density(pr.nelx, SEED_C2)
    assert np.array_equal(rho, icx.random_density(job.nelx, SEED_C2))
    job.set_design(rho)
    b = pr.assemble(rho)
    uo, ito, _ = pr.cg(b, tol=1e-13, maxit=100000)
    yield dict(job=job, pr=pr, rho=rho, b=b, uo=uo, ito=ito)
    job.close()


@pytest.mark.timeout(600)
def test_c2_setup_bit_exact(c2):
    job, pr = c2["job"], c2["pr"]
    assert job.numeq == pr.numeq == 691488
    assert np.array_equal(job.dof_map(), pr.dof)
    x, conn = job.mesh()
    assert np.array_equal(x, pr.x) and np.array_equal(conn, pr.conn)
    assert np.abs(job.fext() - pr.fext).max() <= 1e-15


@pytest.mark.timeout(600)
def test_c2_plain_form_against_oracle_cg(c2):
    job, pr, rho, uo = c2["job"], c2["pr"], c2["rho"], c2["uo"]
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=100000)
    assert info["converged"]
    du = _rel(u, uo)
    co = float(uo @ pr.fext)
    dcrel = abs(job.compliance() - co) / co
    dc, _ = job.sensitivity()
    dco, _ = pr.sensitivity(rho, pr.element_energy(pr.node_values(uo)), pr.volumes())
    ddc = np.abs(dc - dco).max() / np.abs(dco).max()
    print("C2 plain: iters %d (oracle %d), |du|/|u| %.3e, dc %.3e, ddc %.3e"
          % (info["iterations"], c2["ito"], du, dcrel, ddc))
    assert abs(info["iterations"] - c2["ito"]) <= max(3, int(0.01 * c2["ito"]))
    assert du <= 1e-10
    assert dcrel <= 1e-10
    assert ddc <= 1e-8


@pytest.mark.timeout(600)
def test_c2_mixed_form_true_residual(c2):
    job, pr, b, uo = c2["job"], c2["pr"], c2["b"], c2["uo"]
    u, info = job.solve(icx.SOLVE_MIXED, tol=1e-8, maxit=50000)
    assert info["converged"] and info["refinements"] >= 1
    true = float(np.linalg.norm(b - pr.spmv(u)) / np.linalg.norm(b))
    du = _rel(u, uo)
    co = float(uo @ pr.fext)
    dcrel = abs(float(u @ pr.fext) - co) / co
    print("C2 mixed: inner %d + %d replacements, true relres %.3e (device %.3e), "
          "|du|/|u| %.3e, dc %.3e" % (info["iterations"], info["refinements"], true,
                                      info["relres"], du, dcrel))
    # the device's FP64 residual and the oracle's CSR residual agree to rounding
    assert true <= 1e-8 * (1 + 1e-6)
    assert abs(true - info["relres"]) <= 1e-6 * true
    assert du <= 1e-5 and dcrel <= 1e-7


# ------------------------------------------------------------------------------- C3
def _k0_hex8():
    p1 = O.Problem(8, (1, 1, 1), (1.0, 1.0, 1.0), O.SOLID3D)
    K0, _ = p1.element_stiffness(1.0)
    return K0


@pytest.mark.timeout(900)
def test_c3_design_loop_and_full_size_properties(ctx):
    nx, ny, nz = 192, 96, 96
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    pr = O.cantilever_3d(nx, ny, nz)
    assert job.numeq == pr.numeq == 5419584
    assert np.array_equal(job.dof_map(), pr.dof)
    job.opt_init(np.full(job.nelx, 0.3), driver=0, filter=icx.FILTER_HELMHOLTZ, radius=3.0,
                 projection=True, form=icx.SOLVE_MIXED, tol=1e-8, maxit=50000, helm_tol=1e-10)
    recs = [job.opt_step() for _ in range(20)]
    obj = np.array([r["objective"] for r in recs])
    assert np.all(np.isfinite(obj)) and obj[-1] < obj[0]
    assert all(r["cg_iterations"] > 0 for r in recs)
    x = job.opt_design()
    # Helmholtz filter at R = 3: a constant is preserved and mass is conserved
    # (sum_e v_e rho~_e = sum_e v_e rho_e, unit volumes)
    one, _ = job.helmholtz(np.ones(job.nelx), tol=1e-12)
    assert np.abs(one - 1.0).max() <= 1e-9
    rt, _ = job.helmholtz(x, tol=1e-12)
    assert abs(rt.sum() - x.sum()) <= 1e-10 * x.sum()
    # the loop's physical design at its last step (beta = 1 for steps below 40)
    rho, _ = job.project(rt, 1.0)
    job.set_design(rho)
    u_m, info_m = job.solve(icx.SOLVE_MIXED, tol=1e-8, maxit=50000)
    c_m = job.compliance()
    dc_m, _ = job.sensitivity()
    u_p, info_p = job.solve(icx.SOLVE_PLAIN, tol=1e-12, maxit=200000)
    assert info_m["converged"] and info_p["converged"]
    c_p = job.compliance()
    dc_p, _ = job.sensitivity()
    du, dcr = _rel(u_m, u_p), abs(c_m - c_p) / c_p
    ddc = np.abs(dc_m - dc_p).max() / np.abs(dc_p).max()
    print("C3: mixed %d+%d vs plain %d: |du|/|u| %.3e, dc %.3e, ddc %.3e"
          % (info_m["iterations"], info_m["refinements"], info_p["iterations"], du, dcr, ddc))
    assert du <= 1e-8 and dcr <= 1e-8 and ddc <= 1e-8
    # energy identity: u . f = sum_e E_e u_e^T K0 u_e, with u_e^T K0 u_e recovered from
    # dc_e = -p rho^(p-1) (E0 - Emin) u_e^T K0 u_e
    E0, Emin, p = 1.0, 1e-9, 3.0
    Ee = Emin + rho ** p * (E0 - Emin)
    energy = -dc_p / (p * rho ** (p - 1) * (E0 - Emin))
    assert abs(np.sum(Ee * energy) - c_p) <= 1e-9 * c_p
    # CPU spot checks with the oracle's element matrix: element energies of sampled
    # elements and rows of the FP64 operator at sampled free nodes
    K0 = _k0_hex8()
    nv = job.node_values().reshape(-1, 3)
    conn = pr.conn.reshape(-1, 8)
    rng = np.random.default_rng(3)
    els = rng.choice(job.nelx, 2000, replace=False)
    ue = nv[conn[els]].reshape(len(els), 24)
    en_cpu = np.einsum("ei,ij,ej->e", ue, K0, ue)
    assert np.abs(en_cpu - energy[els]).max() <= 1e-10 * np.abs(energy).max()
    Ku = job.matvec(u_p, prec=0)
    sx, sy = nx + 1, ny + 1
    checked = 0
    for _ in range(300):
        i, j, k = int(rng.integers(1, nx + 1)), int(rng.integers(0, ny + 1)), int(rng.integers(0, nz + 1))
        n = i + sx * (j + sy * k)
        row = np.zeros(3)
        for ei in (i - 1, i):
            for ej in (j - 1, j):
                for ek in (k - 1, k):
                    if not (0 <= ei < nx and 0 <= ej < ny and 0 <= ek < nz):
                        continue
                    e = ei + nx * (ej + ny * ek)
                    a = int(np.nonzero(conn[e] == n)[0][0])
                    row += Ee[e] * (K0[3 * a:3 * a + 3] @ nv[conn[e]].reshape(24))
        for c in range(3):
            d = pr.dof[3 * n + c]
            assert d < pr.numeq
            assert abs(Ku[d] - row[c]) <= 1e-12 * np.abs(row).max() + 1e-15
            checked += 1
    assert checked == 900
    job.close()


# ------------------------------------------------------------------------------- C4
@pytest.mark.timeout(900)
def test_c4_64_jobs_batch_equals_alone_and_true_residuals(ctx):
    rng = np.random.default_rng(np.random.PCG64(SEED_C4))
    jobs = [icx.random_job(ctx, rng) for _ in range(64)]
    alone = []
    for j in jobs:
        u, info = j.solve(icx.SOLVE_MIXED, tol=1e-8, maxit=50000)
        assert info["converged"]
        alone.append((u, info))
    infos = ctx.solve_batch(jobs, tol=1e-8)
    worst = 0.0
    for j, (u, si), inf in zip(jobs, alone, infos):
        assert np.array_equal(j.displacement(), u)
        assert inf["iterations"] == si["iterations"] and inf["refinements"] == si["refinements"]
        # the oracle's own system for this job
        hexa = j.div[2] > 1
        st = O.SOLID3D if hexa else O.PLANE_STRESS
        pr = O.Problem(8 if hexa else 4, tuple(j.div), tuple(float(v) for v in j.div), st)
        pr.add_dirichlet(j.spec["flag"], [0] * 6, j.spec["clamp"])
        assert pr.make_dof() == j.numeq
        assert np.array_equal(pr.dof, j.dof_map())
        node, comp, val = j.spec["load"]
        pr.add_nodal_load(node, comp, val)
        assert np.array_equal(pr.fext, j.fext())
        b = pr.assemble(np.full(pr.nelx, j.volfrac))
        true = float(np.linalg.norm(b - pr.spmv(u)) / np.linalg.norm(b))
        worst = max(worst, true)
        assert true <= 1e-8 * (1 + 1e-6), (j.div, true)
        del pr
    print("C4: 64 jobs, worst true relres %.3e" % worst)
    for j in jobs:
        j.close()
