"""GPU parity of heat conduction (src/problem/heat.hpp, heat.cpp; opt-heat/opt_heat.cpp)
against the oracle on the same seeded inputs: the FP64 operator to rounding, the solve
to 1e-10 of the direct solution, the sensitivity and the VTK fields."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu

QUAD4, HEXA8 = 4, 8
K1 = (100.0, 80.0, 60.0, 5.0, 3.0, 2.0)      # anisotropic material 1
K2 = (0.1, 0.2, 0.3, 0.01, 0.02, 0.03)


def _pair(ctx, ne, div, seed=7):
    L = tuple(float(d) for d in div)
    job = icx.Job(ctx, ne, div, L, icx.HEAT)
    pr = O.Problem(ne, div, L, O.HEAT)
    if ne == QUAD4:
        bc = ([1, 0, 0, 0, 0, 0], [0] * 6, [0.4 * div[0], 0.6 * div[0], div[1], div[1], 0, 0])
        nm = ([0.1, 0, 0], [0, div[0], 0, div[1], 0, 0])
    else:
        bc = ([1, 0, 0, 0, 0, 0], [0] * 6, [0, div[0], 0, div[1], 0, 0])
        nm = ([0.1, 0, 0], [0, div[0], 0, div[1], div[2], div[2]])
    job.add_dirichlet(*bc)
    pr.add_dirichlet(*bc)
    assert job.make_dof() == pr.make_dof()
    job.add_neumann(*nm)
    pr.add_neumann(*nm)
    assert np.abs(job.fext() - pr.fext).max() <= 1e-15 * np.abs(pr.fext).max()
    job.set_conductivity(K1, K2)
    pr.set_conductivity(K1, K2)
    rho = O.random_density(pr.nelx, seed, lo=0.05, hi=1.0)
    job.set_design(rho)
    b = pr.assemble(rho)
    return job, pr, rho, b


@pytest.mark.parametrize("ne,div", [(QUAD4, (12, 8)), (HEXA8, (6, 4, 5))])
def test_heat_operator_solve_and_sensitivity(ctx, ne, div):
    job, pr, rho, b = _pair(ctx, ne, div)
    x = np.random.default_rng(3).standard_normal(pr.numeq)
    y = job.matvec(x)
    yo = pr.spmv(x)
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    T, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    Td = pr.direct(b)
    assert info["converged"]
    assert np.linalg.norm(T - Td) / np.linalg.norm(Td) < 1e-10
    c = job.compliance()
    assert abs(c - Td @ pr.fext) / abs(Td @ pr.fext) < 1e-10
    nv = job.node_values()
    dc, dv = job.sensitivity()
    dco = pr.heat_sensitivity(nv)
    assert np.abs(dc - dco).max() <= 1e-12 * np.abs(dco).max()
    assert np.abs(dv - pr.volumes()).max() <= 1e-15
    # heat flux over all DOFs balances: K T summed over every node is zero
    R = job.reactions().reshape(-1, 3)
    assert np.abs(R[:, 1:]).max() == 0.0
    assert abs(R[:, 0].sum()) <= 1e-10 * np.abs(R[:, 0]).max()


def test_heat_vtk_fields_and_restrictions(ctx, tmp_path):
    job, pr, rho, b = _pair(ctx, QUAD4, (10, 6))
    job.solve(icx.SOLVE_PLAIN, tol=1e-12, maxit=20000)
    path = str(tmp_path / "res1.vtk")
    job.write_vtk(path)
    data = open(path, "rb").read()
    def block(tag, n):
        i = data.index(tag) + len(tag)
        i = data.index(b"\n", i) + 1
        if data[i:i + 20].startswith(b"LOOKUP_TABLE"):
            i = data.index(b"\n", i) + 1
        return np.frombuffer(data, dtype=">f4", count=n, offset=i)
    nn, ne = pr.numnp, pr.nelx
    T = job.node_values().reshape(-1, 3)[:, 0]
    assert np.array_equal(block(b"SCALARS temperature float", nn), T.astype(np.float32))
    conn = pr.conn.reshape(-1, 4)
    nds = np.zeros(nn)
    for e in range(ne):
        for k in range(4):
            nds[conn[e, k]] += rho[e] / 4
    assert np.array_equal(block(b"SCALARS threshold_s float", nn), (nds > 0.5).astype(np.float32))
    assert np.array_equal(block(b"SCALARS design_s float", ne), rho.astype(np.float32))
    for form in (icx.SOLVE_FUSED, icx.SOLVE_MIXED):
        with pytest.raises(icx.IcxError):
            job.solve(form, tol=1e-8)
    with pytest.raises(icx.IcxError):
        job.mises()


def _oracle_heat_driver(steps):
    """opt-heat/opt_heat.cpp:71-150 composed from oracle pieces: Helmholtz filter (CG,
    1e-10), projection, direct heat solve, T^T K T, sensitivity, MMA."""
    k1, k2 = (50.0, 40.0, 0.0, 2.0, 0.0, 0.0), (0.05, 0.05, 0.0, 0.0, 0.0, 0.0)
    pr = O.Problem(QUAD4, (24, 16, 1), (1.5, 1.0, 1.0), O.HEAT, penalty=3.0)
    pr.add_dirichlet([1, 0, 0, 0, 0, 0], [0] * 6, [0.6, 0.9, 1.0, 1.0, 0, 0])
    pr.make_dof()
    pr.add_neumann([0.05, 0, 0], [0.0, 1.5, 0.0, 1.0, 0, 0])
    pr.set_conductivity(k1, k2)
    helm = O.Helmholtz(pr, 0.12)
    mma = O.MMA(pr.nelx, 1)
    vol = pr.volumes()
    x = np.full(pr.nelx, 0.4)
    xmin, xmax = np.full(pr.nelx, 1e-4), np.full(pr.nelx, 1.0)
    V0, obj, con = 0.0, [], []
    for step in range(1, steps + 1):
        fil, _ = helm.apply(x, sens=False, tol=1e-10)
        beta = min(2.0 ** (step // 40), 16.0)
        thr = O.threshold(fil, beta, 0.5)
        b = pr.assemble(thr)
        T = pr.direct(b)
        if step == 1:
            V0 = thr @ vol
        obj.append(T @ (pr.csr() @ T))
        con.append(thr @ vol - V0)
        thd = O.threshold_derivative(fil, beta, 0.5)
        dfds = pr.heat_sensitivity(pr.node_values(T)) * thd
        dgds = vol * thd
        dfds, _ = helm.apply(dfds, sens=True, tol=1e-10)
        dgds, _ = helm.apply(dgds, sens=True, tol=1e-10)
        x, _ = mma.update(x, dfds, [con[-1]], dgds, xmin, xmax)
    return np.array(obj), np.array(con)


def test_cxx_heat_driver_matches_oracle():
    import os
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "icarat-contact_amd", "_build", "opt_heat")
    toml = os.path.join(root, "icarat-contact_amd", "host", "examples", "heat_q4.toml")
    steps = 6
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([exe, toml, str(steps)], cwd=d, capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        rows = np.loadtxt(os.path.join(d, "function.csv"), delimiter=",", ndmin=2)
        head = open(os.path.join(d, "res%d.vtk" % steps), "rb").read(200)
        assert b"BINARY" in head and b"POINTS 425 float" in head
    obj, con = _oracle_heat_driver(steps)
    assert rows.shape[0] == steps
    assert np.abs(rows[:, 1] - obj).max() <= 1e-7 * np.abs(obj).max(), (rows[:, 1], obj)
    assert np.abs(rows[:, 2] - con).max() <= 1e-7 * abs(con[0] + 1.0) + 1e-9


@pytest.mark.parametrize("ne,div", [(QUAD4, (12, 8)), (HEXA8, (6, 4, 5)), (HEXA8, (9, 7, 19)),
                                    (HEXA8, (33, 17, 9))])
def test_heat_haar_operator_for_diagonal_tensors(ctx, ne, div):
    """Diagonal conductivity tensors (as in opt-heat's examples) make both element
    matrices diagonal in the Haar basis: that operator against the oracle, against the
    dense-row operator (ICX_OPT_HEAT_DENSE) on the same job, and through a solve; the
    grids cover one and several 15 x 15 node tiles and z chunks that do not divide."""
    L = tuple(float(d) for d in div)
    k1, k2 = (100.0, 70.0, 40.0, 0, 0, 0), (0.1, 0.3, 0.2, 0, 0, 0)
    job = icx.Job(ctx, ne, div, L, icx.HEAT)
    pr = O.Problem(ne, div, L, O.HEAT)
    bc = ([1, 0, 0, 0, 0, 0], [0] * 6, [0, 0, 0, div[1], 0, div[2] if ne == HEXA8 else 0])
    nm = ([0.1, 0, 0], [0, div[0], 0, div[1], 0, div[2] if ne == HEXA8 else 0])
    for o in (job, pr):
        o.add_dirichlet(*bc)
        o.make_dof()
        o.add_neumann(*nm)
        o.set_conductivity(k1, k2)
    rho = O.random_density(pr.nelx, 5, lo=0.05, hi=1.0)
    job.set_design(rho)
    b = pr.assemble(rho)
    x = np.random.default_rng(4).standard_normal(pr.numeq)
    y = job.matvec(x)
    yo = pr.spmv(x)
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    ctx.set_option(icx.OPT_HEAT_DENSE, 1)
    try:
        yd = job.matvec(x)
        Tdense, _ = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    finally:
        ctx.set_option(icx.OPT_HEAT_DENSE, 0)
    assert np.abs(y - yd).max() <= 1e-13 * np.abs(yd).max()
    T, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    Td = pr.direct(b)
    assert info["converged"]
    assert np.linalg.norm(T - Td) / np.linalg.norm(Td) < 1e-10
    assert np.linalg.norm(T - Tdense) / np.linalg.norm(Td) < 1e-11
