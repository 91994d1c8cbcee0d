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
