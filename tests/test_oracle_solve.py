"""Oracle pins for assembly, the Eigen-CG restatement, the direct solve, compliance and
the sensitivity (reference: src/analysis/linear.hpp:54-166,
src/analysis/solver/solvers.cpp:64-92, opt-macro-oc/opt_macro_oc.cpp:110, 182-212)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O


def _small3d(seed=1):
    pr = O.cantilever_3d(10, 4, 4)
    x = O.random_density(pr.nelx, seed)
    return pr, x


def test_cg_and_direct_match_superlu_2d():
    pr = O.cantilever_2d()
    b = pr.assemble(np.full(pr.nelx, 0.5))
    K = pr.csr()
    assert abs(K - K.T).max() < 1e-15
    us = spla.spsolve(K.tocsc(), b)
    ud = pr.direct(b)
    u, it, err = pr.cg(b, tol=1e-12)
    assert np.linalg.norm(ud - us) / np.linalg.norm(us) < 1e-10
    assert np.linalg.norm(u - us) / np.linalg.norm(us) < 1e-9
    assert err < 1e-12 and it > 0
    # energy identity u^T K u = f^T u
    assert abs(us @ (K @ us) - b @ us) / abs(b @ us) < 1e-12


def test_eigen_cg_stopping_rule():
    """Eigen stops on |r|^2 < tol^2 |b|^2 of the recursive residual and reports the
    number of passes that did not converge."""
    pr, x = _small3d()
    b = pr.assemble(x)
    u1, it1, e1 = pr.cg(b, tol=1e-6)
    u2, it2, e2 = pr.cg(b, tol=1e-6, maxit=it1 + 1)
    assert e1 < 1e-6 and it2 == it1 and np.array_equal(u1, u2)
    _, it3, e3 = pr.cg(b, tol=1e-6, maxit=it1)   # one pass short: not converged
    assert it3 == it1 and e3 >= 1e-6
    # zero right-hand side returns x = 0 with no iteration
    z, itz, ez = pr.cg(np.zeros_like(b))
    assert itz == 0 and ez == 0 and not z.any()


def test_random_density_3d_solution():
    pr, x = _small3d()
    b = pr.assemble(x)
    K = pr.csr()
    us = spla.spsolve(K.tocsc(), b)
    ud = pr.direct(b)
    assert np.linalg.norm(ud - us) / np.linalg.norm(us) < 1e-10
    # residual of the direct solution
    assert np.linalg.norm(b - K @ ud) / np.linalg.norm(b) < 1e-12


def test_csr_equals_dense_sum_of_element_matrices():
    pr = O.Problem(8, (2, 2, 1), (2.0, 2.0, 1.0), O.SOLID3D, young=(1e-3, 1.0))
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, 2, 0, 1])
    pr.make_dof()
    x = np.array([0.2, 0.9, 0.5, 0.7])
    pr.assemble(x)
    K = pr.csr().toarray()
    Kd = np.zeros((pr.numeq, pr.numeq))
    for e in range(pr.nelx):
        E = O.lib().icxo_simp(1e-3, 1.0, x[e], 3.0)
        Ke, _ = pr.element_stiffness(E, e)
        idx = [pr.dof[3 * n + c] for n in pr.conn[8 * e:8 * e + 8] for c in range(3)]
        for a, ia in enumerate(idx):
            for bb, ib in enumerate(idx):
                if ia < pr.numeq and ib < pr.numeq:
                    Kd[ia, ib] += Ke[a, bb]
    assert np.abs(K - Kd).max() < 1e-15


def test_prescribed_displacement_rhs():
    """fres = fext - fint with fint = K_full u_D on the free rows (linear.hpp:62-80)."""
    pr = O.Problem(4, (4, 2, 1), (4.0, 2.0, 1.0), O.PLANE_STRESS, young=(1e-3, 1.0))
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 2, 0, 0])
    pr.add_dirichlet([1, 0, 0, 0, 0, 0], [0.01, 0, 0, 0, 0, 0], [4, 4, 0, 2, 0, 0])
    pr.make_dof()
    x = np.full(pr.nelx, 0.6)
    b = pr.assemble(x)
    u = pr.direct(b)
    # the same problem solved on all DOFs with the constraint rows replaced
    full = np.zeros(pr.numnp * 2)
    nv = pr.node_values(u)
    for n in range(pr.numnp):
        full[2 * n:2 * n + 2] = nv[3 * n:3 * n + 2]
    right = [pr.node_index(4, j) for j in range(3)]
    assert all(abs(nv[3 * n] - 0.01) < 1e-15 for n in right)
    # equilibrium at free DOFs of the full assembled operator
    E = O.lib().icxo_simp(1e-3, 1.0, 0.6, 3.0)
    Kfull = np.zeros((2 * pr.numnp, 2 * pr.numnp))
    for e in range(pr.nelx):
        Ke, _ = pr.element_stiffness(E, e)
        idx = [2 * n + c for n in pr.conn[4 * e:4 * e + 4] for c in range(2)]
        Kfull[np.ix_(idx, idx)] += Ke
    r = Kfull @ full
    for n in range(pr.numnp):
        for c in range(2):
            if pr.dof[2 * n + c] < pr.numeq:
                assert abs(r[2 * n + c]) < 1e-12


def test_compliance_energy_identity_and_flush():
    pr, x = _small3d(3)
    b = pr.assemble(x)
    u = pr.direct(b)
    nv = pr.node_values(u)
    energy = pr.element_energy(nv)
    c = u @ pr.fext
    assert abs(energy.sum() - c) / c < 1e-12
    tiny = u.copy()
    tiny[0] = 1e-16
    assert pr.node_values(tiny)[3 * 1] == 0.0 or pr.dof[3] != 0


@pytest.mark.parametrize("ne", [4, 8])
def test_sensitivity_central_differences(ne):
    """dc/drho_hat by central differences, as opt-mma-HelmholtzFilter/fdm.cpp:17-47
    checks the analytic sensitivity (there with forward differences)."""
    if ne == 4:
        pr = O.cantilever_2d(8, 4)
    else:
        pr = O.cantilever_3d(4, 2, 2)
    x = O.random_density(pr.nelx, 7, 0.2, 0.9)
    b = pr.assemble(x)
    u = pr.direct(b)
    vol = pr.volumes()
    dc, dv = pr.sensitivity(x, pr.element_energy(pr.node_values(u)), vol)
    assert np.allclose(dv, vol)
    h = 1e-6
    for e in range(0, pr.nelx, max(1, pr.nelx // 7)):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        cp = pr.direct(pr.assemble(xp)) @ pr.fext
        cm = pr.direct(pr.assemble(xm)) @ pr.fext
        fd = (cp - cm) / (2 * h)
        assert abs(fd - dc[e]) <= 1e-6 * max(abs(dc).max(), 1e-30)
