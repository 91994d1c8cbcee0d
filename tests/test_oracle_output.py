"""Oracle pins for the output step: the element-mean von Mises stress
(src/problem/elasticity.cpp:165-208) against closed-form stress states, and the
reactions (src/analysis/linear.hpp:90-91, src/analysis/analysis_utility.hpp:33-51)
against equilibrium."""
import numpy as np

import oracle as O


def _field3d(pr, fn):
    """node values 3 per node from fn(x, y, z) -> (ux, uy, uz)"""
    x = pr.x.reshape(-1, 3)
    return np.stack(fn(x[:, 0], x[:, 1], x[:, 2]), axis=1).ravel()


def _unit_design(pr):
    pr.assemble(np.ones(pr.nelx))  # E = young2 = 1 in every element


def test_mises_uniaxial_strain_and_shear_3d():
    pr = O.cantilever_3d(4, 3, 2, nu=0.3)
    _unit_design(pr)
    E, nu = 1.0, 0.3
    mu = E / (2 * (1 + nu))
    eps = 1e-3
    m = pr.mises(_field3d(pr, lambda x, y, z: (eps * x, 0 * y, 0 * z)))
    assert np.allclose(m, 2 * mu * eps, rtol=1e-13, atol=0)       # |s_xx - s_yy|
    g = 2e-3
    m = pr.mises(_field3d(pr, lambda x, y, z: (g * y, 0 * y, 0 * z)))
    assert np.allclose(m, np.sqrt(3.0) * mu * g, rtol=1e-13, atol=0)
    # rigid rotation about z and a translation: no strain
    m = pr.mises(_field3d(pr, lambda x, y, z: (-0.01 * y + 1.0, 0.01 * x, 0 * z + 2.0)))
    assert np.abs(m).max() < 1e-15


def test_mises_plane_stress_uniaxial_and_simp_scaling():
    pr = O.cantilever_2d(6, 3, nu=0.3)
    eps, nu = 1e-3, 0.3
    xy = pr.x.reshape(-1, 3)
    u = np.stack([eps * xy[:, 0], -nu * eps * xy[:, 1], 0 * xy[:, 0]], axis=1).ravel()
    _unit_design(pr)
    assert np.allclose(pr.mises(u), eps, rtol=1e-12, atol=0)      # s_xx = E eps, s_yy = 0
    rho = O.random_density(pr.nelx, 3)
    pr.assemble(rho)
    E = np.array([O.lib().icxo_simp(1e-9, 1.0, r, 3.0) for r in rho])
    assert np.allclose(pr.mises(u), E * eps, rtol=1e-12, atol=0)


def test_reactions_balance_the_load():
    """After a solve, the reactions equal the external load at free DOFs and balance
    it in total: sum over all nodes of K u is zero per component (rigid translations)."""
    pr = O.cantilever_3d(8, 4, 4)
    rho = O.random_density(pr.nelx, 9)
    b = pr.assemble(rho)
    u = pr.direct(b)
    nv = pr.node_values(u)
    R = pr.reactions(nv).reshape(-1, 3)
    dof = pr.dof.reshape(-1, pr.dofnp)
    free = dof < pr.numeq
    f_nodes = np.zeros_like(R)
    f_nodes[free] = b[dof[free]]
    scale = np.abs(b).max()
    assert np.abs(R[free] - f_nodes[free]).max() < 1e-9 * scale
    total = R.sum(axis=0)
    assert np.abs(total).max() < 1e-9 * scale
    assert np.abs(R[~free]).sum() > 0.5 * np.abs(b).sum()         # the clamp carries the load
    # the flush: nothing nonzero below 1e-15
    nz = R[R != 0]
    assert np.abs(nz).min() >= 1e-15
