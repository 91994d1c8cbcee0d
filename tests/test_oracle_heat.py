"""Oracle pins for heat conduction (src/problem/heat.hpp, heat.cpp; the heat TO driver
opt-heat/opt_heat.cpp): element conductivity matrices against closed forms, the
patch test, and the sensitivity against finite differences of the compliance."""
import numpy as np

import oracle as O


def _element(ne, ndim, D, X):
    Ke = np.zeros(ne * ne)
    vol = np.zeros(1)
    O.lib().icxo_heat_element(ne, ndim, np.ascontiguousarray(D, dtype=np.float64).ravel(),
                              np.ascontiguousarray(X, dtype=np.float64).ravel(), Ke, vol)
    return Ke.reshape(ne, ne), vol[0]


def test_quad4_rectangle_closed_form():
    a, b, k = 2.0, 0.5, 3.0
    X = [[0, 0], [a, 0], [a, b], [0, b]]          # the reference's node order
    K, v = _element(4, 2, k * np.eye(2), X)
    Kx = np.array([[2, -2, -1, 1], [-2, 2, 1, -1], [-1, 1, 2, -2], [1, -1, -2, 2]]) / 6.0
    Ky = np.array([[2, 1, -1, -2], [1, 2, -2, -1], [-1, -2, 2, 1], [-2, -1, 1, 2]]) / 6.0
    assert np.allclose(K, k * (b / a * Kx + a / b * Ky), rtol=0, atol=1e-14)
    assert abs(v - a * b) < 1e-15
    # anisotropic diagonal tensor: linear in kx, ky
    K2, _ = _element(4, 2, np.diag([5.0, 0.25]), X)
    assert np.allclose(K2, 5.0 * b / a * Kx + 0.25 * a / b * Ky, rtol=0, atol=1e-14)


def test_hexa8_unit_cube_closed_form():
    X = [[i & 1, (i >> 1) & 1, i >> 2] for i in range(8)]
    X = [X[q] for q in (0, 1, 3, 2, 4, 5, 7, 6)]   # z-low face then z-high face, each CCW
    K, v = _element(8, 3, 2.0 * np.eye(3), X)
    P = np.array(X, dtype=float)
    for i in range(8):
        for j in range(8):
            d = int(np.abs(P[i] - P[j]).sum())      # 0 self, 1 edge, 2 face diagonal, 3 body diagonal
            ref = {0: 1 / 3, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[d] * 2.0
            assert abs(K[i, j] - ref) < 1e-14
    assert abs(v - 1.0) < 1e-15
    assert np.abs(K.sum(axis=1)).max() < 1e-14          # constants in the null space


def _heat_problem(nx=6, ny=4, nz=None, k1=(10.0, 10.0, 10.0, 0, 0, 0), k2=(0.1, 0.1, 0.1, 0, 0, 0)):
    if nz is None:
        pr = O.Problem(4, (nx, ny, 1), (float(nx), float(ny), 1.0), O.HEAT, penalty=3.0)
        pr.add_dirichlet([1, 0, 0, 0, 0, 0], [0] * 6, [0, 0, 0, ny, 0, 0])
    else:
        pr = O.Problem(8, (nx, ny, nz), (float(nx), float(ny), float(nz)), O.HEAT, penalty=3.0)
        pr.add_dirichlet([1, 0, 0, 0, 0, 0], [0] * 6, [0, 0, 0, ny, 0, nz])
    pr.make_dof()
    pr.set_conductivity(k1, k2)
    return pr


def test_patch_test_linear_temperature():
    """A linear temperature field on a uniform design has zero residual at interior
    nodes (the element matrices reproduce constant gradients)."""
    for nz in (None, 3):
        pr = O.Problem(4 if nz is None else 8, (5, 4, nz or 1), (5.0, 4.0, float(nz or 1)), O.HEAT)
        pr.make_dof()                                   # no constraints: all DOFs free
        pr.set_conductivity((2.0, 3.0, 4.0, 0, 0, 0), (0.1, 0.1, 0.1, 0, 0, 0))
        pr.assemble(np.full(pr.nelx, 0.7))
        x = pr.x.reshape(-1, 3)
        T = 1.0 + 0.3 * x[:, 0] - 0.2 * x[:, 1] + 0.1 * x[:, 2]
        r = pr.spmv(T[pr.dof])                          # dof order = node order here
        d = pr.div
        i, j, k = x[:, 0], x[:, 1], x[:, 2]
        inner = (i > 0) & (i < d[0]) & (j > 0) & (j < d[1])
        if nz is not None:
            inner &= (k > 0) & (k < d[2])
        assert np.abs(r[pr.dof[inner]]).max() < 1e-12


def test_heat_sensitivity_finite_differences():
    for nz in (None, 2):
        pr = _heat_problem(nz=nz)
        pr.add_neumann([1.0, 0, 0], [pr.div[0], pr.div[0], 0, pr.div[1], 0, pr.div[2] if nz else 0])
        rho = O.random_density(pr.nelx, 5, lo=0.2, hi=0.9)

        def comp(r):
            b = pr.assemble(r)
            T = pr.direct(b)
            return T @ b, T

        c0, T = comp(rho)
        nv = pr.node_values(T)
        pr.assemble(rho)
        dc = pr.heat_sensitivity(nv)
        h = 1e-6
        for e in (0, pr.nelx // 2, pr.nelx - 1):
            rp, rm = rho.copy(), rho.copy()
            rp[e] += h
            rm[e] -= h
            fd = (comp(rp)[0] - comp(rm)[0]) / (2 * h)
            assert abs(fd - dc[e]) <= 1e-6 * np.abs(dc).max(), (e, fd, dc[e])


def test_heat_problem_volumes():
    """element volumes of a heat problem (one DOF per node) come from the heat element"""
    for ne, div, L in ((4, (6, 4, 1), (3.0, 1.0, 1.0)), (8, (3, 2, 4), (3.0, 1.0, 2.0))):
        pr = O.Problem(ne, div, L, O.HEAT)
        pr.make_dof()
        h = np.array(L) / np.array(div)
        ref = h[0] * h[1] * (h[2] if ne == 8 else 1.0)
        assert np.allclose(pr.volumes(), ref, rtol=1e-14, atol=0)
