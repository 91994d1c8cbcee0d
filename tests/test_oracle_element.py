"""Oracle pins for the element matrices (reference: src/problem/elasticity.cpp:34-98,
210-254; src/basis/bmatrix.hpp:28-162; src/optimization/filter_helm.hpp:178-222)."""
import itertools

import numpy as np
import pytest

import oracle as O


def top88_KE(nu):
    """Closed-form Q4 plane-stress stiffness (Andreassen et al. 2011, top88)."""
    A11 = np.array([[12, 3, -6, -3], [3, 12, 3, 0], [-6, 3, 12, -3], [-3, 0, -3, 12]])
    A12 = np.array([[-6, -3, 0, 3], [-3, -6, -3, -6], [0, -3, -6, 3], [3, -6, 3, -6]])
    B11 = np.array([[-4, 3, -2, 9], [3, -4, -9, 4], [-2, -9, -4, -3], [9, 4, -3, -4]])
    B12 = np.array([[2, -3, 4, -9], [-3, 2, 9, -2], [4, 9, 2, 3], [-9, -2, 3, 2]])
    return 1 / (1 - nu ** 2) / 24 * (np.block([[A11, A12], [A12.T, A11]]) +
                                     nu * np.block([[B11, B12], [B12.T, B11]]))


@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_quad4_plane_stress_matches_top88(nu):
    pr = O.Problem(4, (1, 1, 1), (1.0, 1.0, 1.0), O.PLANE_STRESS, nu=nu)
    K, v = pr.element_stiffness(1.0)
    assert np.abs(K - top88_KE(nu)).max() < 1e-15
    assert abs(v - 1.0) < 1e-15


def test_hexa8_top3d_coefficients():
    # Liu & Tovar (2014) top3d: k1 = (32 - 48 nu)/144 on the diagonal, k2 = 6/144 for
    # the x-y coupling of one node, both times 1/((1+nu)(1-2nu)) for E = 1.
    nu = 0.3
    pr = O.Problem(8, (1, 1, 1), (1.0, 1.0, 1.0), O.SOLID3D, nu=nu)
    K, v = pr.element_stiffness(1.0)
    f = 1 / ((1 + nu) * (1 - 2 * nu))
    assert np.allclose(np.diag(K), (32 - 48 * nu) / 144 * f, rtol=1e-15, atol=0)
    ref = [(0, 0), (1, 0), (1, 1), (0, 1)] * 2   # (a, b) of the reference's local nodes
    for q, (a, b) in enumerate(ref):
        sign = 1.0 if a == b else -1.0           # d2N/dxdy changes sign with the corner
        assert abs(K[3 * q, 3 * q + 1] - sign * 6 / 144 * f) < 1e-15
    assert abs(v - 1.0) < 1e-15


def independent_K(ndim, h, state, nu, E=1.0, order=3):
    """Element stiffness by a 3-point Gauss rule, written independently of the oracle."""
    pts, wts = np.polynomial.legendre.leggauss(order)
    nq = 2 ** ndim
    corners = [[(q >> d) & 1 for d in range(ndim)] for q in range(nq)]
    if ndim == 2:
        if state == O.PLANE_STRESS:
            D = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
        else:
            c = E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
            D = c * np.array([[1, nu / (1 - nu), 0], [nu / (1 - nu), 1, 0],
                              [0, 0, (1 - 2 * nu) / (2 * (1 - nu))]])
    else:
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
        mu = E / (2 * (1 + nu))
        D = np.zeros((6, 6))
        D[:3, :3] = lam
        D[np.arange(3), np.arange(3)] = lam + 2 * mu
        D[3, 3] = D[4, 4] = D[5, 5] = mu
    nd = nq * ndim
    K = np.zeros((nd, nd))
    for idx in itertools.product(range(order), repeat=ndim):
        r = [pts[i] for i in idx]
        w = np.prod([wts[i] for i in idx])
        dN = np.zeros((ndim, nq))
        for q, cq in enumerate(corners):
            s = [1 if c else -1 for c in cq]
            for d in range(ndim):
                v = s[d] / 2.0 ** ndim
                for e in range(ndim):
                    if e != d:
                        v *= 1 + s[e] * r[e]
                dN[d, q] = v * 2.0 / h[d]
        B = np.zeros((3 if ndim == 2 else 6, nd))
        for q in range(nq):
            if ndim == 2:
                B[0, 2 * q], B[1, 2 * q + 1] = dN[0, q], dN[1, q]
                B[2, 2 * q], B[2, 2 * q + 1] = dN[1, q], dN[0, q]
            else:
                B[0, 3 * q], B[1, 3 * q + 1], B[2, 3 * q + 2] = dN[0, q], dN[1, q], dN[2, q]
                B[3, 3 * q], B[3, 3 * q + 1] = dN[1, q], dN[0, q]
                B[4, 3 * q + 1], B[4, 3 * q + 2] = dN[2, q], dN[1, q]
                B[5, 3 * q], B[5, 3 * q + 2] = dN[2, q], dN[0, q]
        K += B.T @ D @ B * w * np.prod(h) / 2.0 ** ndim
    # corner order (a,b,c) -> reference local order
    ref = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    if ndim == 2:
        ref = [(0, 0), (1, 0), (1, 1), (0, 1)]
    perm = [corners.index(list(c)) for c in ref]
    idx = [ndim * p + c for p in perm for c in range(ndim)]
    return K[np.ix_(idx, idx)]


@pytest.mark.parametrize("ne,state,h", [(4, O.PLANE_STRESS, (1.0, 1.0)), (4, O.PLANE_STRAIN, (2.0, 0.5)),
                                        (8, O.SOLID3D, (1.0, 1.0, 1.0)), (8, O.SOLID3D, (2.0, 1.5, 0.5))])
def test_two_point_rule_is_exact(ne, state, h):
    nd = 2 if ne == 4 else 3
    div = (2,) * nd
    L = tuple(2 * v for v in h)
    pr = O.Problem(ne, div + (1,) * (3 - nd), L + (1.0,) * (3 - nd), state, nu=0.3)
    K, v = pr.element_stiffness(2.5)
    Kr = independent_K(nd, h, state, 0.3, E=2.5)
    assert np.abs(K - Kr).max() <= 1e-14 * np.abs(Kr).max()
    assert abs(v - np.prod(h)) < 1e-14


@pytest.mark.parametrize("ne", [4, 8])
def test_rigid_body_null_space_and_rank(ne):
    nd = 2 if ne == 4 else 3
    pr = O.Problem(ne, (1, 1, 1), (1.0, 1.3, 0.7), O.PLANE_STRAIN if nd == 2 else O.SOLID3D)
    K, _ = pr.element_stiffness(1.0)
    X = pr.x.reshape(-1, 3)[pr.conn[:ne]][:, :nd]
    modes = []
    for d in range(nd):
        m = np.zeros((ne, nd)); m[:, d] = 1; modes.append(m.ravel())
    pairs = [(0, 1)] if nd == 2 else [(0, 1), (1, 2), (0, 2)]
    for a, b in pairs:
        m = np.zeros((ne, nd)); m[:, a] = -X[:, b]; m[:, b] = X[:, a]; modes.append(m.ravel())
    for m in modes:
        assert np.abs(K @ m).max() < 1e-14
    ev = np.linalg.eigvalsh(K)
    nrigid = 3 if nd == 2 else 6
    assert np.all(ev[:nrigid] < 1e-13) and ev[nrigid] > 1e-3
    assert np.abs(K - K.T).max() < 1e-15


def test_simp_formula():
    L = O.lib()
    assert L.icxo_simp(1e-9, 1.0, 0.5, 3.0) == (1 - 0.125) * 1e-9 + 0.125
    # reversed materials (young2 < young1): (1-s)^p E1 + (1 - (1-s)^p) E2
    assert abs(L.icxo_simp(2.0, 1.0, 0.25, 2.0) - (0.5625 * 2.0 + 0.4375 * 1.0)) < 1e-15
