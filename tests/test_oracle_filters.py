"""Oracle pins for the cone filter, projection and Helmholtz filter (reference:
src/optimization/filter.hpp:102-234, src/optimization/filter_helm.hpp:75-256)."""
import math

import numpy as np
import pytest

import oracle as O


def test_cone_neighbours_brute_force_equals_cell_search():
    pr = O.Problem(8, (7, 5, 4), (7.0, 5.0, 4.0), O.SOLID3D)
    a = pr.cone_filter(2.3, brute_force=True).neighbours()
    b = pr.cone_filter(2.3, brute_force=False).neighbours()
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_cone_weights_c1_interior():
    pr = O.cantilever_2d()
    ptr, ids, dist = pr.cone_filter(1.5).neighbours()
    e = 60 + 120 * 20  # interior element
    w = sorted(1.5 - dist[ptr[e]:ptr[e + 1]])
    expect = sorted([1.5] + [0.5] * 4 + [1.5 - math.sqrt(2)] * 4)
    assert np.allclose(w, expect, rtol=0, atol=1e-15)
    assert ptr[1] - ptr[0] == 4  # corner element: self, 2 edge neighbours, 1 diagonal


def test_cone_preserves_uniform_field_and_adjoint_identity():
    pr = O.Problem(8, (6, 5, 4), (6.0, 5.0, 4.0), O.SOLID3D)
    f = pr.cone_filter(1.8)
    out = f.apply(np.full(pr.nelx, 0.37))
    assert np.allclose(out, 0.37, rtol=1e-15, atol=0)
    rng = np.random.default_rng(0)
    x, y = rng.random(pr.nelx), rng.random(pr.nelx)
    Fx = f.apply(x)
    FTy = f.apply(y, sens=True)   # uses the normalisation cached by the forward call
    assert abs(Fx @ y - x @ FTy) < 1e-12 * abs(Fx @ y)


@pytest.mark.parametrize("beta", [1.0, 2.0, 16.0])
def test_projection_closed_form_and_derivative(beta):
    x = np.linspace(0.0, 1.0, 101)
    h = O.threshold(x, beta)
    ref = 0.5 * (np.tanh(0.5 * beta) + np.tanh(beta * (x - 0.5))) / np.tanh(0.5 * beta)
    assert np.allclose(h, ref, rtol=0, atol=1e-15)
    assert abs(h[0]) < 1e-15 and abs(h[-1] - 1.0) < 1e-15 and abs(h[50] - 0.5) < 1e-15
    d = O.threshold_derivative(x, beta)
    eps = 1e-6
    fd = (O.threshold(x + eps, beta) - O.threshold(x - eps, beta)) / (2 * eps)
    assert np.allclose(d, fd, rtol=1e-7, atol=1e-7)


@pytest.mark.parametrize("ne", [4, 8])
def test_helmholtz_invariants(ne):
    if ne == 4:
        pr = O.Problem(4, (12, 6, 1), (12.0, 6.0, 1.0), O.PLANE_STRESS)
    else:
        pr = O.Problem(8, (6, 4, 3), (6.0, 4.0, 3.0), O.SOLID3D)
    H = pr.helmholtz(2.0)
    # a constant field is preserved (natural boundary condition)
    out, it = H.apply(np.full(pr.nelx, 0.3))
    assert np.allclose(out, 0.3, rtol=1e-9, atol=0)
    rng = np.random.default_rng(1)
    x, y = rng.random(pr.nelx), rng.random(pr.nelx)
    Fx, _ = H.apply(x, tol=1e-14)
    Fy, _ = H.apply(y, tol=1e-14)
    # mass conservation sum v rho~ = sum v rho (unit volumes)
    assert abs(Fx.sum() - x.sum()) < 1e-11 * x.sum()
    # the forward map is symmetric on a uniform mesh; sensitivity mode is the same map
    assert abs(Fx @ y - x @ Fy) < 1e-11 * abs(Fx @ y)
    Sx, _ = H.apply(x, sens=True, tol=1e-14)
    assert np.allclose(Sx, Fx, rtol=1e-10, atol=1e-14)
    # CG at 1e-14 agrees with the banded direct solve
    Dx, _ = H.apply(x, direct=True)
    assert np.abs(Dx - Fx).max() < 1e-11
    # smoothing: the filtered field lies within the range of the input
    assert Fx.min() >= x.min() - 1e-12 and Fx.max() <= x.max() + 1e-12


def test_helmholtz_element_matrix_closed_form():
    # 1D-product structure: for a unit square, N N^T integrates to the Q4 mass matrix
    # (1/36)[4 2 1 2; ...] and grad-grad to (1/6)[4 -1 -2 -1; ...]
    K = np.zeros(16)
    N = np.zeros(4)
    X = np.array([0.0, 0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0])
    R = 2 * math.sqrt(3)   # r_h = 1
    O.lib().icxo_helm_element(4, 2, R, X, K, N)
    M = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36.0
    L = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6.0
    assert np.abs(K.reshape(4, 4) - (M + L)).max() < 1e-15
    assert np.allclose(N, 0.25, rtol=0, atol=1e-16)


def test_sensitivity_filter_pins():
    """Filter::sensitivityFilter (filter.hpp:72-100): with uniform densities and volumes it
    is the density filter of dfds; multiplying x by a constant changes nothing; and on a
    distorted 2D mesh it equals an O(n^2) evaluation with element centres and distances
    computed here."""
    pr = O.Problem(4, (9, 6, 1), (9.0, 6.0, 1.0), O.PLANE_STRESS)
    f = pr.cone_filter(1.7)
    rng = np.random.default_rng(4)
    df = -rng.random(pr.nelx)
    vol = pr.volumes()
    x = np.full(pr.nelx, 0.4)
    assert np.allclose(f.sensitivity(x, df, vol), f.apply(df), rtol=1e-14, atol=0)
    x = rng.uniform(0.1, 1.0, pr.nelx)
    s1 = f.sensitivity(x, df, vol)
    assert np.allclose(f.sensitivity(3.0 * x, df, vol), s1, rtol=1e-14, atol=0)
    # brute force on a distorted mesh
    from meshgen import box_mesh, distort
    x0, conn = box_mesh(4, (7, 5, 1), (7.0, 5.0, 1.0))
    X = distort(x0, conn, 4, 0.2, 2, (7.0, 5.0))
    pm = O.Problem.from_mesh(4, X, conn, O.PLANE_STRESS)
    r = 1.6
    fm = pm.cone_filter(r)
    vm = pm.volumes()
    xm = rng.uniform(0.1, 1.0, pm.nelx)
    dm = -rng.random(pm.nelx)
    Xn = np.asarray(X).reshape(-1, 3)
    cen = Xn[np.asarray(conn).reshape(-1, 4)].mean(axis=1)[:, :2]
    d = np.sqrt(((cen[:, None, :] - cen[None, :, :]) ** 2).sum(-1))
    w = np.where(d <= r, r - d, 0.0)
    aa = w.sum(axis=1)
    bb = (w * xm[None, :] * dm[None, :] / (xm * vm)[None, :]).sum(axis=1)
    ref = (xm * vm) / (xm * aa) * bb
    assert np.allclose(fm.sensitivity(xm, dm, vm), ref, rtol=1e-12, atol=0)


def _helm_element(ne, X, R):
    K = np.zeros(ne * ne)
    N = np.zeros(ne)
    O.lib().icxo_helm_element(ne, 2 if ne == 4 else 3, R, np.ascontiguousarray(X, dtype=np.float64).ravel(), K, N)
    return K.reshape(ne, ne), N


@pytest.mark.parametrize("ne", [4, 8])
def test_helmholtz_element_on_distorted_shapes(ne):
    """The Helmholtz element matrix on non-rectangular elements, by what holds for every
    shape: N sums to 1 and its gradients to 0, so the rows of Ke sum to int N; int N sums
    to the element volume; for a linear field phi = g.X the gradient part gives
    phi^T L phi = |g|^2 V exactly (bilinear / trilinear elements reproduce linear fields,
    and 2-point Gauss integrates det J exactly); on a parallelogram det J is constant, so
    the mass part is the area times the unit square's mass matrix."""
    from meshgen import box_mesh, distort
    nd = 2 if ne == 4 else 3
    div, L = ((3, 3), (3.0, 3.0)) if ne == 4 else ((3, 3, 3), (3.0, 3.0, 3.0))
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.24, 11, L).reshape(-1, 3)
    E = conn.reshape(-1, ne)
    centre = E[len(E) // 2]           # the element whose nodes all moved
    X = xd[centre, :nd]
    R = 1.7
    rh2 = (R / (2 * math.sqrt(3))) ** 2
    K, N = _helm_element(ne, X, R)
    K0, N0 = _helm_element(ne, X, 0.0)
    assert np.abs(K - K.T).max() < 1e-15 * np.abs(K).max()
    assert np.allclose(N, N0, rtol=0, atol=0)
    assert np.abs(K.sum(axis=1) - N).max() < 1e-14 * N.max()
    Ke = np.zeros((ne * nd) ** 2)
    v = np.zeros(1)
    O.lib().icxo_element_stiffness(ne, nd, O.PLANE_STRESS if ne == 4 else O.SOLID3D, 1.0, 0.3,
                                   np.ascontiguousarray(X).ravel(), Ke, v)
    assert abs(N.sum() - v[0]) < 1e-14 * v[0]
    Lg = (K - K0) / rh2
    g = np.array([0.3, -1.1, 0.7][:nd])
    phi = X @ g
    assert abs(phi @ Lg @ phi - (g @ g) * v[0]) < 1e-12 * (g @ g) * v[0]
    assert np.abs(Lg.sum(axis=1)).max() < 1e-13 * np.abs(Lg).max()
    assert np.linalg.eigvalsh(K).min() > 0
    if ne == 4:
        P = np.array([[0.0, 0.0], [2.0, 0.0], [2.6, 1.5], [0.6, 1.5]])  # area 3
        M, _ = _helm_element(4, P, 0.0)
        Mu = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36.0
        assert np.abs(M - 3.0 * Mu).max() < 1e-15 * 3


@pytest.mark.parametrize("ne", [4, 8])
def test_helmholtz_on_distorted_mesh(ne):
    """The filter on a distorted mesh: constants pass unchanged (K 1 = M 1 = the
    right-hand side of a constant, so phi is that constant at every node), the
    sensitivity mode is the same map as the forward one (f / ipmax in, no / ipmax out),
    CG agrees with the banded direct solve, and the output lies within the input's range."""
    from meshgen import box_mesh, distort
    div, L = ((12, 6), (12.0, 6.0)) if ne == 4 else ((6, 4, 3), (6.0, 4.0, 3.0))
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 7, L)
    pr = O.Problem.from_mesh(ne, xd, conn, O.PLANE_STRESS if ne == 4 else O.SOLID3D)
    H = pr.helmholtz(2.0)
    out, _ = H.apply(np.full(pr.nelx, 0.3), tol=1e-14)
    assert np.abs(out - 0.3).max() < 1e-12
    rng = np.random.default_rng(3)
    v = rng.random(pr.nelx)
    F, _ = H.apply(v, tol=1e-14)
    S, _ = H.apply(v, sens=True, tol=1e-14)
    D, _ = H.apply(v, direct=True)
    assert np.abs(S - F).max() < 1e-11
    assert np.abs(D - F).max() < 1e-11
    assert F.min() >= v.min() - 1e-12 and F.max() <= v.max() + 1e-12
