"""Oracle pins for GCMMA (src/optimization/opt-solver/GCMMASolver.cpp:41-620), from what
the method fixes rather than from its formulas:

- the approximations equal f and grad f at xval (GenSub's r0 / r and p / q split);
- the dual solve returns the minimum of the convex subproblem (checked with SLSQP);
- RaaUpdate raises the approximation at the rejected point above the true value there
  (the change of the approximation at xmma is exactly delta_raa * raacof);
- the outer / inner loop of opt_macro_gcmma.cpp:154-215 reaches the closed-form optimum
  of a separable problem and ends with conservative approximations."""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle as O


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 0.8, n)
    df = rng.standard_normal(n)
    dg = rng.uniform(0.2, 1.0, n) * rng.choice([-1.0, 1.0], n)
    return x, df, dg, 1.7, 0.05 * n


def _fd_grad(fn, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(len(x)):
        e = np.zeros_like(x)
        e[i] = h
        g[i] = (fn(x + e) - fn(x - e)) / (2 * h)
    return g


@pytest.mark.parametrize("seed", [0, 1])
def test_approximations_are_first_order_exact_at_xval(seed):
    n = 12
    x, df, dg, f0, g0 = _data(n, seed)
    lo, hi = np.full(n, 1e-4), np.ones(n)
    M = O.GCMMA(n)
    xmma, it = M.outer(x, f0, df, g0, dg, lo, hi)
    assert it > 0
    st = M.state()
    assert st["raa0"] == pytest.approx(0.1) and st["raa"][0] == pytest.approx(0.1)   # 1 -> 0.1
    for _ in range(2):  # after the outer update and after an inner update (raa changed)
        a0, a1 = M.approx_at(x)
        assert a0 == pytest.approx(f0, rel=1e-13) and a1[0] == pytest.approx(g0, rel=1e-12)
        np.testing.assert_allclose(_fd_grad(lambda z: M.approx_at(z)[0], x), df, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(_fd_grad(lambda z: M.approx_at(z)[1][0], x), dg, rtol=1e-6, atol=1e-7)
        # ComputeApprox = the approximations at the returned point
        st = M.state()
        b0, b1 = M.approx_at(xmma)
        assert st["f0app"] == pytest.approx(b0, rel=1e-14) and st["fapp"][0] == pytest.approx(b1[0], rel=1e-14)
        assert np.all(np.abs(xmma - x) <= 0.15 + 1e-15)
        xmma, _ = M.inner(xmma, st["f0app"] + 1.0, st["fapp"] + 0.5, x, f0, df, g0, dg, lo, hi)


@pytest.mark.parametrize("seed", [2, 3])
def test_dual_solution_is_the_subproblem_minimum(seed):
    """At the first outer update the box of the subproblem is the move limit (asymptotes
    at xval -/+ 0.2 give alpha = L + 0.2 (x - L) = x - 0.16 < x - 0.15), so SLSQP on the
    separable convex subproblem over that box is an independent solution."""
    n = 10
    x, df, dg, f0, _ = _data(n, seed)
    g0 = -0.1          # a subproblem with feasible points inside the move limits (y = 0)
    lo, hi = np.full(n, 1e-4), np.ones(n)
    M = O.GCMMA(n)
    xmma, _ = M.outer(x, f0, df, g0, dg, lo, hi)
    box = list(zip(np.maximum(lo, x - 0.15), np.minimum(hi, x + 0.15)))
    res = minimize(lambda z: M.approx_at(z)[0], x, method="SLSQP", bounds=box,
                   constraints=[{"type": "ineq", "fun": lambda z: -M.approx_at(z)[1][0]}],
                   options={"ftol": 1e-14, "maxiter": 500})
    assert res.success
    a0, a1 = M.approx_at(xmma)
    assert a1[0] <= 1e-7
    assert a0 <= res.fun + 1e-7 * max(1.0, abs(res.fun))
    np.testing.assert_allclose(xmma, res.x, atol=2e-4)


def test_raa_update_makes_the_rejected_point_conservative():
    """Raising raa by d changes the approximation at xmma by d * raacof; RaaUpdate takes
    d >= 1.1 (fnew - fapp) / raacof (when not capped at 10 raa), so the new approximation
    at the rejected xmma is above the true value there."""
    n = 16
    x, df, dg, f0, g0 = _data(n, 5)
    lo, hi = np.full(n, 1e-4), np.ones(n)
    M = O.GCMMA(n)
    xmma, _ = M.outer(x, f0, df, g0, dg, lo, hi)
    st = M.state()
    r0_before, raa_before = st["raa0"], st["raa"][0]
    f_true, g_true = st["f0app"] + 1e-3, st["fapp"][0] + 2e-3   # both under-estimated
    M.inner(xmma, f_true, np.array([g_true]), x, f0, df, g0, dg, lo, hi)
    st = M.state()
    assert st["raa0"] > r0_before and st["raa"][0] > raa_before
    a0, a1 = M.approx_at(xmma)
    capped0 = st["raa0"] == pytest.approx(10 * r0_before)
    if not capped0:
        assert a0 >= f_true + 0.1 * 1e-3 * 0.99
    assert a1[0] >= g_true or st["raa"][0] == pytest.approx(10 * raa_before)
    # an over-estimate leaves raa alone
    xm2 = M.inner(xmma, st["f0app"] - 1.0, st["fapp"] - 1.0, x, f0, df, g0, dg, lo, hi)[0]
    assert M.state()["raa0"] == st["raa0"]
    assert np.all(np.isfinite(xm2))


def _gcmma_loop(M, x, f, g, lo, hi, steps):
    inner_total = 0
    for _ in range(steps):
        f0, df, g0, dg = *f(x), *g(x)
        xn, _ = M.outer(x, f0, df, g0, dg, lo, hi)
        ok = M.concheck(f(xn)[0], g(xn)[0])
        k = 0
        while not ok and k < 15:
            xn, _ = M.inner(xn, f(xn)[0], g(xn)[0], x, f0, df, g0, dg, lo, hi)
            ok = M.concheck(f(xn)[0], g(xn)[0])
            k += 1
        inner_total += k
        assert ok or k == 15
        x = xn
    return x, inner_total


def test_separable_problem_reaches_its_closed_form_optimum():
    """min sum c_i / x_i s.t. sum x_i <= V: x_i = V sqrt(c_i) / sum sqrt(c_j)."""
    n = 20
    rng = np.random.default_rng(7)
    c = rng.uniform(0.5, 2.0, n)
    V = 0.4 * n
    f = lambda x: (np.sum(c / x), -c / x ** 2)
    g = lambda x: (np.sum(x) - V, np.ones(n))
    lo, hi = np.full(n, 1e-3), np.ones(n)
    x, _ = _gcmma_loop(O.GCMMA(n), np.full(n, 0.5), f, g, lo, hi, 80)
    xs = V * np.sqrt(c) / np.sqrt(c).sum()
    assert np.all((xs > 1e-3) & (xs < 1))
    np.testing.assert_allclose(x, xs, rtol=1e-5)


def test_inner_loop_ends_conservative_on_a_nonconvex_problem():
    n = 12
    f = lambda x: (-np.sum(np.sin(4 * x)) + np.sum(x ** 3), -4 * np.cos(4 * x) + 3 * x ** 2)
    g = lambda x: (np.sum(x) - 0.45 * n, np.ones(n))
    lo, hi = np.full(n, 1e-3), np.ones(n)
    x0 = np.random.default_rng(9).uniform(0.3, 0.6, n)
    M = O.GCMMA(n)
    x, inner = _gcmma_loop(M, x0, f, g, lo, hi, 10)
    assert inner > 0                      # the problem does need inner loops
    assert f(x)[0] < f(x0)[0]
    st = M.state()
    assert f(x)[0] <= st["f0app"] + 1e-8 and g(x)[0] <= st["fapp"][0] + 1e-8


def test_oracle_gcmma_driver_keeps_the_first_volume():
    """The loop of opt_macro_gcmma.cpp:83-216 on a small cantilever: const_h = V - V0 is
    0 at step 1 and stays near or below 0; the compliance falls; the inner loop runs."""
    pr = O.cantilever_2d(40, 14)
    ref = pr.run_driver(2, np.full(pr.nelx, 0.3), 18, 2.0, projection=True, direct=True)
    assert ref["vol"][0] == 0.0
    assert ref["vol"].max() < 1e-3 * pr.nelx
    assert ref["obj"][-1] < 0.5 * ref["obj"][0]
    assert ref["inner"][0] > 0 and ref["inner"].sum() > ref["inner"][0]
    assert np.all(ref["inner"] <= 15)
