"""Oracle pins for the OC and MMA updates (reference:
src/optimization/opt-solver/OCSolver.hpp:20-68, MMASolver.cpp:41-553)."""
import numpy as np

import oracle as O


def test_oc_bisection_against_brute_force():
    rng = np.random.default_rng(3)
    n = 8
    x = rng.uniform(0.2, 0.8, n)
    dc = -rng.uniform(0.1, 2.0, n)
    dv = np.ones(n)
    xmin, xmax = np.full(n, 1e-4), np.ones(n)
    V0 = x.sum()

    def xnew(lam):
        v = np.sqrt(np.maximum(0, -dc / (dv * lam))) * x
        lo = np.maximum(xmin, 0.85 * x)
        hi = np.minimum(xmax, 1.15 * x)
        return np.where(v < lo, lo, np.where(v > hi, hi, v))

    xo, it = O.oc_update(x, dc, V0, dv, xmin, xmax, lambda z: z.sum())
    # brute force: the reference's bisection sequence, replayed in Python
    l0, l1, cnt = 0.0, 1e4, 0
    while (l1 - l0) / (l1 + l0) > 1e-5 and cnt < 100:
        lam = 0.5 * (l1 + l0)
        if xnew(lam).sum() - V0 > 0:
            l0 = lam
        else:
            l1 = lam
        cnt += 1
    assert it == cnt
    assert np.array_equal(xo, xnew(lam))
    # move limits and volume within the bisection tolerance
    assert np.all(xo >= np.maximum(xmin, 0.85 * x) - 1e-15)
    assert np.all(xo <= np.minimum(xmax, 1.15 * x) + 1e-15)
    assert abs(xo.sum() - V0) < 1e-3 * V0


def test_mma_converges_on_convex_separable_problem():
    """min sum c_i / x_i  s.t. sum x_i <= V: optimum x_i proportional to sqrt(c_i)."""
    rng = np.random.default_rng(5)
    n = 20
    c = rng.uniform(0.5, 2.0, n)
    V = 0.4 * n
    x = np.full(n, 0.4)
    mma = O.MMA(n, 1)
    xmin, xmax = np.full(n, 0.01), np.ones(n)
    for _ in range(60):
        df = -c / x ** 2
        g = x.sum() / V - 1.0
        dg = np.full(n, 1.0 / V)
        x, _ = mma.update(x, df, g, dg, xmin, xmax)
    opt = np.sqrt(c) / np.sqrt(c).sum() * V
    opt = np.clip(opt, 0.01, 1.0)
    assert np.abs(x - opt).max() < 1e-3
    assert x.sum() <= V * (1 + 1e-4)


def test_mma_subproblem_solution_is_dual_optimal():
    """After one update, x(lambda) minimises the separable MMA Lagrangian on [alpha, beta]:
    here with m = 1 the inactive constraint (g << 0) drives lambda -> 0 and x to the
    unconstrained subproblem minimiser, which moves each variable toward lower f."""
    n = 10
    x0 = np.full(n, 0.5)
    df = np.linspace(-1.0, 1.0, n)
    mma = O.MMA(n, 1)
    x1, steps = mma.update(x0, df, -1.0, np.full(n, 1e-3), np.zeros(n), np.ones(n))
    assert steps > 0
    assert np.all(np.sign(x1 - x0)[df < -0.05] > 0) and np.all(np.sign(x1 - x0)[df > 0.05] < 0)
    # move limit 0.15 * (xmax - xmin) and albefa bound
    assert np.all(np.abs(x1 - x0) <= 0.15 + 1e-12)
