"""Closed-form pins of the oracle's OC and MMA updates.

OC (src/optimization/opt-solver/OCSolver.hpp:20-68): on a separable problem with a
linear volume the bisection's limit lambda* has a closed form, so the exponent (iota),
the move limit and the bracket are fixed by numbers worked out by hand.

MMA (src/optimization/opt-solver/MMASolver.cpp:41-129, 385-512): the asymptote history
(the init rule for the first two updates, then gamma = asymdec / asyminc / 1 from the sign
of (x - xold1)(xold1 - xold2), the 1e-5 and 100 times (xmax - xmin) clamps and the
out-of-box rule), the move bounds alpha/beta, the GenSub coefficients (first-order
exactness at xval, raa0 and the 0.001 terms by literal values) and the dual interior-point
solution (KKT conditions of the m = 1 subproblem with c = 1e5)."""
import numpy as np
import pytest

import oracle as O


# --------------------------------------------------------------------------------- OC
def _oc(x, c, V, xmin=1e-4, xmax=1.0):
    n = len(x)
    return O.oc_update(x, -np.asarray(c, float), V, np.ones(n), np.full(n, xmin),
                       np.full(n, xmax), lambda z: z.sum())


def test_oc_unclamped_limit_is_closed_form():
    """x_new = x (c / lambda)^(1/2) with sum x_new = V gives
    lambda* = (sum x sqrt(c) / V)^2.  Here x = 0.5, c = 0.81 ... 1.21 and V = sum x,
    so every candidate stays inside the 15% move limit."""
    n = 9
    x = np.full(n, 0.5)
    c = np.array([0.81, 0.9, 0.95, 1.0, 1.0, 1.05, 1.1, 1.15, 1.21])
    V = x.sum()
    xo, it = _oc(x, c, V)
    lam_star = (np.sum(x * np.sqrt(c)) / V) ** 2
    # every coordinate implies the same lambda (exponent 1/2), equal to lambda* within
    # the bisection's relative bracket 1e-5
    lam_i = c * (x / xo) ** 2
    assert np.ptp(lam_i) <= 1e-14 * lam_i.mean()
    assert abs(lam_i.mean() - lam_star) <= 2e-5 * lam_star
    assert np.all(np.abs(xo - x * np.sqrt(c / lam_star)) <= 1e-5 * x)
    # bracket [0, 1e4] halved until (l1 - l0) / (l1 + l0) <= 1e-5: lambda* ~ 1 needs
    # log2(1e4 / 2e-5 / 1) ~ 29 halvings
    assert 27 <= it <= 31


def test_oc_move_limits_and_lower_bound_enter_lambda():
    """Two variables hit the upper move limit 1.15 x (large c), one has c = 0 and falls to
    max(xmin, 0.85 x); the others solve sum_free x sqrt(c / lambda) = V - sum_clamped."""
    x = np.array([0.4, 0.4, 0.4, 0.4, 0.4, 0.4])
    c = np.array([1.0, 1.0, 1.0, 9.0, 9.0, 0.0])
    V = 2.4
    xo, _ = _oc(x, c, V)
    hi = 1.15 * 0.4
    lo = 0.85 * 0.4
    assert xo[3] == hi and xo[4] == hi and xo[5] == lo
    rest = V - 2 * hi - lo                       # = 1.14, shared by the three c = 1 variables
    lam_star = (3 * 0.4 / rest) ** 2             # sqrt(lambda*) = 3 * 0.4 / 1.14
    assert np.all(np.abs(xo[:3] - 0.38) <= 1e-5 * 0.38)          # 1.14 / 3
    lam_i = c[:3] * (x[:3] / xo[:3]) ** 2
    assert np.all(np.abs(lam_i - lam_star) <= 2e-5 * lam_star)
    # the clamped candidates are consistent with lambda*: sqrt(9 / lambda*) > 1.15
    assert np.sqrt(9 / lam_star) > 1.15


def test_oc_xmin_bound():
    """A candidate below xmin is raised to xmin when xmin > 0.85 x."""
    x = np.array([0.5, 0.5, 0.001])
    c = np.array([1.0, 1.0, 1e-12])
    xo, _ = _oc(x, c, 1.0, xmin=0.01)
    assert xo[2] == 0.01


# -------------------------------------------------------------------------------- MMA
def _update(mma, xval, n, df=None, g=-0.5, dg=None, xmin=0.0, xmax=1.0):
    df = np.zeros(n) if df is None else df
    dg = np.full(n, 1e-3) if dg is None else dg
    return mma.update(np.asarray(xval, float), df, g, dg, np.full(n, xmin), np.full(n, xmax))


def test_mma_asymptote_history_closed_form():
    """Asymptote distance d_k = x_k - low_k = upp_k - x_k (xmax - xmin = 1):
    d_1 = d_2 = asyminit = 0.2, then d_k = gamma_k d_{k-1} with gamma 0.7 when the last
    two moves change sign, 1.2 when they keep it and 1 when a move is zero."""
    seq = np.array([[0.5, 0.3, 0.5, 0.5],
                    [0.6, 0.4, 0.5, 0.6],
                    [0.5, 0.5, 0.5, 0.5],
                    [0.6, 0.6, 0.5, 0.4]])
    expect = [[0.2, 0.2, 0.2, 0.2],
              [0.2, 0.2, 0.2, 0.2],
              [0.14, 0.24, 0.2, 0.14],        # oscillate, monotone, still, oscillate
              [0.098, 0.288, 0.2, 0.168]]     # 0.7^2, 1.2^2, 1, 0.7 * 1.2
    n = 4
    mma = O.MMA(n, 1)
    for k in range(4):
        _update(mma, seq[k], n)
        s = mma.subproblem()
        assert np.allclose(seq[k] - s["low"], expect[k], rtol=0, atol=1e-15), k
        assert np.allclose(s["upp"] - seq[k], expect[k], rtol=0, atol=1e-15), k
    # move bounds at update 4: alpha = max(xmin, low + 0.2 (x - low), x - 0.15):
    # variable 0 (d = 0.098) takes the albefa bound x - 0.8 * 0.098, variable 1
    # (d = 0.288) the move limit x - 0.15
    assert abs(s["alpha"][0] - (0.6 - 0.0784)) <= 1e-15
    assert abs(s["beta"][0] - (0.6 + 0.0784)) <= 1e-15
    assert abs(s["alpha"][1] - 0.45) <= 1e-15 and abs(s["beta"][1] - 0.75) <= 1e-15


def test_mma_asymptote_clamps_after_many_updates():
    """40 updates: a variable that keeps moving the same way grows its distance by 1.2
    per update until the cap 100 (xmax - xmin); one that oscillates shrinks it by 0.7 to
    the floor 1e-5 (xmax - xmin)."""
    n = 2
    mma = O.MMA(n, 1)
    for k in range(1, 41):
        xv = [0.1 + 0.02 * k, 0.5 + 0.1 * (k % 2)]
        _update(mma, xv, n)
    s = mma.subproblem()
    d_up = np.array([s["upp"][0] - xv[0], xv[0] - s["low"][0]])
    d_osc = np.array([s["upp"][1] - xv[1], xv[1] - s["low"][1]])
    assert np.allclose(d_up, 100.0, rtol=1e-15, atol=0)       # 0.2 * 1.2^38 = 202 > 100
    assert np.allclose(d_osc, 1e-5, rtol=1e-9, atol=0)        # 0.2 * 0.7^38 < 1e-5


def test_mma_asymptotes_for_a_point_outside_the_box():
    """xval above xmax + 1e-6: low/upp = x -/+ (x - (xmin - 1e-6)) / 0.9."""
    n = 1
    mma = O.MMA(n, 1)
    for xv in (0.5, 0.6, 1.2):
        _update(mma, [xv], n)
    s = mma.subproblem()
    half = (1.2 + 1e-6) / 0.9
    assert abs(s["low"][0] - (1.2 - half)) <= 1e-15
    assert abs(s["upp"][0] - (1.2 + half)) <= 1e-15
    assert s["alpha"][0] == 1.0 and s["beta"][0] == 1.0


def test_mma_gensub_coefficients():
    """At the first update (d = 0.2, xmax - xmin = 1):
    p0 = d^2 (max(df, 0) + 0.001 |df| + raa0), q0 = d^2 (max(-df, 0) + 0.001 |df| + raa0)
    with raa0 = 1e-5 -- literal values for df = 0, 1 and -2; the same form for the
    constraint; and b makes the constraint approximation equal gx at xval."""
    n = 3
    mma = O.MMA(n, 1)
    df = np.array([0.0, 1.0, -2.0])
    dg = np.array([0.5, 0.0, -0.25])
    _update(mma, [0.5, 0.5, 0.5], n, df=df, g=0.125, dg=dg)
    s = mma.subproblem()
    assert np.allclose(s["p0"], [4e-7, 0.04 * 1.00101, 0.04 * 0.00201], rtol=1e-14, atol=0)
    assert np.allclose(s["q0"], [4e-7, 0.04 * 0.00101, 0.04 * 2.00201], rtol=1e-14, atol=0)
    assert np.allclose(s["pij"][:, 0], [0.04 * 0.50051, 4e-7, 0.04 * 0.00026], rtol=1e-14)
    assert np.allclose(s["qij"][:, 0], [0.04 * 0.00051, 4e-7, 0.04 * 0.25026], rtol=1e-14)
    x = np.full(n, 0.5)
    gt = np.sum(s["pij"][:, 0] / (s["upp"] - x) + s["qij"][:, 0] / (x - s["low"])) - s["b"][0]
    assert abs(gt - 0.125) <= 1e-15


def test_mma_gensub_first_order_exact_with_history():
    """After several updates (asymptotes from the history) the approximations still
    match f and g to first order at xval: p/(U-x)^2 - q/(x-L)^2 = df, and both are convex
    (p, q > 0)."""
    rng = np.random.default_rng(8)
    n = 40
    mma = O.MMA(n, 1)
    x = rng.uniform(0.2, 0.8, n)
    for _ in range(5):
        df = rng.standard_normal(n)
        dg = rng.standard_normal(n) * 0.1
        xn, _ = mma.update(x, df, -0.1, dg, np.zeros(n), np.ones(n))
        s = mma.subproblem()
        U, L = s["upp"] - x, x - s["low"]
        assert np.allclose(s["p0"] / U ** 2 - s["q0"] / L ** 2, df, rtol=1e-12, atol=1e-14)
        assert np.allclose(s["pij"][:, 0] / U ** 2 - s["qij"][:, 0] / L ** 2, dg,
                           rtol=1e-12, atol=1e-14)
        assert np.all(s["p0"] > 0) and np.all(s["q0"] > 0)
        assert np.all(s["alpha"] <= xn) and np.all(xn <= s["beta"])
        x = xn


@pytest.mark.parametrize("active", [True, False])
def test_mma_subproblem_kkt(active):
    """KKT of the m = 1 subproblem (a = 0, c = 1e5, d = 0 as the reference constructs it)
    at the dual interior point's answer: x minimises the Lagrangian on [alpha, beta]
    (stationary where free, the right sign at a bound), y = max(0, lambda - c), z = 0,
    the approximated constraint holds, and lambda times its slack is at the interior
    point's final epsilon (sqrt(n + m) 1e-9)."""
    rng = np.random.default_rng(21)
    n = 30
    cst = rng.uniform(0.5, 2.0, n)
    # active: start on the constraint; slack: V = 2n can never be reached (x <= 1)
    V = 0.4 * n if active else 2.0 * n
    x = np.full(n, 0.4)
    mma = O.MMA(n, 1)
    for _ in range(3):
        df = -cst / x ** 2
        g = x.sum() / V - 1.0
        xn, _ = mma.update(x, df, g, np.full(n, 1.0 / V), np.full(n, 0.01), np.ones(n))
        s = mma.subproblem()
        lam = s["lam"][0]
        U, L = s["upp"] - xn, xn - s["low"]
        P = s["p0"] + lam * s["pij"][:, 0]
        Q = s["q0"] + lam * s["qij"][:, 0]
        phi = P / U ** 2 - Q / L ** 2
        scale = P / U ** 2 + Q / L ** 2
        free = (xn > s["alpha"]) & (xn < s["beta"])
        # with the constraint active some variables are interior; with it slack every
        # variable runs to its upper move bound (f decreases in every x)
        assert free.sum() > 0 if active else np.all(xn == s["beta"])
        assert np.all(np.abs(phi[free]) <= 1e-12 * scale[free])
        assert np.all(phi[xn == s["alpha"]] >= -1e-12 * scale[xn == s["alpha"]])
        assert np.all(phi[xn == s["beta"]] <= 1e-12 * scale[xn == s["beta"]])
        gt = np.sum(s["pij"][:, 0] / U + s["qij"][:, 0] / L) - s["b"][0]
        assert s["y"][0] == max(0.0, lam - 1e5) and s["z"] == 0.0
        # the barrier parameter runs 1, 0.1, ..., 1e-8 (the next, 1e-9, is below
        # sqrt(n + 1) 1e-9 = 5.6e-9 and ends the loop); a stage ends with
        # |constraint + mu| and |mu lambda - eps| below 0.9 eps, and a stage whose
        # tolerance the previous stage's residual already meets takes no step
        # (MMASolver.cpp:102-149), so mu lambda may stay at an earlier eps
        slack = gt - s["y"][0]
        assert slack <= 0.9e-8
        assert abs(lam * slack) <= (1.9e-8 + 0.9e-8 * lam if active else 2e-7)
        if active:
            assert lam > 1e-3
        x = xn
    if not active:
        # a slack constraint ends with lambda at the interior point's floor
        assert lam < 1e-6
