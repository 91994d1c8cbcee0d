"""Oracle pins for the stress-constrained driver (constraint/opt-stress-constraint/
sensitivity.cpp:8-155, opt_stress_constraint.cpp:85-158): the adjoint derivative of the
qp-relaxed p-norm stress against central finite differences through the solve, the
p-norm against its definition over the element stresses, and the loop's behaviour."""
import numpy as np
import pytest

import oracle as O


def _eval(pr, rho, qp=0.7, pn=8.0, slim=1.0):
    u = pr.direct(pr.assemble(rho))
    nv = pr.node_values(u)
    return pr.stress_constraint(nv, qp, pn, slim), pr.mises(nv)


@pytest.mark.parametrize("state", [O.PLANE_STRESS, O.PLANE_STRAIN])
def test_stress_derivative_matches_finite_differences(state):
    pr = O.cantilever_2d(12, 6, state=state)
    rho = np.random.default_rng(1).uniform(0.3, 1.0, pr.nelx)
    (g, dg, sqp), vm = _eval(pr, rho)
    scale = np.abs(dg).max()
    h = 1e-6
    for e in (0, 5, 17, 40, 71):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        fd = (_eval(pr, rp)[0][0] - _eval(pr, rm)[0][0]) / (2 * h)
        assert abs(dg[e] - fd) <= 1e-6 * scale, (e, dg[e], fd)


def test_stress_pnorm_is_its_definition():
    """s_e = rho^qp E2 vm_e / E_e from the element-mean von Mises stress, and
    g = ||s / slim||_P - 1; s_e <= vm_e E2 / E_e, and g tends to max(s)/slim - 1 as P grows."""
    pr = O.cantilever_2d(10, 5)
    rho = np.random.default_rng(2).uniform(0.2, 1.0, pr.nelx)
    (g, _, sqp), vm = _eval(pr, rho, slim=0.5)
    E = 1e-9 + rho ** 3 * (1.0 - 1e-9)
    np.testing.assert_allclose(sqp, rho ** 0.7 * vm / E, rtol=1e-12)
    assert g == pytest.approx(np.sum((sqp / 0.5) ** 8) ** (1 / 8) - 1, rel=1e-12)
    (g64, _, _), _ = _eval(pr, rho, pn=64.0, slim=0.5)
    assert abs(g64 - (sqp.max() / 0.5 - 1)) < 0.1 * (sqp.max() / 0.5)


def test_stress_driver_runs_and_limits_the_stress():
    """20 steps of driver 3 on a small cantilever whose point load violates the stress
    limit at the start: MMA adds material (the volume objective rises) and drives the
    constraint down by more than an order of magnitude."""
    pr = O.cantilever_2d(24, 12)
    ref = pr.run_driver(3, np.full(pr.nelx, 0.5), 20, 2.0, projection=True, direct=True,
                        stress=(1.0, 8.0, 0.7))
    assert np.all(np.isfinite(ref["obj"])) and np.all(np.isfinite(ref["vol"]))
    assert ref["vol"][0] > 1.0
    assert ref["vol"][-1] < 0.1 * ref["vol"][0]
    assert ref["obj"][-1] > ref["obj"][0]
