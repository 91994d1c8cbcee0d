"""The oracle against the golden fixtures of tests/golden (computed without the
oracle: numpy 3-point quadrature, closed-form DOF numbering, SciPy SuperLU)."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = ["q4_cantilever_24x8.npz", "hex8_cantilever_10x4x4.npz"]


def load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


def oracle_problem(g):
    div = [int(v) for v in g["div"]]
    return O.cantilever_2d(*div) if len(div) == 2 else O.cantilever_3d(*div)


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_golden(name):
    g = load(name)
    pr = oracle_problem(g)
    b = pr.assemble(g["rho"])
    u = pr.direct(b)
    assert np.linalg.norm(u - g["u"]) / np.linalg.norm(g["u"]) < 1e-10
    assert abs(u @ pr.fext - g["c"]) / g["c"] < 1e-11
    dc, _ = pr.sensitivity(g["rho"], pr.element_energy(pr.node_values(u)), pr.volumes())
    assert np.abs(dc - g["dc"]).max() <= 1e-9 * np.abs(g["dc"]).max()
    assert np.abs(pr.cone_filter(1.5).apply(g["rho"]) - g["cone"]).max() < 1e-14
    h, _ = pr.helmholtz(2.0).apply(g["rho"], direct=True)
    assert np.abs(h - g["helm"]).max() < 1e-10
    hc, _ = pr.helmholtz(2.0).apply(g["rho"], tol=1e-12)
    assert np.abs(hc - g["helm"]).max() < 1e-10
