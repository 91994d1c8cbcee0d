"""Oracle pins for the integer set-up: box mesh, Dirichlet flags, DOF numbering,
Neumann tractions (reference: src/io/mesh_icarat.hpp:325-494,
src/io/boundary_condition_icarat.hpp:68-396, src/io/boundary_condition_base.hpp:18-88,
src/io/boundary_condition_base.cpp:16-180).  Expected values are worked out by hand
from the reference's rules, not by re-running the oracle's formulas."""
import numpy as np

import oracle as O


def test_quad4_2x1_mesh_and_dofs():
    pr = O.Problem(4, (2, 1, 1), (2.0, 1.0, 1.0), O.PLANE_STRESS)
    X = pr.x.reshape(-1, 3)
    assert X[:, :2].tolist() == [[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]]
    assert pr.conn.reshape(-1, 4).tolist() == [[0, 1, 4, 3], [1, 2, 5, 4]]
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 1, 0, 0])
    assert pr.make_dof() == 8
    assert pr.dof.reshape(-1, 2).tolist() == [[8, 9], [0, 1], [2, 3], [10, 11], [4, 5], [6, 7]]


def test_hexa8_2x1x1_mesh_and_dofs():
    pr = O.Problem(8, (2, 1, 1), (2.0, 1.0, 1.0), O.SOLID3D)
    assert pr.conn.reshape(-1, 8).tolist() == [[0, 1, 4, 3, 6, 7, 10, 9],
                                               [1, 2, 5, 4, 7, 8, 11, 10]]
    X = pr.x.reshape(-1, 3)
    assert X[11].tolist() == [2, 1, 1] and X[6].tolist() == [0, 0, 1]
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, 1, 0, 1])
    assert pr.make_dof() == 24
    d = pr.dof.reshape(-1, 3)
    free_nodes = [1, 2, 4, 5, 7, 8, 10, 11]
    for q, n in enumerate(free_nodes):
        assert d[n].tolist() == [3 * q, 3 * q + 1, 3 * q + 2]
    for q, n in enumerate([0, 3, 6, 9]):
        assert d[n].tolist() == [24 + 3 * q, 25 + 3 * q, 26 + 3 * q]


def test_flags_are_or_ed_and_partial_constraints():
    pr = O.Problem(4, (2, 1, 1), (2.0, 1.0, 1.0), O.PLANE_STRESS)
    pr.add_dirichlet([1, 0, 0, 0, 0, 0], [0.5, 0, 0, 0, 0, 0], [0, 0, 0, 1, 0, 0])
    pr.add_dirichlet([0, 1, 0, 0, 0, 0], [0, 0.25, 0, 0, 0, 0], [0, 2, 0, 0, 0, 0])
    assert pr.make_dof() == 7
    assert pr.dof.reshape(-1, 2).tolist() == [[7, 8], [0, 9], [1, 10], [11, 2], [3, 4], [5, 6]]
    # the last matching entry's values win (setDirichValue overwrites val)
    nv = pr.node_val.reshape(-1, 6)
    assert nv[0, :2].tolist() == [0.0, 0.25] and nv[3, :2].tolist() == [0.5, 0.0]


def test_clamped_face_dof_closed_form():
    nx, ny, nz = 8, 4, 4
    pr = O.cantilever_3d(nx, ny, nz)
    sx, sy = nx + 1, ny + 1
    d = pr.dof.reshape(-1, 3)
    for n in range(pr.numnp):
        i, row = n % sx, n // sx
        if i == 0:
            base = pr.numeq + 3 * row
        else:
            base = 3 * (n - (row + 1))
        assert d[n].tolist() == [base, base + 1, base + 2]


def test_numeq_at_benchmark_sizes():
    assert O.cantilever_2d().numeq == 9840
    pr = O.cantilever_3d(96, 48, 48)
    assert pr.numeq == 691488 and pr.numnp == 232897 and pr.nelx == 221184
    # 192x96x96: 3*193*97*97 DOFs minus the clamped face
    assert 3 * 193 * 97 * 97 - 3 * 97 * 97 == 5419584


def test_edge_traction_consistent_load():
    pr = O.Problem(4, (2, 1, 1), (2.0, 1.0, 1.0), O.PLANE_STRESS)
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 1, 0, 0])
    pr.make_dof()
    pr.add_neumann([0.0, -1.0, 0.0], [2, 2, 0, 1, 0, 0])
    f = pr.fext
    assert np.allclose(f[[3, 7]], [-0.5, -0.5], rtol=0, atol=1e-15)
    assert np.count_nonzero(f) == 2


def test_traction_box_covering_whole_element_follows_judgeline():
    # e1 has all four nodes in the box: its 2nd hit is local node 1 -> side 1
    # (bottom edge); e0 hits local nodes 1 and 2 -> side 2 (right edge).
    pr = O.Problem(4, (2, 1, 1), (2.0, 1.0, 1.0), O.PLANE_STRESS)
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 1, 0, 0])
    pr.make_dof()
    pr.add_neumann([1.0, 0.0, 0.0], [1, 2, 0, 1, 0, 0])
    d = pr.dof.reshape(-1, 2)
    f = pr.fext
    assert abs(f[d[1, 0]] - 1.0) < 1e-15   # node 1: e1 bottom + e0 right
    assert abs(f[d[2, 0]] - 0.5) < 1e-15   # node 2: e1 bottom
    assert abs(f[d[4, 0]] - 0.5) < 1e-15   # node 4: e0 right
    assert abs(f.sum() - 2.0) < 1e-14


def test_face_traction_quarter_weights():
    pr = O.Problem(8, (2, 1, 1), (2.0, 1.0, 1.0), O.SOLID3D)
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, 1, 0, 1])
    pr.make_dof()
    pr.add_neumann([0.0, 0.0, -2.0], [2, 2, 0, 1, 0, 1])
    d = pr.dof.reshape(-1, 3)
    for n in (2, 5, 8, 11):
        assert abs(pr.fext[d[n, 2]] + 0.5) < 1e-15
    assert abs(pr.fext.sum() + 2.0) < 1e-14


def test_nodal_load_convention():
    pr = O.cantilever_2d()
    n = pr.node_index(120, 20)
    assert pr.fext[pr.dof[2 * n + 1]] == -1.0 and pr.fext.sum() == -1.0
    p3 = O.cantilever_3d(8, 4, 4)
    assert abs(p3.fext.sum() + 1.0) < 1e-15
