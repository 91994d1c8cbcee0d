"""The oracle on explicit (unstructured) meshes: the general element matrices
(src/problem/elasticity.cpp:210-254 with the Jacobian of each element) pinned by the
patch test, the rigid-body null space and exact volumes; the MDPA reader of libicx
(host code, no device) against the writer."""
import numpy as np
import pytest

import icx
import oracle as O
from meshgen import annulus_quad, box_mesh, distort, write_mdpa

CASES = [
    (4, O.PLANE_STRESS, (6, 4), (6.0, 4.0)),
    (4, O.PLANE_STRAIN, (5, 3), (5.0, 6.0)),
    (8, O.SOLID3D, (4, 3, 3), (4.0, 3.0, 3.0)),
]


def _free_problem(ne, state, x, conn):
    pr = O.Problem.from_mesh(ne, x, conn, state, young=(1e-9, 1.0))
    assert pr.make_dof() == pr.numnp * pr.dofnp
    pr.assemble(np.ones(pr.nelx))
    return pr


def _field(pr, f):
    """u[dof] = f(x) node by node"""
    X = pr.x.reshape(-1, 3)
    u = np.zeros(pr.numeq)
    for n in range(pr.numnp):
        v = f(X[n])
        for c in range(pr.dofnp):
            u[pr.dof[n * pr.dofnp + c]] = v[c]
    return u


@pytest.mark.parametrize("ne,state,div,L", CASES)
def test_patch_test_and_rigid_modes_on_distorted_meshes(ne, state, div, L):
    """A linear displacement field has constant strain, so the internal force K u is zero
    at every interior node of any mesh (patch test); rigid motions give zero everywhere.
    A wrong Jacobian, inverse or B-matrix entry fails the first."""
    x, conn = box_mesh(ne, div, L)
    xd = distort(x, conn, ne, 0.2, 7, L)
    pr = _free_problem(ne, state, xd, conn)
    nd = pr.ndim
    X = xd.reshape(-1, 3)
    lo, hi = X[:, :nd].min(axis=0), X[:, :nd].max(axis=0)
    interior = np.all((X[:, :nd] > lo + 1e-9) & (X[:, :nd] < hi - 1e-9), axis=1)
    assert interior.any()
    A = np.random.default_rng(3).standard_normal((nd, nd))
    u = _field(pr, lambda p: A @ p[:nd] + 0.5)
    f = pr.spmv(u)
    scale = np.abs(pr.csr()).max() * np.abs(u).max()
    for n in np.nonzero(interior)[0]:
        for c in range(pr.dofnp):
            assert abs(f[pr.dof[n * pr.dofnp + c]]) <= 1e-12 * scale
    assert np.abs(f).max() > 1e-3 * scale  # boundary nodes carry the tractions
    # rigid motions: translations and infinitesimal rotations
    modes = [lambda p, c=c: np.eye(nd)[c] for c in range(nd)]
    if nd == 2:
        modes.append(lambda p: np.array([-p[1], p[0]]))
    else:
        modes += [lambda p: np.array([-p[1], p[0], 0.0]), lambda p: np.array([0.0, -p[2], p[1]]),
                  lambda p: np.array([p[2], 0.0, -p[0]])]
    for m in modes:
        assert np.abs(pr.spmv(_field(pr, m))).max() <= 1e-12 * scale


@pytest.mark.parametrize("ne,state,div,L", CASES)
def test_volumes_tile_the_box(ne, state, div, L):
    """det J is integrated exactly by the 2 (x 2 x 2) Gauss rule, and the distorted
    elements tile the box."""
    x, conn = box_mesh(ne, div, L)
    pr = O.Problem.from_mesh(ne, distort(x, conn, ne, 0.2, 11, L), conn, state)
    v = pr.volumes()
    assert (v > 0).all()
    assert abs(v.sum() - np.prod(L)) <= 1e-12 * np.prod(L)


def test_annulus_volumes_equal_the_polygon_areas():
    x, conn = annulus_quad(4, 6, 1.0, 2.0)
    pr = O.Problem.from_mesh(4, x, conn, O.PLANE_STRESS)
    X = x.reshape(-1, 3)
    v = pr.volumes()
    for e, nodes in enumerate(conn.reshape(-1, 4)):
        P = X[nodes, :2]
        shoelace = 0.5 * sum(P[a, 0] * P[(a + 1) % 4, 1] - P[(a + 1) % 4, 0] * P[a, 1] for a in range(4))
        assert abs(v[e] - shoelace) <= 1e-14 * shoelace


@pytest.mark.parametrize("ne", [4, 8])
def test_mdpa_reader_round_trip(tmp_path, ne):
    """icx_mdpa_read (host code of libicx) reads back what the writer wrote: coordinates
    bit for bit, connectivity 0-based; blocks of other element types are skipped."""
    L = (3.0, 2.0) if ne == 4 else (3.0, 2.0, 2.0)
    div = (3, 2) if ne == 4 else (3, 2, 2)
    x, conn = box_mesh(ne, div, L)
    x = distort(x, conn, ne, 0.15, 5, L)
    path = tmp_path / "mesh.mdpa"
    write_mdpa(path, x, conn, ne)
    with open(path, "a") as f:  # a block of another element type
        f.write("\nBegin Elements SmallDisplacementElement2D3N\n1 0 1 2 3\nEnd Elements\n")
    xr, cr = icx.read_mdpa(path, ne)
    assert np.array_equal(xr, x) and np.array_equal(cr, conn)
    other = 8 if ne == 4 else 4
    with pytest.raises(icx.IcxError):
        icx.read_mdpa(path, other)          # no block of that type
    with pytest.raises(icx.IcxError):
        icx.read_mdpa(tmp_path / "missing.mdpa", ne)
