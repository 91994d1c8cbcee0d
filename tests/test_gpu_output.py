"""GPU parity of the output step (reference: src/problem/elasticity.cpp:165-208,
src/analysis/linear.hpp:90-91, src/analysis/analysis_utility.hpp:33-51,
src/io/export_vtk.hpp:47-156): element-mean von Mises stress and reactions against
the oracle on the same node values (1e-12 relative to the largest entry), and the
binary VTK file read back against the mesh and the device fields."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu

QUAD4, HEXA8 = 4, 8


def _pair(ctx, ne, div, length, state, bcs, loads=(), seed=4):
    job = icx.Job(ctx, ne, div, length, state)
    pr = O.Problem(ne, div, length, state)
    for f, v, r in bcs:
        job.add_dirichlet(f, v, r)
        pr.add_dirichlet(f, v, r)
    assert job.make_dof() == pr.make_dof()
    for node, comp, val in loads:
        job.add_nodal_load(node, comp, val)
        pr.add_nodal_load(node, comp, val)
    rho = O.random_density(pr.nelx, seed)
    job.set_design(rho)
    pr.assemble(rho)
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    assert info["converged"]
    return job, pr, rho


CASES = {
    "hexa8": (HEXA8, (6, 3, 4), (6.0, 1.5, 2.0), icx.SOLID3D,
              [([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, 1.5, 0, 2])], [((6, 1, 0), 2, -1.0)]),
    "quad4_stress": (QUAD4, (8, 4, 1), (8.0, 4.0, 1.0), icx.PLANE_STRESS,
                     [([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 4, 0, 0])], [((8, 2), 1, -1.0)]),
    "quad4_strain": (QUAD4, (8, 4, 1), (8.0, 2.0, 1.0), icx.PLANE_STRAIN,
                     [([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 2, 0, 0])], [((8, 2), 0, 0.5)]),
    # prescribed displacement on the right edge: node.val carries the Dirichlet values
    "quad4_prescribed": (QUAD4, (6, 3, 1), (6.0, 3.0, 1.0), icx.PLANE_STRESS,
                         [([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 3, 0, 0]),
                          ([1, 0, 0, 0, 0, 0], [0.01, 0, 0, 0, 0, 0], [6, 6, 0, 3, 0, 0])], []),
}


def _node(div, ijk):
    """node id i + sx (j + sy k) (x fastest, the reference's numbering)"""
    sx, sy = div[0] + 1, div[1] + 1
    i, j, k = (tuple(ijk) + (0,))[:3]
    return i + sx * (j + sy * k)


def _build(ctx, name):
    ne, div, length, state, bcs, loads = CASES[name]
    return _pair(ctx, ne, div, length, state, bcs, [(_node(div, n), c, v) for n, c, v in loads])


@pytest.mark.parametrize("name", sorted(CASES))
def test_mises_and_reactions_match_oracle(ctx, name):
    job, pr, rho = _build(ctx, name)
    nv = job.node_values()
    m = job.mises()
    mo = pr.mises(nv)
    assert np.abs(m - mo).max() <= 1e-12 * np.abs(mo).max()
    R = job.reactions()
    Ro = pr.reactions(nv)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
    # K u over all DOFs sums to zero per component (rigid translations)
    assert np.abs(R.reshape(-1, 3).sum(axis=0)).max() <= 1e-10 * np.abs(Ro).max()


def _read_vtk(path):
    """minimal reader of the binary legacy VTK unstructured grid written by icx_write_vtk"""
    data = open(path, "rb").read()
    out, pos = {}, 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    def block(n, dt):
        nonlocal pos
        a = np.frombuffer(data, dtype=dt, count=n, offset=pos)
        pos += 4 * n
        assert data[pos:pos + 1] == b"\n"
        pos += 1
        return a

    assert line() == "# vtk DataFile Version 4.1"
    line()
    assert line() == "BINARY"
    assert line() == "DATASET UNSTRUCTURED_GRID"
    while pos < len(data):
        w = line().split()
        if w[0] == "POINTS":
            out["points"] = block(3 * int(w[1]), ">f4").reshape(-1, 3)
        elif w[0] == "CELLS":
            out["cells"] = block(int(w[2]), ">i4")
        elif w[0] == "CELL_TYPES":
            out["types"] = block(int(w[1]), ">i4")
        elif w[0] in ("POINT_DATA", "CELL_DATA"):
            out["n_" + w[0]] = int(w[1])
        elif w[0] == "VECTORS":
            out[w[1]] = block(3 * out["n_POINT_DATA"], ">f4").reshape(-1, 3)
        elif w[0] == "SCALARS":
            assert line() == "LOOKUP_TABLE default"
            out[w[1]] = block(out["n_CELL_DATA"], ">f4")
    return out


@pytest.mark.parametrize("name", ["hexa8", "quad4_prescribed"])
def test_vtk_file_holds_the_state(ctx, name, tmp_path):
    job, pr, rho = _build(ctx, name)
    path = str(tmp_path / "res1.vtk")
    job.write_vtk(path)
    v = _read_vtk(path)
    x, conn = job.mesh()
    nq = 8 if name == "hexa8" else 4
    conn = conn.reshape(-1, nq)
    assert np.array_equal(v["points"], x.reshape(-1, 3).astype(np.float32))
    cells = v["cells"].reshape(-1, nq + 1)
    assert np.all(cells[:, 0] == nq) and np.array_equal(cells[:, 1:], conn)
    assert np.all(v["types"] == (12 if nq == 8 else 9))
    assert np.array_equal(v["displacement"], job.node_values().reshape(-1, 3).astype(np.float32))
    assert np.array_equal(v["density"], rho.astype(np.float32))
    assert np.array_equal(v["mises"], job.mises().astype(np.float32))


def test_output_begin_captures_the_state_at_begin(ctx, tmp_path):
    """icx_output_begin queues its copies on the job's stream: a solve for another design
    issued before icx_output_write does not change the file."""
    job, pr, rho = _build(ctx, "hexa8")
    m0, nv0 = job.mises(), job.node_values()
    job.output_begin()
    job.set_design(np.full(job.nelx, 0.9))
    job.solve(icx.SOLVE_PLAIN, tol=1e-10, maxit=20000)
    path = str(tmp_path / "async.vtk")
    job.output_write(path)
    v = _read_vtk(path)
    assert np.array_equal(v["mises"], m0.astype(np.float32))
    assert np.array_equal(v["displacement"], nv0.reshape(-1, 3).astype(np.float32))
    assert np.array_equal(v["density"], rho.astype(np.float32))
    assert not np.array_equal(job.mises().astype(np.float32), v["mises"])


def test_sensitivity_reads_prescribed_displacements(ctx):
    """The drivers' sensitivity() gathers node.val, which holds the prescribed values on
    constrained DOFs (opt-macro-oc/opt_macro_oc.cpp:193-203): with a prescribed
    displacement the device sensitivity matches the oracle on the same node values."""
    job, pr, rho = _build(ctx, "quad4_prescribed")
    nv = job.node_values()
    assert np.abs(nv).max() >= 0.01 - 1e-15        # the prescribed edge displacement
    dc, _ = job.sensitivity()
    dco, _ = pr.sensitivity(rho, pr.element_energy(nv), pr.volumes())
    assert np.abs(dc - dco).max() <= 1e-12 * np.abs(dco).max()
