"""Edge cases through the C ABI: single-element meshes in every solve form (structured
and unstructured jobs), jobs with no free DOF, zero loads, and the size limits the
library states (at most 2^27 nodes per job; positive divisions)."""
import numpy as np
import pytest

import icx
import oracle as O
from meshgen import box_mesh

pytestmark = pytest.mark.gpu

FORMS = (icx.SOLVE_PLAIN, icx.SOLVE_FUSED, icx.SOLVE_MIXED)


def _one_element(ne):
    nd = 2 if ne == 4 else 3
    div = (1, 1, 1) if nd == 3 else (1, 1)
    L = (2.0, 1.0, 1.5) if nd == 3 else (2.0, 1.0)
    st = O.SOLID3D if nd == 3 else O.PLANE_STRESS
    clamp = [0.0, 0.0, 0.0, L[1], 0.0, L[2] if nd == 3 else 0.0]
    load = [L[0], L[0], 0.0, L[1], 0.0, L[2] if nd == 3 else 0.0]
    val = [0.1, -0.3, 0.2] if nd == 3 else [0.1, -0.3, 0.0]
    return nd, div, L, st, clamp, load, val


@pytest.mark.parametrize("ne", [4, 8])
@pytest.mark.parametrize("mesh", [False, True])
def test_single_element_every_form(ctx, ne, mesh):
    nd, div, L, st, clamp, load, val = _one_element(ne)
    if mesh:
        x, conn = box_mesh(ne, div, L)
        job = icx.Job.from_mesh(ctx, ne, x, conn, st)
        forms = (icx.SOLVE_PLAIN, icx.SOLVE_MIXED)
    else:
        job = icx.Job(ctx, ne, div, L, st)
        forms = FORMS
    pr = O.Problem(ne, tuple(div) + (1,) * (3 - len(div)), tuple(L) + (1.0,) * (3 - len(L)), st)
    for o in (job, pr):
        o.add_dirichlet([1] * nd + [0] * (6 - nd), [0] * 6, clamp)
    assert job.make_dof() == pr.make_dof() == (ne // 2) * nd
    for o in (job, pr):
        o.add_neumann(val, load)
    assert np.abs(job.fext() - pr.fext).max() <= 1e-15
    rho = np.array([0.7])
    job.set_design(rho)
    ud = pr.direct(pr.assemble(rho))
    for form in forms:
        u, info = job.solve(form, tol=1e-12, maxit=1000)
        assert info["converged"], form
        assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-9, form
        c = job.compliance()
        assert abs(c - ud @ pr.fext) <= 1e-9 * abs(ud @ pr.fext), form


@pytest.mark.parametrize("mesh", [False, True])
def test_no_free_dof_and_zero_load(ctx, mesh):
    """Every node clamped: no free DOF, every form returns at once with an empty solution
    and zero compliance and sensitivity.  Then a free job with a zero load: x = 0 at once."""
    div, L = (4, 2, 3), (4.0, 2.0, 3.0)
    if mesh:
        x, conn = box_mesh(8, div, L)
        mk = lambda: icx.Job.from_mesh(ctx, 8, x, conn, icx.SOLID3D)  # noqa: E731
        forms = (icx.SOLVE_PLAIN, icx.SOLVE_MIXED)
    else:
        mk = lambda: icx.Job(ctx, 8, div, L, icx.SOLID3D)  # noqa: E731
        forms = FORMS
    job = mk()
    job.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, L[0], 0, L[1], 0, L[2]])
    assert job.make_dof() == 0
    job.set_design(np.full(job.nelx, 0.5))
    for form in forms:
        u, info = job.solve(form, tol=1e-10, maxit=100)
        assert info["converged"] and u.size == 0 and info["iterations"] == 0, form
        assert job.compliance() == 0.0
        dc, dv = job.sensitivity()
        assert not dc.any() and np.all(dv > 0)
    free = mk()
    free.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, L[1], 0, L[2]])
    n = free.make_dof()
    assert n > 0
    free.set_fext(np.zeros(n))
    free.set_design(np.full(free.nelx, 0.5))
    for form in forms:
        u, info = free.solve(form, tol=1e-10, maxit=100)
        assert info["converged"] and info["iterations"] == 0 and not u.any(), form


def test_size_limits(ctx):
    """Refused before any allocation: more than 2^27 nodes (node vectors are addressed
    with 32-bit byte offsets) and non-positive divisions."""
    with pytest.raises(icx.IcxError) as e:
        icx.Job(ctx, 8, (512, 512, 512), (512.0, 512.0, 512.0), icx.SOLID3D)  # 513^3 nodes
    assert e.value.code == 1
    with pytest.raises(icx.IcxError) as e:
        icx.Job(ctx, 4, (0, 4), (1.0, 4.0), icx.PLANE_STRESS)
    assert e.value.code == 1
