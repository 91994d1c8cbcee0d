"""GPU parity of libicx against the CPU oracle on the same seeded inputs.

Bars: integer set-up bit-exact; FP64 operator and vector kernels to rounding
(1e-13 relative); solutions to 1e-10 relative against the oracle's direct solve;
sensitivities to 1e-8 relative (scaled by max |dc|)."""
import numpy as np
import pytest

import icx
import oracle as O

pytestmark = pytest.mark.gpu


def _pair(ctx, ne, div, length, state, bcs, young=(1e-9, 1.0)):
    job = icx.Job(ctx, ne, div, length, state, young)
    pr = O.Problem(ne, div, length, state, young)
    for f, v, r in bcs:
        job.add_dirichlet(f, v, r)
        pr.add_dirichlet(f, v, r)
    assert job.make_dof() == pr.make_dof()
    return job, pr


BC2 = [([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 5, 0, 0]),
       ([0, 1, 0, 0, 0, 0], [0] * 6, [7, 7, 0, 2, 0, 0])]
BC3 = [([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, 3, 0, 4]),
       ([0, 0, 1, 0, 0, 0], [0] * 6, [0, 5, 0, 3, 0, 0])]


def test_mesh_dof_map_bit_exact(ctx):
    for ne, div, L, st, bcs in [(4, (7, 5, 1), (7.0, 5.0, 1.0), O.PLANE_STRAIN, BC2),
                                (8, (5, 3, 4), (5.0, 3.0, 4.0), O.SOLID3D, BC3)]:
        job, pr = _pair(ctx, ne, div, L, st, bcs)
        x, conn = job.mesh()
        assert np.array_equal(x, pr.x) and np.array_equal(conn, pr.conn)
        assert np.array_equal(job.dof_map(), pr.dof)


def test_neumann_tractions(ctx):
    job, pr = _pair(ctx, 8, (6, 3, 4), (6.0, 3.0, 4.0), O.SOLID3D, BC3)
    for val, rng in [([0.0, 0.0, -1.0], [6, 6, 0, 3, 0, 4]), ([0.3, 0.1, 0.0], [4, 6, 0, 3, 0, 1]),
                     ([0.0, -2.0, 0.0], [0, 6, 3, 3, 2, 4])]:
        job.add_neumann(val, rng)
        pr.add_neumann(val, rng)
    assert np.abs(job.fext() - pr.fext).max() <= 1e-15
    j2, p2 = _pair(ctx, 4, (7, 5, 1), (7.0, 5.0, 1.0), O.PLANE_STRESS, BC2)
    j2.add_neumann([0.0, -1.0, 0.0], [7, 7, 0, 5, 0, 0])
    p2.add_neumann([0.0, -1.0, 0.0], [7, 7, 0, 5, 0, 0])
    j2.add_neumann([0.5, 0.0, 0.0], [3, 7, 0, 1, 0, 0])
    p2.add_neumann([0.5, 0.0, 0.0], [3, 7, 0, 1, 0, 0])
    assert np.abs(j2.fext() - p2.fext).max() <= 1e-15


CASES = [
    (4, (9, 5, 1), (9.0, 5.0, 1.0), O.PLANE_STRESS, BC2),
    (4, (7, 5, 1), (10.5, 3.0, 1.0), O.PLANE_STRAIN, BC2),
    (8, (5, 3, 4), (5.0, 3.0, 4.0), O.SOLID3D, BC3),
    (8, (17, 19, 9), (17.0, 38.0, 4.5), O.SOLID3D, BC3),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_matvec_fp64_and_fp32(ctx, case):
    ne, div, L, st, bcs = CASES[case]
    job, pr = _pair(ctx, ne, div, L, st, bcs)
    rho = O.random_density(pr.nelx, 100 + case)
    job.set_design(rho)
    pr.assemble(rho)
    xv = np.random.default_rng(case).standard_normal(pr.numeq)
    y_ref = pr.spmv(xv)
    y64 = job.matvec(xv, prec=0)
    scale = np.abs(y_ref).max()
    assert np.abs(y64 - y_ref).max() <= 1e-13 * scale
    y32 = job.matvec(xv, prec=1)
    assert np.abs(y32 - y_ref).max() <= 2e-5 * scale
    d = job.diagonal()
    K = pr.csr()
    assert np.abs(d - K.diagonal()).max() <= 1e-14 * np.abs(K.diagonal()).max()


@pytest.mark.parametrize("div,depth", [((5, 3, 4), 0), ((33, 17, 9), 0), ((33, 17, 9), 3), ((16, 31, 20), 2)])
def test_matvec_fp32_on_matrix_cores(ctx, div, depth):
    """The dense-K0 MFMA matvec (ICX_OPT_MFMA_MATVEC) against the oracle's CSR product
    and the Haar VALU kernel: one and several 15x15 node tiles, several z bricks,
    constrained nodes on two faces."""
    L = tuple(float(d) for d in div)
    job, pr = _pair(ctx, 8, div, L, O.SOLID3D,
                    [([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, div[1], 0, div[2]]),
                     ([0, 0, 1, 0, 0, 0], [0] * 6, [0, div[0], 0, div[1], 0, 0])])
    rho = O.random_density(pr.nelx, 7 + div[0])
    job.set_design(rho)
    pr.assemble(rho)
    xv = np.random.default_rng(div[0] + depth).standard_normal(pr.numeq)
    y_ref = pr.spmv(xv)
    scale = np.abs(y_ref).max()
    ctx.set_option(icx.OPT_TILE_DEPTH, depth)
    try:
        y_valu = job.matvec(xv, prec=1)
        ctx.set_option(icx.OPT_MFMA_MATVEC, 1)
        y_mfma = job.matvec(xv, prec=1)
        y_mfma_na = job.matvec_noalloc(xv, prec=1)
    finally:
        ctx.set_option(icx.OPT_MFMA_MATVEC, 0)
        ctx.set_option(icx.OPT_TILE_DEPTH, 0)
    assert np.abs(y_mfma - y_ref).max() <= 2e-5 * scale
    assert np.abs(y_mfma - y_valu).max() <= 2e-5 * scale
    assert np.array_equal(y_mfma, y_mfma_na)
    # constrained DOFs are not in the numeq vector; a free DOF next to the clamped face
    # and one on the loaded far face are nonzero
    assert np.count_nonzero(y_mfma) > 0.9 * pr.numeq


def _solve_case(ctx, nx=12, ny=6, nz=6, seed=5):
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    pr = O.cantilever_3d(nx, ny, nz)
    rho = O.random_density(pr.nelx, seed)
    job.set_design(rho)
    ud = pr.direct(pr.assemble(rho))
    return job, pr, rho, ud


@pytest.mark.parametrize("tol", [1e-10, 1e-13])
@pytest.mark.parametrize("dims", [(12, 6, 6, 5), (20, 10, 10, 7), (17, 19, 9, 3)])
def test_plain_form_follows_eigen_loop(ctx, dims, tol):
    """The plain form is Eigen's loop in FP64; only the summation order of the dot
    products and of the operator differs from the oracle's, so the iteration counts agree
    within one and the iterates within rounding (measured: 0-1 iterations apart,
    |du|/|u| <= 8e-14)."""
    nx, ny, nz, seed = dims
    job, pr, rho, ud = _solve_case(ctx, nx, ny, nz, seed)
    u, info = job.solve(icx.SOLVE_PLAIN, tol=tol, maxit=100000)
    uo, ito, erro = pr.cg(pr.assemble(rho), tol=tol, maxit=100000)
    assert info["converged"]
    assert abs(info["iterations"] - ito) <= 1
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-12


@pytest.mark.parametrize("form", [icx.SOLVE_PLAIN, icx.SOLVE_FUSED])
def test_fp64_forms_match_direct_solution(ctx, form):
    job, pr, rho, ud = _solve_case(ctx)
    u, info = job.solve(form, tol=1e-13, maxit=20000)
    assert info["converged"]
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-10
    c_ref = ud @ pr.fext
    assert abs(job.compliance() - c_ref) / c_ref < 1e-10


@pytest.mark.parametrize("dims,depth", [((33, 17, 9, 11), 0), ((33, 17, 9, 11), 3), ((16, 31, 20, 12), 0)])
def test_fused_form_on_several_node_tiles(ctx, dims, depth):
    """The FP64 fused form (node-interleaved vectors, 15 x 7 node tiles) on grids of
    several tiles in x and y and several z bricks: the same PCG as the plain form up to
    rounding (iteration counts within one), the direct solution to 1e-10 and the
    compliance to 1e-10."""
    nx, ny, nz, seed = dims
    job, pr, rho, ud = _solve_case(ctx, nx, ny, nz, seed)
    ctx.set_option(icx.OPT_TILE_DEPTH, depth)
    try:
        u, info = job.solve(icx.SOLVE_FUSED, tol=1e-13, maxit=100000)
        up, pinfo = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=100000)
    finally:
        ctx.set_option(icx.OPT_TILE_DEPTH, 0)
    assert info["converged"] and pinfo["converged"]
    assert abs(info["iterations"] - pinfo["iterations"]) <= 1
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-10
    assert np.linalg.norm(u - up) / np.linalg.norm(up) < 1e-11
    job.solve(icx.SOLVE_FUSED, tol=1e-13, maxit=100000)
    c_ref = ud @ pr.fext
    assert abs(job.compliance() - c_ref) / c_ref < 1e-10


def test_mixed_form_meets_true_residual(ctx):
    job, pr, rho, ud = _solve_case(ctx)
    u, info = job.solve(icx.SOLVE_MIXED, tol=1e-10, maxit=20000)
    assert info["converged"] and info["refinements"] >= 1
    K = pr.csr()
    b = pr.assemble(rho)
    assert np.linalg.norm(b - K @ u) / np.linalg.norm(b) < 1e-10
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-7


@pytest.mark.parametrize("depth", [0, 1, 2, 3, 5, 40])
def test_mixed_form_every_tile_depth(ctx, depth):
    """The FP32 kernel A tiles the element layers along z (seam split: each seam layer
    is completed by kernel B).  Depth 1 makes every interior node layer a seam, 3 and 5
    leave a shorter top tile on 7 element layers, 40 a single tile; every depth meets the
    true residual and the direct solution."""
    job, pr, rho, ud = _solve_case(ctx, nx=9, ny=5, nz=7, seed=11)
    ctx.set_option(icx.OPT_TILE_DEPTH, depth)
    try:
        u, info = job.solve(icx.SOLVE_MIXED, tol=1e-10, maxit=20000)
    finally:
        ctx.set_option(icx.OPT_TILE_DEPTH, 0)
    assert info["converged"]
    b = pr.assemble(rho)
    assert np.linalg.norm(b - pr.csr() @ u) / np.linalg.norm(b) < 1e-10
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-7


@pytest.mark.parametrize("depth", [0, 1, 3])
@pytest.mark.parametrize("dims", [(9, 5, 7), (20, 9, 8)])
def test_single_reduction_matches_two_kernel_form(ctx, depth, dims):
    """Switched on, the hexa8 mixed form runs one kernel per iteration (the next r.z from
    an expansion reduced with the previous iteration's dot products); the iterates are
    those of the two-kernel iteration up to rounding.  Both meet the true residual
    after several replacements, agree with the direct solution and take nearly the same
    number of iterations, at every tile depth."""
    job, pr, rho, ud = _solve_case(ctx, *dims, seed=13)
    ctx.set_option(icx.OPT_TILE_DEPTH, depth)
    try:
        u2, i2 = job.solve(icx.SOLVE_MIXED, tol=1e-10, maxit=20000)
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 1)
        u1, i1 = job.solve(icx.SOLVE_MIXED, tol=1e-10, maxit=20000)
    finally:
        ctx.set_option(icx.OPT_TILE_DEPTH, 0)
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 0)
    b = pr.assemble(rho)
    K = pr.csr()
    for u, info in ((u1, i1), (u2, i2)):
        assert info["converged"] and info["refinements"] >= 2
        assert np.linalg.norm(b - K @ u) / np.linalg.norm(b) < 1e-10
        assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-7
    assert abs(i1["iterations"] - i2["iterations"]) <= max(3, int(0.05 * i2["iterations"]))


def test_single_reduction_iteration_limit_and_zero_load(ctx):
    job, pr, rho, ud = _solve_case(ctx, 10, 6, 5, seed=17)
    ctx.set_option(icx.OPT_SINGLE_REDUCTION, 1)
    try:
        _single_reduction_limits(job, pr, rho, ud)
    finally:
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 0)


def _single_reduction_limits(job, pr, rho, ud):
    u, info = job.solve(icx.SOLVE_MIXED, tol=1e-12, maxit=9)
    assert not info["converged"] and info["iterations"] == 9
    K = pr.csr()
    # CG from x = 0 lowers the energy norm of the error at every step
    e = u - ud
    assert np.isfinite(u).all() and e @ (K @ e) < 0.5 * (ud @ (K @ ud))
    job.set_fext(np.zeros(job.numeq))
    u, info = job.solve(icx.SOLVE_MIXED, tol=1e-8)
    assert info["converged"] and info["iterations"] == 0 and not np.any(u)


def test_2d_solve_and_node_values(ctx):
    job = icx.cantilever_2d(ctx, 40, 14)
    pr = O.cantilever_2d(40, 14)
    rho = O.random_density(pr.nelx, 9)
    job.set_design(rho)
    ud = pr.direct(pr.assemble(rho))
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-10
    nv = job.node_values()
    assert np.abs(nv - pr.node_values(u)).max() == 0.0


def test_sensitivity(ctx):
    for nx, ny, nz in [(12, 6, 6), (7, 3, 5)]:
        job, pr, rho, ud = _solve_case(ctx, nx, ny, nz, seed=nx)
        u, _ = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
        dc, dv = job.sensitivity()
        dco, dvo = pr.sensitivity(rho, pr.element_energy(pr.node_values(ud)), pr.volumes())
        assert np.abs(dc - dco).max() <= 1e-8 * np.abs(dco).max()
        # element volume: exact 1.0 here, 1 - 1.1e-16 from the reference-style quadrature
        assert np.abs(dv - dvo).max() <= 1e-15
        dc2, dv2 = job.sensitivity(vscale=7.5)
        assert np.array_equal(dc2, dc) and np.abs(dv2 - dvo / 7.5).max() <= 2.3e-16
    j2 = icx.cantilever_2d(ctx, 30, 10)
    p2 = O.cantilever_2d(30, 10)
    rho = O.random_density(p2.nelx, 2)
    j2.set_design(rho)
    ud = p2.direct(p2.assemble(rho))
    j2.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    dc, _ = j2.sensitivity()
    dco, _ = p2.sensitivity(rho, p2.element_energy(p2.node_values(ud)), p2.volumes())
    assert np.abs(dc - dco).max() <= 1e-8 * np.abs(dco).max()


@pytest.mark.parametrize("ne,div,r", [(4, (30, 11, 1), 1.5), (4, (20, 9, 1), 3.0),
                                      (8, (9, 7, 6), 1.5), (8, (8, 6, 7), 2.3)])
def test_cone_filter_forward_adjoint(ctx, ne, div, r):
    st = O.PLANE_STRESS if ne == 4 else O.SOLID3D
    L = tuple(float(d) for d in div)
    job = icx.Job(ctx, ne, div, L, st)
    pr = O.Problem(ne, div, L, st)
    job.cone_setup(r)
    f = pr.cone_filter(r)
    x = O.random_density(pr.nelx, 4)
    s = np.random.default_rng(2).standard_normal(pr.nelx)
    assert np.abs(job.cone(x) - f.apply(x)).max() <= 1e-15
    assert np.abs(job.cone(s, adjoint=True) - f.apply(s, sens=True)).max() <= 1e-14


def test_sensitivity_filter(ctx):
    """Filter::sensitivityFilter (filter.hpp:72-100) on box jobs and a distorted mesh job
    against the oracle."""
    from meshgen import box_mesh, distort
    rng = np.random.default_rng(12)
    for ne, div, r in [(4, (20, 9, 1), 1.5), (8, (7, 6, 5), 1.8)]:
        st = O.PLANE_STRESS if ne == 4 else O.SOLID3D
        L = tuple(float(d) for d in div)
        job = icx.Job(ctx, ne, div, L, st)
        pr = O.Problem(ne, div, L, st)
        job.cone_setup(r)
        x, df = O.random_density(pr.nelx, 5), -rng.random(pr.nelx)
        ref = pr.cone_filter(r).sensitivity(x, df, pr.volumes())
        assert np.abs(job.sensitivity_filter(x, df) - ref).max() <= 1e-14 * np.abs(ref).max()
        job.close()
    x0, conn = box_mesh(8, (6, 5, 4), (6.0, 5.0, 4.0))
    xd = distort(x0, conn, 8, 0.2, 3, (6.0, 5.0, 4.0))
    job = icx.Job.from_mesh(ctx, 8, xd, conn, icx.SOLID3D)
    pr = O.Problem.from_mesh(8, xd, conn, O.SOLID3D)
    job.cone_setup(1.7)
    x, df = O.random_density(pr.nelx, 6), -rng.random(pr.nelx)
    ref = pr.cone_filter(1.7).sensitivity(x, df, pr.volumes())
    assert np.abs(job.sensitivity_filter(x, df) - ref).max() <= 1e-13 * np.abs(ref).max()
    job.close()


def test_projection(ctx):
    job = icx.Job(ctx, 8, (5, 4, 3), (5.0, 4.0, 3.0), O.SOLID3D)
    x = np.linspace(0, 1, job.nelx)
    for beta in (1.0, 4.0, 16.0):
        h, dh = job.project(x, beta)
        assert np.abs(h - O.threshold(x, beta)).max() <= 1e-15
        assert np.abs(dh - O.threshold_derivative(x, beta)).max() <= 1e-14 * beta


@pytest.mark.parametrize("ne,div,R,depth", [(4, (24, 10, 1), 2.0, 0), (8, (9, 6, 5), 2.0, 0), (8, (10, 7, 6), 3.0, 0),
                                            (8, (20, 17, 11), 2.0, 0), (8, (20, 17, 11), 2.0, 4),
                                            (8, (10, 7, 6), 3.0, 8)])
def test_helmholtz_filter(ctx, ne, div, R, depth):
    """Helmholtz filter against the oracle.  hexa8 at tile depth 0 takes the shared-element
    tile kernel (20 x 17 x 11: several 15 x 15 node tiles and 8-layer z chunks); depths
    4 and 8 take the per-node gather."""
    st = O.PLANE_STRESS if ne == 4 else O.SOLID3D
    L = tuple(float(d) for d in div)
    ctx.set_option(icx.OPT_TILE_DEPTH, depth)
    try:
        _helmholtz_case(ctx, ne, div, R, st, L)
    finally:
        ctx.set_option(icx.OPT_TILE_DEPTH, 0)


def _helmholtz_case(ctx, ne, div, R, st, L):
    job = icx.Job(ctx, ne, div, L, st)
    pr = O.Problem(ne, div, L, st)
    job.helmholtz_setup(R)
    H = pr.helmholtz(R)
    x = O.random_density(pr.nelx, 8)
    for sens in (False, True):
        g, it = job.helmholtz(x, sens=sens, tol=1e-13)
        o, _ = H.apply(x, sens=sens, tol=1e-13)
        assert np.abs(g - o).max() <= 1e-10 * np.abs(o).max()
    # reference tolerance 1e-10: same iteration count as the oracle's Eigen loop (+-2)
    g, it = job.helmholtz(x, tol=1e-10)
    o, ito = H.apply(x, tol=1e-10)
    assert abs(it - ito) <= 2


def test_helmholtz_new_radius_on_one_job(ctx):
    """A second radius on the same job: the filter solves replayed from HIP graphs must
    take the new operator (the captured launches of the first radius are dropped)."""
    ne, div = 8, (10, 6, 5)
    L = tuple(float(d) for d in div)
    job = icx.Job(ctx, ne, div, L, O.SOLID3D)
    pr = O.Problem(ne, div, L, O.SOLID3D)
    x = O.random_density(pr.nelx, 12)
    for R in (1.5, 3.0, 1.5):
        job.helmholtz_setup(R)
        g, _ = job.helmholtz(x, tol=1e-13)
        o, _ = pr.helmholtz(R).apply(x, tol=1e-13)
        assert np.abs(g - o).max() <= 1e-10 * np.abs(o).max(), R


def test_oc_update_matches(ctx):
    job = icx.Job(ctx, 8, (12, 6, 6), (12.0, 6.0, 6.0), O.SOLID3D)
    pr = O.Problem(8, (12, 6, 6), (12.0, 6.0, 6.0), O.SOLID3D)
    job.cone_setup(1.5)
    f = pr.cone_filter(1.5)
    rng = np.random.default_rng(6)
    x = rng.uniform(0.2, 0.8, pr.nelx)
    dc = -rng.uniform(0.01, 1.0, pr.nelx)
    dv = np.ones(pr.nelx)
    beta = 2.0
    V0 = O.threshold(f.apply(np.full(pr.nelx, 0.5)), beta).sum()
    xg, nb = job.oc_update(x, dc, V0, dv, filter=icx.FILTER_CONE, beta=beta)
    xo, no = O.oc_update(x, dc, V0, dv, np.full(pr.nelx, 1e-4), np.ones(pr.nelx),
                         lambda z: O.threshold(f.apply(z), beta).sum())
    assert nb == no
    assert np.abs(xg - xo).max() <= 1e-13


def test_mma_updates_match(ctx):
    n = 300
    job = icx.Job(ctx, 8, (10, 6, 5), (10.0, 6.0, 5.0), O.SOLID3D)
    assert job.nelx == n
    mma = O.MMA(n, 1)
    rng = np.random.default_rng(7)
    xg = rng.uniform(0.2, 0.6, n)
    xo = xg.copy()
    c = rng.uniform(0.5, 2.0, n)
    for it in range(4):
        df = -c / xo ** 2
        g = xo.sum() / (0.4 * n) - 1.0
        dg = np.full(n, 1.0 / (0.4 * n))
        xg, sg = job.mma_update(xg, df, g, dg, 1e-4, 1.0)
        xo, so = mma.update(xo, df, g, dg, np.full(n, 1e-4), np.ones(n))
        assert np.abs(xg - xo).max() <= 1e-9, it
        xg = xo.copy()   # teacher forcing: continue both from the oracle's iterate


def test_mma_dual_on_device_equals_host_loop(ctx):
    """The dual's Newton loop run on the device (one read of its state per 16 steps)
    against the same loop run on the host (one read per step): the same Newton step
    count and the same bits, for MMA updates and for driver B's device design loop."""
    n = 300
    rng = np.random.default_rng(17)
    c = rng.uniform(0.5, 2.0, n)
    runs = {}
    for host in (1, 0):
        ctx.set_option(icx.OPT_MMA_HOST_DUAL, host)
        try:
            job = icx.Job(ctx, 8, (10, 6, 5), (10.0, 6.0, 5.0), O.SOLID3D)
            x = np.random.default_rng(7).uniform(0.2, 0.6, n)
            hist = []
            for it in range(5):
                df = -c / x ** 2
                g = x.sum() / (0.4 * n) - 1.0
                x, steps = job.mma_update(x, df, g, np.full(n, 1.0 / (0.4 * n)), 1e-4, 1.0)
                hist.append((x.copy(), steps))
            drv = icx.cantilever_2d(ctx, 30, 10)
            drv.opt_init(np.full(drv.nelx, 0.4), driver=1, filter=icx.FILTER_HELMHOLTZ, radius=2.0,
                         form=icx.SOLVE_PLAIN, tol=1e-12, maxit=50000, helm_tol=1e-12)
            objs = [drv.opt_step()["objective"] for _ in range(4)]
            runs[host] = (hist, objs, drv.opt_design())
        finally:
            ctx.set_option(icx.OPT_MMA_HOST_DUAL, 0)
    for (xh, sh), (xd, sd) in zip(runs[1][0], runs[0][0]):
        assert sh == sd and sh > 0
        assert np.array_equal(xh, xd)
    assert runs[1][1] == runs[0][1]
    assert np.array_equal(runs[1][2], runs[0][2])


@pytest.mark.parametrize("name", ["q4_cantilever_24x8.npz", "hex8_cantilever_10x4x4.npz"])
def test_gpu_matches_golden(ctx, name):
    import os
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", name), allow_pickle=False))
    div = [int(v) for v in g["div"]]
    job = icx.cantilever_2d(ctx, *div) if len(div) == 2 else icx.cantilever_3d(ctx, *div)
    job.set_design(g["rho"])
    u, info = job.solve(icx.SOLVE_PLAIN, tol=1e-13, maxit=20000)
    assert np.linalg.norm(u - g["u"]) / np.linalg.norm(g["u"]) < 1e-10
    assert abs(job.compliance() - g["c"]) / g["c"] < 1e-10
    dc, _ = job.sensitivity()
    assert np.abs(dc - g["dc"]).max() <= 1e-8 * np.abs(g["dc"]).max()
    job.cone_setup(1.5)
    assert np.abs(job.cone(g["rho"]) - g["cone"]).max() < 1e-14
    job.helmholtz_setup(2.0)
    h, _ = job.helmholtz(g["rho"], tol=1e-13)
    assert np.abs(h - g["helm"]).max() < 1e-10


@pytest.mark.parametrize("dims", [(24, 10), (10, 5, 5)])
def test_solve_forms_in_any_order_on_one_job(ctx, dims):
    """Plain, fused and mixed solves interleaved on one job (each form's captured graphs
    and polling state reused across the others), a new design before every solve; each
    result against the oracle's direct solve."""
    if len(dims) == 2:
        job, pr = icx.cantilever_2d(ctx, *dims), O.cantilever_2d(*dims)
    else:
        job, pr = icx.cantilever_3d(ctx, *dims), O.cantilever_3d(*dims)
    rng = np.random.default_rng(11)
    order = [icx.SOLVE_FUSED, icx.SOLVE_PLAIN, icx.SOLVE_FUSED, icx.SOLVE_MIXED, icx.SOLVE_PLAIN,
             icx.SOLVE_PLAIN, icx.SOLVE_FUSED, icx.SOLVE_MIXED, icx.SOLVE_FUSED]
    for k, form in enumerate(order):
        rho = rng.uniform(0.2, 1.0, pr.nelx)
        job.set_design(rho)
        u_ref = pr.direct(pr.assemble(rho))
        tol = 1e-10 if form == icx.SOLVE_MIXED else 1e-12
        u, info = job.solve(form, tol=tol, maxit=50000)
        assert info["converged"], (k, form)
        err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
        assert err < (1e-7 if form == icx.SOLVE_MIXED else 1e-9), (k, form, err)
