"""Python front end of libicx (ctypes over the C ABI in include/icx.h).

The product path: every call goes to the HIP library built for gfx950.  If the
library is missing, importing this package still works but `lib()` raises --
there is no CPU fallback.
"""
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "_build", "libicx.so"))

QUAD4, HEXA8 = 4, 8
PLANE_STRESS, PLANE_STRAIN, SOLID3D, HEAT = 0, 1, 2, 3
SOLVE_PLAIN, SOLVE_FUSED, SOLVE_MIXED = 0, 1, 2
FILTER_NONE, FILTER_CONE, FILTER_HELMHOLTZ = 0, 1, 2
ERR_NOCONV = 5

EXPORTS = [
    "icx_ctx_create", "icx_ctx_destroy", "icx_last_error", "icx_ctx_synchronize",
    "icx_job_create", "icx_job_create_toml", "icx_job_create_mesh", "icx_mdpa_read", "icx_job_destroy", "icx_job_sizes", "icx_job_get_problem", "icx_get_mesh",
    "icx_add_dirichlet", "icx_make_dof", "icx_get_dof_map", "icx_add_neumann",
    "icx_set_fext", "icx_get_fext", "icx_add_nodal_load",
    "icx_set_design", "icx_set_conductivity", "icx_matvec", "icx_matvec_noalloc", "icx_diagonal", "icx_solve", "icx_get_displacement",
    "icx_get_node_values", "icx_compliance", "icx_sensitivity", "icx_element_volume",
    "icx_mises", "icx_reactions", "icx_output_begin", "icx_output_write", "icx_write_vtk",
    "icx_cone_filter_setup", "icx_cone_filter", "icx_sensitivity_filter", "icx_helmholtz_setup", "icx_helmholtz_filter",
    "icx_project", "icx_oc_update", "icx_mma_reset", "icx_mma_update",
    "icx_gcmma_reset", "icx_gcmma_outer_update", "icx_gcmma_inner_update", "icx_gcmma_concheck",
    "icx_mma_set_asymptotes", "icx_stress_constraint",
    "icx_opt_init", "icx_opt_step", "icx_opt_step_begin", "icx_opt_step_end", "icx_opt_get_design",
    "icx_submit", "icx_submit_form", "icx_poll", "icx_job_last_info", "icx_cancel", "icx_solve_batch", "icx_job_footprint",
    "icx_profile_begin", "icx_profile_end", "icx_set_option",
]
OPT_GRAPHS, OPT_TILE_DEPTH, OPT_PIPELINE, OPT_SINGLE_REDUCTION, OPT_HEAT_DENSE, OPT_MMA_HOST_DUAL = 1, 2, 3, 4, 5, 6
OPT_MFMA_MATVEC = 7
PROF_MATVEC64, PROF_FUSED_A, PROF_FUSED_B, PROF_MATVEC32 = 1, 2, 3, 4


class Problem(C.Structure):
    _fields_ = [("eltype", C.c_int), ("div", C.c_int * 3), ("length", C.c_double * 3),
                ("state", C.c_int), ("young1", C.c_double), ("young2", C.c_double),
                ("poisson", C.c_double), ("penalty", C.c_double)]


class MeshDesc(C.Structure):
    _fields_ = [("eltype", C.c_int), ("numnp", C.c_int), ("nelx", C.c_int),
                ("x", C.POINTER(C.c_double)), ("conn", C.POINTER(C.c_int)), ("state", C.c_int),
                ("young1", C.c_double), ("young2", C.c_double), ("poisson", C.c_double),
                ("penalty", C.c_double)]


def read_mdpa(path, eltype):
    """Nodes (numnp x 3, flat) and 0-based connectivity (flat) of an MDPA file's
    elements of one type (MeshMDPA, src/io/mesh_mdpa.hpp)."""
    n, m = C.c_int(), C.c_int()
    p = str(path).encode()
    rc = lib().icx_mdpa_read(p, int(eltype), C.byref(n), C.byref(m), None, None)
    if rc:
        raise IcxError(rc, "cannot read %s as an MDPA mesh of element type %d" % (path, eltype))
    x = np.zeros(3 * n.value)
    conn = np.zeros(m.value * int(eltype), dtype=np.int32)
    rc = lib().icx_mdpa_read(p, int(eltype), C.byref(n), C.byref(m), _dp(x), _ip(conn))
    if rc:
        raise IcxError(rc, "cannot read %s" % path)
    return x, conn


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("matvecs", C.c_int), ("fp64_matvecs", C.c_int),
                ("refinements", C.c_int), ("relres", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OptParams(C.Structure):
    _fields_ = [("driver", C.c_int), ("filter", C.c_int), ("filter_radius", C.c_double),
                ("projection", C.c_int), ("form", C.c_int), ("tol", C.c_double),
                ("maxit", C.c_int), ("helm_tol", C.c_double), ("xmin", C.c_double),
                ("xmax", C.c_double), ("move", C.c_double)]


class IterRecord(C.Structure):
    _fields_ = [("step", C.c_int), ("objective", C.c_double), ("constraint", C.c_double),
                ("volume", C.c_double), ("cg_iterations", C.c_int), ("update_steps", C.c_int),
                ("seconds", C.c_double), ("inner_loops", C.c_int),
                ("trial_cg_iterations", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libicx.so (raises if it has not been built: no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libicx.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.icx_last_error.restype = C.c_char_p
        L.icx_last_error.argtypes = [C.c_void_p]
        _lib = L
    return _lib


class IcxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("icx error %d: %s" % (code, msg))
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        self.device = device
        self.jobs = weakref.WeakSet()
        self.options = {}  # values set through set_option
        rc = lib().icx_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise IcxError(rc, "icx_ctx_create failed (no GPU?)")

    def check(self, rc, allow=()):
        if rc != 0 and rc not in allow:
            raise IcxError(rc, lib().icx_last_error(self.h).decode())
        return rc

    def synchronize(self):
        self.check(lib().icx_ctx_synchronize(self.h))

    def set_option(self, option, value):
        self.check(lib().icx_set_option(self.h, int(option), C.c_longlong(int(value))))
        self.options[int(option)] = int(value)

    def profile_begin(self, kernel):
        self.check(lib().icx_profile_begin(self.h, int(kernel)))

    def poll(self, chunks=1 << 20, cap=4096):
        """Advance every submitted solve (icx_poll); returns the jobs that finished."""
        arr = (C.c_void_p * cap)()
        n = C.c_int()
        self.check(lib().icx_poll(self.h, int(chunks), arr, int(cap), C.byref(n)))
        by_handle = {j.h.value: j for j in list(self.jobs) if j.h}
        return [by_handle[arr[i]] for i in range(n.value)]

    def solve_batch(self, jobs, form=None, tol=1e-8, maxit=50000):
        """icx_solve_batch over jobs; returns one info dict per job."""
        form = SOLVE_MIXED if form is None else form
        arr = (C.c_void_p * max(len(jobs), 1))(*[j.h.value for j in jobs])
        infos = (SolveInfo * max(len(jobs), 1))()
        rc = lib().icx_solve_batch(self.h, arr, len(jobs), int(form), C.c_double(tol), int(maxit), infos)
        self.check(rc, allow=(ERR_NOCONV,))
        return [infos[i].as_dict() for i in range(len(jobs))]

    def profile_end(self):
        s, n = C.c_double(), C.c_longlong()
        self.check(lib().icx_profile_end(self.h, C.byref(s), C.byref(n)))
        return s.value, n.value

    def close(self):
        """Destroys the context and, with it, every job still open on it."""
        if self.h:
            for j in list(self.jobs):
                j.h = C.c_void_p()
            lib().icx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """One structured box problem resident on the device."""

    def __init__(self, ctx, eltype, div, length, state, young=(1e-9, 1.0), nu=0.3, penalty=3.0):
        self.ctx = ctx
        p = Problem()
        p.eltype = eltype
        d = list(div) + [1] * (3 - len(div))
        L = list(length) + [1.0] * (3 - len(length))
        for i in range(3):
            p.div[i] = int(d[i])
            p.length[i] = float(L[i])
        p.state, p.young1, p.young2 = state, float(young[0]), float(young[1])
        p.poisson, p.penalty = float(nu), float(penalty)
        self.h = C.c_void_p()
        ctx.check(lib().icx_job_create(ctx.h, C.byref(p), C.byref(self.h)))
        ctx.jobs.add(self)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().icx_job_sizes(self.h, C.byref(a), C.byref(b), C.byref(c))
        self.numnp, self.nelx, self.neq = a.value, b.value, c.value
        self.ndim = 2 if eltype == QUAD4 else 3
        self.dofnp = 1 if state == HEAT else self.ndim
        self.div = d
        self.numeq = None

    @classmethod
    def from_mesh(cls, ctx, eltype, x, conn, state, young=(1e-9, 1.0), nu=0.3, penalty=3.0):
        """A job on an unstructured mesh (icx_job_create_mesh): x numnp x 3, conn nelx x ne
        (0-based, the reference's local node order)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1)
        m = MeshDesc()
        m.eltype, m.numnp, m.nelx = int(eltype), len(x) // 3, len(conn) // int(eltype)
        m.x, m.conn = _dp(x), _ip(conn)
        m.state, m.young1, m.young2 = state, float(young[0]), float(young[1])
        m.poisson, m.penalty = float(nu), float(penalty)
        self.h = C.c_void_p()
        ctx.check(lib().icx_job_create_mesh(ctx.h, C.byref(m), C.byref(self.h)))
        ctx.jobs.add(self)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().icx_job_sizes(self.h, C.byref(a), C.byref(b), C.byref(c))
        self.numnp, self.nelx, self.neq = a.value, b.value, c.value
        self.ndim = 2 if eltype == QUAD4 else 3
        self.dofnp = self.ndim
        self.div = None
        self.numeq = None
        return self

    @classmethod
    def from_toml(cls, ctx, text):
        """A box job from the text of a reference job file (icx_job_create_toml): mesh,
        material, Dirichlet ranges, DOF map and tractions; `design0` holds designs_0."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = C.c_void_p()
        n, d0 = C.c_int(), C.c_double()
        ctx.check(lib().icx_job_create_toml(ctx.h, text.encode(), C.byref(self.h), C.byref(n), C.byref(d0)))
        ctx.jobs.add(self)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().icx_job_sizes(self.h, C.byref(a), C.byref(b), C.byref(c))
        self.numnp, self.nelx, self.neq = a.value, b.value, c.value
        p = Problem()
        ctx.check(lib().icx_job_get_problem(self.h, C.byref(p)))
        self.ndim = 2 if p.eltype == QUAD4 else 3
        self.dofnp = 1 if p.state == HEAT else self.ndim
        self.div = [p.div[0], p.div[1], p.div[2] if self.ndim == 3 else 1]
        self.numeq = n.value
        self.design0 = d0.value
        return self

    def close(self):
        if self.h and self.ctx.h:
            lib().icx_job_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up --------------------------------------------------------------
    def mesh(self):
        x = np.zeros(self.numnp * 3)
        conn = np.zeros(self.nelx * (4 if self.ndim == 2 else 8), dtype=np.int32)
        self.ctx.check(lib().icx_get_mesh(self.h, _dp(x), _ip(conn)))
        return x, conn

    def add_dirichlet(self, flag6, val6, range6):
        f = np.asarray(flag6, dtype=np.int32)
        v, r = _f64(val6), _f64(range6)
        self.ctx.check(lib().icx_add_dirichlet(self.h, _ip(f), _dp(v), _dp(r)))

    def make_dof(self):
        n = C.c_int()
        self.ctx.check(lib().icx_make_dof(self.h, C.byref(n)))
        self.numeq = n.value
        return self.numeq

    def dof_map(self):
        d = np.zeros(self.numnp * self.dofnp, dtype=np.int32)
        self.ctx.check(lib().icx_get_dof_map(self.h, _ip(d)))
        return d

    def add_neumann(self, val3, range6):
        v, r = _f64(val3), _f64(range6)
        self.ctx.check(lib().icx_add_neumann(self.h, _dp(v), _dp(r)))

    def set_fext(self, f):
        f = _f64(f)
        self.ctx.check(lib().icx_set_fext(self.h, _dp(f)))

    def fext(self):
        f = np.zeros(self.numeq)
        self.ctx.check(lib().icx_get_fext(self.h, _dp(f)))
        return f

    def add_nodal_load(self, node, comp, value):
        self.ctx.check(lib().icx_add_nodal_load(self.h, int(node), int(comp), C.c_double(value)))

    def node_index(self, i, j, k=0):
        d0, d1 = self.div[0], self.div[1]
        return i + (d0 + 1) * (j + (d1 + 1) * k)

    # -- analysis ---------------------------------------------------------------
    def set_design(self, x):
        x = _f64(x)
        self.ctx.check(lib().icx_set_design(self.h, _dp(x)))

    def set_conductivity(self, k1, k2):
        """heat jobs: conductivity rows (xx, yy, zz, xy, yz, xz) of materials 1 and 2"""
        self.ctx.check(lib().icx_set_conductivity(self.h, _dp(_f64(k1)), _dp(_f64(k2))))

    def matvec(self, x, prec=0):
        x = _f64(x)
        y = np.zeros(self.numeq)
        self.ctx.check(lib().icx_matvec(self.h, int(prec), _dp(x), _dp(y)))
        return y

    def matvec_noalloc(self, x, prec=0):
        """icx_matvec_noalloc: the same product with the library's per-GPU scratch array
        as device staging (nothing allocated)"""
        x = _f64(x)
        y = np.zeros(self.numeq)
        self.ctx.check(lib().icx_matvec_noalloc(self.h, int(prec), _dp(x), _dp(y)))
        return y

    def diagonal(self):
        d = np.zeros(self.numeq)
        self.ctx.check(lib().icx_diagonal(self.h, _dp(d)))
        return d

    def solve(self, form=SOLVE_PLAIN, tol=1e-10, maxit=10000, want_u=True):
        u = np.zeros(self.numeq) if want_u else None
        info = SolveInfo()
        rc = lib().icx_solve(self.h, int(form), C.c_double(tol), int(maxit),
                             _dp(u) if want_u else None, C.byref(info))
        self.ctx.check(rc, allow=(ERR_NOCONV,))
        d = info.as_dict()
        d["converged"] = rc == 0
        return u, d

    def displacement(self):
        u = np.zeros(self.numeq)
        self.ctx.check(lib().icx_get_displacement(self.h, _dp(u)))
        return u

    def node_values(self):
        v = np.zeros(self.numnp * 3)
        self.ctx.check(lib().icx_get_node_values(self.h, _dp(v)))
        return v

    def compliance(self):
        c = C.c_double()
        self.ctx.check(lib().icx_compliance(self.h, C.byref(c)))
        return c.value

    def sensitivity(self, vscale=1.0):
        a, b = np.zeros(self.nelx), np.zeros(self.nelx)
        self.ctx.check(lib().icx_sensitivity(self.h, C.c_double(vscale), _dp(a), _dp(b)))
        return a, b

    def volumes(self):
        v = np.zeros(self.nelx)
        self.ctx.check(lib().icx_element_volume(self.h, _dp(v)))
        return v

    # -- output step ---------------------------------------------------------------
    def mises(self):
        """element-mean von Mises stress of the last solve (element order)"""
        out = np.zeros(self.nelx)
        self.ctx.check(lib().icx_mises(self.h, _dp(out)))
        return out

    def reactions(self):
        """K u over all DOFs, numnp x 3 (flattened), |r| < 1e-15 flushed"""
        out = np.zeros(self.numnp * 3)
        self.ctx.check(lib().icx_reactions(self.h, _dp(out)))
        return out

    def output_begin(self):
        self.ctx.check(lib().icx_output_begin(self.h))

    def output_write(self, path):
        self.ctx.check(lib().icx_output_write(self.h, os.fsencode(path)))

    def write_vtk(self, path):
        """binary legacy VTK: points, cells, displacement, density, mises"""
        self.ctx.check(lib().icx_write_vtk(self.h, os.fsencode(path)))

    # -- filters -----------------------------------------------------------------
    def cone_setup(self, radius):
        self.ctx.check(lib().icx_cone_filter_setup(self.h, C.c_double(radius)))

    def cone(self, v, adjoint=False):
        v = _f64(v)
        out = np.zeros(self.nelx)
        self.ctx.check(lib().icx_cone_filter(self.h, int(adjoint), _dp(v), _dp(out)))
        return out

    def sensitivity_filter(self, design, dfds):
        """Filter::sensitivityFilter (icx_sensitivity_filter; cone_setup first)"""
        x, df = _f64(design), _f64(dfds)
        out = np.zeros(self.nelx)
        self.ctx.check(lib().icx_sensitivity_filter(self.h, _dp(x), _dp(df), _dp(out)))
        return out

    def helmholtz_setup(self, radius):
        self.ctx.check(lib().icx_helmholtz_setup(self.h, C.c_double(radius)))

    def helmholtz(self, v, sens=False, tol=1e-10):
        v = _f64(v)
        out = np.zeros(self.nelx)
        it = C.c_int()
        self.ctx.check(lib().icx_helmholtz_filter(self.h, int(sens), C.c_double(tol), _dp(v),
                                                  _dp(out), C.byref(it)))
        return out, it.value

    def project(self, v, beta, T=0.5):
        v = _f64(v)
        out, dout = np.zeros(self.nelx), np.zeros(self.nelx)
        self.ctx.check(lib().icx_project(self.h, C.c_double(beta), C.c_double(T), _dp(v),
                                         _dp(out), _dp(dout)))
        return out, dout

    def oc_update(self, x, dfdx, g, dgdx, xmin=1e-4, xmax=1.0, move=0.15,
                  filter=FILTER_CONE, beta=1.0, T=0.5):
        x = _f64(x).copy()
        df, dg = _f64(dfdx), _f64(dgdx)
        n = C.c_int()
        self.ctx.check(lib().icx_oc_update(self.h, _dp(x), _dp(df), C.c_double(g), _dp(dg),
                                           C.c_double(xmin), C.c_double(xmax), C.c_double(move),
                                           int(filter), C.c_double(beta), C.c_double(T),
                                           C.byref(n)))
        return x, n.value

    def mma_reset(self):
        self.ctx.check(lib().icx_mma_reset(self.h))

    def mma_update(self, x, dfdx, gx, dgdx, xmin=1e-4, xmax=1.0):
        x = _f64(x).copy()
        df, dg = _f64(dfdx), _f64(dgdx)
        n = C.c_int()
        self.ctx.check(lib().icx_mma_update(self.h, _dp(x), _dp(df), C.c_double(gx), _dp(dg),
                                            C.c_double(xmin), C.c_double(xmax), C.byref(n)))
        return x, n.value

    def mma_set_asymptotes(self, init, decrease=0.7, increase=1.2):
        self.ctx.check(lib().icx_mma_set_asymptotes(self.h, C.c_double(init), C.c_double(decrease),
                                                    C.c_double(increase)))

    def stress_constraint(self, stress_limit, p_norm, qp, tol=1e-12, maxit=100000, sens=True):
        """(g, dg/drho_hat or None, adjoint PCG iterations) of the stress-constrained driver"""
        g = C.c_double()
        it = C.c_int()
        dg = np.zeros(self.nelx) if sens else None
        self.ctx.check(lib().icx_stress_constraint(self.h, C.c_double(stress_limit), C.c_double(p_norm),
                                                   C.c_double(qp), C.c_double(tol), int(maxit), C.byref(g),
                                                   _dp(dg) if sens else None, C.byref(it)))
        return g.value, dg, it.value

    def gcmma_reset(self):
        self.ctx.check(lib().icx_gcmma_reset(self.h))

    def gcmma_outer(self, xval, f0x, df0dx, fx, dfdx, xmin=1e-4, xmax=1.0):
        """GCMMASolver::OuterUpdate (m = 1): returns (xmma, dual Newton steps)."""
        xmma = np.zeros(self.nelx)
        xv, df, dg = _f64(xval), _f64(df0dx), _f64(dfdx)
        n = C.c_int()
        self.ctx.check(lib().icx_gcmma_outer_update(
            self.h, _dp(xmma), _dp(xv), C.c_double(f0x), _dp(df), C.c_double(fx), _dp(dg),
            C.c_double(xmin), C.c_double(xmax), C.byref(n)))
        return xmma, n.value

    def gcmma_inner(self, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin=1e-4, xmax=1.0):
        """GCMMASolver::InnerUpdate (m = 1): xmma is the last returned point."""
        xm = _f64(xmma).copy()
        xv, df, dg = _f64(xval), _f64(df0dx), _f64(dfdx)
        n = C.c_int()
        self.ctx.check(lib().icx_gcmma_inner_update(
            self.h, _dp(xm), C.c_double(f0xnew), C.c_double(fxnew), _dp(xv), C.c_double(f0x),
            _dp(df), C.c_double(fx), _dp(dg), C.c_double(xmin), C.c_double(xmax), C.byref(n)))
        return xm, n.value

    def gcmma_concheck(self, f0xnew, fxnew):
        ok = C.c_int()
        self.ctx.check(lib().icx_gcmma_concheck(self.h, C.c_double(f0xnew), C.c_double(fxnew),
                                                C.byref(ok)))
        return bool(ok.value)

    # -- device-resident design loop ------------------------------------------------
    def opt_init(self, x0, driver=0, filter=FILTER_CONE, radius=1.5, projection=True,
                 form=SOLVE_PLAIN, tol=1e-10, maxit=10000, helm_tol=1e-10, xmin=1e-4,
                 xmax=1.0, move=0.15):
        p = OptParams(driver, filter, radius, int(projection), form, tol, maxit, helm_tol,
                      xmin, xmax, move)
        x0 = _f64(x0)
        self.ctx.check(lib().icx_opt_init(self.h, C.byref(p), _dp(x0)))

    def opt_step(self):
        r = IterRecord()
        self.ctx.check(lib().icx_opt_step(self.h, C.byref(r)))
        return r.as_dict()

    def opt_step_begin(self):
        self.ctx.check(lib().icx_opt_step_begin(self.h))

    def opt_step_end(self, info=None):
        """info: the solve's info dict (its iteration count goes into the record)."""
        r = IterRecord()
        si = SolveInfo()
        if info:
            for k, _ in SolveInfo._fields_:
                setattr(si, k, info[k])
        self.ctx.check(lib().icx_opt_step_end(self.h, C.byref(si) if info else None, C.byref(r)))
        return r.as_dict()

    def submit(self, tol=1e-8, maxit=50000, form=None):
        """Start a batched solve (icx_submit: mixed form; icx_submit_form with form
        SOLVE_MIXED or SOLVE_FUSED); it advances in Context.poll."""
        if form is None:
            self.ctx.check(lib().icx_submit(self.h, C.c_double(tol), int(maxit)))
        else:
            self.ctx.check(lib().icx_submit_form(self.h, int(form), C.c_double(tol), int(maxit)))

    def cancel(self):
        """icx_cancel: stop this job's submitted solve"""
        self.ctx.check(lib().icx_cancel(self.h))

    def last_info(self):
        """Info dict of the last finished submitted solve (with 'converged')."""
        si = SolveInfo()
        rc = lib().icx_job_last_info(self.h, C.byref(si))
        self.ctx.check(rc, allow=(ERR_NOCONV,))
        d = si.as_dict()
        d["converged"] = rc == 0
        return d

    def opt_design(self):
        x = np.zeros(self.nelx)
        self.ctx.check(lib().icx_opt_get_design(self.h, _dp(x)))
        return x

    def footprint(self):
        b = C.c_size_t()
        self.ctx.check(lib().icx_job_footprint(self.h, C.byref(b)))
        return b.value


# -- the benchmark problems (same conventions as the oracle's) -------------------------

def random_density(n, seed, lo=0.01, hi=1.0):
    """Element densities i.i.d. U[lo, hi) in element-id order (numpy PCG64 stream)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    return lo + (hi - lo) * rng.random(n)


def cantilever_2d(ctx, nx=120, ny=40, state=PLANE_STRESS, young=(1e-9, 1.0), nu=0.3, p=3.0):
    """Q4 nx x ny, h = 1, left edge clamped, unit downward nodal load at the right
    mid-edge node."""
    j = Job(ctx, QUAD4, (nx, ny, 1), (float(nx), float(ny), 1.0), state, young, nu, p)
    j.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0.0, 0.0, 0.0, float(ny), 0.0, 0.0])
    j.make_dof()
    j.add_nodal_load(j.node_index(nx, ny // 2), 1, -1.0)
    return j


def cantilever_3d(ctx, nx, ny, nz, young=(1e-9, 1.0), nu=0.3, p=3.0):
    """hex8 nx x ny x nz, h = 1, x = 0 face clamped, total F_z = -1 on the edge
    {x = L, z = 0}, end nodes carrying half weight."""
    j = Job(ctx, HEXA8, (nx, ny, nz), (float(nx), float(ny), float(nz)), SOLID3D, young, nu, p)
    j.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0.0, 0.0, 0.0, float(ny), 0.0, float(nz)])
    j.make_dof()
    f = np.zeros(j.numeq)
    dof = j.dof_map()
    for jj in range(ny + 1):
        w = 0.5 if jj in (0, ny) else 1.0
        d = dof[3 * j.node_index(nx, jj, 0) + 2]
        if d < j.numeq:
            f[d] += -w / ny
    j.set_fext(f)
    return j


def random_job(ctx, rng, dims=None, size2d=(60, 240), size3d=(32, 96), young=(1e-9, 1.0), nu=0.3, p=3.0):
    """One job of the batched workload: Q4 d0 x d0/3 (plane stress) or hex8 d0 x d0/2 x
    d0/2, h = 1, d0 uniform in the given range; the x = 0 or the x = L face clamped; a
    unit point load (-y in 2D, -z in 3D) at a random node of the opposite face; uniform
    density = volume fraction ~ U[0.2, 0.5].  rng: numpy Generator.  Returns the job, with
    its clamped box and load recorded in `spec` (flag, clamp range, (node, comp, value))."""
    ndim = dims if dims is not None else (2 if rng.random() < 0.5 else 3)
    side = int(rng.integers(0, 2))
    if ndim == 2:
        d0 = int(rng.integers(size2d[0], size2d[1] + 1))
        d1 = max(1, d0 // 3)
        j = Job(ctx, QUAD4, (d0, d1, 1), (float(d0), float(d1), 1.0), PLANE_STRESS, young, nu, p)
        x = 0.0 if side == 0 else float(d0)
        j.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [x, x, 0.0, float(d1), 0.0, 0.0])
        j.make_dof()
        node = j.node_index(d0 if side == 0 else 0, int(rng.integers(0, d1 + 1)))
        j.add_nodal_load(node, 1, -1.0)
        j.spec = {"clamp": [x, x, 0.0, float(d1), 0.0, 0.0], "flag": [1, 1, 0, 0, 0, 0],
                  "load": (node, 1, -1.0)}
        ne = d0 * d1
    else:
        d0 = int(rng.integers(size3d[0], size3d[1] + 1))
        d1 = max(1, d0 // 2)
        j = Job(ctx, HEXA8, (d0, d1, d1), (float(d0), float(d1), float(d1)), SOLID3D, young, nu, p)
        x = 0.0 if side == 0 else float(d0)
        j.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [x, x, 0.0, float(d1), 0.0, float(d1)])
        j.make_dof()
        node = j.node_index(d0 if side == 0 else 0, int(rng.integers(0, d1 + 1)), int(rng.integers(0, d1 + 1)))
        j.add_nodal_load(node, 2, -1.0)
        j.spec = {"clamp": [x, x, 0.0, float(d1), 0.0, float(d1)], "flag": [1, 1, 1, 0, 0, 0],
                  "load": (node, 2, -1.0)}
        ne = d0 * d1 * d1
    j.volfrac = float(rng.uniform(0.2, 0.5))
    j.set_design(np.full(ne, j.volfrac))
    return j
