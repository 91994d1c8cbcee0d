"""ctypes front end of the CPU oracle (oracle/icx_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the parity checker.  The product library never imports it.

`Problem` mirrors the reference's problem set-up sequence
(opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:27-65): MeshIcarat box mesh,
BCIcarat Dirichlet ranges, countDOF numbering, Neumann tractions into
Force::fext_org, plus direct nodal loads written into fext_org.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ICX_ORACLE_LIB selects another build of the same source (the sanitizer build)
_LIB_PATH = os.environ.get("ICX_ORACLE_LIB") or os.path.join(_HERE, "_build", "libicx_oracle.so")

PLANE_STRESS, PLANE_STRAIN, SOLID3D, HEAT = 0, 1, 2, 3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "icxo_mesh_box": (i, [i, _ip, _dp, _dp, _ip]),
            "icxo_dirichlet": (None, [i, i, i, _dp, _ip, _dp, _dp, _ip, _ip, _dp]),
            "icxo_count_dof": (i, [i, i, _ip, _ip, _ip]),
            "icxo_neumann": (None, [i, i, i, i, _dp, _ip, _ip, i, _dp, _dp, _dp]),
            "icxo_simp": (d, [d, d, d, d]),
            "icxo_element_stiffness": (None, [i, i, i, d, d, _dp, _dp, _dp]),
            "icxo_helm_element": (None, [i, i, d, _dp, _dp, _dp]),
            "icxo_sys_create": (vp, [i, i, i, i, i, i, _dp, _ip, _ip, i, d, d, d, d]),
            "icxo_sys_destroy": (None, [vp]),
            "icxo_sys_nnz": (C.c_longlong, [vp]),
            "icxo_sys_assemble": (None, [vp, _dp, vp, vp, vp, vp]),
            "icxo_sys_csr": (None, [vp, _ip, _ip, _dp]),
            "icxo_sys_spmv": (None, [vp, _dp, _dp]),
            "icxo_sys_cg": (i, [vp, _dp, _dp, d, i, C.POINTER(d)]),
            "icxo_sys_direct": (i, [vp, _dp, _dp]),
            "icxo_sys_element_energy": (None, [vp, _dp, _dp]),
            "icxo_sys_mises": (None, [vp, _dp, _dp]),
            "icxo_sys_set_conductivity": (None, [vp, _dp, _dp]),
            "icxo_sys_heat_sensitivity": (None, [vp, _dp, _dp]),
            "icxo_heat_element": (None, [C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
            "icxo_sys_reactions": (None, [vp, _dp, _dp]),
            "icxo_stress_constraint": (i, [vp, _dp, vp, d, d, d, i, d, C.POINTER(d), _dp, _dp]),
            "icxo_csr_cg": (i, [i, _ip, _ip, _dp, _dp, _dp, d, i, C.POINTER(d)]),
            "icxo_divide_val_into_node": (None, [i, i, i, _ip, _dp, vp, _dp]),
            "icxo_sensitivity": (None, [i, i, d, d, d, _dp, _dp, _dp, d, _dp, _dp]),
            "icxo_filter_create": (vp, [i, i, i, _dp, _ip, d, i]),
            "icxo_filter_destroy": (None, [vp]),
            "icxo_filter_nnz": (C.c_longlong, [vp]),
            "icxo_filter_neighbours": (None, [vp, _ip, _ip, _dp]),
            "icxo_filter_apply": (None, [vp, _dp, _dp, i]),
            "icxo_filter_sensitivity": (None, [vp, _dp, _dp, _dp, _dp]),
            "icxo_threshold": (None, [i, _dp, d, d, _dp]),
            "icxo_threshold_derivative": (None, [i, _dp, d, d, _dp]),
            "icxo_helm_create": (vp, [i, i, i, i, _dp, _ip, d]),
            "icxo_helm_destroy": (None, [vp]),
            "icxo_helm_apply": (i, [vp, _dp, _dp, i, d, i, i]),
            "icxo_mma_create": (vp, [i, i, d, d, d]),
            "icxo_mma_destroy": (None, [vp]),
            "icxo_mma_update": (i, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.icxo_oc_update.restype = i
        L.icxo_run_driver.restype = i
        _lib = L
    return _lib


VOLFN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Problem:
    """Structured box problem following the reference's set-up order."""

    def __init__(self, ne, div, length, state, young=(1e-9, 1.0), nu=0.3, penalty=3.0):
        self.ne = int(ne)
        self.ndim = 2 if ne == 4 else 3
        self.dofnp = 1 if state == HEAT else self.ndim
        self.div = np.array(list(div) + [1] * (3 - len(div)), dtype=np.int32)[:3]
        if self.ndim == 2:
            self.div[2] = 1
        self.length = _f64(list(length) + [1.0] * (3 - len(length)))[:3].copy()
        self.state = state
        self.young1, self.young2 = float(young[0]), float(young[1])
        self.nu, self.p = float(nu), float(penalty)
        d0, d1, d2 = (int(v) for v in self.div)
        if self.ndim == 2:
            self.numnp, self.nelx = (d0 + 1) * (d1 + 1), d0 * d1
        else:
            self.numnp, self.nelx = (d0 + 1) * (d1 + 1) * (d2 + 1), d0 * d1 * d2
        self.x = np.zeros(self.numnp * 3)
        self.conn = np.zeros(self.nelx * self.ne, dtype=np.int32)
        lib().icxo_mesh_box(self.ne, self.div, self.length, self.x, self.conn)
        self.has = np.zeros(self.numnp, dtype=np.int32)
        self.node_flag = np.zeros(self.numnp * 6, dtype=np.int32)
        self.node_val = np.zeros(self.numnp * 6)
        self.dof = None
        self.numeq = None
        self.fext = None
        self._sys = None

    @classmethod
    def from_mesh(cls, ne, x, conn, state, young=(1e-9, 1.0), nu=0.3, penalty=3.0):
        """Problem on an explicit mesh (MeshMDPA, src/io/mesh_mdpa.hpp:84-238): x numnp x 3,
        conn nelx x ne (0-based, the reference's local node order)."""
        self = cls.__new__(cls)
        self.ne = int(ne)
        self.ndim = 2 if ne == 4 else 3
        self.dofnp = 1 if state == HEAT else self.ndim
        self.div = None
        self.length = None
        self.state = state
        self.young1, self.young2 = float(young[0]), float(young[1])
        self.nu, self.p = float(nu), float(penalty)
        self.x = _f64(x).reshape(-1).copy()
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1).copy()
        self.numnp, self.nelx = len(self.x) // 3, len(self.conn) // self.ne
        self.has = np.zeros(self.numnp, dtype=np.int32)
        self.node_flag = np.zeros(self.numnp * 6, dtype=np.int32)
        self.node_val = np.zeros(self.numnp * 6)
        self.dof = None
        self.numeq = None
        self.fext = None
        self._sys = None
        return self

    # -- boundary conditions ------------------------------------------------
    def add_dirichlet(self, flag6, val6, range6):
        assert self.dof is None, "Dirichlet ranges must precede make_dof"
        lib().icxo_dirichlet(self.ndim, self.dofnp, self.numnp, self.x,
                             np.asarray(flag6, dtype=np.int32), _f64(val6), _f64(range6),
                             self.has, self.node_flag, self.node_val)

    def make_dof(self):
        self.dof = np.zeros(self.numnp * self.dofnp, dtype=np.int32)
        self.numeq = lib().icxo_count_dof(self.numnp, self.dofnp, self.has, self.node_flag,
                                          self.dof)
        self.fext = np.zeros(self.numeq)
        return self.numeq

    def add_neumann(self, val3, range6):
        lib().icxo_neumann(self.ndim, self.dofnp, self.ne, self.nelx, self.x, self.conn,
                           self.dof, self.numeq, _f64(val3), _f64(range6), self.fext)

    def add_nodal_load(self, node, comp, value):
        d = self.dof[self.dofnp * node + comp]
        if d < self.numeq:
            self.fext[d] += value

    def node_index(self, i, j, k=0):
        d0, d1 = int(self.div[0]), int(self.div[1])
        return i + (d0 + 1) * (j + (d1 + 1) * k)

    # -- structural system ---------------------------------------------------
    def system(self):
        if self._sys is None:
            self._sys = lib().icxo_sys_create(self.ne, self.ndim, self.dofnp, self.state,
                                              self.numnp, self.nelx, self.x, self.conn,
                                              self.dof, self.numeq, self.young1, self.young2,
                                              self.nu, self.p)
        return self._sys

    def __del__(self):
        if getattr(self, "_sys", None) and _lib is not None:
            _lib.icxo_sys_destroy(self._sys)
            self._sys = None

    def _bc_ptrs(self):
        if self.has.any():
            return (self.has.ctypes.data, self.node_flag.ctypes.data, self.node_val.ctypes.data)
        return (None, None, None)

    def assemble(self, design):
        """Assemble K(design) on the free DOFs; returns rhs = fext - fint."""
        s = self.system()
        fint = np.zeros(self.numeq)
        h, f, v = self._bc_ptrs()
        lib().icxo_sys_assemble(s, _f64(design), h, f, v, fint.ctypes.data)
        return self.fext - fint

    def csr(self):
        import scipy.sparse as sp
        s = self.system()
        nnz = lib().icxo_sys_nnz(s)
        rp = np.zeros(self.numeq + 1, dtype=np.int32)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        lib().icxo_sys_csr(s, rp, col, val)
        return sp.csr_matrix((val, col, rp), shape=(self.numeq, self.numeq))

    def spmv(self, x):
        y = np.zeros(self.numeq)
        lib().icxo_sys_spmv(self.system(), _f64(x), y)
        return y

    def cg(self, b, tol=1e-10, maxit=10000):
        x = np.zeros(self.numeq)
        err = C.c_double()
        it = lib().icxo_sys_cg(self.system(), _f64(b), x, tol, maxit, C.byref(err))
        return x, it, err.value

    def direct(self, b):
        x = np.zeros(self.numeq)
        rc = lib().icxo_sys_direct(self.system(), _f64(b), x)
        if rc != 0:
            raise RuntimeError("banded Cholesky failed (%d)" % rc)
        return x

    def node_values(self, u):
        out = np.zeros(self.numnp * 3)
        lib().icxo_divide_val_into_node(self.numnp, self.dofnp, self.numeq, self.dof, _f64(u),
                                        self.node_val.ctypes.data, out)
        return out

    def element_energy(self, nodeval):
        out = np.zeros(self.nelx)
        lib().icxo_sys_element_energy(self.system(), _f64(nodeval), out)
        return out

    def set_conductivity(self, k1, k2):
        """heat: conductivity rows (xx, yy, zz, xy, yz, xz) of materials 1 and 2"""
        lib().icxo_sys_set_conductivity(self.system(), _f64(k1), _f64(k2))

    def heat_sensitivity(self, nodeval):
        """-T_e^T dKe/ds T_e for the design of the last assembly"""
        out = np.zeros(self.nelx)
        lib().icxo_sys_heat_sensitivity(self.system(), _f64(nodeval), out)
        return out

    def mises(self, nodeval):
        """element-mean von Mises stress of the last assembly's design"""
        out = np.zeros(self.nelx)
        lib().icxo_sys_mises(self.system(), _f64(nodeval), out)
        return out

    def stress_constraint(self, nodeval, qp, p_norm, stress_limit, direct=True, cg_tol=1e-12):
        """g and dg/drho_hat of the stress driver (constraint/opt-stress-constraint/
        sensitivity.cpp:8-155) for the last assembly's design and its solution; also the
        qp-relaxed stresses"""
        g = C.c_double()
        dg, sqp = np.zeros(self.nelx), np.zeros(self.nelx)
        rc = lib().icxo_stress_constraint(self.system(), _f64(nodeval), self.node_val.ctypes.data, qp,
                                          p_norm, stress_limit, int(direct), cg_tol, C.byref(g), dg, sqp)
        if rc != 0:
            raise RuntimeError("stress constraint failed (%d)" % rc)
        return g.value, dg, sqp

    def reactions(self, nodeval):
        """K_full u over all DOFs, 3 per node, |r| < 1e-15 flushed"""
        out = np.zeros(3 * self.numnp)
        lib().icxo_sys_reactions(self.system(), _f64(nodeval), out)
        return out

    def volumes(self):
        X = self.x.reshape(-1, 3)[:, :self.ndim]
        vol = np.zeros(self.nelx)
        Ke = np.zeros(24 * 24)  # the largest element matrix (hexa8, 3 DOFs per node)
        v = np.zeros(1)
        eye = np.eye(self.ndim).ravel()
        # all elements of a box mesh share one shape; volumes via quadrature per element
        for e in range(self.nelx):
            Xe = _f64(X[self.conn[e * self.ne:(e + 1) * self.ne]])
            if self.state == HEAT:
                lib().icxo_heat_element(self.ne, self.ndim, eye, Xe, Ke, v)
            else:
                lib().icxo_element_stiffness(self.ne, self.ndim, self.state, 1.0, self.nu, Xe, Ke, v)
            vol[e] = v[0]
        return vol

    def sensitivity(self, design, energy, volume, V0=1.0, driver_b=False):
        dfds = np.zeros(self.nelx)
        dgds = np.zeros(self.nelx)
        lib().icxo_sensitivity(self.nelx, int(driver_b), self.young1, self.young2, self.p,
                               _f64(design), _f64(energy), _f64(volume), V0, dfds, dgds)
        return dfds, dgds

    def element_stiffness(self, E=1.0, e=0):
        """elastic element stiffness E K0 of element e (elastic problems only)"""
        if self.state == HEAT:
            raise ValueError("element_stiffness: elastic problems only (heat: icxo_heat_element)")
        X = _f64(self.x.reshape(-1, 3)[self.conn[e * self.ne:(e + 1) * self.ne], :self.ndim])
        n = self.ne * self.dofnp
        Ke = np.zeros(n * n)
        v = np.zeros(1)
        lib().icxo_element_stiffness(self.ne, self.ndim, self.state, E, self.nu, X, Ke, v)
        return Ke.reshape(n, n), v[0]

    # -- filters ---------------------------------------------------------------
    def cone_filter(self, radius, brute_force=False):
        return ConeFilter(self, radius, brute_force)

    def helmholtz(self, radius):
        return Helmholtz(self, radius)

    # -- drivers ---------------------------------------------------------------
    def run_driver(self, driver, x0, iterations, radius, projection=True, direct=False,
                   cg_tol=1e-10, cg_maxit=10000, helm_tol=1e-10, helm_direct=False,
                   xmin=1e-4, xmax=1.0, move=0.15, stress=(1.0, 8.0, 0.7)):
        """driver 0 (OC), 1 (MMA + Helmholtz), 2 (GCMMA), 3 (stress constraint; stress =
        (stress_limit, p_norm, qp), obj = volume and vol = the stress constraint)"""
        class Cfg(C.Structure):
            _fields_ = [("driver", C.c_int), ("ne", C.c_int), ("ndim", C.c_int),
                        ("dofnp", C.c_int), ("state", C.c_int), ("numnp", C.c_int),
                        ("nelx", C.c_int), ("numeq", C.c_int),
                        ("x", C.c_void_p), ("conn", C.c_void_p), ("dof", C.c_void_p),
                        ("has", C.c_void_p), ("node_flag", C.c_void_p), ("node_val", C.c_void_p),
                        ("fext", C.c_void_p),
                        ("young1", C.c_double), ("young2", C.c_double), ("nu", C.c_double),
                        ("p", C.c_double), ("filter_radius", C.c_double),
                        ("xmin", C.c_double), ("xmax", C.c_double), ("move", C.c_double),
                        ("iterations", C.c_int), ("projection", C.c_int), ("direct", C.c_int),
                        ("cg_tol", C.c_double), ("cg_maxit", C.c_int),
                        ("helm_tol", C.c_double), ("helm_direct", C.c_int),
                        ("inner_hist", C.c_void_p), ("stress_limit", C.c_double),
                        ("p_norm", C.c_double), ("qp", C.c_double)]
        h, f, v = self._bc_ptrs()
        inner = np.zeros(iterations, dtype=np.int32)
        cfg = Cfg(driver, self.ne, self.ndim, self.dofnp, self.state, self.numnp, self.nelx,
                  self.numeq, self.x.ctypes.data, self.conn.ctypes.data, self.dof.ctypes.data,
                  h, f, v, self.fext.ctypes.data, self.young1, self.young2, self.nu, self.p,
                  radius, xmin, xmax, move, iterations, int(projection), int(direct), cg_tol,
                  cg_maxit, helm_tol, int(helm_direct), inner.ctypes.data, stress[0], stress[1],
                  stress[2])
        x = _f64(x0).copy()
        obj = np.zeros(iterations)
        vol = np.zeros(iterations)
        dfds = np.zeros(self.nelx)
        u = np.zeros(self.numeq)
        its = np.zeros(iterations, dtype=np.int32)
        L = lib()
        L.icxo_run_driver.argtypes = [C.POINTER(Cfg), _dp, _dp, _dp, _dp, _dp, _ip]
        rc = L.icxo_run_driver(C.byref(cfg), x, obj, vol, dfds, u, its)
        if rc != 0:
            raise RuntimeError("driver failed %d" % rc)
        return dict(design=x, obj=obj, vol=vol, dfds=dfds, u=u, cg_iters=its, inner=inner)


class ConeFilter:
    def __init__(self, prob, radius, brute_force=False):
        self.n = prob.nelx
        self.h = lib().icxo_filter_create(prob.ne, prob.ndim, prob.nelx, prob.x, prob.conn,
                                          float(radius), int(brute_force))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.icxo_filter_destroy(self.h)

    def neighbours(self):
        nnz = lib().icxo_filter_nnz(self.h)
        ptr = np.zeros(self.n + 1, dtype=np.int32)
        ids = np.zeros(nnz, dtype=np.int32)
        dist = np.zeros(nnz)
        lib().icxo_filter_neighbours(self.h, ptr, ids, dist)
        return ptr, ids, dist

    def apply(self, v, sens=False):
        out = np.zeros(self.n)
        lib().icxo_filter_apply(self.h, _f64(v), out, int(sens))
        return out

    def sensitivity(self, design, dfds, volume):
        """Filter::sensitivityFilter (filter.hpp:72-100)"""
        out = np.zeros(self.n)
        lib().icxo_filter_sensitivity(self.h, _f64(design), _f64(dfds), _f64(volume), out)
        return out


class Helmholtz:
    def __init__(self, prob, radius):
        self.n = prob.nelx
        self.prob = prob  # icxo_helm keeps a pointer to prob.conn
        self.h = lib().icxo_helm_create(prob.ne, prob.ndim, prob.numnp, prob.nelx, prob.x,
                                        prob.conn, float(radius))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.icxo_helm_destroy(self.h)

    def apply(self, v, sens=False, tol=1e-10, maxit=10000, direct=False):
        out = np.zeros(self.n)
        it = lib().icxo_helm_apply(self.h, _f64(v), out, int(sens), tol, maxit, int(direct))
        return out, it


def threshold(v, beta, T=0.5):
    out = np.zeros(len(v))
    lib().icxo_threshold(len(v), _f64(v), beta, T, out)
    return out


def threshold_derivative(v, beta, T=0.5):
    out = np.zeros(len(v))
    lib().icxo_threshold_derivative(len(v), _f64(v), beta, T, out)
    return out


def oc_update(x, dfdx, g, dgdx, xmin, xmax, volume_fn, move=0.15):
    x = _f64(x).copy()
    n = len(x)

    def cb(ptr, _user):
        arr = np.ctypeslib.as_array(ptr, shape=(n,)).copy()
        return float(volume_fn(arr))

    fn = VOLFN(cb)
    L = lib()
    L.icxo_oc_update.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp, C.c_double,
                                 VOLFN, C.c_void_p]
    it = L.icxo_oc_update(n, x, _f64(dfdx), float(g), _f64(dgdx), _f64(xmin), _f64(xmax),
                          move, fn, None)
    return x, it


class MMA:
    def __init__(self, n, m=1, a=0.0, c=1e5, d=0.0):
        self.n, self.m = n, m
        self.h = lib().icxo_mma_create(n, m, a, c, d)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.icxo_mma_destroy(self.h)

    def update(self, x, dfdx, gx, dgdx, xmin, xmax):
        x = _f64(x).copy()
        it = lib().icxo_mma_update(self.h, x, _f64(dfdx), _f64(np.atleast_1d(gx)), _f64(dgdx),
                                   _f64(xmin), _f64(xmax))
        return x, it

    def set_asymptotes(self, init, decrease=0.7, increase=1.2):
        L = lib()
        L.icxo_mma_set_asymptotes.restype = None
        L.icxo_mma_set_asymptotes.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.icxo_mma_set_asymptotes(self.h, init, decrease, increase)

    def subproblem(self):
        """the last update's subproblem: low, upp, alpha, beta, p0, q0, pij, qij (n x m),
        b, and the dual solution lam, y, z"""
        n, m = self.n, self.m
        out = {k: np.zeros(n) for k in ("low", "upp", "alpha", "beta", "p0", "q0")}
        out.update(pij=np.zeros(n * m), qij=np.zeros(n * m), b=np.zeros(m), lam=np.zeros(m),
                   y=np.zeros(m))
        z = C.c_double()
        L = lib()
        L.icxo_mma_subproblem.restype = None
        L.icxo_mma_subproblem.argtypes = [C.c_void_p] + [_dp] * 11 + [C.POINTER(C.c_double)]
        L.icxo_mma_subproblem(self.h, out["low"], out["upp"], out["alpha"], out["beta"],
                              out["p0"], out["q0"], out["pij"], out["qij"], out["b"],
                              out["lam"], out["y"], C.byref(z))
        out["pij"] = out["pij"].reshape(n, m)
        out["qij"] = out["qij"].reshape(n, m)
        out["z"] = z.value
        return out


class GCMMA:
    """GCMMASolver(n, m, a=0, c=1000, d=0) (src/optimization/opt-solver/GCMMASolver.cpp:41-129)."""

    def __init__(self, n, m=1, a=0.0, c=1000.0, d=0.0):
        self.n, self.m = n, m
        L = lib()
        for name, res, args in (
                ("icxo_gcmma_create", C.c_void_p, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]),
                ("icxo_gcmma_outer", C.c_int, [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp]),
                ("icxo_gcmma_inner", C.c_int, [C.c_void_p, _dp, C.c_double, _dp, _dp, C.c_double, _dp, _dp,
                                               _dp, _dp, _dp]),
                ("icxo_gcmma_concheck", C.c_int, [C.c_void_p, C.c_double, _dp]),
                ("icxo_gcmma_state", None, [C.c_void_p, _dp, _dp, _dp, _dp]),
                ("icxo_gcmma_approx_at", None, [C.c_void_p, _dp, _dp, _dp])):
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        self.h = L.icxo_gcmma_create(n, m, a, c, d)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.icxo_mma_destroy(self.h)

    def outer(self, xval, f0x, df0dx, fx, dfdx, xmin, xmax):
        xmma = np.zeros(self.n)
        it = lib().icxo_gcmma_outer(self.h, xmma, _f64(xval), float(f0x), _f64(df0dx),
                                    _f64(np.atleast_1d(fx)), _f64(dfdx), _f64(xmin), _f64(xmax))
        return xmma, it

    def inner(self, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax):
        xmma = _f64(xmma).copy()
        it = lib().icxo_gcmma_inner(self.h, xmma, float(f0xnew), _f64(np.atleast_1d(fxnew)), _f64(xval),
                                    float(f0x), _f64(df0dx), _f64(np.atleast_1d(fx)), _f64(dfdx),
                                    _f64(xmin), _f64(xmax))
        return xmma, it

    def concheck(self, f0xnew, fxnew):
        return bool(lib().icxo_gcmma_concheck(self.h, float(f0xnew), _f64(np.atleast_1d(fxnew))))

    def approx_at(self, x):
        f0, fj = np.zeros(1), np.zeros(self.m)
        lib().icxo_gcmma_approx_at(self.h, _f64(x), f0, fj)
        return f0[0], fj

    def state(self):
        f0, fa, r0, ra = np.zeros(1), np.zeros(self.m), np.zeros(1), np.zeros(self.m)
        lib().icxo_gcmma_state(self.h, f0, fa, r0, ra)
        return dict(f0app=f0[0], fapp=fa, raa0=r0[0], raa=ra)


# -- the benchmark problems ----------------------------------------------------

def cantilever_2d(nx=120, ny=40, state=PLANE_STRESS, young=(1e-9, 1.0), nu=0.3, p=3.0):
    """Q4 nx x ny, h = 1, left edge clamped, unit downward nodal load at the right
    mid-edge node (written straight into fext_org)."""
    pr = Problem(4, (nx, ny, 1), (float(nx), float(ny), 1.0), state, young, nu, p)
    pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0.0, 0.0, 0.0, float(ny), 0.0, 0.0])
    pr.make_dof()
    pr.add_nodal_load(pr.node_index(nx, ny // 2), 1, -1.0)
    return pr


def cantilever_3d(nx, ny, nz, young=(1e-9, 1.0), nu=0.3, p=3.0):
    """hex8 nx x ny x nz, h = 1, x = 0 face clamped, total F_z = -1 on the edge
    {x = L, z = 0} with consistent line-load weights (end nodes 1/2)."""
    pr = Problem(8, (nx, ny, nz), (float(nx), float(ny), float(nz)), SOLID3D, young, nu, p)
    pr.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0.0, 0.0, 0.0, float(ny), 0.0, float(nz)])
    pr.make_dof()
    for j in range(ny + 1):
        w = 0.5 if j in (0, ny) else 1.0
        pr.add_nodal_load(pr.node_index(nx, j, 0), 2, -w / ny)
    return pr


def random_density(n, seed, lo=0.01, hi=1.0):
    """Element densities i.i.d. U[lo, hi) in element-id order from numpy's PCG64
    stream with the given seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return lo + (hi - lo) * rng.random(n)
