"""Writes the golden fixtures of tests/golden/*.npz.

The fixtures are computed here without the oracle: element matrices by a 3-point
Gauss rule in numpy, assembly by the closed-form DOF numbering of a clamped edge /
face (free DOFs in node order, reference countDOF), and the solve by SciPy's SuperLU.
They pin the oracle and the GPU path on small cantilevers:
  u (free DOFs), compliance c = f.u, sensitivity dc_e = -dE_e u_e^T K0 u_e,
  cone-filter output for r = 1.5, and the Helmholtz filter output (R = 2).
Run: python tests/golden/make_golden.py
"""
import itertools
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def gauss3(ndim):
    p, w = np.polynomial.legendre.leggauss(3)
    for idx in itertools.product(range(3), repeat=ndim):
        yield [p[i] for i in idx], np.prod([w[i] for i in idx])


def shape_grads(ndim, r, h):
    nq = 2 ** ndim
    N = np.zeros(nq)
    dN = np.zeros((ndim, nq))
    for q in range(nq):
        s = [1 if (q >> d) & 1 else -1 for d in range(ndim)]
        N[q] = np.prod([(1 + s[d] * r[d]) / 2 for d in range(ndim)])
        for d in range(ndim):
            v = s[d] / 2.0
            for e in range(ndim):
                if e != d:
                    v *= (1 + s[e] * r[e]) / 2
            dN[d, q] = v * 2.0 / h[d]
    return N, dN


def elastic_K0(ndim, h, nu, plane_stress=True):
    if ndim == 2:
        D = 1 / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    else:
        lam, mu = nu / ((1 + nu) * (1 - 2 * nu)), 1 / (2 * (1 + nu))
        D = np.zeros((6, 6)); D[:3, :3] = lam
        D[range(3), range(3)] = lam + 2 * mu; D[3, 3] = D[4, 4] = D[5, 5] = mu
    nq = 2 ** ndim
    K = np.zeros((nq * ndim, nq * ndim))
    for r, w in gauss3(ndim):
        _, dN = shape_grads(ndim, r, h)
        B = np.zeros((3 if ndim == 2 else 6, nq * ndim))
        for q in range(nq):
            if ndim == 2:
                B[0, 2 * q], B[1, 2 * q + 1], B[2, 2 * q], B[2, 2 * q + 1] = dN[0, q], dN[1, q], dN[1, q], dN[0, q]
            else:
                B[0, 3 * q], B[1, 3 * q + 1], B[2, 3 * q + 2] = dN[0, q], dN[1, q], dN[2, q]
                B[3, 3 * q], B[3, 3 * q + 1] = dN[1, q], dN[0, q]
                B[4, 3 * q + 1], B[4, 3 * q + 2] = dN[2, q], dN[1, q]
                B[5, 3 * q], B[5, 3 * q + 2] = dN[2, q], dN[0, q]
        K += B.T @ D @ B * w * np.prod(h) / 2 ** ndim
    return K


def helm_K0(ndim, h, R):
    rh2 = (R / (2 * np.sqrt(3))) ** 2
    nq = 2 ** ndim
    K = np.zeros((nq, nq))
    for r, w in gauss3(ndim):
        N, dN = shape_grads(ndim, r, h)
        K += (np.outer(N, N) + rh2 * dN.T @ dN) * w * np.prod(h) / 2 ** ndim
    return K


def case(ndim, div, seed, p=3.0, Emin=1e-9, nu=0.3):
    h = [1.0] * ndim
    sx = div[0] + 1
    shape_n = [d + 1 for d in div]
    nn = int(np.prod(shape_n))
    ne = int(np.prod(div))
    nq = 2 ** ndim
    # node / element ids, x fastest; corner q = a + 2b + 4c
    def node(ix):
        n, m = 0, 1
        for d in range(ndim):
            n += ix[d] * m
            m *= shape_n[d]
        return n
    conn = np.zeros((ne, nq), dtype=np.int64)
    for e, ix in enumerate(itertools.product(*[range(d) for d in reversed(div)])):
        ix = list(reversed(ix))
        for q in range(nq):
            conn[e, q] = node([ix[d] + ((q >> d) & 1) for d in range(ndim)])
    # DOF map: nodes with x index 0 clamped; free DOFs numbered in node order
    free_node = np.array([(n % sx) != 0 for n in range(nn)])
    dof = -np.ones((nn, ndim), dtype=np.int64)
    k = 0
    for n in range(nn):
        if free_node[n]:
            dof[n] = np.arange(k, k + ndim)
            k += ndim
    numeq = k
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 1.0, ne)
    E = Emin + rho ** p * (1.0 - Emin)
    K0 = elastic_K0(ndim, h, nu)
    rows, cols, vals = [], [], []
    for e in range(ne):
        idx = np.array([dof[n, c] for n in conn[e] for c in range(ndim)])
        for a in range(nq * ndim):
            for b in range(nq * ndim):
                if idx[a] >= 0 and idx[b] >= 0:
                    rows.append(idx[a]); cols.append(idx[b]); vals.append(E[e] * K0[a, b])
    K = sp.csc_matrix((vals, (rows, cols)), shape=(numeq, numeq))
    f = np.zeros(numeq)
    if ndim == 2:
        f[dof[node([div[0], div[1] // 2]), 1]] = -1.0
    else:
        for j in range(div[1] + 1):
            w = 0.5 if j in (0, div[1]) else 1.0
            f[dof[node([div[0], j, 0]), 2]] += -w / div[1]
    u = spla.spsolve(K, f)
    c = f @ u
    full = np.zeros((nn, ndim))
    for n in range(nn):
        if free_node[n]:
            full[n] = u[dof[n]]
    dc = np.array([-p * rho[e] ** (p - 1) * (1.0 - Emin) *
                   (full[conn[e]].ravel() @ K0 @ full[conn[e]].ravel()) for e in range(ne)])
    # cone filter r = 1.5 (weights r - d over element centres, truncated at the boundary)
    cen = np.zeros((ne, ndim))
    for e, ix in enumerate(itertools.product(*[range(d) for d in reversed(div)])):
        cen[e] = np.array(list(reversed(ix))) + 0.5
    d = np.sqrt(((cen[:, None, :] - cen[None, :, :]) ** 2).sum(-1))
    W = np.where(d <= 1.5, 1.5 - d, 0.0)
    cone = W @ rho / W.sum(1)
    # Helmholtz filter R = 2 (natural BC): phi = KF^-1 sum_e rho_e int N; rho~ = mean phi
    KF0 = helm_K0(ndim, h, 2.0)
    rr, cc, vv = [], [], []
    fh = np.zeros(nn)
    for e in range(ne):
        for a in range(nq):
            fh[conn[e, a]] += rho[e] / nq
            for b in range(nq):
                rr.append(conn[e, a]); cc.append(conn[e, b]); vv.append(KF0[a, b])
    phi = spla.spsolve(sp.csc_matrix((vv, (rr, cc)), shape=(nn, nn)), fh)
    helm = phi[conn].mean(1)
    return dict(rho=rho, u=u, c=c, dc=dc, cone=cone, helm=helm, div=np.array(div))


if __name__ == "__main__":
    np.savez(os.path.join(HERE, "q4_cantilever_24x8.npz"), **case(2, [24, 8], 31))
    np.savez(os.path.join(HERE, "hex8_cantilever_10x4x4.npz"), **case(3, [10, 4, 4], 32))
    print("wrote fixtures")
