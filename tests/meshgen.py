"""Unstructured test meshes in the reference's local node order (quad4 counter-clockwise;
hexa8 z-low face then z-high face, src/problem/shape_function.cpp:97-119, 403-447), and
an MDPA writer in the layout MeshMDPA reads (src/io/mesh_mdpa.hpp:84-238)."""
import numpy as np

import oracle as O


def box_mesh(ne, div, length):
    """The structured box mesh of MeshIcarat as explicit arrays (x: numnp x 3, conn)."""
    pr = O.Problem(ne, div, length, O.PLANE_STRESS if ne == 4 else O.SOLID3D)
    return pr.x.copy(), pr.conn.copy()


def distort(x, conn, ne, amplitude, seed, lengths):
    """Interior nodes moved by up to `amplitude` (in units of the smallest element edge)
    in every direction; boundary nodes stay on the box faces, so box-shaped boundary
    conditions select the same nodes.  Jacobians stay positive for amplitude < 1/4."""
    rng = np.random.default_rng(seed)
    X = x.reshape(-1, 3).copy()
    ndim = 2 if ne == 4 else 3
    lo = X[:, :ndim].min(axis=0)
    hi = X[:, :ndim].max(axis=0)
    E = conn.reshape(-1, ne)
    edge = np.inf
    for a in range(ne):
        for b in range(a + 1, ne):
            d = np.linalg.norm(X[E[:, a]] - X[E[:, b]], axis=1)
            edge = min(edge, d.min())
    for d in range(ndim):
        inner = (X[:, d] > lo[d] + 1e-9) & (X[:, d] < hi[d] - 1e-9)
        X[inner, d] += amplitude * edge * rng.uniform(-1.0, 1.0, inner.sum())
    return X.reshape(-1)


def annulus_quad(nr, nt, r0, r1, angle=np.pi / 2):
    """Quarter annulus r0 <= r <= r1, 0 <= theta <= angle: nr x nt quad4 elements,
    nodes radial-fastest, counter-clockwise elements."""
    r = np.linspace(r0, r1, nr + 1)
    t = np.linspace(0.0, angle, nt + 1)
    X = np.zeros(((nr + 1) * (nt + 1), 3))
    for j in range(nt + 1):
        for i in range(nr + 1):
            X[i + (nr + 1) * j] = (r[i] * np.cos(t[j]), r[i] * np.sin(t[j]), 0.0)
    conn = []
    for j in range(nt):
        for i in range(nr):
            n = i + (nr + 1) * j
            conn += [n, n + 1, n + 1 + (nr + 1), n + (nr + 1)]
    return X.reshape(-1), np.asarray(conn, dtype=np.int32)


def write_mdpa(path, x, conn, ne):
    """Nodes and one element block; ids 1-based, coordinates round-trip exactly."""
    tag = "SmallDisplacementElement2D4N" if ne == 4 else "SmallDisplacementElement3D8N"
    X = np.asarray(x).reshape(-1, 3)
    E = np.asarray(conn).reshape(-1, ne)
    with open(path, "w") as f:
        f.write("Begin ModelPartData\nEnd ModelPartData\n\nBegin Properties 0\nEnd Properties\n\n")
        f.write("Begin Nodes\n")
        for i, p in enumerate(X):
            f.write("%d %s %s %s\n" % (i + 1, repr(float(p[0])), repr(float(p[1])), repr(float(p[2]))))
        f.write("End Nodes\n\n")
        f.write("Begin Elements %s\n" % tag)
        for e, nodes in enumerate(E):
            f.write("%d 0 %s\n" % (e + 1, " ".join(str(int(n) + 1) for n in nodes)))
        f.write("End Elements\n")
