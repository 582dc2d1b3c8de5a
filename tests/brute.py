"""Independent numpy constructions used to pin the oracle and the product.

Everything here is built from first principles (numpy's Legendre module,
barycentric Lagrange polynomials, the analytic derivative of the trilinear
element map) and shares no code with the oracle or the HIP library.
"""
import numpy as np
import numpy.polynomial.legendre as L


def gll_nodes(N):
    c = np.zeros(N + 1)
    c[N] = 1
    interior = np.sort(L.Legendre(c).deriv().roots().real) if N > 1 else np.zeros(0)
    x = np.concatenate([[-1.0], interior, [1.0]])
    PN = L.Legendre(c)(x)
    w = 2.0 / (N * (N + 1) * PN**2)
    return x, w


def lagrange(nodes, pts):
    """Values and derivatives of the Lagrange basis of `nodes` at `pts`:
    V[q, j] = l_j(pts[q]), Vd[q, j] = l_j'(pts[q])."""
    n = len(nodes)
    V = np.zeros((len(pts), n))
    Vd = np.zeros((len(pts), n))
    for j in range(n):
        others = np.delete(nodes, j)
        poly = np.poly1d(others, r=True) / np.prod(nodes[j] - others)
        V[:, j] = poly(pts)
        Vd[:, j] = poly.deriv()(pts)
    return V, Vd


def trilinear_jacobian(X, Y, Z, r, s, t):
    """Analytic Jacobian d(x,y,z)/d(r,s,t) of the trilinear map with vertex
    coordinates X,Y,Z[8] (vertex order of libs/mesh/meshSetupBoxHex3D.cpp)."""
    sr = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
    ss = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
    st = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)
    dr = 0.125 * sr * (1 + ss * s) * (1 + st * t)
    ds = 0.125 * ss * (1 + sr * r) * (1 + st * t)
    dt = 0.125 * st * (1 + sr * r) * (1 + ss * s)
    Jm = np.zeros((3, 3))
    for c, V in enumerate((X, Y, Z)):
        Jm[c] = [dr @ V, ds @ V, dt @ V]
    return Jm


def trilinear_point(X, Y, Z, r, s, t):
    sr = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
    ss = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
    st = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)
    c = 0.125 * (1 + sr * r) * (1 + ss * s) * (1 + st * t)
    return np.array([c @ X, c @ Y, c @ Z])


def element_matrix(N, Xv, Yv, Zv, qpts, qw, lam, stiff=True, kappa=None):
    """Dense element matrix (Np x Np) of lam*M + K for the GLL Lagrange basis
    of degree N, integrated with the tensor rule (qpts, qw), on the trilinear
    element with vertices Xv, Yv, Zv. Node index n = i + j*Nq + k*Nq^2.
    kappa (optional): coefficient of the stiffness term at each quadrature
    point, indexed a + b*nq + c*nq^2."""
    x, _ = gll_nodes(N)
    V, Vd = lagrange(x, qpts)
    nq = len(qpts)
    Np = (N + 1) ** 3
    A = np.zeros((Np, Np))
    for c in range(nq):
        for b in range(nq):
            for a in range(nq):
                r, s, t = qpts[a], qpts[b], qpts[c]
                w = qw[a] * qw[b] * qw[c]
                Jm = trilinear_jacobian(Xv, Yv, Zv, r, s, t)
                J = np.linalg.det(Jm)
                phi = np.einsum("k,j,i->kji", V[c], V[b], V[a]).ravel()
                A += lam * w * J * np.outer(phi, phi)
                if stiff:
                    gr = np.einsum("k,j,i->kji", V[c], V[b], Vd[a]).ravel()
                    gs = np.einsum("k,j,i->kji", V[c], Vd[b], V[a]).ravel()
                    gt = np.einsum("k,j,i->kji", Vd[c], V[b], V[a]).ravel()
                    Gref = np.stack([gr, gs, gt])  # 3 x Np
                    Jinv = np.linalg.inv(Jm)       # rows = grad r, grad s, grad t
                    G = w * J * (Jinv @ Jinv.T)
                    if kappa is not None:
                        G = G * kappa[a + b * nq + c * nq * nq]
                    A += Gref.T @ G @ Gref
    return A


def assembled_operator(N, EX, EY, EZ, G2L, Ngather, qpts, qw, lam, stiff=True, nfields=1, kappaL=None):
    """Assembled dense operator Z^T blockdiag(A_e) Z using the local->global
    map G2L (-1 = masked). kappaL (optional): stiffness coefficient per local
    node (quadrature at the GLL nodes)."""
    Np = (N + 1) ** 3
    Nel = len(EX) // 8
    A = np.zeros((Ngather, Ngather))
    for e in range(Nel):
        Ae = element_matrix(N, EX[8 * e:8 * e + 8], EY[8 * e:8 * e + 8], EZ[8 * e:8 * e + 8],
                            qpts, qw, lam, stiff,
                            kappa=None if kappaL is None else kappaL[e * Np:(e + 1) * Np])
        for f in range(nfields):
            ids = G2L[e * Np * nfields + f * Np:e * Np * nfields + (f + 1) * Np]
            keep = ids >= 0
            A[np.ix_(ids[keep], ids[keep])] += Ae[np.ix_(keep, keep)]
    return A


def first_appearance_maps(ids):
    """Rows numbered by first appearance of each non-zero id in local order,
    members ascending: returns rowStarts, colIds, GlobalToLocal."""
    ids = np.asarray(ids)
    nz = np.nonzero(ids)[0]
    _, first, inv = np.unique(ids[nz], return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    row_of = rank[inv]
    G2L = -np.ones(len(ids), dtype=np.int64)
    G2L[nz] = row_of
    counts = np.bincount(row_of, minlength=len(order))
    rowStarts = np.concatenate([[0], np.cumsum(counts)])
    colIds = nz[np.argsort(row_of, kind="stable")]
    return rowStarts, colIds, G2L
