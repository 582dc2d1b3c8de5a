"""Pins the oracle's forcing (okl/rhsHex3D.okl:28-128) in all of its
branches: one field and three fields, affine (per-element J times the GLL
weights) and non-affine (stored wJ per node).

The independent construction builds every local node's position from the
analytic trilinear map of its element's 8 vertices at numpy's GLL nodes,
and its weight from the determinant of the analytic Jacobian times the
tensor GLL weights; it shares no code with the oracle. The forcing formulas
follow okl/rhsHex3D.okl:30-41. The one-field constant is the manufactured
solution's: -lap u + lam u = (3 pi^2 + lam) u for u = sin(pi x) sin(pi y)
sin(pi z), checked below by a finite-difference Laplacian. The three-field
constant is the reference's text, (4 pi^2 + lam), which is NOT -lap u_d +
lam u_d for its field u_d = sin(2 pi x_d) prod_{c != d} sin(pi x_c) (that
is (6 pi^2 + lam) u_d, also checked below): the reference's BP2/BP4/BP6
forcing is a right-hand side, not a manufactured solution, so its constant is
pinned by the reference source alone, while the independent assembly pins
the field layout [e][f][n], the node positions and the weights.
"""
import numpy as np
import pytest

import brute


def independent_forcing(P, N, nfields, lam):
    EX, EY, EZ = P.array("EX"), P.array("EY"), P.array("EZ")
    r, w = brute.gll_nodes(N)
    Nq = N + 1
    Np = Nq**3
    Nel = len(EX) // 8
    out = np.zeros(Nel * nfields * Np)
    for e in range(Nel):
        X, Y, Z = EX[8 * e:8 * e + 8], EY[8 * e:8 * e + 8], EZ[8 * e:8 * e + 8]
        for k in range(Nq):
            for j in range(Nq):
                for i in range(Nq):
                    n = i + j * Nq + k * Nq * Nq
                    x, y, z = brute.trilinear_point(X, Y, Z, r[i], r[j], r[k])
                    W = np.linalg.det(brute.trilinear_jacobian(X, Y, Z, r[i], r[j], r[k])) * w[i] * w[j] * w[k]
                    if nfields == 1:
                        out[e * Np + n] = W * (3 * np.pi**2 + lam) * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
                    else:
                        c = 4 * np.pi**2 + lam
                        s1 = np.sin(np.pi * np.array([x, y, z]))
                        s2 = np.sin(2 * np.pi * np.array([x, y, z]))
                        for d in range(3):
                            f = c * np.prod([s2[a] if a == d else s1[a] for a in range(3)])
                            out[e * nfields * Np + d * Np + n] = W * f
    return out


@pytest.mark.parametrize("bp", [5, 6, 2])
@pytest.mark.parametrize("affine", [False, True])
def test_forcing_matches_independent_assembly(oracle, bp, affine):
    N, lam = 3, 1.3
    P = oracle.Problem(bp=bp, N=N, nx=2, ny=2, nz=1, affine=affine, lam=lam,
                       perturb=0.0 if affine else 0.12, seed=5)
    nf = 3 if bp in (2, 4, 6) else 1
    got = P.rhs_forcing_local()
    ref = independent_forcing(P, N, nf, lam)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("bp", [5, 6])
def test_affine_forcing_equals_nonaffine_on_box(oracle, bp):
    """On the undistorted box the affine branch (J per element times the
    GLL weights) and the non-affine branch (stored wJ per node) agree."""
    kw = dict(bp=bp, N=5, nx=3, ny=2, nz=2, lam=0.7)
    a = oracle.Problem(affine=True, **kw).rhs_forcing_local()
    b = oracle.Problem(affine=False, **kw).rhs_forcing_local()
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("nfields", [1, 3])
def test_forcing_constants_against_the_laplacian(nfields):
    """-lap u + lam u by a 4th-order finite-difference Laplacian at random
    points: (3 pi^2 + lam) u for the one-field u (the reference's forcing
    constant), (6 pi^2 + lam) u_d for the three-field u_d (the reference
    uses 4 pi^2 + lam there, so its three-field forcing is not this u's)."""
    lam = 0.9
    rng = np.random.default_rng(3)
    h = 1e-3

    def u(p, d):
        s1 = np.sin(np.pi * p)
        if nfields == 1:
            return np.prod(s1)
        s2 = np.sin(2 * np.pi * p)
        return np.prod([s2[a] if a == d else s1[a] for a in range(3)])

    for _ in range(5):
        p = rng.uniform(-1, 1, 3)
        for d in range(nfields):
            lap = 0.0
            for a in range(3):
                e = np.zeros(3)
                e[a] = h
                lap += (-u(p + 2 * e, d) + 16 * u(p + e, d) - 30 * u(p, d) + 16 * u(p - e, d) - u(p - 2 * e, d)) / (12 * h * h)
            c = (3 if nfields == 1 else 6) * np.pi**2 + lam
            assert abs((-lap + lam * u(p, d)) - c * u(p, d)) <= 1e-6
