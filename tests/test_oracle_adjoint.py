"""Pins the oracle's variable-coefficient BP5 operator and its adjoint
gradient (an extension with no reference counterpart) against dense
brute-force assembly on the trilinear map, central finite differences of
the dense objective, and the homogeneity identity of a linear coefficient."""
import numpy as np
import pytest

import brute

RNG = np.random.default_rng


def lattice_kappa(P, seed, spread=0.5):
    return np.exp(spread * RNG(seed).standard_normal(P.lattice_nodes()))


def test_lattice_index_matches_node_connectivity(oracle):
    """Two local nodes share a lattice point iff the reference's node
    connectivity (ConnectNodes ids) makes them one node; on the unperturbed
    box the lattice point is where the node sits."""
    N, nx, ny, nz = 3, 2, 3, 2
    P = oracle.Problem(bp=5, N=N, nx=nx, ny=ny, nz=nz, perturb=0.15, seed=2)
    lat = P.lattice_index()
    gid = P.array("globalIds")
    pairs = set(zip(lat.tolist(), gid.tolist()))
    assert len(pairs) == len(set(lat.tolist())) == len(set(gid.tolist()))
    assert P.lattice_nodes() == (nx * N + 1) * (ny * N + 1) * (nz * N + 1) == len(set(lat.tolist()))

    Q = oracle.Problem(bp=5, N=N, nx=nx, ny=ny, nz=nz)
    r, _ = brute.gll_nodes(N)
    LX, LY = nx * N + 1, ny * N + 1
    lat = Q.lattice_index()
    X, Y, Z = lat % LX, (lat // LX) % LY, lat // (LX * LY)
    for coord, I, n in ((Q.array("x"), X, nx), (Q.array("y"), Y, ny), (Q.array("z"), Z, nz)):
        h = 2.0 / n
        expect = -1.0 + h * (I // N + (1 + r[I % N]) / 2)
        np.testing.assert_allclose(coord, expect, rtol=0, atol=1e-14)


@pytest.mark.parametrize("N,affine", [(2, False), (3, False), (3, True)])
def test_kappa_operator_matches_dense_assembly(oracle, N, affine):
    P = oracle.Problem(bp=5, N=N, nx=2, ny=2, nz=3, affine=affine, perturb=0.0 if affine else 0.2, seed=5,
                       lam=0.7)
    kap = lattice_kappa(P, 11)
    P.set_kappa(kap)
    kL = kap[P.lattice_index()]
    x, w = brute.gll_nodes(N)
    A = brute.assembled_operator(N, P.array("EX"), P.array("EY"), P.array("EZ"), P.array("GlobalToLocal"),
                                 P.Ngather, x, w, 0.7, kappaL=kL)
    q = RNG(3).standard_normal(P.Ngather)
    ref = A @ q
    got = P.operator(q)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    P.set_kappa(None)  # back to the constant coefficient
    A1 = brute.assembled_operator(N, P.array("EX"), P.array("EY"), P.array("EZ"), P.array("GlobalToLocal"),
                                  P.Ngather, x, w, 0.7)
    assert np.abs(P.operator(q) - A1 @ q).max() <= 1e-12 * np.abs(A1 @ q).max()


@pytest.mark.parametrize("affine", [False, True])
def test_kappa_gradient_matches_dense_finite_differences(oracle, affine):
    """L = 1/2 (u - u_obs)^T W (u - u_obs), A(kappa) u = b, on a mesh small
    enough for dense solves; every lattice direction by central differences."""
    N = 2
    P = oracle.Problem(bp=5, N=N, nx=2, ny=2, nz=2, affine=affine, perturb=0.0 if affine else 0.2, seed=7,
                       lam=0.3)
    kap = lattice_kappa(P, 13)
    x, w = brute.gll_nodes(N)
    lat = P.lattice_index()
    EX, EY, EZ, G2L = P.array("EX"), P.array("EY"), P.array("EZ"), P.array("GlobalToLocal")

    def dense(k):
        return brute.assembled_operator(N, EX, EY, EZ, G2L, P.Ngather, x, w, 0.3, kappaL=k[lat])

    b = RNG(1).standard_normal(P.Ngather)
    u_obs = RNG(2).standard_normal(P.Ngather)
    W = np.diag(RNG(4).uniform(0.5, 1.5, P.Ngather))

    def L(k):
        u = np.linalg.solve(dense(k), b)
        return 0.5 * (u - u_obs) @ W @ (u - u_obs)

    A = dense(kap)
    u = np.linalg.solve(A, b)
    lam = np.linalg.solve(A, W @ (u - u_obs))
    P.set_kappa(kap)
    grad = P.kappa_grad(u, lam)
    h = 1e-6
    fd = np.zeros_like(kap)
    for p in range(kap.size):
        e = np.zeros_like(kap)
        e[p] = h * kap[p]
        fd[p] = (L(kap + e) - L(kap - e)) / (2 * h * kap[p])
    assert np.abs(grad - fd).max() <= 1e-6 * np.abs(fd).max()
    # lattice points on the boundary still carry sensitivity (kappa there
    # weights the stiffness of the boundary elements)
    assert np.count_nonzero(np.abs(fd) > 1e-8 * np.abs(fd).max()) > kap.size // 2


@pytest.mark.parametrize("affine", [False, True])
def test_kappa_gradient_homogeneity(oracle, affine):
    """The stiffness part K(kappa) is linear in kappa, so for any u and lam
    sum_p kappa_p dL/dkappa_p = -lam^T K(kappa) u."""
    P = oracle.Problem(bp=5, N=4, nx=3, ny=2, nz=3, affine=affine, perturb=0.0 if affine else 0.15, seed=9,
                       lam=1.0)
    kap = lattice_kappa(P, 17)
    P.set_kappa(kap)
    u = RNG(5).standard_normal(P.Ngather)
    lam = RNG(6).standard_normal(P.Ngather)
    grad = P.kappa_grad(u, lam)
    oracle.lib().orc_set_lambda(P.h, 0.0)
    Ku = P.operator(u)
    lhs, rhs = kap @ grad, -(lam @ Ku)
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(kap * grad).sum())
