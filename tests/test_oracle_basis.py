"""Pins the oracle's 1D basis (restating libs/mesh/meshBasis1D.cpp) against
closed forms and numpy's independent Legendre routines."""
import numpy as np
import numpy.polynomial.legendre as L
import pytest


def legendre(n):
    c = np.zeros(n + 1)
    c[n] = 1
    return L.Legendre(c)


@pytest.mark.parametrize("N", range(1, 16))
def test_gll_nodes_weights_closed_form(oracle, N):
    x, w = oracle.gll(N)
    interior = np.sort(legendre(N).deriv().roots().real) if N > 1 else np.zeros(0)
    ref = np.concatenate([[-1.0], interior, [1.0]])
    np.testing.assert_allclose(x, ref, atol=2e-15)
    PN = legendre(N)(ref)
    np.testing.assert_allclose(w, 2.0 / (N * (N + 1) * PN**2), rtol=1e-12)
    assert abs(w.sum() - 2) < 1e-14


@pytest.mark.parametrize("n", range(1, 18))
def test_gauss_matches_numpy(oracle, n):
    x, w = oracle.gauss(n - 1)  # JacobiGQ(0,0,N) has N+1 points
    xr, wr = L.leggauss(n)
    np.testing.assert_allclose(x, xr, atol=2e-15)
    np.testing.assert_allclose(w, wr, rtol=1e-12)


@pytest.mark.parametrize("N", range(1, 16))
def test_dmatrix_closed_form_and_exactness(oracle, N):
    x, _ = oracle.gll(N)
    D = oracle.dmatrix(N, x, x)
    PN = legendre(N)(x)
    Dc = np.zeros((N + 1, N + 1))
    for i in range(N + 1):
        for j in range(N + 1):
            if i != j:
                Dc[i, j] = PN[i] / (PN[j] * (x[i] - x[j]))
    Dc[0, 0] = -N * (N + 1) / 4.0
    Dc[N, N] = N * (N + 1) / 4.0
    scale = np.abs(Dc).max()
    assert np.abs(D - Dc).max() / scale < 1e-12
    # exact on monomials of degree <= N
    for k in range(N + 1):
        d = D @ x**k
        ref = k * x ** (k - 1) if k > 0 else np.zeros_like(x)
        assert np.abs(d - ref).max() < 1e-11 * max(1.0, N**2)


@pytest.mark.parametrize("N", range(1, 16))
def test_interp_gll_to_gauss_exact(oracle, N):
    x, _ = oracle.gll(N)
    g, _ = oracle.gauss(N + 1)  # p+2 Gauss points (meshCubatureSetupHex3D.cpp:35-38)
    I = oracle.interp(N, x, g)
    assert I.shape == (N + 2, N + 1)
    rng = np.random.default_rng(N)
    for _ in range(3):
        c = rng.standard_normal(N + 1)
        p = np.polynomial.Polynomial(c)
        np.testing.assert_allclose(I @ p(x), p(g), atol=1e-12 * np.abs(c).sum())
    # GLL->Gauss interpolation is centro-symmetric (the OKL half-loops use it)
    np.testing.assert_allclose(I, I[::-1, ::-1], atol=1e-14)


@pytest.mark.parametrize("N", range(1, 16))
def test_mass_matrix_is_exact_gram_matrix(oracle, N):
    x, w = oracle.gll(N)
    MM = oracle.mass1d(N, x)
    g, gw = L.leggauss(N + 2)
    # Lagrange basis of the GLL nodes evaluated at Gauss points (barycentric)
    lam = np.array([1.0 / np.prod([x[j] - x[m] for m in range(N + 1) if m != j]) for j in range(N + 1)])
    Lg = np.zeros((len(g), N + 1))
    for a, ga in enumerate(g):
        d = ga - x
        Lg[a] = lam / d / np.sum(lam / d)
    Mref = (Lg * gw[:, None]).T @ Lg
    np.testing.assert_allclose(MM, Mref, atol=1e-13)
    np.testing.assert_allclose(MM.sum(axis=1), w, rtol=1e-12)


@pytest.mark.parametrize("N", [0, 1, 4, 9])
def test_jacobi_p_orthonormal(oracle, N):
    g, gw = L.leggauss(N + 2)
    vals = np.array([[oracle.lib().orc_jacobi_p(t, 0.0, 0.0, n) for t in g] for n in range(N + 1)])
    G = (vals * gw) @ vals.T
    np.testing.assert_allclose(G, np.eye(N + 1), atol=1e-13)


def test_splitmix64_published_vectors(oracle):
    # Vigna's splitmix64 reference, seeds 1234567 and 0
    assert [oracle.rng_u64(1234567, i) for i in range(5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]
    assert oracle.rng_u64(0, 0) == 0xE220A8397B1DCDAF
    assert oracle.rng_u64(0, 1) == 0x6E789E6AA1B965F4


def test_rng_normal_moments(oracle):
    v = np.array([oracle.rng_normal(3, i) for i in range(20000)])
    assert abs(v.mean()) < 0.03 and abs(v.std() - 1) < 0.03
