"""Pins the oracle's Ax loop nests (okl/bp{1..6}Ax{,Affine}Hex3D.okl), the
gather, the reductions and CG (libs/core/linearSolverCG.cpp) against dense
brute-force assembly, invariants and the README's sample run."""
import json
import os

import numpy as np
import numpy.polynomial.legendre as L
import pytest

import brute

LAM = 1.3
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dense_reference(P, N, bp):
    EX, EY, EZ = P.array("EX"), P.array("EY"), P.array("EZ")
    G2L = P.array("GlobalToLocal")
    nf = 3 if bp in (2, 4, 6) else 1
    if bp in (5, 6):
        q, w = brute.gll_nodes(N)
    else:
        q, w = L.leggauss(N + 2)
    stiff = bp in (3, 4, 5, 6)
    lam = LAM if stiff else 1.0
    return brute.assembled_operator(N, EX, EY, EZ, G2L, P.Ngather, q, w, lam, stiff, nf)


@pytest.mark.parametrize("bp", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("affine", [False, True])
def test_ax_matches_dense_bruteforce(oracle, bp, affine):
    N = 2 if bp in (2, 4, 6) else 3
    perturb = 0.0 if affine else 0.1
    P = oracle.Problem(bp=bp, N=N, nx=2, ny=2, nz=1 if bp in (2, 4, 6) else 2, affine=affine,
                       lam=LAM, perturb=perturb, seed=4)
    A = dense_reference(P, N, bp)
    rng = np.random.default_rng(bp)
    for _ in range(2):
        q = rng.standard_normal(P.Ngather)
        ref = A @ q
        got = P.operator(q)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("bp", [1, 3, 5])
def test_affine_equals_nonaffine_on_box(oracle, bp):
    P = oracle.Problem(bp=bp, N=7, nx=3, lam=LAM)
    Q = oracle.Problem(bp=bp, N=7, nx=3, lam=LAM, affine=True)
    q = np.random.default_rng(0).standard_normal(P.Ngather)
    a, b = P.operator(q), Q.operator(q)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("bp", [2, 4, 6])
def test_three_field_equals_three_single_fields(oracle, bp):
    P3 = oracle.Problem(bp=bp, N=3, nx=2, lam=LAM, perturb=0.1, seed=2)
    P1 = oracle.Problem(bp=bp - 1, N=3, nx=2, lam=LAM, perturb=0.1, seed=2)
    rng = np.random.default_rng(1)
    q1 = [rng.standard_normal(P1.Ngather) for _ in range(3)]
    # rows follow first appearance over the [e][f][n] local order, so map
    # field-f row r of the 1-field system to its 3-field row through the
    # two GlobalToLocal maps.
    Np = P1.Np
    g1 = P1.array("GlobalToLocal").reshape(-1, Np)
    g3 = P3.array("GlobalToLocal").reshape(-1, 3, Np)
    row3 = np.zeros((3, P1.Ngather), dtype=np.int64)
    for f in range(3):
        keep = g1 >= 0
        row3[f, g1[keep]] = g3[:, f, :][keep]
    q3 = np.zeros(P3.Ngather)
    for f in range(3):
        q3[row3[f]] = q1[f]
    out = P3.operator(q3)
    for f in range(3):
        ref = P1.operator(q1[f])
        assert np.abs(out[row3[f]] - ref).max() <= 1e-13 * np.abs(ref).max()


def test_operator_symmetric_positive(oracle):
    P = oracle.Problem(bp=5, N=5, nx=3, lam=LAM, perturb=0.1, seed=8)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(P.Ngather), rng.standard_normal(P.Ngather)
    a, b = y @ P.operator(x), x @ P.operator(y)
    assert abs(a - b) <= 1e-12 * abs(a)
    assert x @ P.operator(x) > 0


def test_unmasked_constant_gives_lambda_mass(oracle):
    # A*1 = lambda*M*1 without the Dirichlet mask (the stiffness kills constants)
    P = oracle.Problem(bp=5, N=4, nx=3, lam=LAM, perturb=0.1, seed=1)
    P.set_mask(False)
    one = np.ones(P.Ngather)
    Aone = P.operator(one)
    M1 = P.gather(P.array("wJ"))
    np.testing.assert_allclose(Aone, LAM * M1, rtol=1e-10, atol=1e-12)


def test_reductions_and_axpy(oracle):
    rng = np.random.default_rng(5)
    for n in [1, 255, 256, 257, 70000, 300001]:
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        assert abs(oracle.inner_prod(x, y) - np.dot(x, y)) <= 1e-13 * np.abs(x * y).sum()
        assert abs(oracle.norm2(x) - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(x)


def test_cg_textbook_properties(oracle):
    P = oracle.Problem(bp=5, N=3, nx=3, lam=1.0, perturb=0.1, seed=6)
    n = P.Ngather
    A = np.stack([P.operator(e) for e in np.eye(n)], 1)
    b = np.random.default_rng(2).standard_normal(n)
    x, it, hist, alpha = P.cg(b, 400, tol=1e-13)
    assert it < 400
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)
    # the recurrence residual equals the true residual
    assert abs(np.linalg.norm(b - A @ x) - hist[-1]) < 1e-9 * hist[0]
    assert hist[0] == pytest.approx(np.linalg.norm(b), rel=1e-14)


def test_cg_converges_to_exact_solution_reference_semantics(oracle):
    # BP5 with the reference forcing (okl/rhsHex3D.okl:33): -lap u + u = f has
    # u = sin(pi x) sin(pi y) sin(pi z) on [-1,1]^3; error falls with p.
    errs = []
    for N in (2, 4, 6):
        P = oracle.Problem(bp=5, N=N, nx=4, lam=1.0)
        x, it, hist, _ = P.cg(P.rhs_forcing(), 2000, tol=1e-12)
        uL = P.scatter(x)
        G2L = P.array("GlobalToLocal")
        X, Y, Z = P.array("x"), P.array("y"), P.array("z")
        exact = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
        errs.append(np.abs(uL - exact)[G2L >= 0].max())
    assert errs[0] > 10 * errs[1] > 100 * errs[2]
    assert errs[2] < 1e-4


@pytest.mark.slow
def test_cg_readme_sample(oracle):
    """README.md:109-113: BP5, p=4, 10^3 elements, 100 CG iterations. The
    sample counted 68,921 = 41^3 DOFs, i.e. it ran without the Dirichlet mask
    (README.md:115). On that problem rounding alone (FMA on/off) moves the
    iteration-100 norm by ~2e-3, so the match is soft."""
    g = json.load(open(os.path.join(GOLDEN, "bp5_p4_nx10_unmasked_cg100.json")))
    P = oracle.Problem(bp=g["bp"], N=g["N"], nx=g["nx"], lam=g["lambda"])
    P.set_mask(g["mask"])
    assert P.Ngather == g["Ndofs"]
    x, it, hist, alpha = P.cg(P.rhs_forcing(), g["iterations"])
    its = np.array(g["it"])
    assert it == g["iterations"]
    # hist[k] is the norm after iteration k; alpha[k-1] is iteration k's step
    np.testing.assert_allclose(hist[its], g["r_norm"], rtol=5e-3)
    np.testing.assert_allclose(alpha[its - 1], g["alpha"], rtol=3e-3)
