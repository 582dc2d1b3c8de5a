"""Pins of the oracle's quadrilateral, triangle and tetrahedron problems
(oracle/other.cpp) by independent numpy computations: the warp & blend node
sets (edge nodes are the GLL nodes, the sets are symmetric), the mass and
derivative matrices (exact on polynomials, against closed-form monomial
integrals), the packed stiffness matrices (= D_a^T MM D_b by numpy), areas
and volumes, the node grouping (= coincident coordinates), and the
operators (energies of linear functions, the mass of the domain, the
multi-field forms, and the spectral convergence of the forced solve)."""
import math

import numpy as np
import pytest

import oracle_lib

TRI, QUAD, TET = 3, 4, 6


def mono_tri(a, b):
    """integral of r^a s^b over the reference triangle (r, s >= -1, r + s <= 0)"""
    # r = 2x - 1, s = 2y - 1 on x, y >= 0, x + y <= 1, dr ds = 4 dx dy
    tot = 0.0
    for i in range(a + 1):
        for j in range(b + 1):
            c = math.comb(a, i) * math.comb(b, j) * (2 ** i) * (2 ** j) * (-1) ** (a - i + b - j)
            tot += c * math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)
    return 4 * tot


def mono_tet(a, b, c):
    tot = 0.0
    for i in range(a + 1):
        for j in range(b + 1):
            for k in range(c + 1):
                co = (math.comb(a, i) * math.comb(b, j) * math.comb(c, k) * 2 ** (i + j + k)
                      * (-1) ** (a - i + b - j + c - k))
                tot += co * math.factorial(i) * math.factorial(j) * math.factorial(k) / math.factorial(i + j + k + 3)
    return 8 * tot


def problem(el, N, bp=5, perturb=0.0, **kw):
    if el == TET:
        kw.setdefault("nx", 2), kw.setdefault("ny", 2), kw.setdefault("nz", 2)
    else:
        kw.setdefault("nx", 3), kw.setdefault("ny", 2)
    return oracle_lib.Problem(bp=bp, N=N, element=el, perturb=perturb, seed=5, **kw)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 6, 9])
def test_tri_nodes_edges_are_gll_and_symmetric(N):
    P = problem(TRI, N)
    r, s = P.array("r"), P.array("s")
    assert len(r) == (N + 1) * (N + 2) // 2
    gll = np.polynomial.legendre.Legendre.basis(N).deriv().roots() if N > 1 else np.array([])
    gll = np.sort(np.concatenate([[-1.0, 1.0], gll]))
    edge = np.sort(r[np.abs(s + 1) < 1e-12])
    np.testing.assert_allclose(edge, gll, atol=1e-13)
    # barycentric coordinates; the set is invariant under swapping them
    L = np.stack([-(r + s) / 2, (1 + r) / 2, (1 + s) / 2], axis=1)
    key = lambda A: np.round(np.sort(A.view([("", A.dtype)] * A.shape[1]), axis=0).view(A.dtype).reshape(A.shape), 12)
    for perm in ([1, 0, 2], [0, 2, 1], [2, 1, 0]):
        a = np.round(L, 12)
        b = np.round(L[:, perm], 12)
        assert {tuple(v) for v in a} == {tuple(v) for v in b}


@pytest.mark.parametrize("N", [1, 2, 3, 5, 8])
def test_tet_nodes_edges_are_gll_and_symmetric(N):
    """Edge nodes are the GLL nodes; the node set is invariant under the
    permutations of the barycentric coordinates. (The face nodes are the
    triangle's warp & blend construction with the tetrahedron's alpha, so
    they match the triangle's own nodes only where alpha plays no part.)"""
    P = problem(TET, N)
    r, s, t = P.array("r"), P.array("s"), P.array("t")
    assert len(r) == (N + 1) * (N + 2) * (N + 3) // 6
    gll = np.polynomial.legendre.Legendre.basis(N).deriv().roots() if N > 1 else np.array([])
    gll = np.sort(np.concatenate([[-1.0, 1.0], gll]))
    edge = np.sort(r[(np.abs(s + 1) < 1e-12) & (np.abs(t + 1) < 1e-12)])
    np.testing.assert_allclose(edge, gll, atol=1e-13)
    if N <= 3:  # one or no interior face node: the faces are the triangle's nodes
        T = problem(TRI, N)
        face = np.abs(t + 1) < 1e-12
        fa = {tuple(v) for v in np.round(np.stack([r[face], s[face]], 1), 12)}
        tb = {tuple(v) for v in np.round(np.stack([T.array("r"), T.array("s")], 1), 12)}
        assert fa == tb
    L = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2], axis=1)
    base = {tuple(v) for v in np.round(L, 12)}
    for perm in ([1, 0, 2, 3], [0, 2, 1, 3], [0, 1, 3, 2], [3, 2, 1, 0]):
        assert {tuple(v) for v in np.round(L[:, perm], 12)} == base


@pytest.mark.parametrize("el,N", [(TRI, 1), (TRI, 3), (TRI, 6), (TET, 1), (TET, 2), (TET, 4)])
def test_simplex_mass_and_derivatives_exact(el, N):
    P = problem(el, N)
    Np = P.Np
    MM = P.array("MM").reshape(Np, Np)
    Dsim = P.array("Dsim").reshape(-1, Np, Np)
    r, s = P.array("r"), P.array("s")
    t = P.array("t") if el == TET else None
    if el == TRI:
        exps = [(a, b) for a in range(N + 1) for b in range(N + 1 - a)]
        val = lambda e: r ** e[0] * s ** e[1]
        integ = lambda e, f: mono_tri(e[0] + f[0], e[1] + f[1])
    else:
        exps = [(a, b, c) for a in range(N + 1) for b in range(N + 1 - a) for c in range(N + 1 - a - b)]
        val = lambda e: r ** e[0] * s ** e[1] * t ** e[2]
        integ = lambda e, f: mono_tet(e[0] + f[0], e[1] + f[1], e[2] + f[2])
    for e in exps:
        for f in exps:
            assert abs(val(e) @ MM @ val(f) - integ(e, f)) <= 1e-11 * max(1.0, abs(integ(e, f)))
        # D_d applied to the nodal values of the monomial: its derivative
        for d in range(len(e)):
            de = list(e)
            if e[d] == 0:
                exact = np.zeros(Np)
            else:
                de[d] -= 1
                exact = e[d] * val(de)
            np.testing.assert_allclose(Dsim[d] @ val(e), exact, atol=1e-9 * max(1, N ** 2))


@pytest.mark.parametrize("el,N", [(TRI, 2), (TRI, 5), (TET, 2), (TET, 3)])
def test_simplex_stiffness_is_dt_m_d(el, N):
    P = problem(el, N)
    Np = P.Np
    MM = P.array("MM").reshape(Np, Np)
    D = P.array("Dsim").reshape(-1, Np, Np)
    S = P.array("Ssim").reshape(-1, Np, Np)
    dim = D.shape[0]
    pairs = [(a, b) for a in range(dim) for b in range(a, dim)]
    for c, (a, b) in enumerate(pairs):
        ref = D[a].T @ MM @ D[b]
        if a != b:
            ref = ref + D[b].T @ MM @ D[a]
        np.testing.assert_allclose(S[c], ref, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("el", [TRI, QUAD, TET])
@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_geometry_measures_and_grouping(el, perturb):
    P = problem(el, 3, bp=3, perturb=perturb)
    wJ = P.array("wJ")
    ref = {TRI: 2.0, TET: 4.0 / 3.0, QUAD: 1.0}[el]
    area = wJ.sum() * ref
    assert abs(area - (8.0 if el == TET else 4.0)) <= 1e-12
    if el == QUAD:
        assert abs(P.array("cubwJ").sum() - 4.0) <= 1e-12
    # nodes with the same global id are the nodes at the same point
    x, y = P.array("x"), P.array("y")
    pts = np.stack([x, y] + ([P.array("z")] if el == TET else []), 1)
    g = P.array("globalIds")
    keys = [tuple(v) for v in np.round(pts, 9)]
    first = {}
    for k, gid in zip(keys, g):
        first.setdefault(k, gid)
        assert first[k] == gid
    assert len(set(keys)) == len(set(g.tolist()))


def assembled(P):
    """the unmasked assembled operator as a dense matrix (small problems)"""
    n = P.Ngather
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = P.operator(e)
    return A


@pytest.mark.parametrize("el", [TRI, QUAD, TET])
def test_energy_of_linear_functions_and_domain_mass(el):
    """On the unperturbed box (affine elements) with no mask: u^T A u for a
    linear u is |grad u|^2 |Omega| + lambda int u^2 (exactly integrated), and
    the mass operator's total is |Omega|."""
    lam = 0.7
    P = problem(el, 3, bp=5, lam=lam)
    P.set_mask(False)
    vol = 8.0 if el == TET else 4.0
    x, y = P.array("x"), P.array("y")
    z = P.array("z") if el == TET else np.zeros_like(x)
    # node value by global id (the first copy; copies agree)
    G2L = P.array("GlobalToLocal")
    n = P.Ngather
    X = np.zeros(n)
    Y = np.zeros(n)
    Z = np.zeros(n)
    X[G2L] = x
    Y[G2L] = y
    Z[G2L] = z
    A = assembled(P)
    for c in ((1.0, 0.0, 0.0), (0.3, -1.2, 0.0), (0.5, 0.25, -0.75)):
        if el != TET and c[2] != 0:
            continue
        u = c[0] * X + c[1] * Y + c[2] * Z
        grad2 = c[0] ** 2 + c[1] ** 2 + c[2] ** 2
        # int u^2 over the box [-1,1]^d: sum_i c_i^2 * (2/3) * vol / 2 (cross terms vanish)
        int_u2 = (c[0] ** 2 + c[1] ** 2 + c[2] ** 2) * vol / 3.0
        assert abs(u @ A @ u - (grad2 * vol + lam * int_u2)) <= 1e-10 * (grad2 * vol + lam * int_u2)
    M = problem(el, 3, bp=1)
    M.set_mask(False)
    one = np.ones(M.Ngather)
    assert abs(one @ M.operator(one) - vol) <= 1e-11


@pytest.mark.parametrize("el", [TRI, QUAD, TET])
@pytest.mark.parametrize("bp", [2, 4, 6])
def test_vector_forms_are_per_field(el, bp):
    P = problem(el, 2, bp=bp, perturb=0.1)
    S = problem(el, 2, bp=bp - 1, perturb=0.1)
    nf = P.Nfields
    assert nf == (3 if el == TET else 2)
    rng = np.random.default_rng(bp)
    qs = rng.standard_normal(S.Ngather)
    # field f of the vector problem is the scalar problem on ids * nf + f
    q = np.zeros(P.Ngather)
    Ls, Lv = S.array("GlobalToLocal"), P.array("GlobalToLocal")
    Np, Nel = S.Np, S.Nelements
    for f in range(nf):
        for e in range(Nel):
            for k in range(Np):
                gs, gv = Ls[e * Np + k], Lv[(e * nf + f) * Np + k]
                if gs >= 0:
                    q[gv] = qs[gs] * (f + 1)
    Aq = P.operator(q)
    As = S.operator(qs)
    for f in range(nf):
        for e in range(Nel):
            for k in range(Np):
                gs, gv = Ls[e * Np + k], Lv[(e * nf + f) * Np + k]
                if gs >= 0:
                    assert abs(Aq[gv] - (f + 1) * As[gs]) <= 1e-12 * np.abs(As).max()


def test_forced_solve_converges_spectrally_on_quads():
    """Quad BP5 with the forcing: the CG solution at the nodes approaches the
    manufactured solution sin(pi x) sin(pi y), faster with the degree (pins
    the forcing and the operator together). The simplex boxes are left out:
    the reference masks only nodes on an element's own boundary face, and a
    triangle or tetrahedron that meets the boundary at a vertex or an edge
    leaves that copy free (libs/mesh/meshMaskedGatherScatterSetup.cpp:44-58
    assumes a brick), so their discrete problem is not the Dirichlet one."""
    errs = []
    for N in (2, 4):
        P = problem(QUAD, N, bp=5, perturb=0.1, lam=1.0)
        b = P.rhs_forcing()
        x, its, _, _ = P.cg(b, 2000, tol=1e-12)
        xs, ys = P.array("x"), P.array("y")
        G2L = P.array("GlobalToLocal")
        u = np.sin(np.pi * xs) * np.sin(np.pi * ys)
        m = G2L >= 0
        errs.append(np.abs(x[G2L[m]] - u[m]).max())
    assert errs[1] < errs[0] / 10 and errs[1] < 5e-3, errs


@pytest.mark.parametrize("el", [TRI, TET])
def test_simplex_forcing_integrates_spectrally(el):
    """The local forcing is MM (J f) per element, so its dot with the nodal
    values of u = sin sin (sin) is the integral of the interpolants of f u:
    it approaches (2 pi^2 + lambda) in 2D, (3 pi^2 + lambda) in 3D."""
    lam = 1.3
    exact = (2 if el == TRI else 3) * np.pi ** 2 + lam
    errs = []
    for N in (3, 7):
        P = problem(el, N, bp=5, perturb=0.1, lam=lam)
        rl = P.rhs_forcing_local()
        xs, ys = P.array("x"), P.array("y")
        u = np.sin(np.pi * xs) * np.sin(np.pi * ys)
        if el == TET:
            u = u * np.sin(np.pi * P.array("z"))
        errs.append(abs(rl @ u - exact) / exact)
    assert errs[1] < errs[0] / 20 and errs[1] < 1e-3, errs
