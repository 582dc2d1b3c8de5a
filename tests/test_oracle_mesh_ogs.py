"""Pins the oracle's mesh, geometric factors and ogs T maps (restating
libs/mesh/*Hex3D*.cpp, libs/mesh/meshConnectNodes.cpp,
libs/mesh/meshMaskedGatherScatterSetup.cpp, libs/ogs/ogsSetup.cpp and
libs/ogs/ogsOperator.cpp)."""
import numpy as np
import numpy.polynomial.legendre as L
import pytest

import brute


@pytest.mark.parametrize("N,nx,ny,nz", [(1, 2, 3, 2), (3, 2, 2, 2), (4, 3, 2, 1), (7, 2, 2, 3)])
def test_undistorted_geofactors_closed_form(oracle, N, nx, ny, nz):
    P = oracle.Problem(bp=5, N=N, nx=nx, ny=ny, nz=nz)
    _, w = brute.gll_nodes(N)
    hx, hy, hz = 2.0 / nx, 2.0 / ny, 2.0 / nz
    J = hx * hy * hz / 8
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    wJ = P.array("wJ").reshape(P.Nelements, -1)
    g = P.array("ggeo").reshape(P.Nelements, P.Np, 6)
    np.testing.assert_allclose(wJ, np.tile(J * W, (P.Nelements, 1)), rtol=1e-13)
    for c, h in zip((0, 3, 5), (hx, hy, hz)):
        np.testing.assert_allclose(g[:, :, c], np.tile(J * W * (2 / h) ** 2, (P.Nelements, 1)), rtol=1e-12)
    for c in (1, 2, 4):
        assert np.abs(g[:, :, c]).max() < 1e-12


@pytest.mark.parametrize("N", [2, 5, 7])
def test_perturbed_geofactors_match_analytic_trilinear_map(oracle, N):
    P = oracle.Problem(bp=5, N=N, nx=3, perturb=0.1, seed=5)
    x, w = brute.gll_nodes(N)
    EX, EY, EZ = P.array("EX"), P.array("EY"), P.array("EZ")
    wJ = P.array("wJ").reshape(P.Nelements, -1)
    g = P.array("ggeo").reshape(P.Nelements, P.Np, 6)
    xs = P.array("x").reshape(P.Nelements, -1)
    Nq = N + 1
    for e in [0, 13, 26]:
        for n in [0, 5, P.Np // 2, P.Np - 1]:
            i, j, k = n % Nq, (n // Nq) % Nq, n // (Nq * Nq)
            sl = slice(8 * e, 8 * e + 8)
            Jm = brute.trilinear_jacobian(EX[sl], EY[sl], EZ[sl], x[i], x[j], x[k])
            J = np.linalg.det(Jm)
            Ji = np.linalg.inv(Jm)
            G = J * w[i] * w[j] * w[k] * (Ji @ Ji.T)
            assert abs(wJ[e, n] - J * w[i] * w[j] * w[k]) < 1e-13 * abs(wJ[e, n])
            ref = [G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2]]
            np.testing.assert_allclose(g[e, n], ref, atol=1e-12 * np.abs(ref).max())
            p = brute.trilinear_point(EX[sl], EY[sl], EZ[sl], x[i], x[j], x[k])
            assert abs(xs[e, n] - p[0]) < 1e-14


def test_perturbation_keeps_boundary_and_volume(oracle):
    P = oracle.Problem(bp=5, N=4, nx=4, ny=3, nz=5, perturb=0.1, seed=9)
    assert abs(P.array("wJ").sum() - 8.0) < 1e-12  # GLL exact for the trilinear J
    Q = oracle.Problem(bp=5, N=4, nx=4, ny=3, nz=5)
    EX, EXu = P.array("EX"), Q.array("EX")
    moved = np.abs(EX - EXu) > 0
    assert moved.any()
    assert np.all(np.abs(EX[np.isclose(np.abs(EXu), 1.0)] - EXu[np.isclose(np.abs(EXu), 1.0)]) == 0)
    assert np.abs(EX - EXu).max() <= 0.1 * 0.5 + 1e-15


@pytest.mark.parametrize("N", [2, 4, 7])
def test_cubature_geofactors_match_analytic(oracle, N):
    P = oracle.Problem(bp=3, N=N, nx=2, perturb=0.1, seed=3)
    g, gw = L.leggauss(N + 2)
    cw = P.array("cubwJ").reshape(P.Nelements, -1)
    cg = P.array("cubggeo").reshape(P.Nelements, -1, 6)
    EX, EY, EZ = P.array("EX"), P.array("EY"), P.array("EZ")
    cNq = N + 2
    for e in range(P.Nelements):
        sl = slice(8 * e, 8 * e + 8)
        for n in [0, 7, cNq**3 // 2, cNq**3 - 1]:
            i, j, k = n % cNq, (n // cNq) % cNq, n // (cNq * cNq)
            Jm = brute.trilinear_jacobian(EX[sl], EY[sl], EZ[sl], g[i], g[j], g[k])
            J = np.linalg.det(Jm)
            Ji = np.linalg.inv(Jm)
            W = gw[i] * gw[j] * gw[k]
            G = J * W * (Ji @ Ji.T)
            assert abs(cw[e, n] - J * W) < 1e-12 * J * W
            ref = [G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2]]
            np.testing.assert_allclose(cg[e, n], ref, atol=1e-12 * np.abs(ref).max())
    assert abs(cw.sum() - 8.0) < 1e-12


def test_global_ids_group_coincident_nodes(oracle):
    P = oracle.Problem(bp=5, N=3, nx=3, ny=2, nz=2)
    gid = P.array("globalIds")
    xyz = np.round(np.stack([P.array("x"), P.array("y"), P.array("z")], 1), 9)
    _, inv_xyz = np.unique(xyz, axis=0, return_inverse=True)
    _, inv_gid = np.unique(gid, return_inverse=True)
    # same partition: equal coordinates <=> equal global id
    pairs = set(zip(inv_xyz.ravel().tolist(), inv_gid.ravel().tolist()))
    assert len(pairs) == len(set(inv_xyz.ravel().tolist())) == len(set(inv_gid.ravel().tolist()))
    # vmapP matches physical coordinates across faces
    vM, vP = P.array("vmapM"), P.array("vmapP")
    assert np.abs(xyz[vM] - xyz[vP]).max() < 1e-8


@pytest.mark.parametrize("bp,N,nx,ny,nz", [(5, 1, 3, 3, 3), (5, 3, 3, 2, 4), (5, 7, 2, 2, 2), (6, 2, 2, 3, 2)])
def test_ogs_counts_and_multiplicity(oracle, bp, N, nx, ny, nz):
    P = oracle.Problem(bp=bp, N=N, nx=nx, ny=ny, nz=nz)
    Nf = 3 if bp in (2, 4, 6) else 1
    assert P.Ngather == Nf * (nx * N - 1) * (ny * N - 1) * (nz * N - 1)
    Nq = N + 1
    assert P.nnzT == Nf * (nx * Nq - 2) * (ny * Nq - 2) * (nz * Nq - 2)
    rs = P.array("rowStartsT")
    mult = np.diff(rs)
    assert set(np.unique(mult)) <= {1, 2, 4, 8}
    # number of rows with multiplicity 8 = interior mesh vertices
    assert np.sum(mult == 8) == Nf * (nx - 1) * (ny - 1) * (nz - 1)


@pytest.mark.parametrize("bp,N,nx,mask", [(5, 2, 2, True), (5, 3, 3, True), (6, 2, 2, True), (5, 2, 3, False)])
def test_ogs_maps_equal_first_appearance_bruteforce(oracle, bp, N, nx, mask):
    P = oracle.Problem(bp=bp, N=N, nx=nx)
    if not mask:
        P.set_mask(False)
    rs, ci, g2l = brute.first_appearance_maps(P.array("maskedIds"))
    np.testing.assert_array_equal(P.array("rowStartsT"), rs)
    np.testing.assert_array_equal(P.array("colIdsT"), ci)
    np.testing.assert_array_equal(P.array("GlobalToLocal"), g2l)


def test_block_rows_hand_derived(oracle):
    # 3 rows of 300: bisect [0,3) at the first row start > 450 -> 2, then
    # [0,2) at first start > 300 -> 2 == end -> 1. Blocks: {0},{1},{2}.
    np.testing.assert_array_equal(oracle.block_rows([0, 300, 600, 900]), [0, 1, 2, 3])
    # fits in one block
    np.testing.assert_array_equal(oracle.block_rows([0, 100, 200]), [0, 2])
    # 4 rows of 200: total 800 -> mid start > 400 -> row 3 ([0,3) 600, [3,4) 200);
    # [0,3): > 0+300 -> row 2 ([0,2) 400, [2,3) 200).
    np.testing.assert_array_equal(oracle.block_rows([0, 200, 400, 600, 800]), [0, 2, 3, 4])
    with pytest.raises(ValueError):
        oracle.block_rows([0, 513])


def test_block_rows_invariants_on_mesh(oracle):
    P = oracle.Problem(bp=5, N=7, nx=4)
    rs = P.array("rowStartsT")
    bs = P.array("blockRowStartsT")
    assert bs[0] == 0 and bs[-1] == P.Ngather and np.all(np.diff(bs) > 0)
    assert np.max(rs[bs[1:]] - rs[bs[:-1]]) <= 512
    assert len(bs) - 1 == P.NrowBlocksT
