"""Output and verification on the GPU: the solution written as a VTU file
(BP/BP5/src/bp5Run.cpp:220-238) and the error norms against the forcing's
manufactured solution, checked against numpy on the same solution and
for spectral convergence in p."""
import numpy as np
import pytest

import oracle_lib
from benchparanumal_amd import capi
from test_output_vtu import read_vtu

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plat():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return capi.Platform(0)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def solve_forcing(plat, bp, N, nx, affine=False, lam=1.0):
    m = capi.Mesh(N=N, nx=nx, affine=affine, cubature=bp <= 4)
    b = capi.BP(plat, m, bp, lam=lam)
    rhs = torch.zeros(b.Ngather, dtype=torch.float64, device="cuda")
    b.rhs_forcing(rhs)
    x = torch.zeros_like(rhs)
    its, _ = b.solve(x, rhs.clone(), tol=1e-13, maxit=20000, history=False)
    assert its < 20000
    return m, b, x


@pytest.mark.parametrize("bp,affine", [(5, False), (5, True), (3, False), (1, False)])
def test_error_norms_match_numpy_and_converge(plat, bp, affine):
    errs = []
    for N in (2, 4, 6):
        m, b, x = solve_forcing(plat, bp, N, 3, affine=affine)
        got = b.error_norms(x)
        P = oracle_lib.Problem(bp=bp, N=N, nx=3, affine=affine, lam=1.0)
        uL = P.scatter(host(x))
        X, Y, Z = P.array("x"), P.array("y"), P.array("z")
        c = 3 * np.pi**2 if bp <= 2 else 1.0
        ex = c * np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
        if affine:
            w = P.array("gllw")
            Nq = N + 1
            W = np.repeat(P.array("wJ"), Nq**3) * np.einsum("k,j,i->kji", w, w, w).ravel()[None, :].repeat(
                P.Nelements, 0).ravel()
        else:
            W = P.array("wJ")
        d = uL - ex
        assert got["max"] == pytest.approx(np.abs(d).max(), rel=1e-12)
        assert got["l2"] == pytest.approx(np.sqrt((W * d * d).sum()), rel=1e-10)
        assert got["rel_l2"] == pytest.approx(np.sqrt((W * d * d).sum() / (W * ex * ex).sum()), rel=1e-10)
        errs.append(got["rel_l2"])
    assert errs[0] > 20 * errs[1] > 400 * errs[2]


def test_error_norms_three_fields_unsupported(plat):
    b = capi.BP(plat, capi.Mesh(N=2, nx=2), 6, lam=1.0)
    x = torch.zeros(b.Ngather, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.BpxError):
        b.error_norms(x)


@pytest.mark.parametrize("bp", [5, 6])
def test_plot_solution_vtu(plat, tmp_path, bp):
    m = capi.Mesh(N=3, nx=2, ny=2, nz=1, perturb=0.1, seed=3)
    b = capi.BP(plat, m, bp, lam=1.0)
    u = torch.from_numpy(capi.random_normal(5, b.Ngather)).cuda()
    path = tmp_path / "sol.vtu"
    b.plot_solution(u, str(path))
    npts, ncells, arr = read_vtu(path)
    assert npts == m.Nelements * m.Np and ncells == m.Nelements * 6 * 27
    P = oracle_lib.Problem(bp=bp, N=3, nx=2, ny=2, nz=1, perturb=0.1, seed=3, lam=1.0)
    uL = P.scatter(host(u))  # Dirichlet nodes stay 0
    nf = b.Nfields
    fld = arr["Fields"][0].reshape(m.Nelements, m.Np, nf)
    np.testing.assert_allclose(fld, uL.reshape(m.Nelements, nf, m.Np).transpose(0, 2, 1), rtol=1e-5, atol=1e-6)
