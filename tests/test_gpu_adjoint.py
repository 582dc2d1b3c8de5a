"""Variable-coefficient BP5 and the adjoint dL/dkappa on the GPU: the
kappa operator and the gradient kernel against the oracle (FP64, same
inputs), the whole pipeline (solve, adjoint solve, gradient) against the
oracle's own pipeline, central finite differences along 16 seeded
directions on the 16^3 p=7 problem, and the torch.autograd wrapper."""
import numpy as np
import pytest

import oracle_lib
from benchparanumal_amd import adjoint, capi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plat():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return capi.Platform(0)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_err(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


def pair(plat, N, nx, affine=False, lam=0.0, seed=5):
    kw = dict(N=N, nx=nx, ny=nx, nz=nx, affine=affine, perturb=0.0 if affine else 0.1, seed=seed)
    m = capi.Mesh(**kw)
    b = capi.BP(plat, m, 5, lam=lam)
    P = oracle_lib.Problem(bp=5, lam=lam, **kw)
    assert b.Nnodes == P.lattice_nodes()
    return m, b, P


def kappa_field(m, seed):
    return np.exp(adjoint.smooth_lattice_field(adjoint.lattice_shape(m), seed).ravel())


@pytest.mark.parametrize("N,affine", [(3, False), (5, True), (7, False), (7, True)])
def test_kappa_operator_matches_oracle(plat, N, affine):
    m, b, P = pair(plat, N, 3, affine=affine, lam=0.4)
    kap = kappa_field(m, 21)
    b.set_kappa(dev(kap))
    P.set_kappa(kap)
    q = np.random.default_rng(N).standard_normal(b.Ngather)
    Aq = torch.zeros(b.Ngather, dtype=torch.float64, device="cuda")
    b.operator(dev(q), Aq)
    assert rel_err(host(Aq), P.operator(q)) <= 1e-12
    b.set_kappa(None)  # constant coefficient again
    P.set_kappa(None)
    b.operator(dev(q), Aq)
    assert rel_err(host(Aq), P.operator(q)) <= 1e-12


@pytest.mark.parametrize("N,affine", [(3, False), (7, False), (7, True), (8, False)])
def test_kappa_gradient_kernel_matches_oracle(plat, N, affine):
    m, b, P = pair(plat, N, 3, affine=affine)
    kap = kappa_field(m, 22)
    b.set_kappa(dev(kap))
    P.set_kappa(kap)
    rng = np.random.default_rng(100 + N)
    u, lam = rng.standard_normal(b.Ngather), rng.standard_normal(b.Ngather)
    g = torch.zeros(b.Nnodes, dtype=torch.float64, device="cuda")
    b.kappa_gradient(dev(u), dev(lam), g)
    assert rel_err(host(g), P.kappa_grad(u, lam)) <= 1e-12


def objective_parts(b, m):
    """Lumped mass W (gathered GLL-weighted Jacobians) on the device."""
    o = capi.Ogs(mesh=m)
    wJ = torch.from_numpy(np.ascontiguousarray(m.array("wJ"))).cuda()
    W = torch.zeros(b.Ngather, dtype=torch.float64, device="cuda")
    o.gather(b.platform, wJ, W)
    return W


def test_pipeline_matches_oracle_pipeline(plat):
    """Solve, adjoint solve and gradient on the GPU against the oracle doing
    the same with its own CG (both to 1e-13)."""
    m, b, P = pair(plat, 5, 4, lam=0.0)
    kap = kappa_field(m, 23)
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal(b.Ngather)
    u_obs = rng.standard_normal(b.Ngather)
    W = host(objective_parts(b, m))
    kt = dev(kap).requires_grad_(True)
    u = adjoint.kappa_solve(b, kt, dev(rhs), tol=1e-13)
    L = 0.5 * ((u - dev(u_obs)) ** 2 * dev(W)).sum()
    L.backward()
    P.set_kappa(kap)
    u_ref, _, _, _ = P.cg(rhs, 50000, tol=1e-13)
    lam_ref, _, _, _ = P.cg(W * (u_ref - u_obs), 50000, tol=1e-13)
    g_ref = P.kappa_grad(u_ref, lam_ref)
    assert rel_err(host(u.detach()), u_ref) <= 1e-10
    assert rel_err(host(kt.grad), g_ref) <= 1e-8


def test_gradient_central_differences_16cube_p7(plat):
    """16^3 elements, p = 7: kappa = exp(g), g smoothed N(0,1) with std 0.5;
    u_obs = solve(kappa_2) + 1% noise; L = 1/2 (u - u_obs)^T W (u - u_obs);
    16 seeded directions d (relative: kappa*z), central differences with
    step h = 1e-6, against the adjoint gradient to 1e-6 relative.

    L(kappa +- h d) is evaluated from the converged u (A(kappa) u = f - r0):
    u+- = u + w+- with A(kappa +- h d) w+- = r0 -+ h A(d) u (lambda = 0, so
    A is linear in kappa and A(d) is the operator with coefficient d), which
    makes u+- the exact solutions of the perturbed systems; then
    L(u + w) - L(u) = (u - u_obs)^T W w + 1/2 w^T W w exactly (L is
    quadratic). The two large values of L are never subtracted, so the
    difference carries the corrections' solve error (relative 1e-13 of an
    O(h) quantity) instead of the rounding of L divided by 2h. The h^2
    truncation is checked by the same quotient at 2h."""
    m, b, _ = pair(plat, 7, 16, lam=0.0, seed=1)
    assert b.Ngather == 111 ** 3
    kap = kappa_field(m, 31)
    kap2 = kappa_field(m, 32)
    W = objective_parts(b, m)
    f = W.clone()  # b = M f with f = 1
    tol = 1e-13
    u2 = adjoint.kappa_solve(b, dev(kap2), f, tol=tol).detach()
    noise = torch.from_numpy(capi.random_normal(33, b.Ngather)).cuda()
    u_obs = u2 + 0.01 * u2.abs().mean() * noise

    kt = dev(kap).requires_grad_(True)
    u = adjoint.kappa_solve(b, kt, f, tol=tol)
    L = 0.5 * ((u - u_obs) ** 2 * W).sum()
    L.backward()
    grad = kt.grad.detach()
    u = u.detach()
    k0 = kt.detach()

    def apply(kappa, x):
        b.set_kappa(kappa)
        y = torch.zeros_like(x)
        b.operator(x, y)
        return y

    r0 = f - apply(k0, u)
    res = u - u_obs

    def central_difference(d, h):
        Adu = apply(d, u)
        diff = 0.0
        for sgn in (1.0, -1.0):
            b.set_kappa(k0 + sgn * h * d)
            w, its = adjoint._solve_relative(b, r0 - sgn * h * Adu, 1e-13, 100000)
            assert its < 100000
            diff += sgn * float((res * W * w).sum() + 0.5 * (w * W * w).sum())
        return diff / (2 * h)

    h = 1e-6
    errs, trunc = [], []
    for dseed in range(16):
        z = torch.from_numpy(capi.random_normal(1000 + dseed, b.Nnodes)).cuda()
        d = z * k0
        ad = float((grad * d).sum())
        fd = central_difference(d, h)
        errs.append(abs(fd - ad) / abs(ad))
        if dseed < 2:
            trunc.append(abs(central_difference(d, 2 * h) - fd) / abs(ad))
    print("central differences vs adjoint, relative:", ["%.2e" % e for e in errs])
    print("quotient at 2h vs h, relative:", ["%.2e" % e for e in trunc])
    assert max(errs) <= 1e-6, errs
    assert max(trunc) <= 1e-6, trunc


def test_autograd_gradient_in_b_is_adjoint(plat):
    m, b, P = pair(plat, 4, 3, lam=0.2)
    kap = kappa_field(m, 41)
    rng = np.random.default_rng(9)
    rhs = dev(rng.standard_normal(b.Ngather)).requires_grad_(True)
    c = dev(rng.standard_normal(b.Ngather))
    u = adjoint.kappa_solve(b, dev(kap), rhs, tol=1e-13)
    (u * c).sum().backward()
    # d(c.u)/db = A^{-1} c (A symmetric)
    P.set_kappa(kap)
    lam_ref, _, _, _ = P.cg(host(c), 50000, tol=1e-13)
    assert rel_err(host(rhs.grad), lam_ref) <= 1e-9


def test_kappa_errors(plat):
    m = capi.Mesh(N=3, nx=2, cubature=True)
    b3 = capi.BP(plat, m, 3, lam=1.0)
    k = torch.ones(b3.Nnodes, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.BpxError):
        b3.set_kappa(k)
    b5 = capi.BP(plat, capi.Mesh(N=3, nx=2), 5, lam=1.0)
    g = torch.zeros_like(k)
    u = torch.zeros(b5.Ngather, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.BpxError):
        b5.kappa_gradient(u, u, g)  # no coefficient set yet
