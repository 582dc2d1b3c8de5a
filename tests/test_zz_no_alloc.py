"""bpx_cg_solve_noalloc (the CG solve that allocates nothing, its work memory a
fixed device array) against bpx_cg_solve (the cg object with its own work
vectors, libs/core/linearSolverCG.cpp:62-156) on the same operator: the BP5
operator of a perturbed box, called back through the C ABI. Both run the
same kernels in the same order, so iterations, residual histories and
solutions agree bit for bit; the history is also checked against the
oracle's CG (1e-9 relative per iteration). Two calls in a row, the second on
a smaller problem, then the size limits. The file is collected last, so a
failure here hides no other test."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from benchparanumal_amd import capi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

OPFN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def solve_pair(plat, L, N, nx, maxit, tol, seed):
    m = capi.Mesh(N=N, nx=nx, perturb=0.1, seed=seed)
    bp = capi.BP(plat, m, 5, lam=1.0)
    n = bp.Ngather

    @OPFN
    def op(ctx, q, Aq):
        return L.bpx_bp_operator(bp.h, q, Aq)

    rhs = capi.random_normal(100 + seed, n)
    stream = L.bpx_platform_stream(plat.h)
    out = {}
    # the cg object (allocates its work vectors)
    h = C.c_void_p()
    capi.check(L.bpx_cg_create(plat.h, n, 0, C.byref(h)))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    r = torch.from_numpy(rhs).cuda()
    its = C.c_int()
    hist = np.zeros(maxit + 1)
    capi.check(L.bpx_cg_solve(h, op, None, capi.ptr(x), capi.ptr(r), tol, maxit, 0, C.byref(its),
                              hist.ctypes.data_as(C.c_void_p)))
    capi.check(L.bpx_cg_destroy(h))
    out["obj"] = (its.value, hist, host(x), host(r))
    # the allocation-free solve on the same stream
    x2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    r2 = torch.from_numpy(rhs).cuda()
    its2 = C.c_int()
    hist2 = np.zeros(maxit + 1)
    capi.check(L.bpx_cg_solve_noalloc(stream, n, op, None, capi.ptr(x2), capi.ptr(r2), tol, maxit, C.byref(its2),
                                      hist2.ctypes.data_as(C.c_void_p)))
    out["noalloc"] = (its2.value, hist2, host(x2), host(r2))
    P = oracle_lib.Problem(bp=5, N=N, nx=nx, lam=1.0, perturb=0.1, seed=seed)
    _, it_ref, hist_ref, _ = P.cg(rhs, maxit, tol=tol)
    return n, out, (it_ref, hist_ref)


def test_noalloc_cg_matches_cg_object_twice():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    oracle_lib.set_threads(16)
    plat = capi.Platform(0)
    L = capi.lib()
    sizes = []
    # first call 55^3 unknowns (p=7, 8^3 elements), then 19^3 (p=4, 5^3):
    # the second runs on the array the first left behind
    for N, nx, maxit, tol, seed in ((7, 8, 120, 0.0, 1), (4, 5, 400, 1e-10, 2)):
        n, out, (it_ref, hist_ref) = solve_pair(plat, L, N, nx, maxit, tol, seed)
        sizes.append(n)
        its, hist, x, r = out["obj"]
        its2, hist2, x2, r2 = out["noalloc"]
        print(f"n={n}: cg object {its} its, no-alloc {its2} its, oracle {it_ref}")
        assert its2 == its == it_ref
        np.testing.assert_array_equal(hist2, hist)
        np.testing.assert_array_equal(x2, x)
        np.testing.assert_array_equal(r2, r)
        # the oracle's history to the CG bar over the first 100 iterations of
        # the unconverged solve; the converged one (6,859 unknowns, 157
        # iterations to 1e-10) over its first 40, after which the two
        # orders of summation drift apart by more than 1e-9 (1e-5 measured
        # at a residual of 1e-6 of the start), then its count of iterations
        n = 100 if tol == 0 else 40
        np.testing.assert_allclose(hist2[:n], hist_ref[:n], rtol=1e-9)
        if tol > 0:
            assert its2 < maxit and hist2[its2] <= tol * hist2[0]
    assert sizes[1] < sizes[0]


def test_noalloc_cg_rejects_sizes_beyond_its_array():
    plat = capi.Platform(0)
    L = capi.lib()
    stream = L.bpx_platform_stream(plat.h)

    @OPFN
    def op(ctx, q, Aq):
        return 0

    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    its = C.c_int()
    big = 2097152 + 1
    assert L.bpx_cg_solve_noalloc(stream, big, op, None, capi.ptr(x), capi.ptr(x), 0.0, 1, C.byref(its),
                                  None) == -1
    assert b"N must be in 0..2097152" in L.bpx_last_error()
    assert L.bpx_cg_solve_noalloc(stream, 8, op, None, capi.ptr(x), capi.ptr(x), 0.0, 65536, C.byref(its),
                                  None) == -1
    assert b"maxit must be in 0..65535" in L.bpx_last_error()
    assert L.bpx_cg_solve_noalloc(stream, 8, None, None, capi.ptr(x), capi.ptr(x), 0.0, 1, C.byref(its),
                                  None) == -1
