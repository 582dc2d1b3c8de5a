"""k right-hand sides at once (BP5): every right-hand side of the batched
operator equals the single-vector operator to 1e-15 relative (same
arithmetic), and the single-vector operator is itself pinned against the
oracle; one case is also checked against the oracle directly. The FP32
k-RHS kernel is checked against the FP32 single-vector kernel and against
the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from benchparanumal_amd import capi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

KRHS_RTOL = 1e-15  # sweep and sheet forms: the single-vector kernel's operations, in its order
PENCIL_RTOL = 1e-13  # pencil form (chunk 0 up to N = 9, or -2): the r and s divergence parts summed apart


def form_rtol(N, chunk):
    pencil = chunk == -2 or (chunk == 0 and N <= 9)
    return PENCIL_RTOL if pencil else KRHS_RTOL


@pytest.fixture(scope="module")
def plat():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return capi.Platform(0)


def rel_err(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def single_columns(b, qk, K, local):
    """Apply the single-vector path to each column of qk ([n][K])."""
    n_out = b.NLocal if local else b.Ngather
    out = np.zeros((n_out, K))
    for r in range(K):
        q = torch.from_numpy(np.ascontiguousarray(qk[:, r])).cuda()
        A = torch.zeros(n_out, dtype=torch.float64, device="cuda")
        (b.ax_local if local else b.operator)(q, A)
        out[:, r] = host(A)
    return out


def oracle_arrays(P):
    return {k: torch.from_numpy(np.ascontiguousarray(P.array(k))).cuda()
            for k in ("elementList", "GlobalToLocal", "wJ", "ggeo", "gllw", "D")}


def krhs_kernel(plat, P, arr, qk, K, chunk, affine, lam):
    """bpx_bp5_ax[_affine]_hex_krhs on the oracle problem's arrays."""
    out = torch.zeros(K * P.NLocal, dtype=torch.float64, device="cuda")
    qd = torch.from_numpy(np.ascontiguousarray(qk.T.ravel())).cuda()  # right-hand side after right-hand side
    fn = capi.lib().bpx_bp5_ax_affine_hex_krhs if affine else capi.lib().bpx_bp5_ax_hex_krhs
    capi.check(fn(plat.h, P.N, K, chunk, P.Nelements, capi.ptr(arr["elementList"]), capi.ptr(arr["GlobalToLocal"]),
                  capi.ptr(arr["wJ"]), None, capi.ptr(arr["ggeo"]), capi.ptr(arr["gllw"]), capi.ptr(arr["D"]),
                  None, None, lam, capi.ptr(qd), P.Ngather, capi.ptr(out), P.NLocal))
    return host(out).reshape(K, P.NLocal).T


def single_kernel(plat, P, arr, qk, K, affine, lam):
    """bpx_bp5_ax[_affine]_hex once per right-hand side, same arrays."""
    out = np.zeros((P.NLocal, K))
    fn = capi.lib().bpx_bp5_ax_affine_hex if affine else capi.lib().bpx_bp5_ax_hex
    for r in range(K):
        qd = torch.from_numpy(np.ascontiguousarray(qk[:, r])).cuda()
        o = torch.zeros(P.NLocal, dtype=torch.float64, device="cuda")
        capi.check(fn(plat.h, P.N, P.Nelements, capi.ptr(arr["elementList"]), capi.ptr(arr["GlobalToLocal"]),
                      capi.ptr(arr["wJ"]), None, capi.ptr(arr["ggeo"]), capi.ptr(arr["gllw"]), capi.ptr(arr["D"]),
                      None, None, lam, capi.ptr(qd), capi.ptr(o)))
        out[:, r] = host(o)
    return out


@pytest.mark.parametrize("affine", [False, True])
@pytest.mark.parametrize("N,K,chunk", [(7, 1, 1), (7, 2, 2), (7, 4, 4), (7, 8, 2), (7, 8, 4), (7, 3, 1),
                                       (3, 4, 2), (4, 6, 2), (10, 4, 4), (15, 2, 2),
                                       # chunk 0: the pencil form up to N = 9, else the sheet form
                                       (7, 1, 0), (7, 3, 0), (7, 9, 0), (1, 5, 0), (4, 6, 0), (9, 4, 0),
                                       (10, 8, 0), (12, 3, 0), (15, 2, 0), (15, 5, 0),
                                       # chunk -2: the pencil form; -1: the sheet workgroup form
                                       (2, 3, -2), (5, 4, -2), (6, 7, -2), (7, 33, -2), (8, 2, -2), (9, 5, -2),
                                       (7, 3, -1), (7, 9, -1), (4, 6, -1), (9, 4, -1)])
def test_krhs_ax_local_equals_single(plat, N, K, chunk, affine):
    """Kernel entry: every right-hand side equals the single-vector kernel on
    the same arrays to rounding: the sweep and sheet forms run the same fused
    multiply-adds in the same order (except at N = 15, where the
    single-vector kernel runs its r and s contractions on the matrix cores);
    the pencil form sums the divergence's r and s parts apart."""
    if chunk == -1 and affine:
        pytest.skip("the sheet workgroup form serves non-affine meshes")
    nx = 2 if N >= 10 else 3
    lam = 1.3
    P = oracle_lib.Problem(bp=5, N=N, nx=nx, affine=affine, perturb=0.0 if affine else 0.1, seed=3, lam=lam)
    arr = oracle_arrays(P)
    qk = np.random.default_rng(N * 10 + K).standard_normal((P.Ngather, K))
    got = krhs_kernel(plat, P, arr, qk, K, chunk, affine, lam)
    ref = single_kernel(plat, P, arr, qk, K, affine, lam)
    assert rel_err(got, ref) <= form_rtol(N, chunk)
    # and against the oracle for the first and last right-hand side
    for r in (0, K - 1):
        assert rel_err(got[:, r], P.ax_local(np.ascontiguousarray(qk[:, r]))) <= 1e-12


F32_RTOL = 2e-5  # FP32 operator against the FP64 oracle (max-norm relative)
F32_SAME = 5e-6  # FP32 k-RHS against the FP32 single-vector kernel (the same operations; the pencil form groups two sums apart)


@pytest.mark.parametrize("affine", [False, True])
@pytest.mark.parametrize("N,K,chunk", [(7, 4, 0), (7, 9, 0), (7, 8, 2), (7, 4, 4), (3, 5, 0), (3, 6, 2),
                                       (9, 4, 0), (12, 3, 0), (15, 2, 0), (15, 4, 2), (7, 5, -1), (5, 3, -2)])
def test_krhs_f32_equals_single_f32(plat, N, K, chunk, affine):
    """bpx_bp5_ax[_affine]_hex_krhs_f32 on the oracle problem's arrays rounded
    to float: non-affine, every right-hand side equals the FP32 single-vector
    kernel bpx_bp5_ax_hex_f32 to FP32 rounding; both forms agree with the
    FP64 oracle to F32_RTOL."""
    if chunk == -1 and affine:
        pytest.skip("the sheet workgroup form serves non-affine meshes")
    nx = 2 if N >= 10 else 3
    lam = 1.3
    P = oracle_lib.Problem(bp=5, N=N, nx=nx, affine=affine, perturb=0.0 if affine else 0.1, seed=5, lam=lam)
    arr = {k: (v if v.dtype == torch.int32 else v.float()) for k, v in oracle_arrays(P).items()}
    qk = np.random.default_rng(N * 7 + K).standard_normal((P.Ngather, K))
    out = torch.zeros(K * P.NLocal, dtype=torch.float32, device="cuda")
    qd = torch.from_numpy(np.ascontiguousarray(qk.T.ravel())).float().cuda()
    L = capi.lib()
    fn = L.bpx_bp5_ax_affine_hex_krhs_f32 if affine else L.bpx_bp5_ax_hex_krhs_f32
    capi.check(fn(plat.h, P.N, K, chunk, P.Nelements, capi.ptr(arr["elementList"]), capi.ptr(arr["GlobalToLocal"]),
                  capi.ptr(arr["wJ"]), None, capi.ptr(arr["ggeo"]), capi.ptr(arr["gllw"]), capi.ptr(arr["D"]),
                  None, None, lam, capi.ptr(qd), P.Ngather, capi.ptr(out), P.NLocal))
    got = host(out).reshape(K, P.NLocal).T.astype(np.float64)
    for r in range(K):
        ref64 = P.ax_local(np.ascontiguousarray(qk[:, r]))
        assert rel_err(got[:, r], ref64) <= F32_RTOL, r
        if not affine:
            o = torch.zeros(P.NLocal, dtype=torch.float32, device="cuda")
            capi.check(L.bpx_bp5_ax_hex_f32(plat.h, P.N, P.Nelements, capi.ptr(arr["elementList"]),
                                            capi.ptr(arr["GlobalToLocal"]), capi.ptr(arr["wJ"]), None,
                                            capi.ptr(arr["ggeo"]), capi.ptr(arr["gllw"]), capi.ptr(arr["D"]),
                                            None, None, lam, capi.ptr(qd[r * P.Ngather:]), capi.ptr(o)))
            assert rel_err(got[:, r], host(o).astype(np.float64)) <= F32_SAME, r


@pytest.mark.parametrize("K", [2, 4, 8, 16, 32])
def test_krhs_operator_equals_single_20cube(plat, K):
    """BP5 operator (Ax + gather) on K right-hand sides at 20^3 p=7 (K <= 8)
    or 8^3 (K > 8, to bound the single-vector reference loop)."""
    nx = 20 if K <= 8 else 8
    m = capi.Mesh(N=7, nx=nx, perturb=0.1, seed=1)
    b = capi.BP(plat, m, 5, lam=1.0)
    qk = np.random.default_rng(K).standard_normal((b.Ngather, K))
    Aq = torch.zeros(K * b.Ngather, dtype=torch.float64, device="cuda")
    b.operator_krhs(torch.from_numpy(np.ascontiguousarray(qk.T.ravel())).cuda(), Aq, K)
    got = host(Aq).reshape(K, b.Ngather).T
    ref = single_columns(b, qk, K, local=False)
    assert rel_err(got, ref) <= form_rtol(7, 0)


def test_krhs_argument_errors(plat):
    m = capi.Mesh(N=3, nx=2)
    b = capi.BP(plat, m, 5, lam=1.0)
    q = torch.zeros(b.Ngather * 3, dtype=torch.float64, device="cuda")
    Aq = torch.zeros_like(q)
    with pytest.raises(capi.BpxError):
        b.operator_krhs(q, Aq, 3, chunk=2)  # chunk must divide K
    with pytest.raises(capi.BpxError):
        b.operator_krhs(q, Aq, 0)
    with pytest.raises(capi.BpxError):
        b.operator_krhs(q, Aq, 3, chunk=-3)
    b3 = capi.BP(plat, capi.Mesh(N=3, nx=2, cubature=True), 3, lam=1.0)
    with pytest.raises(capi.BpxError):
        b3.operator_krhs(q, Aq, 1)


def test_time_krhs_stamped(plat):
    """bpx_bp_time_krhs: dispatch-stamped mean times of the k-RHS and the
    single-vector Ax are positive and finite, and the argument checks hold."""
    b = capi.BP(plat, capi.Mesh(N=3, nx=3, perturb=0.1, seed=1), 5, lam=1.0)
    tk, t1 = b.time_krhs(4, 0, 3)
    assert 0 < tk < 1e6 and 0 < t1 < 1e6
    q = torch.from_numpy(np.random.default_rng(0).standard_normal(4 * b.Ngather)).cuda()
    tk, t1 = b.time_krhs(4, -1, 3, q)
    assert 0 < tk < 1e6 and 0 < t1 < 1e6
    with pytest.raises(capi.BpxError):
        b.time_krhs(0, 0, 3)
