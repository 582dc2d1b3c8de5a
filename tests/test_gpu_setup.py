"""GPU-side setup: the mesh nodes, global ids and geometric and cubature
factors formed on the device, the ogs T maps built by a device sort, and the
fused BP5 operator's solve-space maps, each against the host setup (bit for
bit) and against the oracle (maps bit for bit, factors to 2e-13 relative),
plus solves on a device-built problem that repeat the host-built problem's
bits. Every call goes through libbpx.so's C ABI (bpx_mesh_create_dev,
bpx_ogs_setup_dev, bpx_bp_array for the problems' device arrays)."""
import numpy as np
import pytest

import brute
import oracle_lib
from benchparanumal_amd import capi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MESH_ARRAYS = ("x", "y", "z", "globalIds", "wJ", "ggeo", "cubwJ", "cubggeo", "EX", "EY", "EZ")
BP_ARRAYS = ("rowStartsT", "colIdsT", "blockRowStartsT", "GlobalToLocal", "wJ", "ggeo", "solvePerm",
             "solveGlobalToLocal", "solveRowStarts", "solveColIds", "solveBlocks", "sharedRowStarts",
             "sharedColIds", "sharedBlocks")


@pytest.fixture(scope="module")
def plat():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return capi.Platform(0)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = [(N, 3 if N <= 6 else 2, False, True, 0.1) for N in range(1, 16)] + [
    (3, 3, True, True, 0.0), (7, 2, True, False, 0.0), (15, 1, True, True, 0.0),
    (5, 3, False, True, 0.0), (2, 4, False, False, 0.3)]


@pytest.mark.parametrize("N,nx,affine,cubature,perturb", CASES)
def test_device_mesh_equals_host_mesh(plat, N, nx, affine, cubature, perturb):
    """Every array of the device-built mesh equals the host-built one bit for
    bit (same operations in the same order, no contraction on either side),
    on a box of nx x 2 x (nx-1 or 1) elements."""
    kw = dict(N=N, nx=nx, ny=2, nz=max(nx - 1, 1), affine=affine, cubature=cubature, perturb=perturb, seed=9)
    print(f"device mesh N={N} nx={nx} affine={affine} cubature={cubature}", flush=True)
    mh = capi.Mesh(**kw)
    md = capi.Mesh(platform=plat, **kw)
    assert md.info == mh.info
    for name in MESH_ARRAYS:
        a, b = md.array(name), mh.array(name)
        assert a.dtype == b.dtype and a.shape == b.shape, name
        np.testing.assert_array_equal(a, b, err_msg=name)


@pytest.mark.parametrize("affine,perturb", [(False, 0.0), (False, 0.1), (True, 0.0)])
def test_device_mesh_geometry_matches_oracle(plat, affine, perturb):
    """The device factors against the oracle's restatement of
    meshGeometricFactorsHex3D / meshCubatureSetupHex3D."""
    kw = dict(N=5, nx=3, ny=2, nz=4, affine=affine, perturb=perturb, seed=11)
    m = capi.Mesh(cubature=True, platform=plat, **kw)
    P = oracle_lib.Problem(bp=3, lam=1.0, **kw)
    for name in ("x", "y", "z"):
        np.testing.assert_allclose(m.array(name), P.array(name), atol=1e-15)
    for name in ("wJ", "ggeo") + (() if affine else ("cubwJ", "cubggeo")):
        a, b = m.array(name), P.array(name)
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 2e-13 * np.abs(b).max(), name


def test_device_mesh_full_size_equals_host(plat):
    """The bench problem's mesh (24^3 elements, p = 7, perturbed): nodes,
    ids and factors bit for bit."""
    kw = dict(N=7, nx=24, perturb=0.1, seed=1)
    mh = capi.Mesh(**kw)
    md = capi.Mesh(platform=plat, **kw)
    for name in ("x", "y", "z", "globalIds", "wJ", "ggeo"):
        np.testing.assert_array_equal(md.array(name), mh.array(name), err_msg=name)


def test_device_mesh_rejects_negative_jacobian(plat):
    with pytest.raises(capi.BpxError):
        capi.Mesh(N=2, nx=2, dim=(2.0, 0.0, 2.0), platform=plat)


@pytest.mark.parametrize("N,nx,ny,nz,nf", [(1, 2, 2, 2, 1), (4, 3, 2, 2, 1), (7, 3, 3, 3, 1),
                                           (2, 2, 3, 2, 3), (3, 1, 1, 1, 1), (15, 2, 1, 1, 1)])
def test_device_ogs_maps_bit_exact_vs_oracle(plat, N, nx, ny, nz, nf):
    """Masked ids and T maps built on the device for a device mesh equal the
    oracle's (libs/ogs/ogsSetup.cpp order: rows by first appearance, each
    row ascending, 512-entry blocks by bisection)."""
    m = capi.Mesh(N=N, nx=nx, ny=ny, nz=nz, platform=plat)
    o = capi.Ogs(mesh=m, nfields=nf)
    P = oracle_lib.Problem(bp=5 if nf == 1 else 6, N=N, nx=nx, ny=ny, nz=nz)
    assert o.Ngather == P.Ngather and o.nnzT == P.nnzT and o.NrowBlocksT == P.NrowBlocksT
    for name in ("rowStartsT", "colIdsT", "blockRowStartsT", "GlobalToLocal"):
        np.testing.assert_array_equal(o.array(name), P.array(name), err_msg=name)
    mh = capi.Mesh(N=N, nx=nx, ny=ny, nz=nz)
    oh = capi.Ogs(mesh=mh, nfields=nf)
    np.testing.assert_array_equal(o.array("maskedIds"), oh.array("maskedIds"))


def test_device_ogs_general_ids_vs_bruteforce(plat):
    """Signed ids with masked (0) entries: the device maps equal the
    brute-force first-appearance maps and the host maps."""
    rng = np.random.default_rng(7)
    for n, span in [(1, 3), (50, 10), (5000, 700), (9000, 30), (200000, 60000)]:
        ids = rng.integers(-span, span + 1, size=n).astype(np.int64)
        o = capi.Ogs(ids=dev(ids), platform=plat)
        oh = capi.Ogs(ids=ids)
        rs, ci, g2l = brute.first_appearance_maps(np.abs(ids))
        np.testing.assert_array_equal(o.array("rowStartsT"), rs)
        np.testing.assert_array_equal(o.array("colIdsT"), ci)
        np.testing.assert_array_equal(o.array("GlobalToLocal"), g2l)
        for name in ("rowStartsT", "colIdsT", "blockRowStartsT", "GlobalToLocal"):
            np.testing.assert_array_equal(o.array(name), oh.array(name), err_msg=name)


def test_device_ogs_large_ids_and_many_blocks(plat):
    """Ids above 2^40 (the sort runs over all their bits) and groups of up
    to 512 copies, so the bisection runs many levels."""
    rng = np.random.default_rng(3)
    sizes = rng.integers(1, 513, size=3000)
    ids = np.repeat((np.arange(1, len(sizes) + 1, dtype=np.int64) << 41) + 17, sizes)
    rng.shuffle(ids)
    ids[rng.random(len(ids)) < 0.05] = 0
    o = capi.Ogs(ids=dev(ids), platform=plat)
    oh = capi.Ogs(ids=ids)
    assert o.NrowBlocksT == oh.NrowBlocksT > 100
    for name in ("rowStartsT", "colIdsT", "blockRowStartsT", "GlobalToLocal"):
        np.testing.assert_array_equal(o.array(name), oh.array(name), err_msg=name)
    np.testing.assert_array_equal(o.array("blockRowStartsT"), oracle_lib.block_rows(o.array("rowStartsT")))


def test_device_ogs_empty_all_masked_and_limits(plat):
    o = capi.Ogs(ids=dev(np.zeros(10, dtype=np.int64)), platform=plat)
    assert o.Ngather == 0 and o.nnzT == 0 and o.NrowBlocksT == 0
    np.testing.assert_array_equal(o.array("GlobalToLocal"), -np.ones(10))
    np.testing.assert_array_equal(o.array("rowStartsT"), [0])
    with pytest.raises(capi.BpxError) as ei:
        capi.Ogs(ids=dev(np.full(513, 5, dtype=np.int64)), platform=plat)
    assert ei.value.code == -6
    o = capi.Ogs(ids=dev(np.full(512, -5, dtype=np.int64)), platform=plat)
    assert o.Ngather == 1 and o.NrowBlocksT == 1


@pytest.mark.parametrize("bp,N,nx,affine", [(5, 7, 3, False), (5, 2, 4, False), (5, 15, 2, False), (5, 4, 3, True),
                                            (6, 3, 2, False), (3, 5, 2, False), (1, 3, 2, False), (4, 2, 2, False),
                                            (2, 3, 2, True), (3, 4, 2, True)])
def test_device_problem_equals_host_problem(plat, bp, N, nx, affine):
    """bpx_bp_create on a device mesh builds every map and factor array on
    the device: they equal the host-built problem's, and the CG solve (and
    for BP5 the mixed solve) repeats its iterations, residual history and
    solution bit for bit."""
    kw = dict(N=N, nx=nx, affine=affine, perturb=0.0 if affine else 0.1, seed=5, cubature=bp <= 4)
    print(f"device problem bp={bp} N={N} nx={nx} affine={affine}", flush=True)
    mh = capi.Mesh(**kw)
    md = capi.Mesh(platform=plat, **kw)
    bh = capi.BP(plat, mh, bp, lam=1.1)
    bd = capi.BP(plat, md, bp, lam=1.1)
    assert (bd.Ngather, bd.NLocal) == (bh.Ngather, bh.NLocal)
    for name in BP_ARRAYS:
        np.testing.assert_array_equal(bd.array(name), bh.array(name), err_msg=name)
    rhs = capi.random_normal(3, bh.Ngather)
    res = []
    for b in (bh, bd):
        x = torch.zeros(b.Ngather, dtype=torch.float64, device="cuda")
        its, hist = b.solve(x, dev(rhs), tol=1e-9, maxit=3000)
        out = [its, hist, host(x)]
        if bp == 5 and not affine:
            xm = torch.zeros_like(x)
            st = b.solve_mixed(xm, dev(rhs), tol=1e-10, interval=16)
            out += [st["iterations"], host(xm)]
        res.append(out)
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        np.testing.assert_array_equal(a, b)


def test_device_problem_full_size_maps(plat):
    """The bench problem (BP5, 24^3, p = 7): the device-built maps and the
    fused operator's solve space equal the host-built ones."""
    kw = dict(N=7, nx=24, perturb=0.1, seed=1)
    bh = capi.BP(plat, capi.Mesh(**kw), 5, lam=1.0)
    bd = capi.BP(plat, capi.Mesh(platform=plat, **kw), 5, lam=1.0)
    for name in BP_ARRAYS:
        np.testing.assert_array_equal(bd.array(name), bh.array(name), err_msg=name)


def test_device_problem_forcing_and_error_norms(plat):
    """Forcing, error norms and the FP32 geometry check read the device
    mesh's coordinates: same right-hand side and norms as the host mesh."""
    kw = dict(N=5, nx=3, perturb=0.1, seed=2)
    out = []
    for m in (capi.Mesh(**kw), capi.Mesh(platform=plat, **kw)):
        b = capi.BP(plat, m, 5, lam=1.0)
        rhs = torch.zeros(b.Ngather, dtype=torch.float64, device="cuda")
        b.rhs_forcing(rhs)
        x = torch.zeros_like(rhs)
        b.solve(x, rhs.clone(), tol=1e-10, maxit=2000)
        xm = torch.zeros_like(rhs)
        b.solve_mixed(xm, rhs, tol=1e-10, interval=16)
        out.append((host(rhs), b.error_norms(x), b.tuning()["fp32_trilinear"], host(xm)))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    assert out[0][2] == out[1][2] == 1
    np.testing.assert_array_equal(out[0][3], out[1][3])
