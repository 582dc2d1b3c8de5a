"""The product's quadrilateral, triangle and tetrahedron setup (host side,
no GPU): reference nodes and matrices, nodes, geometric and cubature
factors against the oracle (to rounding), the masked ogs maps bit for bit,
and the element-type checks of the C ABI."""
import numpy as np
import pytest

import oracle_lib
from benchparanumal_amd import capi

FAMILIES = [(capi.QUAD, dict(nx=3, ny=2)), (capi.TRI, dict(nx=3, ny=2)), (capi.TET, dict(nx=2, ny=2, nz=2))]


def rel(a, b):
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("el,kw", FAMILIES)
@pytest.mark.parametrize("N", [1, 2, 4, 7])
@pytest.mark.parametrize("bp", [1, 3, 5])
def test_family_setup_matches_oracle(el, kw, N, bp):
    P = oracle_lib.Problem(bp=bp, N=N, element=el, perturb=0.1, seed=3, **kw)
    m = capi.Mesh(N=N, element=el, perturb=0.1, seed=3, cubature=bp <= 4, **kw)
    assert m.info["elementType"] == el and m.info["Np"] == P.Np and m.Nelements == P.Nelements
    names = [("x", "x"), ("y", "y"), ("EX", "EX"), ("EY", "EY"), ("wJ", "wJ"), ("ggeo", "ggeo")]
    if el == capi.TET:
        names += [("z", "z"), ("EZ", "EZ")]
    if el == capi.QUAD and bp <= 4:
        names += [("cubwJ", "cubwJ"), ("cubggeo", "cubggeo"), ("cubD", "cubD"), ("cubInterp", "cubInterp")]
    if el != capi.QUAD:
        names += [("rN", "r"), ("sN", "s"), ("MM", "MM"), ("Dsim", "Dsim"), ("Ssim", "Ssim")]
        if el == capi.TET:
            names += [("tN", "t")]
    for a, b in names:
        assert rel(m.array(a), P.array(b)) <= 1e-12 * max(1, N), a
    nf = 1 if bp != 6 else (3 if el == capi.TET else 2)
    o = capi.Ogs(mesh=m, nfields=nf)
    assert o.Ngather == P.Ngather
    for k in ("rowStartsT", "colIdsT", "blockRowStartsT", "GlobalToLocal"):
        np.testing.assert_array_equal(o.array(k), P.array(k), err_msg=k)


@pytest.mark.parametrize("el,kw", FAMILIES)
def test_family_vector_maps_match_oracle(el, kw):
    nf = 3 if el == capi.TET else 2
    P = oracle_lib.Problem(bp=6, N=3, element=el, perturb=0.1, seed=2, **kw)
    m = capi.Mesh(N=3, element=el, perturb=0.1, seed=2, **kw)
    o = capi.Ogs(mesh=m, nfields=nf)
    for k in ("rowStartsT", "colIdsT", "blockRowStartsT", "GlobalToLocal"):
        np.testing.assert_array_equal(o.array(k), P.array(k), err_msg=k)


def test_family_argument_checks():
    with pytest.raises(capi.BpxError):
        capi.Mesh(N=2, nx=2, element=5)
    with pytest.raises(capi.BpxError):
        capi.Mesh(N=16, nx=2, element=capi.TRI)
    with pytest.raises(capi.BpxError):
        capi.Mesh(N=2, nx=2, ny=2, nz=0, element=capi.TET)
