"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the checker and the timed CPU
baseline, never part of the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.orc_last_error.restype = C.c_char_p
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_jacobi_gq.argtypes = [C.c_double, C.c_double, C.c_int, _D, _D]
    L.orc_jacobi_gll.argtypes = [C.c_int, _D, _D]
    L.orc_jacobi_p.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
    L.orc_jacobi_p.restype = C.c_double
    L.orc_vandermonde1d.argtypes = [C.c_int, C.c_int, _D, _D]
    L.orc_mass_matrix1d.argtypes = [C.c_int, _D, _D]
    L.orc_dmatrix1d.argtypes = [C.c_int, C.c_int, _D, _D, _D]
    L.orc_interp1d.argtypes = [C.c_int, C.c_int, _D, _D, _D]
    L.orc_rng_u64.argtypes = [C.c_ulonglong, C.c_ulonglong]
    L.orc_rng_u64.restype = C.c_ulonglong
    L.orc_rng_u01.argtypes = [C.c_ulonglong, C.c_ulonglong]
    L.orc_rng_u01.restype = C.c_double
    L.orc_rng_normal.argtypes = [C.c_ulonglong, C.c_ulonglong]
    L.orc_rng_normal.restype = C.c_double
    L.orc_problem_create.argtypes = [C.c_int] * 6 + [C.c_double, _D, _D, C.c_double, C.c_ulonglong]
    L.orc_problem_create.restype = C.c_void_p
    L.orc_problem_create_family.argtypes = [C.c_int] * 6 + [C.c_double, _D, _D, C.c_double, C.c_ulonglong]
    L.orc_problem_create_family.restype = C.c_void_p
    L.orc_problem_destroy.argtypes = [C.c_void_p]
    L.orc_problem_info.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.orc_problem_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_longlong), _I]
    L.orc_problem_array.restype = C.c_void_p
    for name in ("orc_ax_local", "orc_gather", "orc_scatter", "orc_operator"):
        getattr(L, name).argtypes = [C.c_void_p, _D, _D]
    L.orc_rhs_forcing_local.argtypes = [C.c_void_p, _D]
    L.orc_set_lambda.argtypes = [C.c_void_p, C.c_double]
    L.orc_cg_solve.argtypes = [C.c_void_p, _D, _D, C.c_double, C.c_int, _D, _D]
    L.orc_problem_set_mask.argtypes = [C.c_void_p, C.c_int]
    L.orc_set_kappa.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_lattice_nodes.argtypes = [C.c_void_p]
    L.orc_lattice_nodes.restype = C.c_longlong
    L.orc_lattice_index.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_kappa_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_block_rows.argtypes = [C.c_int, _I, _I]
    L.orc_la_axpy.argtypes = [C.c_int, C.c_double, _D, C.c_double, _D]
    L.orc_la_inner_prod.argtypes = [C.c_int, _D, _D]
    L.orc_la_inner_prod.restype = C.c_double
    L.orc_la_norm2.argtypes = [C.c_int, _D]
    L.orc_la_norm2.restype = C.c_double
    L.orc_la_update_cg.argtypes = [C.c_int, _D, C.c_double, _D]
    L.orc_la_update_cg.restype = C.c_double
    _lib = L
    return L


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def set_threads(n):
    return lib().orc_set_threads(int(n))


def gll(N):
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    assert lib().orc_jacobi_gll(N, dptr(x), dptr(w)) == 0
    return x, w


def gauss(N, alpha=0.0, beta=0.0):
    """Gauss-Jacobi rule with N+1 points (JacobiGQ(alpha, beta, N))."""
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    assert lib().orc_jacobi_gq(alpha, beta, N, dptr(x), dptr(w)) == 0
    return x, w


def dmatrix(N, rin, rout):
    rin = np.ascontiguousarray(rin, dtype=np.float64)
    rout = np.ascontiguousarray(rout, dtype=np.float64)
    D = np.zeros((len(rout), N + 1))
    assert lib().orc_dmatrix1d(N, len(rout), dptr(rin), dptr(rout), dptr(D)) == 0
    return D


def interp(N, rin, rout):
    rin = np.ascontiguousarray(rin, dtype=np.float64)
    rout = np.ascontiguousarray(rout, dtype=np.float64)
    I = np.zeros((len(rout), N + 1))
    assert lib().orc_interp1d(N, len(rout), dptr(rin), dptr(rout), dptr(I)) == 0
    return I


def mass1d(N, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    MM = np.zeros((N + 1, N + 1))
    assert lib().orc_mass_matrix1d(N, dptr(r), dptr(MM)) == 0
    return MM


_INFO = ["bp", "N", "Nq", "Np", "cubNq", "cubNp", "Nelements", "Nfields", "affine", "NLocal",
         "Ngather", "nnzT", "NrowBlocksT"]


class Problem:
    """One BP problem on a box mesh, built by the oracle's setup restatement."""

    def __init__(self, bp=5, N=2, nx=2, ny=None, nz=None, affine=False, lam=1.0,
                 lo=(-1.0, -1.0, -1.0), dim=(2.0, 2.0, 2.0), perturb=0.0, seed=0, element=12):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        lo_a = np.ascontiguousarray(lo, dtype=np.float64)
        dim_a = np.ascontiguousarray(dim, dtype=np.float64)
        if element == 12:
            h = lib().orc_problem_create(bp, N, nx, ny, nz, int(affine), lam, dptr(lo_a),
                                         dptr(dim_a), perturb, seed)
        else:  # quad (4), triangle (3), tetrahedron (6)
            h = lib().orc_problem_create_family(element, bp, N, nx, ny, nz, lam, dptr(lo_a),
                                                dptr(dim_a), perturb, seed)
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.h = C.c_void_p(h)
        self._refresh()
        self.lam = lam

    def _refresh(self):
        buf = (C.c_longlong * 32)()
        n = lib().orc_problem_info(self.h, buf)
        self.info = {k: int(buf[i]) for i, k in enumerate(_INFO[:n])}
        self.__dict__.update(self.info)

    def set_mask(self, mask):
        """Rebuild the ogs maps with (True) or without (False) the Dirichlet mask."""
        if lib().orc_problem_set_mask(self.h, int(mask)):
            raise RuntimeError(lib().orc_last_error().decode())
        self._refresh()

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_problem_destroy(self.h)
            self.h = None

    def array(self, name):
        n = C.c_longlong()
        dt = C.c_int()
        p = lib().orc_problem_array(self.h, name.encode(), C.byref(n), C.byref(dt))
        if dt.value < 0:
            raise KeyError(name)
        ctype = {0: C.c_double, 1: C.c_int32, 2: C.c_int64}[dt.value]
        if n.value == 0:
            return np.zeros(0, dtype={0: np.float64, 1: np.int32, 2: np.int64}[dt.value])
        arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), shape=(n.value,))
        return arr.copy()

    def ax_local(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.NLocal)
        lib().orc_ax_local(self.h, dptr(q), dptr(out))
        return out

    def gather(self, vL):
        vL = np.ascontiguousarray(vL, dtype=np.float64)
        out = np.zeros(self.Ngather)
        lib().orc_gather(self.h, dptr(vL), dptr(out))
        return out

    def scatter(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.NLocal)
        lib().orc_scatter(self.h, dptr(v), dptr(out))
        return out

    def operator(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.Ngather)
        lib().orc_operator(self.h, dptr(q), dptr(out))
        return out

    def rhs_forcing_local(self):
        rL = np.zeros(self.NLocal)
        lib().orc_rhs_forcing_local(self.h, dptr(rL))
        return rL

    def rhs_forcing(self):
        return self.gather(self.rhs_forcing_local())

    # ---- variable coefficient (kappa on the node lattice) and adjoint gradient
    def lattice_nodes(self):
        return int(lib().orc_lattice_nodes(self.h))

    def lattice_index(self):
        out = np.zeros(self.NLocal, dtype=np.int64)
        lib().orc_lattice_index(self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def set_kappa(self, kappa):
        """kappa: lattice-ordered array, or None for the constant coefficient."""
        if kappa is None:
            lib().orc_set_kappa(self.h, None)
            self._kappa = None
            return
        self._kappa = np.ascontiguousarray(kappa, dtype=np.float64)
        assert self._kappa.size == self.lattice_nodes()
        if lib().orc_set_kappa(self.h, dptr(self._kappa)):
            raise RuntimeError(lib().orc_last_error().decode())

    def kappa_grad(self, u, lam):
        """dL/dkappa on the lattice, -sum (D lam).G(D u) over node copies."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(self.lattice_nodes())
        lib().orc_kappa_grad(self.h, dptr(u), dptr(lam), dptr(out))
        return out

    def cg(self, b, maxit, tol=0.0, x0=None):
        x = np.zeros(self.Ngather) if x0 is None else np.array(x0, dtype=np.float64)
        r = np.array(b, dtype=np.float64)
        hist = np.zeros(maxit + 1)
        alph = np.zeros(maxit)
        it = lib().orc_cg_solve(self.h, dptr(x), dptr(r), tol, maxit, dptr(hist), dptr(alph))
        return x, it, hist[: it + 1], alph[:it]


def inner_prod(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_la_inner_prod(len(x), dptr(x), dptr(y))


def norm2(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_la_norm2(len(x), dptr(x))


def rng_u64(seed, counter):
    return lib().orc_rng_u64(seed, counter)


def rng_u01(seed, counter):
    return lib().orc_rng_u01(seed, counter)


def rng_normal(seed, counter):
    return lib().orc_rng_normal(seed, counter)


def block_rows(row_starts):
    rs = np.ascontiguousarray(row_starts, dtype=np.int32)
    out = np.zeros(len(rs), dtype=np.int32)
    nb = lib().orc_block_rows(len(rs) - 1, rs.ctypes.data_as(_I), out.ctypes.data_as(_I))
    if nb < 0:
        raise ValueError("row multiplicity above 512")
    return out[: nb + 1]
