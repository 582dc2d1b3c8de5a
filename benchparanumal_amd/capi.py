"""ctypes binding of libbpx.so (include/bpx.h)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "bpx.h")

BPX_OK = 0
STATUS = {
    0: "BPX_OK", -1: "BPX_ERR_INVALID_ARGUMENT", -2: "BPX_ERR_UNSUPPORTED", -3: "BPX_ERR_HIP",
    -4: "BPX_ERR_OUT_OF_MEMORY", -5: "BPX_ERR_NEGATIVE_JACOBIAN", -6: "BPX_ERR_OGS",
    -7: "BPX_ERR_BREAKDOWN", -8: "BPX_ERR_NO_DEVICE",
}

_lib = None
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_F = C.c_float
_I64P = C.POINTER(C.c_int64)


class BpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def library_path():
    return os.path.join(HERE, "lib", "libbpx.so")


_SIGS = {
    "bpx_last_error": (C.c_char_p, []),
    "bpx_version": (_I, []),
    "bpx_device_count": (_I, [C.POINTER(_I)]),
    "bpx_random_normal": (_I, [C.c_uint64, C.c_int64, C.c_int64, _P]),
    "bpx_platform_create": (_I, [_I, C.POINTER(_P)]),
    "bpx_platform_destroy": (_I, [_P]),
    "bpx_platform_set_stream": (_I, [_P, _P]),
    "bpx_platform_stream": (_P, [_P]),
    "bpx_platform_finish": (_I, [_P]),
    "bpx_rhs_hex": (_I, [_P, _I, _I, _I, _I] + [_P] * 6 + [_D, _P]),
    "bpx_ogs_gather": (_I, [_P, _I, _I] + [_P] * 5),
    "bpx_ogs_gather_f32": (_I, [_P, _I, _I] + [_P] * 5),
    "bpx_ogs_scatter": (_I, [_P, _I, _I] + [_P] * 5),
    "bpx_ogs_gather_scatter": (_I, [_P, _I, _I] + [_P] * 4),
    "bpx_linalg_set": (_I, [_P, _I, _D, _P]),
    "bpx_linalg_axpy": (_I, [_P, _I, _D, _P, _D, _P]),
    "bpx_linalg_inner_prod": (_I, [_P, _I, _P, _P, C.POINTER(_D)]),
    "bpx_linalg_norm2": (_I, [_P, _I, _P, C.POINTER(_D)]),
    "bpx_cg_update": (_I, [_P, _I, _P, _P, _D, _P, _P, C.POINTER(_D)]),
    "bpx_mesh_create": (_I, [_P, C.POINTER(_P)]),
    "bpx_mesh_create_dev": (_I, [_P, _P, C.POINTER(_P)]),
    "bpx_mesh_destroy": (_I, [_P]),
    "bpx_mesh_info": (_I, [_P, _I64P, _I]),
    "bpx_mesh_array": (_I, [_P, C.c_char_p, _P, _I64P, C.POINTER(_I)]),
    "bpx_ogs_setup": (_I, [_I, _P, C.POINTER(_P)]),
    "bpx_ogs_setup_dev": (_I, [_P, _I, _P, C.POINTER(_P)]),
    "bpx_ogs_setup_masked": (_I, [_P, _I, C.POINTER(_P)]),
    "bpx_ogs_destroy": (_I, [_P]),
    "bpx_ogs_info": (_I, [_P, _I64P, _I]),
    "bpx_ogs_array": (_I, [_P, C.c_char_p, _P, _I64P, C.POINTER(_I)]),
    "bpx_ogs_gather_dev": (_I, [_P, _P, _I, _P, _P]),
    "bpx_ogs_scatter_dev": (_I, [_P, _P, _I, _P, _P]),
    "bpx_ogs_gather_scatter_dev": (_I, [_P, _P, _I, _P]),
    "bpx_cg_create": (_I, [_P, _I, _I, C.POINTER(_P)]),
    "bpx_cg_destroy": (_I, [_P]),
    "bpx_cg_solve": (_I, [_P, _P, _P, _P, _P, _D, _I, _I, C.POINTER(_I), _P]),
    "bpx_cg_solve_noalloc": (_I, [_P, _I, _P, _P, _P, _P, _D, _I, C.POINTER(_I), _P]),
    "bpx_bp_create": (_I, [_P, _P, _I, _D, C.POINTER(_P)]),
    "bpx_bp_destroy": (_I, [_P]),
    "bpx_bp_info": (_I, [_P, _I64P, _I]),
    "bpx_bp_operator": (_I, [_P, _P, _P]),
    "bpx_bp_ax_local": (_I, [_P, _P, _P]),
    "bpx_bp_operator_f32": (_I, [_P, _P, _P]),
    "bpx_bp_operator_krhs": (_I, [_P, _I, _I, _P, _P]),
    "bpx_bp_tuning": (_I, [_P, _P, _I]),
    "bpx_bp_array": (_I, [_P, C.c_char_p, _P, _I64P, C.POINTER(_I)]),
    "bpx_bp_operator_dot": (_I, [_P, _P, _P, C.POINTER(_D)]),
    "bpx_bp_set_tuning": (_I, [_P, _I, _I, _I]),
    "bpx_bp_time_operator": (_I, [_P, _I, _I, _P]),
    "bpx_bp_time_krhs": (_I, [_P, _I, _I, _I, _P, _P]),
    "bpx_bp_set_ax_engine": (_I, [_P, _I]),
    "bpx_bp_set_cub_form": (_I, [_P, _I]),
    "bpx_bp_set_cg_fold": (_I, [_P, _I]),
    "bpx_bp_set_fp32_geometry": (_I, [_P, _I]),
    "bpx_bp_time_operator_in_solve": (_I, [_P, _P, _P, _D, _I, _I, _P]),
    "bpx_bp_set_kappa": (_I, [_P, _P]),
    "bpx_batch_create": (_I, [_P, _I, _P, _D, C.POINTER(_P)]),
    "bpx_batch_destroy": (_I, [_P]),
    "bpx_batch_info": (_I, [_P, _I64P, _I]),
    "bpx_batch_offsets": (_I, [_P, _I64P]),
    "bpx_batch_tuning": (_I, [_P, _P, _I]),
    "bpx_batch_solve": (_I, [_P, _P, _P, _D, _I, _P, _P]),
    "bpx_batch_solve_mixed": (_I, [_P, _P, _P, _D, _I, _I, _P, _P, _P]),
    "bpx_bp_kappa_gradient": (_I, [_P, _P, _P, _P]),
    "bpx_bp_rhs_forcing": (_I, [_P, _P]),
    "bpx_mesh_plot_fields": (_I, [_P, _I, _P, C.c_char_p]),
    "bpx_bp_plot_solution": (_I, [_P, _P, C.c_char_p]),
    "bpx_bp_error_norms": (_I, [_P, _P, _P]),
    "bpx_bp_solve": (_I, [_P, _P, _P, _D, _I, _I, C.POINTER(_I), _P]),
    "bpx_bp_solve_mixed": (_I, [_P, _P, _P, _D, _I, _I, _P]),
}
# Ax kernels: (platform, N, Nelements, elementList, G2L, <pointers...>, [lambda], q, Aq)
for _name in ("bpx_bp5_ax_hex", "bpx_bp5_ax_affine_hex", "bpx_bp6_ax_hex", "bpx_bp6_ax_affine_hex"):
    _SIGS[_name] = (_I, [_P, _I, _I, _P, _P] + [_P] * 7 + [_D, _P, _P])
for _name in ("bpx_bp1_ax_hex", "bpx_bp1_ax_affine_hex", "bpx_bp2_ax_hex", "bpx_bp2_ax_affine_hex"):
    _SIGS[_name] = (_I, [_P, _I, _I, _P, _P] + [_P] * 3 + [_P, _P])
for _name in ("bpx_bp3_ax_hex", "bpx_bp3_ax_affine_hex", "bpx_bp4_ax_hex", "bpx_bp4_ax_affine_hex"):
    _SIGS[_name] = (_I, [_P, _I, _I, _P, _P] + [_P] * 7 + [_D, _P, _P])
for _name in ("bpx_bp5_ax_hex_krhs", "bpx_bp5_ax_affine_hex_krhs"):
    _SIGS[_name] = (_I, [_P, _I, _I, _I, _I, _P, _P] + [_P] * 7 + [_D, _P, C.c_int64, _P, C.c_int64])
for _name in ("bpx_bp5_ax_hex_krhs_f32", "bpx_bp5_ax_affine_hex_krhs_f32"):
    _SIGS[_name] = (_I, [_P, _I, _I, _I, _I, _P, _P] + [_P] * 7 + [_F, _P, C.c_int64, _P, C.c_int64])
_SIGS["bpx_bp5_ax_hex_f32"] = (_I, [_P, _I, _I, _P, _P] + [_P] * 7 + [_F, _P, _P])
_SIGS["bpx_bp3_ax_hex_f32"] = (_I, [_P, _I, _I, _P, _P] + [_P] * 7 + [_F, _P, _P])


class _MeshOpts(C.Structure):
    _fields_ = [("N", _I), ("nx", _I), ("ny", _I), ("nz", _I), ("affine", _I), ("cubature", _I),
                ("lo", _D * 3), ("dim", _D * 3), ("perturb", _D), ("seed", C.c_uint64), ("elementType", _I)]

# element families (the reference's codes)
TRI, QUAD, TET, HEX = 3, 4, 6, 12


def lib():
    """Load libbpx.so; raises when it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(f"libbpx.so not built at {path}; run __graft_entry__.build()")
    L = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code):
    if code != BPX_OK:
        raise BpxError(code, lib().bpx_last_error().decode())
    return code


def ptr(t):
    """Device (or host numpy) pointer of a tensor/array; None -> NULL."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def random_normal(seed, n, counter0=0):
    """Seeded N(0,1) vector (SplitMix64 + Box-Muller), computed by libbpx."""
    out = np.zeros(int(n))
    check(lib().bpx_random_normal(seed, counter0, int(n), out.ctypes.data if n else None))
    return out


def device_count():
    n = _I()
    check(lib().bpx_device_count(C.byref(n)))
    return n.value


class Platform:
    """platform_t: one device and the stream every call is queued on. By
    default the stream is torch's current stream, so torch allocations and
    bpx kernels are ordered."""

    def __init__(self, device=0, torch_stream=True):
        h = _P()
        check(lib().bpx_platform_create(device, C.byref(h)))
        self.h = h
        self.device = device
        if torch_stream:
            import torch

            s = torch.cuda.current_stream(device).cuda_stream
            check(lib().bpx_platform_set_stream(self.h, s))

    def finish(self):
        check(lib().bpx_platform_finish(self.h))

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bpx_platform_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _read_array(fn, h, name):
    n = C.c_int64()
    dt = _I()
    check(fn(h, name.encode(), None, C.byref(n), C.byref(dt)))
    dtype = {0: np.float64, 1: np.int32, 2: np.int64}[dt.value]
    out = np.zeros(n.value, dtype=dtype)
    check(fn(h, name.encode(), out.ctypes.data if n.value else None, C.byref(n), C.byref(dt)))
    return out


class Mesh:
    """mesh_t for the hex box: set up on the host, or on the device of
    `platform` when one is given (the same arrays, copied down on request)."""

    def __init__(self, N, nx, ny=None, nz=None, affine=False, cubature=False, lo=(-1.0, -1.0, -1.0),
                 dim=(2.0, 2.0, 2.0), perturb=0.0, seed=0, platform=None, element=HEX):
        o = _MeshOpts()
        o.N, o.nx = N, nx
        o.ny = nx if ny is None else ny
        o.nz = nx if nz is None else nz
        o.affine, o.cubature = int(affine), int(cubature)
        for c in range(3):
            o.lo[c], o.dim[c] = lo[c], dim[c]
        o.perturb, o.seed = perturb, seed
        o.elementType = element
        h = _P()
        if platform is not None:
            check(lib().bpx_mesh_create_dev(platform.h, C.byref(o), C.byref(h)))
        else:
            check(lib().bpx_mesh_create(C.byref(o), C.byref(h)))
        self.h = h
        self.platform = platform
        buf = (C.c_int64 * 9)()
        n = lib().bpx_mesh_info(self.h, buf, 9)
        keys = ["N", "Nq", "Np", "cubNq", "cubNp", "Nelements", "elementType", "dim", "Nverts"]
        self.info = {k: int(buf[i]) for i, k in enumerate(keys[:n])}
        self.__dict__.update(self.info)
        self.nx, self.ny, self.nz = o.nx, o.ny, o.nz

    def array(self, name):
        return _read_array(lib().bpx_mesh_array, self.h, name)

    def plot_fields(self, QL, nfields, filename):
        """VTU file of host data QL ([element][field][node]) at the local nodes."""
        QL = np.ascontiguousarray(QL, dtype=np.float64)
        check(lib().bpx_mesh_plot_fields(self.h, int(nfields), QL.ctypes.data, os.fsencode(filename)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bpx_mesh_destroy(self.h)
            self.h = None


class Ogs:
    """ogs::ogs_t T maps: built on the host from host ids, on the device from
    device ids (`platform` + a device tensor `ids`), or for a mesh where the
    mesh was set up."""

    def __init__(self, ids=None, mesh=None, nfields=1, platform=None):
        h = _P()
        if mesh is not None:
            check(lib().bpx_ogs_setup_masked(mesh.h, nfields, C.byref(h)))
        elif platform is not None:
            check(lib().bpx_ogs_setup_dev(platform.h, int(ids.numel()), ptr(ids), C.byref(h)))
        else:
            ids = np.ascontiguousarray(ids, dtype=np.int64)
            check(lib().bpx_ogs_setup(len(ids), ids.ctypes.data, C.byref(h)))
        self.h = h
        buf = (C.c_int64 * 4)()
        lib().bpx_ogs_info(self.h, buf, 4)
        self.N, self.Ngather, self.nnzT, self.NrowBlocksT = (int(v) for v in buf)

    def array(self, name):
        return _read_array(lib().bpx_ogs_array, self.h, name)

    def gather(self, platform, v, gv, K=1):
        check(lib().bpx_ogs_gather_dev(platform.h, self.h, K, ptr(v), ptr(gv)))

    def scatter(self, platform, gv, v, K=1):
        check(lib().bpx_ogs_scatter_dev(platform.h, self.h, K, ptr(gv), ptr(v)))

    def gather_scatter(self, platform, v, K=1):
        """In place: every copy of a node gets the sum over its copies."""
        check(lib().bpx_ogs_gather_scatter_dev(platform.h, self.h, K, ptr(v)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bpx_ogs_destroy(self.h)
            self.h = None


class BP:
    """bpN_t: masked ogs, operator data on the device, Operator and CG."""

    def __init__(self, platform, mesh, bp, lam=1.0):
        h = _P()
        check(lib().bpx_bp_create(platform.h, mesh.h, bp, lam, C.byref(h)))
        self.h = h
        self.platform = platform
        self.mesh = mesh
        buf = (C.c_int64 * 7)()
        lib().bpx_bp_info(self.h, buf, 7)
        self.bp, self.Nfields, self.Ngather, self.NLocal, self.Nelements, self.N, self.Nnodes = (
            int(v) for v in buf)

    def operator(self, q, Aq):
        check(lib().bpx_bp_operator(self.h, ptr(q), ptr(Aq)))

    def ax_local(self, q, AqL):
        check(lib().bpx_bp_ax_local(self.h, ptr(q), ptr(AqL)))

    def operator_f32(self, q, Aq):
        check(lib().bpx_bp_operator_f32(self.h, ptr(q), ptr(Aq)))

    def set_kappa(self, kappa):
        """Variable coefficient on the node lattice (device tensor of Nnodes), or None."""
        check(lib().bpx_bp_set_kappa(self.h, None if kappa is None else ptr(kappa)))

    def kappa_gradient(self, u, lam, grad):
        """grad (Nnodes) = dL/dkappa from the solution u and the adjoint lam."""
        check(lib().bpx_bp_kappa_gradient(self.h, ptr(u), ptr(lam), ptr(grad)))

    def tuning(self):
        """Launch choices timed at setup (identical results whichever is chosen)."""
        out = np.zeros(10)
        lib().bpx_bp_tuning(self.h, ptr(out), 10)
        f = [float(v) for v in out]
        return {"gather_mode": int(f[0]), "gather_us": f[1:3], "fused": int(f[3]), "fused_us": f[4],
                "unfused_us": f[5], "fused32": int(f[6]), "fused32_us": f[7], "unfused32_us": f[8],
                "fp32_trilinear": int(f[9])}

    def set_fp32_geometry(self, mode):
        """FP32 BP5 geometry: 0 = from the elements' trilinear maps when the
        mesh's nodes are those maps (default), 1 = stored factors rounded to FP32."""
        check(lib().bpx_bp_set_fp32_geometry(self.h, int(mode)))

    def set_tuning(self, gather_mode=-1, fused=-1, fused32=-1):
        """Pin launch choices (-1 keeps the timed one)."""
        check(lib().bpx_bp_set_tuning(self.h, gather_mode, fused, fused32))

    def time_operator_in_solve(self, x, rhs, tol=1e-10, interval=64, maxit=100000):
        """solve_mixed with the FP32 CG operator timed per application inside
        the solve: {"ax_us", "gather_us", "launches", "fused"}."""
        out = np.zeros(4)
        check(lib().bpx_bp_time_operator_in_solve(self.h, ptr(x), ptr(rhs), tol, interval, maxit, ptr(out)))
        return {"ax_us": float(out[0]), "gather_us": float(out[1]), "launches": int(out[2]),
                "fused": bool(out[3])}

    def array(self, name):
        """A device array of the problem copied to the host (bpx_bp_array)."""
        return _read_array(lib().bpx_bp_array, self.h, name)

    def set_cg_fold(self, mask):
        """Where the fused CG operator also does the interior rows' direction
        update: mask bit 0 = FP64 solves (the default, mask 1), bit 1 = the
        FP32 iterations of the mixed solve; elsewhere one update kernel covers
        all rows. Same bits either way."""
        check(lib().bpx_bp_set_cg_fold(self.h, int(mask)))

    def set_ax_engine(self, engine):
        """BP5 at N = 15, FP64: 0 = auto (matrix cores), 1 = vector-ALU Ax,
        2 = matrix-core Ax (the two agree to rounding, not bit for bit)."""
        check(lib().bpx_bp_set_ax_engine(self.h, int(engine)))

    def set_cub_form(self, form):
        """BP1-BP4 non-affine at N <= 7: 0 = auto (pencil kernel), 1 =
        pencil (even-odd contractions), 2 = one element per wavefront, 3 =
        pencil with row-by-row contractions (the forms agree to rounding)."""
        check(lib().bpx_bp_set_cub_form(self.h, int(form)))

    def time_krhs(self, K, chunk=0, launches=10, q=None):
        """(k-RHS Ax us, single-vector Ax us), dispatch-stamped means; q:
        device tensor of K * Ngather values (None: zeros)."""
        out = np.zeros(2)
        check(lib().bpx_bp_time_krhs(self.h, K, chunk, launches, ptr(q), ptr(out)))
        return float(out[0]), float(out[1])

    def time_operator(self, precision=32, launches=50):
        """The CG operator (solver row order, tuned form) timed per kernel:
        {"ax_us", "gather_us", "fused"}."""
        out = np.zeros(3)
        check(lib().bpx_bp_time_operator(self.h, precision, launches, ptr(out)))
        return {"ax_us": float(out[0]), "gather_us": float(out[1]), "fused": bool(out[2])}

    def operator_dot(self, q, Aq):
        """Aq = A q as CG applies it (fused for BP5); returns q.Aq."""
        r = _D()
        check(lib().bpx_bp_operator_dot(self.h, ptr(q), ptr(Aq), C.byref(r)))
        return r.value

    def operator_krhs(self, q, Aq, K, chunk=0):
        """Operator on K right-hand sides stored one after another (K x Ngather), BP5."""
        check(lib().bpx_bp_operator_krhs(self.h, K, chunk, ptr(q), ptr(Aq)))

    def rhs_forcing(self, b):
        check(lib().bpx_bp_rhs_forcing(self.h, ptr(b)))

    def plot_solution(self, u, filename):
        """VTU file of the solution u (device, Ngather) at the local nodes."""
        check(lib().bpx_bp_plot_solution(self.h, ptr(u), os.fsencode(filename)))

    def error_norms(self, u):
        """u (device, Ngather) against the forcing's manufactured solution:
        {"max", "l2", "rel_l2"}."""
        out = np.zeros(3)
        check(lib().bpx_bp_error_norms(self.h, ptr(u), ptr(out)))
        return {"max": float(out[0]), "l2": float(out[1]), "rel_l2": float(out[2])}

    def solve(self, x, r, tol=0.0, maxit=100, verbose=False, history=True):
        its = _I()
        hist = np.zeros(maxit + 1) if history else None
        check(lib().bpx_bp_solve(self.h, ptr(x), ptr(r), tol, maxit, int(verbose), C.byref(its),
                                 ptr(hist)))
        return its.value, (hist[: its.value + 1] if history else None)

    def solve_mixed(self, x, b, tol=1e-10, interval=32, maxit=100000):
        """FP32 CG with FP64 defect corrections every `interval` iterations."""
        stats = np.zeros(3)
        check(lib().bpx_bp_solve_mixed(self.h, ptr(x), ptr(b), tol, interval, maxit, ptr(stats)))
        return {"updates": int(stats[0]), "iterations": int(stats[1]), "rel_res": float(stats[2])}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bpx_bp_destroy(self.h)
            self.h = None


class Batch:
    """Ragged batch of BP5 problems, one CG each, replayed from a hipGraph.

    problems: list of dicts with Mesh arguments (N, nx, ny, nz, perturb, seed, ...)."""

    def __init__(self, platform, problems, lam=1.0):
        n = len(problems)
        arr = (_MeshOpts * n)()
        for i, pr in enumerate(problems):
            o = arr[i]
            o.N, o.nx = pr["N"], pr["nx"]
            o.ny = pr.get("ny", pr["nx"])
            o.nz = pr.get("nz", pr["nx"])
            o.affine, o.cubature = 0, 0
            lo, dim = pr.get("lo", (-1.0, -1.0, -1.0)), pr.get("dim", (2.0, 2.0, 2.0))
            for c in range(3):
                o.lo[c], o.dim[c] = lo[c], dim[c]
            o.perturb, o.seed = pr.get("perturb", 0.0), pr.get("seed", 0)
        h = _P()
        check(lib().bpx_batch_create(platform.h, n, C.cast(arr, _P), lam, C.byref(h)))
        self.h = h
        self.platform = platform
        self.nprob = n
        off = np.zeros(n + 1, dtype=np.int64)
        check(lib().bpx_batch_offsets(self.h, off.ctypes.data_as(_I64P)))
        self.offsets = off
        self.Ntot = int(off[-1])

    def info(self):
        buf = (C.c_int64 * 7)()
        lib().bpx_batch_info(self.h, buf, 7)
        keys = ["nprob", "Ntot", "ngroups", "nchunks", "nsupers", "NLocal", "replays"]
        return {k: int(buf[i]) for i, k in enumerate(keys)}

    def tuning(self):
        n = lib().bpx_batch_tuning(self.h, None, 0)
        out = np.zeros(max(n, 1))
        lib().bpx_batch_tuning(self.h, ptr(out), n)
        return [{"N": int(out[6 * g]), "shape": int(out[6 * g + 1]), "elements_per_workgroup": int(out[6 * g + 2]),
                 "us": [float(v) for v in out[6 * g + 3:6 * g + 6]]} for g in range(n // 6)]

    def solve(self, x, b, tol=1e-8, maxit=10000):
        """Solve every problem; returns (iterations, final residual norms) per problem."""
        its = np.zeros(self.nprob, dtype=np.int32)
        rn = np.zeros(self.nprob)
        check(lib().bpx_batch_solve(self.h, ptr(x), ptr(b), tol, maxit, its.ctypes.data_as(_P), ptr(rn)))
        return its, rn

    def solve_mixed(self, x, b, tol=1e-10, interval=32, maxit=100000):
        """FP32 CG with FP64 corrections for every problem; returns per problem
        (FP32 iterations, corrections, final ||b - A x|| / ||b||)."""
        its = np.zeros(self.nprob, dtype=np.int32)
        ups = np.zeros(self.nprob, dtype=np.int32)
        rel = np.zeros(self.nprob)
        check(lib().bpx_batch_solve_mixed(self.h, ptr(x), ptr(b), tol, interval, maxit, its.ctypes.data_as(_P),
                                          ups.ctypes.data_as(_P), ptr(rel)))
        return its, ups, rel

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bpx_batch_destroy(self.h)
            self.h = None


def header_symbols():
    """Names of the functions include/bpx.h declares."""
    import re

    src = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bpx_[a-z0-9_]+)\s*\(", src)))
