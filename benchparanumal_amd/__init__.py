"""MI355X (gfx950) implementation of benchParanumal's BK/BP hex hot path.

The compute path is libbpx.so (HIP kernels + C++ host layer, C ABI in
include/bpx.h). This package only binds it with ctypes for tests and
benchmarks; device buffers are torch tensors passed by pointer. There is no
Python or CPU fallback: without the built library every call raises.
"""
from .capi import (  # noqa: F401
    BPX_OK,
    BpxError,
    BP,
    Mesh,
    Ogs,
    Platform,
    library_path,
    lib,
)
