"""Shared test helpers: small LPs with known answers and relative errors, and
the ctypes view of device CSR descriptors and the HIP runtime."""
import ctypes as C
import os

import numpy as np


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def capi_example():
    """docs/C_API.md:47-90: min x1+x2 s.t. x1+2x2 = 5, x2 <= 2, 3x1+2x2 <= 8, x free.
    Closed form: x* = (1, 2), y* = (1, -1, 0), objective 3."""
    A = np.array([[1.0, 2.0], [0.0, 1.0], [3.0, 2.0]])
    c = np.array([1.0, 1.0])
    l = np.array([5.0, -np.inf, -np.inf])
    u = np.array([5.0, 2.0, 8.0])
    return A, c, l, u


def random_csr(m, n, lens, seed, long_rows=()):
    """CSR with given row lengths (distinct sorted columns), N(0,1) values."""
    rng = np.random.default_rng(seed)
    lens = np.array(lens, dtype=np.int64)
    for r, L in long_rows:
        lens[r] = L
    lens = np.minimum(lens, n)
    rp = np.zeros(m + 1, dtype=np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.empty(rp[-1], dtype=np.int32)
    for i in range(m):
        ci[rp[i]:rp[i + 1]] = np.sort(rng.choice(n, size=lens[i], replace=False))
    v = rng.standard_normal(rp[-1])
    return rp, ci, v


def csr_matvec(rp, ci, v, x):
    m = len(rp) - 1
    y = np.zeros(m)
    rows = np.repeat(np.arange(m), np.diff(rp))
    np.add.at(y, rows, v * x[ci])
    return y


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DevCsr(C.Structure):
    """dpdlp_csr_t of include/dpdlp_dev.h"""
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int), ("rp", _ip),
                ("ci", _ip), ("v", _dp), ("v32", C.c_void_p), ("blk", _ip), ("nblk", C.c_int),
                ("cap", C.c_int), ("P", C.c_int), ("toff", _ip), ("rofs", C.c_void_p),
                ("pci", _ip), ("pv", _dp), ("pv32", C.c_void_p), ("xpart", _dp),
                ("tick", C.c_void_p)]


def hip_runtime():
    h = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def device_array(hip, ptr, count, dtype):
    """A host copy of `count` elements at a device address."""
    out = np.empty(count, dtype=dtype)
    if count:
        assert hip.hipMemcpy(out.ctypes.data, C.cast(ptr, C.c_void_p), out.nbytes, D2H) == 0
    return out
