"""dpdlp_dev_spmv_noalloc, the SpMV entry point that allocates nothing (its
tiled-SpMV scratch is a fixed per-GPU array): called through the C ABI on
device buffers, it gives the same bits as dpdlp_dev_spmv with the caller's
scratch, on two calls in a row (the second smaller, with a different panel
count, so it must clear the tickets the first left behind)."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import random_csr

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DevCsr(C.Structure):
    """dpdlp_csr_t of include/dpdlp_dev.h"""
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int), ("rp", _ip),
                ("ci", _ip), ("v", _dp), ("v32", C.c_void_p), ("blk", _ip), ("nblk", C.c_int),
                ("cap", C.c_int), ("P", C.c_int), ("toff", _ip), ("rofs", C.c_void_p),
                ("pci", _ip), ("pv", _dp), ("pv32", C.c_void_p), ("xpart", _dp),
                ("tick", C.c_void_p)]


def _hip():
    h = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _noalloc_spmv(dp, hip, csr, x):
    xd, yd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(xd), max(1, x.nbytes)) == 0
    assert hip.hipMalloc(C.byref(yd), max(1, 8 * csr.rows)) == 0
    try:
        assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, H2D) == 0
        rc = dp.lib.dpdlp_dev_spmv_noalloc(C.byref(csr), C.cast(xd, _dp), C.cast(yd, _dp), None)
        assert rc == 0, dp.last_error()
        assert hip.hipDeviceSynchronize() == 0
        y = np.empty(csr.rows)
        assert hip.hipMemcpy(y.ctypes.data, yd, y.nbytes, D2H) == 0
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
    return y


def _setup(dp):
    lib = dp.lib
    lib.dpdlp_dev_spmv_noalloc.restype = C.c_int
    lib.dpdlp_dev_spmv_noalloc.argtypes = [C.POINTER(DevCsr), _dp, _dp, C.c_void_p]
    lib.dpdlp_resident_device_csr.restype = C.c_int
    lib.dpdlp_resident_device_csr.argtypes = [C.c_void_p, C.c_int, C.c_void_p]


def test_noalloc_spmv_equals_spmv_with_caller_scratch(dp):
    _setup(dp)
    hip = _hip()
    rng = np.random.default_rng(41)
    cases = [(30000, 40000, (4, 2), 2048), (6000, 5000, (8, 4), 1024)]   # the second smaller
    for m, n, panels, cap in cases:
        lens = np.minimum((rng.pareto(1.3, m) * 4).astype(int), n - 1)
        rp, ci, v = random_csr(m, n, lens, seed=m, long_rows=((3, n - 1),))
        prob = dp.Problem.from_csr(m, n, rp, ci, v, c=np.zeros(n))
        R = dp.Resident(prob, spmv_cap=cap, panels=panels)
        for transpose in (0, 1):
            csr = DevCsr()
            assert dp.lib.dpdlp_resident_device_csr(R.h, transpose, C.byref(csr)) == 0
            assert csr.P == panels[transpose] and csr.rows == (n if transpose else m)
            x = rng.standard_normal(csr.cols)
            want = R.spmv(x, transpose=bool(transpose))
            got = _noalloc_spmv(dp, hip, csr, x)
            assert np.array_equal(got, want), (m, transpose)
        R.free()


def test_noalloc_spmv_rejects_matrices_beyond_its_scratch(dp):
    """A tiled matrix needing more partial sums than the fixed scratch holds
    is refused before anything is launched."""
    _setup(dp)
    csr = DevCsr(rows=1 << 20, cols=8, nnz=0, nblk=1, cap=1024, P=8)
    rc = dp.lib.dpdlp_dev_spmv_noalloc(C.byref(csr), None, None, None)
    assert rc != 0
    assert "exceed the fixed scratch" in dp.last_error()
