"""dpdlp_dev_spmv_noalloc, the SpMV entry point that allocates nothing (its
tiled-SpMV scratch is a fixed per-GPU array): called through the C ABI on
device buffers, it gives the same bits as dpdlp_dev_spmv with the caller's
scratch, on two calls in a row (the second smaller, with a different panel
count, so it must clear the tickets the first left behind)."""
import ctypes as C

import numpy as np
import pytest

from helpers import D2H, H2D, DevCsr, hip_runtime as _hip, random_csr

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)




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


def test_noalloc_spmv_replays_from_a_graph(dp):
    """The no-allocation SpMV captured into a hipGraph (its ticket clearing is
    a fill kernel, so the graph holds only kernel nodes) and replayed five
    times, with an eager SpMV of another matrix on the same scratch and a
    synchronisation between replays: every replay gives the eager result."""
    _setup(dp)
    hip = _hip()
    vp = C.c_void_p
    hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
    hip.hipStreamDestroy.argtypes = [vp]
    hip.hipStreamSynchronize.argtypes = [vp]
    hip.hipStreamBeginCapture.argtypes = [vp, C.c_int]
    hip.hipStreamEndCapture.argtypes = [vp, C.POINTER(vp)]
    hip.hipGraphInstantiate.argtypes = [C.POINTER(vp), vp, vp, vp, C.c_size_t]
    hip.hipGraphLaunch.argtypes = [vp, vp]
    hip.hipGraphExecDestroy.argtypes = [vp]
    hip.hipGraphDestroy.argtypes = [vp]
    hip.hipMemsetAsync.argtypes = [vp, C.c_int, C.c_size_t, vp]
    rng = np.random.default_rng(43)
    mats = []
    for m, n, panels in ((20000, 30000, (4, 1)), (4000, 6000, (8, 1))):
        lens = np.minimum((rng.pareto(1.3, m) * 4).astype(int), n - 1)
        rp, ci, v = random_csr(m, n, lens, seed=m + 1)
        R = dp.Resident(dp.Problem.from_csr(m, n, rp, ci, v, c=np.zeros(n)), spmv_cap=2048,
                        panels=panels)
        csr = DevCsr()
        assert dp.lib.dpdlp_resident_device_csr(R.h, 0, C.byref(csr)) == 0
        mats.append((R, csr))
    (R1, A1), (R2, A2) = mats
    x1 = rng.standard_normal(A1.cols)
    x2 = rng.standard_normal(A2.cols)
    want1 = _noalloc_spmv(dp, hip, A1, x1)
    want2 = _noalloc_spmv(dp, hip, A2, x2)
    assert np.array_equal(want1, R1.spmv(x1))
    s = vp()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    bufs = {}
    for name, size in (("x", x1.nbytes), ("y", 8 * A1.rows)):
        p = vp()
        assert hip.hipMalloc(C.byref(p), size) == 0
        bufs[name] = p
    graph, exe = vp(), vp()
    try:
        assert hip.hipMemcpy(bufs["x"], x1.ctypes.data, x1.nbytes, H2D) == 0
        assert hip.hipStreamBeginCapture(s, 2) == 0          # relaxed capture
        rc = dp.lib.dpdlp_dev_spmv_noalloc(C.byref(A1), C.cast(bufs["x"], _dp),
                                           C.cast(bufs["y"], _dp), s)
        assert hip.hipStreamEndCapture(s, C.byref(graph)) == 0
        assert rc == 0, dp.last_error()
        assert hip.hipGraphInstantiate(C.byref(exe), graph, None, None, 0) == 0
        for k in range(5):
            assert hip.hipMemsetAsync(bufs["y"], 0, 8 * A1.rows, s) == 0   # outside the graph
            assert hip.hipGraphLaunch(exe, s) == 0
            assert hip.hipStreamSynchronize(s) == 0
            got = np.empty(A1.rows)
            assert hip.hipMemcpy(got.ctypes.data, bufs["y"], got.nbytes, D2H) == 0
            assert np.array_equal(got, want1), k
            # eager work between replays: another tiled SpMV on the same scratch
            assert np.array_equal(_noalloc_spmv(dp, hip, A2, x2), want2), k
    finally:
        if exe:
            hip.hipGraphExecDestroy(exe)
        if graph:
            hip.hipGraphDestroy(graph)
        hip.hipStreamSynchronize(s)
        for p in bufs.values():
            hip.hipFree(p)
        hip.hipStreamDestroy(s)
        R1.free()
        R2.free()
