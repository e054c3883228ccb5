"""GPU parity tests of the core path: SpMV, full solves vs the oracle."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import DevCsr, capi_example, csr_matvec, device_array, hip_runtime, random_csr, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cap", [1024, 2048, 4096])
@pytest.mark.parametrize("shape", ["uniform", "powerlaw", "empty_rows", "rows17to23",
                                   "rows33to64"])
def test_spmv_matches_brute_force(dp, cap, shape):
    """Row lengths that put every row of a block in one lane (rows of 17-23
    entries, about 200 rows per 4096-nonzero block), in lane pairs and
    groups, or on whole wavefronts (power-law rows, long rows)."""
    rng = np.random.default_rng(5)
    m, n = 3000, 2000
    if shape == "uniform":
        lens = rng.poisson(10, m)
        extra = ()
    elif shape == "rows17to23":
        lens = rng.integers(17, 24, m)
        extra = ()
    elif shape == "rows33to64":
        lens = rng.integers(33, 65, m)
        extra = ((1500, 1900),)
    elif shape == "powerlaw":
        lens = np.minimum((rng.pareto(1.5, m) * 3).astype(int) + 1, 1500)
        extra = ((7, 1999), (2500, 1800), (2999, 4096 if cap < 4096 else 1024))
    else:
        lens = rng.integers(0, 3, m) * (rng.random(m) < 0.3)
        extra = ((100, 600),)
    rp, ci, v = random_csr(m, n, lens, seed=11, long_rows=extra)
    c = np.zeros(n)
    prob = dp.Problem.from_csr(m, n, rp, ci, v, c=c)
    R = dp.Resident(prob, spmv_cap=cap)
    x = rng.standard_normal(n)
    y = R.spmv(x)
    ref = csr_matvec(rp, ci, v, x)
    assert np.max(np.abs(y - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    yv = rng.standard_normal(m)
    z = R.spmv(yv, transpose=True)
    # adjointness y'(Ax) = (A'y)'x
    assert abs(yv @ y - z @ x) <= 1e-10 * (np.abs(yv) @ np.abs(ref) + 1)
    refz = np.zeros(n)
    np.add.at(refz, ci, v * np.repeat(yv, np.diff(rp)))
    assert np.max(np.abs(z - refz)) <= 1e-12 * max(1.0, np.max(np.abs(refz)))


def test_capi_example(dp):
    A, c, l, u = capi_example()
    prob = dp.Problem.from_dense(A, c=c, con_lb=l, con_ub=u)
    res = dp.solve_plain(prob, dp.default_params(presolve=False))
    assert res.status == "OPTIMAL"
    assert np.allclose(res.x, [1.0, 2.0], atol=1e-3)
    assert np.allclose(res.y, [1.0, -1.0, 0.0], atol=1e-3)
    assert abs(res.primal_obj - 3.0) < 1e-3


@pytest.mark.parametrize("seed", [0, 1])
def test_parity_trace_vs_oracle(dp, orc, seed, row_order):
    """Iterates within 1e-8 relative of the oracle over the first 1000
    iterations, same status, objective within 1e-6 relative."""
    prob, opt = dp.generate(2000, 5000, 50000, row_rule=0, seed=seed)
    iters = [0, 1, 10, 50, 100, 200, 201, 400, 600, 800, 1000]
    params = dp.default_params(presolve=False)
    res, o, (tx, ty) = dp.solve(prob, params, trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(presolve=False), trace_iters=iters,
                                 n=prob.n, m=prob.m)
    assert res.status == ref["status"]
    k = len(ox)
    assert len(tx) == k > 0
    for i in range(k):
        assert rel(tx[i], ox[i]) < 1e-8, (iters[i], rel(tx[i], ox[i]))
        assert rel(ty[i], oy[i]) < 1e-8, (iters[i], rel(ty[i], oy[i]))
    assert abs(res.primal_obj - ref["primal_objective_value"]) <= 1e-6 * max(
        1.0, abs(ref["primal_objective_value"]))
    assert res.rel_primal < 1e-4 and res.rel_dual < 1e-4 and res.rel_gap < 1e-4
    assert abs(res.primal_obj - opt) <= 1e-3 * max(1.0, abs(opt))


def test_banded_lp_matches_oracle(dp, orc):
    """A banded LP (row i's 20 entries in columns 2i-10 .. 2i+9, so every row
    of a 4096-nonzero block is summed by one lane), feasible by construction
    (ranged rows around A x* with x* in the unit box): the trace of the
    iterates within 1e-8 of the oracle's, the same status and iteration count,
    objective within 1e-6."""
    rng = np.random.default_rng(21)
    m, n, per = 4000, 8000, 20
    start = np.clip(2 * np.arange(m) - per // 2, 0, n - per)
    ci = (start[:, None] + np.arange(per)[None, :]).ravel()
    rp = np.arange(0, m * per + 1, per)
    v = rng.standard_normal(m * per)
    xs = rng.random(n)
    b = np.add.reduceat(v * xs[ci], rp[:-1])
    prob = dp.Problem.from_csr(m, n, rp, ci, v, c=rng.standard_normal(n),
                               con_lb=b - rng.random(m), con_ub=b + rng.random(m),
                               var_lb=np.zeros(n), var_ub=np.ones(n))
    iters = [0, 10, 100, 300, 600]
    kw = dict(presolve=False, termination_criteria__iteration_limit=20000)
    res, o, (tx, ty) = dp.solve(prob, dp.default_params(**kw), trace_iters=iters, spmv_cap=4096)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(**kw), trace_iters=iters, n=n, m=m)
    print(f"banded LP: {res.status} at {res.iterations} (oracle {ref['total_count']})", flush=True)
    assert res.status == ref["status"] == "OPTIMAL"
    assert res.iterations == ref["total_count"]
    assert len(tx) == len(ox) == len(iters)
    for i in range(len(ox)):
        assert rel(tx[i], ox[i]) < 1e-8, (iters[i], rel(tx[i], ox[i]))
        assert rel(ty[i], oy[i]) < 1e-8, (iters[i], rel(ty[i], oy[i]))
    obj = ref["primal_objective_value"]
    assert abs(res.primal_obj - obj) <= 1e-6 * max(1.0, abs(obj))


CAP = dict(presolve=False, termination_criteria__iteration_limit=5000)


def test_primal_infeasible(dp, orc, row_order):
    # x + y >= 2 and x + y <= 1 over the unit box
    A = np.array([[1.0, 1.0], [1.0, 1.0]])
    prob = dp.Problem.from_dense(A, c=np.zeros(2), con_lb=np.array([2.0, -np.inf]),
                                 con_ub=np.array([np.inf, 1.0]), var_lb=np.zeros(2),
                                 var_ub=np.ones(2))
    res = dp.solve_plain(prob, dp.default_params(**CAP))
    ref, _ = orc.solve(prob.ptr, orc.default_params(**CAP))
    assert res.status == ref["status"] == "PRIMAL_INFEASIBLE"
    assert res.iterations == ref["total_count"] == 200


def test_dual_infeasible(dp, orc, row_order):
    # min -x s.t. x - y >= 0 (x, y >= 0): unbounded along x = y
    A = np.array([[1.0, -1.0]])
    prob = dp.Problem.from_dense(A, c=np.array([-1.0, 0.0]), con_lb=np.array([0.0]),
                                 con_ub=np.array([np.inf]), var_lb=np.zeros(2),
                                 var_ub=np.full(2, np.inf))
    res = dp.solve_plain(prob, dp.default_params(**CAP))
    ref, _ = orc.solve(prob.ptr, orc.default_params(**CAP))
    assert res.status == ref["status"] == "DUAL_INFEASIBLE"
    assert abs(res.raw["primal_ray_linear_objective"] - ref["primal_ray_linear_objective"]) < 1e-9


def test_iteration_and_time_limits(dp):
    prob, _ = dp.generate(500, 1200, 8000, row_rule=0, seed=3)
    r0 = dp.solve_plain(prob, dp.default_params(presolve=False,
                                                termination_criteria__iteration_limit=0))
    assert r0.status == "ITERATION_LIMIT" and r0.iterations == 0
    r1 = dp.solve_plain(prob, dp.default_params(presolve=False,
                                                termination_criteria__iteration_limit=37))
    assert r1.status == "ITERATION_LIMIT" and r1.iterations == 37
    r2 = dp.solve_plain(prob, dp.default_params(presolve=False,
                                                termination_criteria__time_sec_limit=0.0))
    assert r2.status == "TIME_LIMIT" and r2.iterations == 0


@pytest.fixture
def panels(request):
    import os
    old = os.environ.get("DPDLP_PANELS")
    os.environ["DPDLP_PANELS"] = request.param
    yield request.param
    if old is None:
        del os.environ["DPDLP_PANELS"]
    else:
        os.environ["DPDLP_PANELS"] = old


@pytest.mark.parametrize("panels", ["2,2", "4,8", "8,4"], indirect=True)
@pytest.mark.parametrize("cap", [1024, 2048])
def test_spmv_column_panels(dp, panels, cap):
    """XCD column-panel tiles: same products as brute force, long rows split
    across panels, empty tiles."""
    rng = np.random.default_rng(17)
    m, n = 2500, 3000
    lens = np.minimum((rng.pareto(1.3, m) * 3).astype(int), 2900)
    rp, ci, v = random_csr(m, n, lens, seed=23, long_rows=((0, 2999), (1234, 2500)))
    prob = dp.Problem.from_csr(m, n, rp, ci, v, c=np.zeros(n))
    R = dp.Resident(prob, spmv_cap=cap)
    x = rng.standard_normal(n)
    ref = csr_matvec(rp, ci, v, x)
    y = R.spmv(x)
    assert np.max(np.abs(y - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    yv = rng.standard_normal(m)
    z = R.spmv(yv, transpose=True)
    refz = np.zeros(n)
    np.add.at(refz, ci, v * np.repeat(yv, np.diff(rp)))
    assert np.max(np.abs(z - refz)) <= 1e-12 * max(1.0, np.max(np.abs(refz)))
    y32 = R.spmv(x, use32=True)
    assert np.max(np.abs(y32 - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("panels", ["2,2", "4,4", "8,8"], indirect=True)
def test_tiled_spmv_one_pass_equals_two_kernels(dp, panels, monkeypatch):
    """The one-pass tiled SpMV (the last of a row block's P workgroups to
    finish adds the panel partials; DPDLP_TILE_ONEPASS=1) gives the same bits
    as the two-kernel form, on the first launch and after many launches on the
    same tickets, with rows longer than a tile, empty tiles and blocks of up to
    1024 rows.  The library reads the switch once per process, so this test
    runs the one-pass form in a child process."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path[:0] = sys.argv[1:]\n"
        "import test_gpu_core as t, dpdlp\n"
        "t._one_pass_vs_two_kernels(dpdlp)\n")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    env = dict(os.environ, DPDLP_TILE_ONEPASS="1")
    r = subprocess.run([sys.executable, "-c", code, here, os.path.join(root, "d-pdlp_amd")],
                       env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr


def _one_pass_vs_two_kernels(dp):
    rng = np.random.default_rng(29)
    m, n = 6000, 5000
    lens = np.minimum((rng.pareto(1.2, m) * 2).astype(int), 4900)
    rp, ci, v = random_csr(m, n, lens, seed=31, long_rows=((7, 4999), (4000, 3000)))
    prob = dp.Problem.from_csr(m, n, rp, ci, v, c=np.zeros(n))
    R = dp.Resident(prob, spmv_cap=1024)
    for transpose, length in ((False, n), (True, m)):
        x = rng.standard_normal(length)
        two = R.spmv(x, transpose=transpose, two_kernel=True)
        one = R.spmv(x, transpose=transpose)
        many = R.spmv(x, transpose=transpose, reps=37)
        assert np.array_equal(one, two) and np.array_equal(many, two), transpose
        for use32 in (True,):
            a32 = R.spmv(x, transpose=transpose, use32=True, reps=5)
            b32 = R.spmv(x, transpose=transpose, use32=True, two_kernel=True)
            assert np.array_equal(a32, b32)
    R.free()


@pytest.mark.parametrize("panels", ["8,4", "2,2"], indirect=True)
def test_parity_trace_with_panels(dp, orc, panels, row_order):
    """Tiled A and A' against the oracle, in the caller's row order and with
    the rows renumbered by length, where a tiled A' deals the length order
    round-robin over its column panels."""
    prob, opt = dp.generate(2000, 5000, 50000, row_rule=0, seed=2)
    iters = [0, 10, 100, 400, 1000]
    res, o, (tx, ty) = dp.solve(prob, dp.default_params(presolve=False), trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(presolve=False), trace_iters=iters,
                                 n=prob.n, m=prob.m)
    assert res.status == ref["status"] == "OPTIMAL"
    for i in range(len(ox)):
        assert rel(tx[i], ox[i]) < 1e-8 and rel(ty[i], oy[i]) < 1e-8, iters[i]
    assert abs(res.primal_obj - ref["primal_objective_value"]) <= 1e-6 * max(
        1.0, abs(ref["primal_objective_value"]))


def _stable_transpose(rp, ci, v, n):
    """A' by a stable sort of the row-major nonzeros on their column."""
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int32), np.diff(rp))
    order = np.argsort(ci, kind="stable")
    t_rp = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(ci, minlength=n), out=t_rp[1:])
    return t_rp, rows[order], v[order]


@pytest.mark.parametrize("case", ["uniform", "zipf", "poisson", "wide", "one_col", "duplicates"])
def test_device_transpose_matches_stable_sort(dp, case):
    """The untiled A' is transposed on the device (radix sort of the nonzero
    positions by column): bit for bit the stable counting-sort transpose,
    including the 1-, 2- and 3-pass key widths, empty rows and columns and a
    repeated (row, column) entry."""
    if case == "uniform":
        prob, _ = dp.generate(700, 300, 6000, row_rule=0, seed=5)        # 2 passes
    elif case == "zipf":
        prob, _ = dp.generate(5000, 70000, 90000, row_rule=1, seed=6)    # 3 passes, long rows
    elif case == "poisson":
        prob, _ = dp.generate(3000, 200, 0, row_rule=2, seed=7)          # 1 pass
    elif case == "wide":
        prob, _ = dp.generate(300, 20000, 0, row_rule=2, poisson_lambda=40.0, seed=8)
    elif case == "one_col":
        prob = dp.Problem.from_csr(4, 1, [0, 1, 1, 2, 3], [0, 0, 0], [1.0, 2.0, 3.0],
                                   c=np.ones(1), var_lb=np.zeros(1), var_ub=np.ones(1))
    else:
        prob = dp.Problem.from_csr(3, 4, [0, 3, 3, 5], [2, 0, 2, 3, 2], [1.0, 2.0, 3.0, 4.0, 5.0],
                                   c=np.ones(4), var_lb=np.zeros(4), var_ub=np.ones(4))
    a = prob.arrays()
    R = dp.Resident(prob, spmv_cap=1024, panels=(1, 1))
    got = R.transpose()
    want = _stable_transpose(a["row_ptr"], a["col_ind"], a["vals"], prob.n)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    R.free()


def _host_tiles(dp, rows, cols, rp, ci, v, P, cap):
    """Block plan and column-panel tiles cut on the host (the planner and
    cutter over a host CSR): blk, toff, rofs, tiled columns, tiled values."""
    Pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    nnz = len(ci)
    blk = np.zeros(rows + 2, np.int32)
    nb = dp.lib.dpdlp_plan_blocks_panels(rows, cols, Pi(rp), Pi(ci), cap, P, 1024, Pi(blk))
    toff = np.zeros(nb * P + 1, np.int32)
    rofs = np.zeros(P * (rows + nb), np.uint16)
    perm = np.zeros(max(1, nnz), np.int32)
    pci = np.zeros(max(1, nnz), np.int32)
    assert dp.lib.dpdlp_build_panels(rows, cols, Pi(rp), Pi(ci), Pi(blk), nb, P, Pi(toff),
                                     rofs.ctypes.data_as(C.POINTER(C.c_ushort)), Pi(perm),
                                     Pi(pci)) == 0
    return blk[:nb + 1], toff, rofs, pci[:nnz], v[perm[:nnz]]


@pytest.mark.parametrize("P", [2, 4, 8])
@pytest.mark.parametrize("case", ["skewed", "dense_column", "sparse"])
def test_device_cut_tiles_equal_host_cut(dp, case, P):
    """Column-panel tiles are cut on the device: A's from the uploaded CSR,
    A''s (planned from A's host CSR, with no host transpose) from the device
    transpose.  Both have the same block plan, tile offsets, tile-local row
    offsets, tiled columns and tiled values as the host planner and cutter
    give on A and on the stable transpose -- with long rows of A split over
    panels, a dense column of A (an A' row above the block capacity, a block
    of its own), and empty rows and columns -- and the same products."""
    rng = np.random.default_rng(71 + P)
    if case == "skewed":
        m, n = 3000, 2500
        lens = np.minimum((rng.pareto(1.3, m) * 3).astype(int), n - 1)
        rp, ci, v = random_csr(m, n, lens, seed=5, long_rows=((0, n - 1), (17, 1800)))
    elif case == "dense_column":
        m, n = 6000, 400
        lens = rng.integers(1, 6, m)
        rp, ci, v = random_csr(m, n, lens, seed=6)
        rows = np.repeat(np.arange(m), np.diff(rp))
        keep = ci != 7
        rows, ci, v = rows[keep], ci[keep], v[keep]
        rows = np.concatenate([rows, np.arange(m)])     # column 7 in every row
        ci = np.concatenate([ci, np.full(m, 7, np.int32)])
        v = np.concatenate([v, rng.standard_normal(m)])
        o = np.lexsort((ci, rows))
        rows, ci, v = rows[o], ci[o].astype(np.int32), v[o]
        rp = np.zeros(m + 1, np.int32)
        np.cumsum(np.bincount(rows, minlength=m), out=rp[1:])
    else:
        m, n = 2000, 3000
        lens = rng.integers(0, 3, m) * (rng.random(m) < 0.5)
        rp, ci, v = random_csr(m, n, lens, seed=7)
    prob = dp.Problem.from_csr(m, n, rp, ci, v, c=np.zeros(n))
    R = dp.Resident(prob, spmv_cap=1024, panels=(P, P))
    assert R.panels == (P, P)
    dp.lib.dpdlp_resident_device_csr.restype = C.c_int
    dp.lib.dpdlp_resident_device_csr.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    hip = hip_runtime()
    t_rp, t_ci, t_v = _stable_transpose(rp, ci, v, n)
    for transpose, mat in ((0, (m, n, rp, ci, v)), (1, (n, m, t_rp, t_ci, t_v))):
        rows, cols = mat[0], mat[1]
        csr = DevCsr()
        assert dp.lib.dpdlp_resident_device_csr(R.h, transpose, C.byref(csr)) == 0
        assert (csr.rows, csr.cols, csr.P) == (rows, cols, P)
        blk, toff, rofs, pci, pv = _host_tiles(dp, *mat, P, csr.cap)
        nb = len(blk) - 1
        assert csr.nblk == nb
        assert np.array_equal(device_array(hip, csr.blk, nb + 1, np.int32), blk)
        assert np.array_equal(device_array(hip, csr.toff, nb * P + 1, np.int32), toff)
        assert np.array_equal(device_array(hip, csr.rofs, P * (rows + nb), np.uint16), rofs)
        assert np.array_equal(device_array(hip, csr.pci, len(ci), np.int32), pci)
        assert np.array_equal(device_array(hip, csr.pv, len(ci), np.float64), pv)
    x = rng.standard_normal(n)
    ref = csr_matvec(rp, ci, v, x)
    assert np.max(np.abs(R.spmv(x) - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    yv = rng.standard_normal(m)
    z = R.spmv(yv, transpose=True)
    refz = np.zeros(n)
    np.add.at(refz, ci, v * np.repeat(yv, np.diff(rp)))
    assert np.max(np.abs(z - refz)) <= 1e-12 * max(1.0, np.max(np.abs(refz)))
    R.free()


def test_device_reductions(dp):
    """Standalone device reductions against exactly rounded host references:
    nrm2 (Blue's scaled sums) without overflow at 1e200 or underflow at
    1e-200, dot, sum, max |x|; n = 0 gives 0; repeated calls give the same
    bits (fixed combination order)."""
    import math
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1_000_003)
    y = rng.standard_normal(1_000_003)
    nrm = dp.reduce(dp.RED_NRM2, x)
    assert abs(nrm - math.sqrt(math.fsum(x * x))) <= 1e-14 * nrm
    dot = dp.reduce(dp.RED_DOT, x, y)
    ref = math.fsum(x * y)
    assert abs(dot - ref) <= 1e-12 * math.fsum(np.abs(x * y))
    sm = dp.reduce(dp.RED_SUM, x)
    assert abs(sm - math.fsum(x)) <= 1e-12 * math.fsum(np.abs(x))
    assert dp.reduce(dp.RED_AMAX, -np.abs(x)) == np.max(np.abs(x))
    assert dp.reduce(dp.RED_NRM2, x) == nrm and dp.reduce(dp.RED_DOT, x, y) == dot
    for scale in (1e200, 1e-200):
        v = np.full(1000, scale)
        r = dp.reduce(dp.RED_NRM2, v)
        assert np.isfinite(r) and r > 0
        assert abs(r - scale * math.sqrt(1000.0)) <= 1e-14 * scale * math.sqrt(1000.0)
    mixed = np.array([1e200, 3.0, -4.0, 1e-200])
    assert abs(dp.reduce(dp.RED_NRM2, mixed) - 1e200) <= 1e-15 * 1e200
    for op in (dp.RED_NRM2, dp.RED_SUM, dp.RED_AMAX):
        assert dp.reduce(op, np.zeros(0)) == 0.0
    assert dp.reduce(dp.RED_DOT, np.zeros(0), np.zeros(0)) == 0.0


def _oracle_rescale(orc, prob, **kw):
    prm = orc.default_params(**kw)
    m, n, nnz = prob.m, prob.n, prob.nnz
    out = dict(val=np.zeros(nnz), c=np.zeros(n), lv=np.zeros(n), uv=np.zeros(n), lc=np.zeros(m),
               uc=np.zeros(m), Dr=np.zeros(m), Dc=np.zeros(n))
    beta, gamma = C.c_double(), C.c_double()
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    orc.lib().orc_rescale(C.byref(prm), prob.ptr, P(out["val"]), P(out["c"]), P(out["lv"]),
                          P(out["uv"]), P(out["lc"]), P(out["uc"]), P(out["Dr"]), P(out["Dc"]),
                          C.byref(beta), C.byref(gamma))
    out["beta"], out["gamma"] = beta.value, gamma.value
    return out


@pytest.mark.parametrize("kw", [dict(), dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False),
                                dict(bound_objective_rescaling=False)],
                         ids=["default", "bound_obj_only", "no_bound_obj"])
@pytest.mark.parametrize("shape", ["uniform", "zipf", "empty_rows"])
def test_device_scaling_bitwise_equals_oracle(dp, orc, kw, shape):
    """The device's Ruiz and Pock-Chambolle scaling gives the oracle's
    (preconditioner.c) scaled matrix, D_r and D_c bit for bit: maxima are
    exact, sums of |a| run in storage order over A and A' as the host
    accumulates them, and the factors are correctly rounded square roots and
    quotients.  The bound/objective step's two sums of squares are reduced in
    the device's fixed tree order instead of the host's sequential one, so
    beta, gamma and the vectors they scale agree to 1e-14 relative (the
    reordered sums of ~1e4 squares; bit for bit when that step is off)."""
    if shape == "uniform":
        prob, _ = dp.generate(600, 1400, 9000, row_rule=0, seed=21)
    elif shape == "zipf":
        prob, _ = dp.generate(2000, 5000, 40000, row_rule=1, pm1_frac=0.1, seed=22)
    else:
        prob = dp.Problem.from_csr(4, 3, [0, 2, 2, 3, 3], [0, 2, 1], [2.0, -3.0, 0.5],
                                   c=np.array([1.0, 0.0, -2.0]),
                                   con_lb=np.array([1.0, -np.inf, 0.0, 2.0]),
                                   con_ub=np.array([4.0, np.inf, 0.0, np.inf]),
                                   var_lb=np.array([0.0, -np.inf, -1.0]),
                                   var_ub=np.array([np.inf, 3.0, 1.0]))
    params = dp.default_params(presolve=False, **kw)
    dev = dp.device_scaling(prob, params)
    ref = _oracle_rescale(orc, prob, presolve=False, **kw)
    for k in ("val", "Dr", "Dc"):
        assert np.array_equal(dev[k], ref[k]), k
    if params.bound_objective_rescaling:
        for k in ("beta", "gamma"):
            assert abs(dev[k] - ref[k]) <= 1e-14 * abs(ref[k]), k
        for k in ("c", "lv", "uv", "lc", "uc"):
            fin = np.isfinite(ref[k])
            assert np.array_equal(np.isfinite(dev[k]), fin), k
            assert np.all(np.abs(dev[k][fin] - ref[k][fin]) <= 1e-14 * np.abs(ref[k][fin])), k
    else:
        for k in ("c", "lv", "uv", "lc", "uc"):
            assert np.array_equal(dev[k], ref[k]), k
        assert dev["beta"] == ref["beta"] == 1.0 and dev["gamma"] == ref["gamma"] == 1.0


@pytest.mark.parametrize("case", ["diagonal", "rank_one", "random"])
def test_power_iteration_step_size(dp, orc, case):
    """The device power iteration (scaling off, so it sees A itself) gives the
    step 0.998 / sigma_max with sigma_max of the closed form (diagonal: max
    |a_ii|; rank one u v': |u| |v|) or of numpy's SVD, and the same step and
    iteration count as the oracle's power iteration from the same start."""
    rng = np.random.default_rng(12)
    if case == "diagonal":
        d = np.array([3.0, -7.0, 2.0, 0.5, 6.0])
        A = np.diag(d)
        sigma = 7.0
    elif case == "rank_one":
        u, w = np.array([1.0, 2.0, 2.0]), np.array([3.0, 4.0])
        A = np.outer(u, w)
        sigma = np.linalg.norm(u) * np.linalg.norm(w)
    else:
        A = rng.standard_normal((30, 20)) * (rng.random((30, 20)) < 0.3)
        sigma = np.linalg.svd(A, compute_uv=False)[0]
    m, n = A.shape
    prob = dp.Problem.from_dense(A, c=np.ones(n), con_lb=np.zeros(m), con_ub=np.ones(m),
                                 var_lb=np.zeros(n), var_ub=np.ones(n))
    params = dp.default_params(presolve=False, l_inf_ruiz_iterations=0,
                               has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                               termination_criteria__iteration_limit=1)
    res, o, _ = dp.solve(prob, params)
    assert abs(o.step_size - 0.998 / sigma) <= 1e-6 * (0.998 / sigma)
    a = prob.arrays()
    nnz = len(a["vals"])
    trp, tci, tv = np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
    P = lambda x, t=C.c_double: x.ctypes.data_as(C.POINTER(t))
    L = orc.lib()
    L.orc_transpose(m, n, P(a["row_ptr"], C.c_int), P(a["col_ind"], C.c_int), P(a["vals"]),
                    P(trp, C.c_int), P(tci, C.c_int), P(tv))
    it = C.c_int()
    s_orc = L.orc_power_iteration(m, n, P(a["row_ptr"], C.c_int), P(a["col_ind"], C.c_int),
                                  P(a["vals"]), P(trp, C.c_int), P(tci, C.c_int), P(tv),
                                  params.sv_max_iter, params.sv_tol, C.byref(it))
    assert o.power_iterations == it.value
    assert abs(o.step_size - 0.998 / s_orc) <= 1e-12 * o.step_size


FUSE_CASES = [(2000, 5000, 50000, None), (2000, 5000, 50000, (1, 2)),
              (100_000, 200_000, 1_000_000, (1, 1))]
FUSE_CHILD = """
import sys, numpy as np
sys.path.insert(0, sys.argv[2])
import dpdlp as dp
sys.path.insert(0, sys.argv[3])
from test_gpu_core import FUSE_CASES, fuse_case
out = {}
for k, case in enumerate(FUSE_CASES):
    res, o = fuse_case(dp, *case)
    out.update({f"x{k}": res.x, f"y{k}": res.y, f"step{k}": o.step_size,
                f"pi{k}": o.power_iterations})
np.savez(sys.argv[1], **out)
"""


def fuse_case(dp, m, n, nnz, panels):
    prob, _ = dp.generate(m, n, nnz, seed=0)
    params = dp.default_params(presolve=False, termination_criteria__iteration_limit=100)
    if panels is None:
        res, o, _ = dp.solve(prob, params)
        return res, o
    r = dp.Resident(prob, spmv_cap=1024, panels=panels)
    assert r.panels == panels
    return r.solve(params)


def test_fused_power_step_gives_the_same_bits(dp, tmp_path):
    """The power iteration folds each step's update into the next A'
    product and totals partial sums inside the kernels that need them (two
    launches per step on a small LP, four on a larger one, three with A' on
    column panels); a process that keeps the five-launch step
    (DPDLP_PI_FUSE=0) gets the same power-iteration count, step size and
    iterates bit for bit in each case."""
    import os
    import subprocess
    import sys
    out = tmp_path / "unfused.npz"
    pkg = os.path.dirname(os.path.dirname(dp.__file__))
    env = dict(os.environ, DPDLP_PI_FUSE="0")
    subprocess.run([sys.executable, "-c", FUSE_CHILD, str(out), pkg,
                    os.path.dirname(os.path.abspath(__file__))], env=env, check=True,
                   timeout=300)
    ref = np.load(out)
    for k, case in enumerate(FUSE_CASES):
        res, o = fuse_case(dp, *case)
        assert o.power_iterations == int(ref[f"pi{k}"]) > 20, case
        assert o.step_size == float(ref[f"step{k}"]), case
        assert np.array_equal(res.x, ref[f"x{k}"]), case
        assert np.array_equal(res.y, ref[f"y{k}"]), case


def test_residuals_vanish_at_known_optimum(dp, orc):
    """Residual closed form (utils.cu:639-719): started at the C API example's
    optimum x* = (1, 2), y* = (1, -1, 0), the evaluation at iteration 0 finds
    zero primal and dual residuals and gap, primal = dual objective = 3, and
    the solve ends there, as the oracle's does."""
    A, c, l, u = capi_example()
    prob = dp.Problem.from_dense(A, c=c, con_lb=l, con_ub=u)
    prob.set_start(np.array([1.0, 2.0]), np.array([1.0, -1.0, 0.0]))
    res = dp.solve_plain(prob, dp.default_params(presolve=False))
    ref, _ = orc.solve(prob.ptr, orc.default_params(presolve=False))
    assert res.status == ref["status"] == "OPTIMAL"
    assert res.iterations == ref["total_count"] == 0
    assert res.rel_primal < 1e-12 and res.rel_dual < 1e-12 and res.rel_gap < 1e-12
    assert abs(res.primal_obj - 3.0) < 1e-12 and abs(res.dual_obj - 3.0) < 1e-12
    assert np.allclose(res.x, [1.0, 2.0], rtol=0, atol=1e-12)
    assert np.allclose(res.y, [1.0, -1.0, 0.0], rtol=0, atol=1e-12)
