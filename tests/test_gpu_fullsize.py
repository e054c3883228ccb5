"""Full-size checks through properties that do not depend on the size: the
large LP (m = 1e6, n = 2e6, nnz = 2e7, Zipf rows) solved through the C API
must return a solution whose feasibility, reduced costs and objective hold
when recomputed independently on the host (scipy), and batch members of the
largest batch-workload sizes must give the same bits in a mix as alone."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def test_large_zipf_lp_solution_properties(dp):
    prob, opt = dp.generate(1_000_000, 2_000_000, 20_000_000, row_rule=1, pm1_frac=0.1,
                            seed=1000)
    a = prob.arrays()
    res = dp.solve_plain(prob, dp.default_params(presolve=False))
    assert res.status == "OPTIMAL"
    A = sp.csr_matrix((a["vals"], a["col_ind"], a["row_ptr"]), shape=(prob.m, prob.n))
    x, y, rc = res.x, res.y, res.reduced_cost
    c, lc, uc, lv, uv = a["c"], a["con_lb"], a["con_ub"], a["var_lb"], a["var_ub"]
    # primal: bounds hold, constraint violation small against the data's scale
    assert np.all(x >= lv - 1e-9 * (1 + np.abs(np.where(np.isfinite(lv), lv, 0))))
    assert np.all(x <= uv + 1e-9 * (1 + np.abs(np.where(np.isfinite(uv), uv, 0))))
    ax = A @ x
    viol = ax - np.clip(ax, lc, uc)
    bnorm = np.linalg.norm(np.concatenate([lc[np.isfinite(lc)], uc[np.isfinite(uc)]]))
    assert np.linalg.norm(viol) <= 1e-3 * (1 + bnorm)
    # dual signs of the projection (pdlp_core_op.cu:420-443): >= rows y >= 0, <= rows y <= 0
    ge = np.isfinite(lc) & ~np.isfinite(uc)
    le = ~np.isfinite(lc) & np.isfinite(uc)
    assert ge.any() and np.all(y[ge] >= 0) and np.all(y[le] <= 0)
    # reduced costs are c - A'y of the returned duals
    r = c - A.T @ y
    assert np.linalg.norm(rc - r) <= 1e-9 * (1 + np.linalg.norm(c))
    # objective: reported value is c'x + c0 of the returned x, near the known optimum
    obj = float(c @ x) + a.get("obj_const", 0.0)
    assert abs(res.primal_obj - obj) <= 1e-9 * max(1.0, abs(obj))
    assert abs(obj - opt) <= 1e-3 * max(1.0, abs(opt))
    assert res.rel_primal < 1e-4 and res.rel_dual < 1e-4 and res.rel_gap < 1e-4


def test_largest_batch_members_alone_equal_mix(dp):
    """Members of the uneven workload with the most rows (up to m = 2e5),
    mixed precision, 4096-nonzero blocks: each equals its solve alone."""
    import ctypes as C
    spec_m = []
    for k in range(512):
        spec = dp.GenSpec()
        dp.lib.dpdlp_batch_member_spec(k, 3000, 500, 200_000, C.byref(spec))
        spec_m.append((spec.m, k))
    ks = [k for _, k in sorted(spec_m)[-6:]]
    probs = [dp.batch_member(k)[0] for k in ks]
    res = [dp.Resident(p, spmv_cap=4096, panels=(1, 1)) for p in probs]
    params = dp.default_params(presolve=False)
    mix = dp.solve_batch(res, params, spmv_cap=4096, precision=dp.PREC_MIXED)
    for i, r in enumerate(res):
        alone = dp.solve_batch([r], params, spmv_cap=4096, precision=dp.PREC_MIXED)[0]
        assert alone.status == mix[i].status == "OPTIMAL"
        assert alone.iterations == mix[i].iterations
        assert np.array_equal(alone.x, mix[i].x) and np.array_equal(alone.y, mix[i].y)
    for r in res:
        r.free()


def test_large_zipf_lp_matches_openmp_oracle(dp, orc):
    """The m = 1e6, n = 2e6, nnz = 2e7 Zipf LP (seed 1000) solved on the device
    through the C API's automatic layout (rows renumbered by length, A on XCD
    column panels) against the OpenMP build of the FP64 oracle on the host (its
    loops over rows and columns are split across threads; each row's sum is
    still taken in order): same status and iteration count, objective within
    1e-6, and x_cur / y_cur within 1e-8 at trace points through the first
    restart (T = 200) and beyond."""
    from helpers import rel
    prob, opt = dp.generate(1_000_000, 2_000_000, 20_000_000, row_rule=1, pm1_frac=0.1,
                            seed=1000)
    iters = [0, 10, 100, 200, 201, 400, 600]
    params = dp.default_params(presolve=False)
    res, opts, (tx, ty) = dp.solve(prob, params, trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(presolve=False), trace_iters=iters,
                                 n=prob.n, m=prob.m, omp=True)
    print(f"device {res.status} T={res.iterations} obj={res.primal_obj!r}; oracle "
          f"{ref['status']} T={ref['total_count']} obj={ref['primal_objective_value']!r}",
          flush=True)
    assert res.status == ref["status"] == "OPTIMAL"
    assert res.iterations == ref["total_count"]
    obj = ref["primal_objective_value"]
    assert abs(res.primal_obj - obj) <= 1e-6 * max(1.0, abs(obj))
    k = min(len(tx), len(ox))
    assert k == len(iters)
    worst = 0.0
    for i in range(k):
        ex, ey = rel(tx[i], ox[i]), rel(ty[i], oy[i])
        worst = max(worst, ex, ey)
        assert ex < 1e-8 and ey < 1e-8, (iters[i], ex, ey)
    print(f"worst iterate deviation over {iters}: {worst:.2e}", flush=True)
    assert rel(res.x, ref["x"]) < 1e-8 and rel(res.y, ref["y"]) < 1e-8


def test_zipf4m_mixed_precision(dp):
    """The m = 4e6, n = 8e6, nnz = 1.2e8 LP (seed 2000) with FP32 matrix values
    and gathers, FP64 state and an FP64 finish, to 1e-4: OPTIMAL, a solution
    whose bounds, constraint violation, dual signs and reduced costs hold when
    recomputed with scipy, the objective near the known optimum; and the FP64
    device solve of the same LP for the error of the mixed result."""
    prob, opt = dp.generate(4_000_000, 8_000_000, 120_000_000, row_rule=1, seed=2000)
    a = prob.arrays()
    params = dp.default_params(presolve=False)
    R = dp.Resident(prob)
    mixed, om = R.solve(params, precision=dp.PREC_MIXED)
    fp64, of = R.solve(params)
    R.free()
    print(f"mixed {mixed.status} T={mixed.iterations} (fp32 {om.iterations_fp32}) "
          f"loop {om.loop_time_sec:.2f}s; fp64 {fp64.status} T={fp64.iterations} "
          f"loop {of.loop_time_sec:.2f}s", flush=True)
    assert mixed.status == fp64.status == "OPTIMAL"
    assert om.iterations_fp32 > 0
    for r in (mixed, fp64):
        assert r.rel_primal < 1e-4 and r.rel_dual < 1e-4 and r.rel_gap < 1e-4
    A = sp.csr_matrix((a["vals"], a["col_ind"], a["row_ptr"]), shape=(prob.m, prob.n))
    x, y, rc = mixed.x, mixed.y, mixed.reduced_cost
    c, lc, uc, lv, uv = a["c"], a["con_lb"], a["con_ub"], a["var_lb"], a["var_ub"]
    assert np.all(x >= lv - 1e-9 * (1 + np.abs(np.where(np.isfinite(lv), lv, 0))))
    assert np.all(x <= uv + 1e-9 * (1 + np.abs(np.where(np.isfinite(uv), uv, 0))))
    ax = A @ x
    viol = ax - np.clip(ax, lc, uc)
    bnorm = np.linalg.norm(np.concatenate([lc[np.isfinite(lc)], uc[np.isfinite(uc)]]))
    assert np.linalg.norm(viol) <= 1e-3 * (1 + bnorm)
    ge = np.isfinite(lc) & ~np.isfinite(uc)
    le = ~np.isfinite(lc) & np.isfinite(uc)
    assert np.all(y[ge] >= 0) and np.all(y[le] <= 0)
    assert np.linalg.norm(rc - (c - A.T @ y)) <= 1e-9 * (1 + np.linalg.norm(c))
    obj = float(c @ x)
    assert abs(mixed.primal_obj - obj) <= 1e-9 * max(1.0, abs(obj))
    assert abs(obj - opt) <= 1e-3 * max(1.0, abs(opt))
    # error of the mixed result against the FP64 device solve (both at 1e-4)
    dobj = abs(mixed.primal_obj - fp64.primal_obj) / max(1.0, abs(fp64.primal_obj))
    ex = np.linalg.norm(mixed.x - fp64.x) / np.linalg.norm(fp64.x)
    ey = np.linalg.norm(mixed.y - fp64.y) / np.linalg.norm(fp64.y)
    print(f"mixed vs fp64: objective {dobj:.2e}, x {ex:.2e}, y {ey:.2e}", flush=True)
    assert dobj <= 1e-4


@pytest.mark.timeout(900)
def test_batch512_every_member_alone_equals_mix(dp):
    """The whole 512-LP workload (m log-uniform in [500, 2e5], n = m U[1.5, 3],
    Poisson(8)+1 nonzeros per row, 1.46e8 nonzeros in all) in one
    mixed-precision batch: every LP OPTIMAL, within 1e-3 of its known
    optimum, and bit for bit the result of solving it alone."""
    from test_gpu_batch import same
    probs = [dp.batch_member(k) for k in range(512)]
    res = [dp.Resident(p, spmv_cap=4096, panels=(1, 1)) for p, _ in probs]
    params = dp.default_params(presolve=False)
    mix = dp.solve_batch(res, params, spmv_cap=4096, precision=dp.PREC_MIXED)
    print(f"batch of {len(res)} solved", flush=True)
    for k, r in enumerate(res):
        assert mix[k].status == "OPTIMAL", k
        opt = probs[k][1]
        assert abs(mix[k].primal_obj - opt) <= 1e-3 * max(1.0, abs(opt)), k
        alone = dp.solve_batch([r], params, spmv_cap=4096, precision=dp.PREC_MIXED)[0]
        assert same(mix[k], alone), k
    for r in res:
        r.free()
