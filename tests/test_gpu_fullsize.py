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
