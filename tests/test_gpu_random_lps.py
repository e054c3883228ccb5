"""Random LPs with every kind of bound and row -- free, one-sided, boxed and
fixed variables; equality, <=, >=, ranged and free rows; empty rows and
columns; objective zeros -- against the oracle, on several layouts (column
panels, row-block capacities, rows renumbered or not).  Many such LPs are
infeasible or unbounded or do not converge within the iteration limit; the
device must agree with the oracle in every case: same status at the same
iteration, iterates within 1e-8 at the trace points (T <= 1000) and at a
converged end point."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu

LAYOUTS = [("1,1", "0"), ("2,2", "0"), ("4,8", "1"), ("8,1", "1"), ("1,2", "0"), ("2,4", "1")]


def random_lp(dp, seed):
    """Mode seed % 3: 0 bounded and feasible (every variable boxed or fixed,
    or one-sided on the side its cost bounds; rows hold at a sampled point),
    1 free and one-sided variables of any cost, 2 a row no variable range
    can meet (infeasible)."""
    rng = np.random.default_rng(seed)
    mode = seed % 3
    m, n = int(rng.integers(5, 400)), int(rng.integers(5, 500))
    dens = rng.uniform(0.005, 0.05)
    lens = np.minimum(rng.binomial(n, dens, m) + (rng.random(m) < 0.02) * (n // 2), n)
    lens[rng.random(m) < 0.05] = 0                      # empty rows
    live = np.flatnonzero(rng.random(n) >= 0.05)        # the rest stay empty columns
    lens = np.minimum(lens, len(live))
    rp = np.zeros(m + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([np.sort(rng.choice(live, L, replace=False)) for L in lens]
                        + [np.zeros(0, np.int64)]).astype(np.int32)
    v = rng.standard_normal(rp[-1])
    v[rng.random(rp[-1]) < 0.1] = 1.0
    c = rng.standard_normal(n)
    c[rng.random(n) < 0.15] = 0.0
    kind = rng.integers(0, 5, n)   # 0 free, 1 lower, 2 upper, 3 boxed, 4 fixed
    if mode == 0:                  # bounded: free -> boxed, one side as the cost bounds it
        kind = np.where(kind == 0, 3, kind)
        kind = np.where((kind == 1) & (c < 0), 3, kind)
        kind = np.where((kind == 2) & (c > 0), 3, kind)
    lo = rng.uniform(-5, 1, n)
    hi = lo + rng.uniform(0, 6, n)
    lv = np.where((kind == 0) | (kind == 2), -np.inf, lo)
    uv = np.where((kind == 0) | (kind == 1), np.inf, np.where(kind == 4, lo, hi))
    x0 = np.where(np.isfinite(lv), lv, np.where(np.isfinite(uv), uv, 0.0))
    x0 = np.where((kind == 3), rng.uniform(lo, hi), x0)
    ax = np.zeros(m)
    np.add.at(ax, np.repeat(np.arange(m), np.diff(rp)), v * x0[ci])
    rk = rng.integers(0, 5, m)     # 0 =, 1 <=, 2 >=, 3 ranged, 4 free
    slack = rng.uniform(0, 2, m)
    lc = np.where(rk == 0, ax, np.where((rk == 1) | (rk == 4), -np.inf, ax - slack))
    uc = np.where(rk == 0, ax, np.where((rk == 2) | (rk == 4), np.inf, ax + slack))
    if mode == 2:                  # row 0 over fixed columns only, asking for more
        cols = np.flatnonzero(kind == 4)[:3]
        if len(cols) == 0:
            cols = np.array([0])
            lv[0] = uv[0] = 1.0
        ci = np.concatenate([cols.astype(np.int32), ci[rp[1]:]])
        v = np.concatenate([np.ones(len(cols)), v[rp[1]:]])
        rp = np.concatenate([[0], rp[1:] - rp[1] + len(cols)]).astype(np.int32)
        lc[0] = np.sum(lv[cols]) + 1.0
        uc[0] = np.inf
    return dp.Problem.from_csr(m, n, rp, ci, v, c=c, con_lb=lc, con_ub=uc, var_lb=lv,
                               var_ub=uv)


@pytest.mark.parametrize("seed", range(18))
def test_random_lp_matches_oracle(dp, orc, seed, monkeypatch):
    panels, order = LAYOUTS[seed % len(LAYOUTS)]
    monkeypatch.setenv("DPDLP_PANELS", panels)
    monkeypatch.setenv("DPDLP_ROW_ORDER", order)
    prob = random_lp(dp, 9000 + seed)
    kw = dict(presolve=False, termination_criteria__iteration_limit=3000)
    iters = [0, 1, 10, 100, 400, 1000]
    res, o, (tx, ty) = dp.solve(prob, dp.default_params(**kw), trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(**kw), trace_iters=iters,
                                 n=prob.n, m=prob.m)
    print(f"seed {seed} m={prob.m} n={prob.n} nnz={prob.nnz} layout {panels}/{order}: "
          f"{res.status} T={res.iterations} (oracle {ref['status']} T={ref['total_count']})",
          flush=True)
    assert res.status == ref["status"]
    assert res.iterations == ref["total_count"]
    assert len(tx) == len(ox)
    for i in range(len(ox)):
        assert rel(tx[i], ox[i]) < 1e-8 and rel(ty[i], oy[i]) < 1e-8, (iters[i],)
    # the final iterates: 1e-8 at a converged point; an LP stopped by the
    # iteration limit (infeasible, unbounded or slow) has drifting iterates on
    # which the two summation orders' rounding differences keep growing
    # (measured up to 2.6e-7 at T = 3000), so 1e-6 there
    tol = 1e-8 if ref["status"] == "OPTIMAL" else 1e-6
    assert rel(res.x, ref["x"]) < tol and rel(res.y, ref["y"]) < tol


def test_random_lps_batched_equal_single(dp):
    """The same random LPs in one FP64 batch (every bound and row kind, some
    stopping at the iteration limit): each result is bit for bit the
    single-LP solve of that LP on the same layout."""
    probs = [random_lp(dp, 9000 + s) for s in range(18)]
    params = dp.default_params(presolve=False, termination_criteria__iteration_limit=3000)
    res = [dp.Resident(p, spmv_cap=1024, panels=(1, 1)) for p in probs]
    mix = dp.solve_batch(res, params, spmv_cap=1024)
    for k, r in enumerate(res):
        one, _ = r.solve(params)
        assert (mix[k].status, mix[k].iterations) == (one.status, one.iterations), k
        assert np.array_equal(mix[k].x, one.x) and np.array_equal(mix[k].y, one.y), k
        assert mix[k].primal_obj == one.primal_obj, k
