"""Serving with continuous batching: every served request's result equals the
same LP solved alone, requests queue for slots and memory and are backfilled
as others retire, and the plain one-at-a-time loop serves the same stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POOL = dict(pool_small=6, pool_large=0, large_frac=0.0, m_lo=300, m_hi=3000, spmv_cap=1024,
            member_seed=3600)


def same(a, b):
    return (a.status == b.status and a.iterations == b.iterations
            and np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
            and a.primal_obj == b.primal_obj and a.dual_obj == b.dual_obj)


def alone_results(dp, params, precision):
    out = []
    for k in range(POOL["pool_small"]):
        prob, _ = dp.batch_member(k, POOL["member_seed"], POOL["m_lo"], POOL["m_hi"])
        r = dp.Resident(prob, spmv_cap=1024, panels=(1, 1))
        out.append(dp.solve_batch([r], params, spmv_cap=1024, precision=precision)[0])
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_served_results_equal_alone(dp, precision):
    params = dp.default_params(presolve=False)
    ref = alone_results(dp, params, precision)
    # ~60 requests in 0.2 s into at most 5 slots: requests queue, join running
    # batches and backfill retired slots
    rep, lat, idx, res = dp.serve(params, want_results=True, rate=300.0, duration_sec=0.2,
                                  drain_sec=60.0, seed=11, max_lps=5, precision=precision,
                                  **POOL)
    assert rep["arrived"] > 20
    assert rep["completed"] == rep["arrived"] and rep["unfinished"] == 0
    assert rep["optimal"] == rep["arrived"]
    assert rep["max_in_flight"] <= 5
    assert np.all(lat > 0) and rep["p99_sec"] >= rep["p50_sec"] > 0
    for k in range(rep["arrived"]):
        assert same(res[k], ref[idx[k]]), (k, idx[k])


def test_plain_loop_serves_same_results(dp):
    params = dp.default_params(presolve=False)
    kw = dict(rate=100.0, duration_sec=0.2, drain_sec=60.0, seed=5, precision=0, **POOL)
    rb, _, ib, resb = dp.serve(params, want_results=True, mode=0, max_lps=8, **kw)
    rp, _, ip, resp = dp.serve(params, want_results=True, mode=1, **kw)
    assert rb["arrived"] == rp["arrived"] > 5 and np.array_equal(ib, ip)
    assert rp["max_in_flight"] == 1 and rp["completed"] == rp["arrived"]
    # small LPs: the single-LP path uses the same row-block layout, so the
    # FP64 results agree bit for bit
    for k in range(rb["arrived"]):
        assert same(resb[k], resp[k]), k


def test_memory_admission_limits_in_flight(dp):
    params = dp.default_params(presolve=False)
    rep, lat, idx, _ = dp.serve(params, rate=200.0, duration_sec=0.1, drain_sec=60.0, seed=3,
                                max_lps=64, hbm_budget_gb=1e-6, precision=0, **POOL)
    # a budget below one LP admits them one at a time (never zero in flight)
    assert rep["max_in_flight"] == 1
    assert rep["completed"] == rep["arrived"] > 3


def test_drain_cap_reports_unfinished(dp):
    params = dp.default_params(presolve=False, termination_criteria__iteration_limit=10**9)
    rep, lat, idx, _ = dp.serve(params, rate=400.0, duration_sec=0.05, drain_sec=0.0, seed=9,
                                max_lps=4, precision=0, **POOL)
    assert rep["arrived"] > 5
    assert rep["completed"] + rep["unfinished"] == rep["arrived"]
    assert np.isinf(lat).sum() == rep["unfinished"]


@pytest.mark.parametrize("precision", [0, 1])
def test_large_lane_results_equal_alone(dp, precision):
    """Requests above the large-LP threshold go to the single-LP lane beside
    the batch; every result still equals the LP solved alone on the path
    that served it: a lane request equals the single-LP solve of the LP
    (automatic layout, as the lane uploads it), a batched one the batch of
    that LP alone.  (In FP64 the two paths agree bit for bit; in mixed
    precision the single-LP path anchors its FP32 shadows and the batch
    does not.)"""
    params = dp.default_params(presolve=False)
    ref = alone_results(dp, params, precision)
    lane = {}
    for k in range(POOL["pool_small"]):
        prob, _ = dp.batch_member(k, POOL["member_seed"], POOL["m_lo"], POOL["m_hi"])
        if prob.nnz >= 9000:
            lane[k] = dp.Resident(prob).solve(params, precision=precision)[0]
    assert lane and len(lane) < POOL["pool_small"], "the pool must feed both paths"
    rep, lat, idx, res = dp.serve(params, want_results=True, rate=300.0, duration_sec=0.2,
                                  drain_sec=60.0, seed=12, max_lps=6, precision=precision,
                                  large_lane_nnz=9000, **POOL)
    assert rep["completed"] == rep["arrived"] > 20
    for k in range(rep["arrived"]):
        want = lane.get(int(idx[k]), ref[idx[k]])
        assert same(res[k], want), (k, idx[k])
    if precision == 0:
        for j, r in lane.items():
            assert same(r, ref[j]), j


def test_large_request_through_the_mix(dp):
    """An m = 1e6, nnz = 1e7 request (the serving workload's large class)
    arrives among small ones, goes to the large-LP lane and overlaps the batch
    on the device; its result equals the same LP solved alone on the
    single-LP path (automatic layout, as the lane uploads it) bit for bit,
    and every small result equals its solve alone."""
    params = dp.default_params(presolve=False)
    pool = dict(POOL, pool_small=3, pool_large=1, large_frac=0.25, large_m=1_000_000,
                large_nnz=10_000_000)
    rep, lat, idx, res = dp.serve(params, want_results=True, rate=40.0, duration_sec=0.2,
                                  drain_sec=120.0, seed=21, max_lps=4, precision=0, **pool)
    print(f"serve with large lane: arrived {rep['arrived']} pool indices {list(idx)}", flush=True)
    assert rep["completed"] == rep["arrived"] and rep["optimal"] == rep["arrived"]
    large = [k for k in range(rep["arrived"]) if idx[k] == 3]
    assert large, "no large request in this stream"
    prob, opt = dp.serve_pool_member(3, pool_small=3, member_seed=POOL["member_seed"],
                                     m_lo=POOL["m_lo"], m_hi=POOL["m_hi"],
                                     large_m=1_000_000, large_nnz=10_000_000)
    assert prob.m == 1_000_000 and abs(prob.nnz - 10_000_000) < 1_000_000
    R = dp.Resident(prob)
    alone, _ = R.solve(params)
    for k in large:
        assert same(res[k], alone), k
    assert abs(alone.primal_obj - opt) <= 1e-3 * max(1.0, abs(opt))
    small_ref = alone_results(dp, params, 0)
    for k in range(rep["arrived"]):
        if idx[k] < 3:
            assert same(res[k], small_ref[idx[k]]), (k, idx[k])


def test_served_request_matches_oracle(dp, orc):
    """A served small request against the FP64 oracle on the same LP: same
    status and iteration count, objective within 1e-6, x and y within 1e-8."""
    from helpers import rel
    params = dp.default_params(presolve=False)
    rep, lat, idx, res = dp.serve(params, want_results=True, rate=200.0, duration_sec=0.05,
                                  drain_sec=60.0, seed=31, max_lps=4, precision=0, **POOL)
    assert rep["completed"] == rep["arrived"] > 0
    checked = set()
    for k in range(rep["arrived"]):
        j = int(idx[k])
        if j in checked:
            continue
        checked.add(j)
        prob, _ = dp.batch_member(j, POOL["member_seed"], POOL["m_lo"], POOL["m_hi"])
        ref, _ = orc.solve(prob.ptr, orc.default_params(presolve=False))
        r = res[k]
        assert r.status == ref["status"] == "OPTIMAL"
        assert r.iterations == ref["total_count"], (j, r.iterations, ref["total_count"])
        obj = ref["primal_objective_value"]
        assert abs(r.primal_obj - obj) <= 1e-6 * max(1.0, abs(obj))
        assert rel(r.x, ref["x"]) < 1e-8 and rel(r.y, ref["y"]) < 1e-8, j
    assert len(checked) >= 2


@pytest.mark.timeout(900)
def test_serving_workload_short_window(dp):
    """The serving workload itself for a 4 s arrival window: a pool of the 512
    batch-workload LPs and 4 m = 1e6, nnz = 1e7 ones, 5 % large requests,
    mixed precision, 50 arrivals/s.  Every request completes OPTIMAL, and each
    result equals the same LP solved alone on the path that served it: the
    batch of that LP alone (4096-nonzero blocks) or, for the large ones, the
    single-LP solve on the automatic layout."""
    params = dp.default_params(presolve=False)
    pool = dict(pool_small=512, pool_large=4, large_frac=0.05, large_m=1_000_000,
                large_nnz=10_000_000, spmv_cap=4096)
    rep, lat, idx, res = dp.serve(params, want_results=True, rate=50.0, duration_sec=4.0,
                                  drain_sec=300.0, seed=4000, precision=dp.PREC_MIXED, **pool)
    n = rep["arrived"]
    print(f"served {rep['completed']} of {n}, large {int(np.sum(idx >= 512))}", flush=True)
    assert n > 100 and rep["completed"] == n and rep["optimal"] == n
    assert np.any(idx >= 512) and np.any(idx < 512)
    alone = {}
    for j in sorted(set(int(i) for i in idx)):
        prob, opt = dp.serve_pool_member(j, pool_small=512, large_m=1_000_000,
                                         large_nnz=10_000_000)
        if j < 512:
            r = dp.Resident(prob, spmv_cap=4096, spmv_cap_At=4096, panels=(1, 1))
            alone[j] = dp.solve_batch([r], params, spmv_cap=4096, precision=dp.PREC_MIXED)[0]
        else:
            alone[j] = dp.Resident(prob).solve(params, precision=dp.PREC_MIXED)[0]
        r = None
        assert abs(alone[j].primal_obj - opt) <= 1e-3 * max(1.0, abs(opt)), j
    for k in range(n):
        assert same(res[k], alone[int(idx[k])]), (k, int(idx[k]))
