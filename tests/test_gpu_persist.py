"""Persistent solving: LPs whose whole solve (power iteration, PDHG loop,
evaluations, restarts, termination) runs on the device in one workgroup each.
Every LP gives the same bits alone and in any mix, matches the FP64 oracle
(same status and iteration count, objective 1e-6, solution 1e-8), and LPs
stepped by windows beside the persistent group are unchanged bit for bit."""
import numpy as np
import pytest

from helpers import rel
from test_gpu_batch import degenerate_lps, same, uneven_lps

pytestmark = pytest.mark.gpu

CAP = 1024
BIG = 1 << 30  # every LP with nonzeros runs persistent
SMALL = [(300, 700), (1200, 2500), (500, 1400), (150, 400), (900, 2000)]


def residents(dp, probs):
    return [dp.Resident(p, spmv_cap=CAP, panels=(1, 1)) for p in probs]


def check_oracle(orc, prob, got, **kw):
    ref, _ = orc.solve(prob.ptr, orc.default_params(presolve=False, **kw))
    assert got.status == ref["status"], (got.status, ref["status"])
    assert got.iterations == ref["total_count"], (got.iterations, ref["total_count"])
    obj = ref["primal_objective_value"]
    assert abs(got.primal_obj - obj) <= 1e-6 * max(1.0, abs(obj)), (got.primal_obj, obj)
    if got.status == "OPTIMAL":
        assert rel(got.x, ref["x"]) < 1e-8, rel(got.x, ref["x"])
        assert rel(got.y, ref["y"]) < 1e-8, rel(got.y, ref["y"])


def test_persistent_alone_equals_mix_and_matches_oracle(dp, orc):
    lps = uneven_lps(dp, SMALL, seed0=3600)
    probs = [p for p, _ in lps]
    res = residents(dp, probs)
    params = dp.default_params(presolve=False)
    mix = dp.solve_batch(res, params, spmv_cap=CAP, persist_max_nnz=BIG)
    for k, r in enumerate(res):
        alone = dp.solve_batch([r], params, spmv_cap=CAP, persist_max_nnz=BIG)[0]
        assert same(mix[k], alone), k
        assert mix[k].status == "OPTIMAL"
        check_oracle(orc, probs[k], mix[k])
        opt = lps[k][1]
        assert abs(mix[k].primal_obj - opt) <= 1e-3 * max(1.0, abs(opt))


def test_windowed_lps_unchanged_beside_the_persistent_group(dp, orc):
    """A mix split by size: the small LPs run persistent, the others in
    windows; the windowed ones equal the batch without a persistent group
    bit for bit, the persistent ones match the oracle."""
    lps = uneven_lps(dp, SMALL + [(4000, 7000), (2500, 6000)], seed0=3700)
    probs = [p for p, _ in lps]
    res = residents(dp, probs)
    params = dp.default_params(presolve=False)
    cut = 12000
    split = dp.solve_batch(res, params, spmv_cap=CAP, persist_max_nnz=cut)
    plain = dp.solve_batch(res, params, spmv_cap=CAP)
    n_persist = 0
    for k, p in enumerate(probs):
        if p.nnz <= cut:
            n_persist += 1
            check_oracle(orc, p, split[k])
        else:
            assert same(split[k], plain[k]), k
    assert 0 < n_persist < len(probs)


def test_persistent_edge_cases_match_oracle(dp, orc):
    """Infeasible and unbounded LPs, an iteration limit that is not a multiple
    of the evaluation cadence, an evaluation frequency that is not a multiple
    of 10, and degenerate shapes (empty rows and columns, fixed variables)."""
    A = np.array([[1.0, 1.0], [1.0, 1.0]])
    infeas = dp.Problem.from_dense(A, c=np.zeros(2), con_lb=np.array([2.0, -np.inf]),
                                   con_ub=np.array([np.inf, 1.0]), var_lb=np.zeros(2),
                                   var_ub=np.ones(2))
    A2 = np.array([[1.0, -1.0]])
    unbdd = dp.Problem.from_dense(A2, c=np.array([-1.0, 0.0]), con_lb=np.array([0.0]),
                                  con_ub=np.array([np.inf]), var_lb=np.zeros(2),
                                  var_ub=np.full(2, np.inf))
    lps = uneven_lps(dp, SMALL[:3], seed0=3800)
    probs = [infeas, unbdd] + [p for p, _ in lps] + degenerate_lps(dp)[1:]
    res = residents(dp, probs)
    cases = [dict(termination_criteria__iteration_limit=5000),
             dict(termination_criteria__iteration_limit=137),
             dict(termination_evaluation_frequency=25,
                  termination_criteria__iteration_limit=5000)]
    for kw in cases:
        params = dp.default_params(presolve=False, **kw)
        got = dp.solve_batch(res, params, spmv_cap=CAP, persist_max_nnz=BIG)
        for k, p in enumerate(probs):
            check_oracle(orc, p, got[k], **kw)
    params = dp.default_params(presolve=False, termination_criteria__iteration_limit=5000)
    got = dp.solve_batch(res[:2], params, spmv_cap=CAP, persist_max_nnz=BIG)
    assert got[0].status == "PRIMAL_INFEASIBLE"
    assert got[1].status == "DUAL_INFEASIBLE"


def test_persistent_polish_and_mixed_precision_batch(dp, orc):
    """Polishing runs after the persistent loop as after a windowed one; in a
    mixed-precision batch the persistent LPs still run FP64 and match the
    FP64 oracle."""
    lps = uneven_lps(dp, SMALL[:3], seed0=3900)
    probs = [p for p, _ in lps]
    res = residents(dp, probs)
    pol = dp.default_params(presolve=False, feasibility_polishing=True,
                            termination_criteria__iteration_limit=5000)
    got = dp.solve_batch(res, pol, spmv_cap=CAP, persist_max_nnz=BIG)
    for k, p in enumerate(probs):
        ref, _ = orc.solve(p.ptr, orc.default_params(
            presolve=False, feasibility_polishing=True,
            termination_criteria__iteration_limit=5000))
        assert got[k].status == ref["status"]
        assert got[k].raw["feasibility_iteration"] == ref["feasibility_iteration"]
    params = dp.default_params(presolve=False)
    mixed = dp.solve_batch(res, params, spmv_cap=CAP, persist_max_nnz=BIG,
                           precision=dp.PREC_MIXED)
    for k, p in enumerate(probs):
        check_oracle(orc, p, mixed[k])


def test_served_small_requests_run_persistent(dp, orc):
    """Serving with a persistent cut: requests at or below it run whole on the
    device beside the windowed batch; every served result equals the same LP
    solved alone with the same cut, and the persistent ones match the oracle."""
    pool = dict(pool_small=6, pool_large=0, large_frac=0.0, m_lo=300, m_hi=3000,
                spmv_cap=1024, member_seed=3600)
    params = dp.default_params(presolve=False)
    cut = 9000
    probs = [dp.batch_member(k, pool["member_seed"], pool["m_lo"], pool["m_hi"])[0]
             for k in range(pool["pool_small"])]
    ref = [dp.solve_batch([dp.Resident(p, spmv_cap=1024, panels=(1, 1))], params,
                          spmv_cap=1024, persist_max_nnz=cut)[0] for p in probs]
    rep, lat, idx, res = dp.serve(params, want_results=True, rate=300.0, duration_sec=0.2,
                                  drain_sec=60.0, seed=11, max_lps=5, precision=0,
                                  persist_max_nnz=cut, **pool)
    assert rep["arrived"] > 20 and rep["completed"] == rep["arrived"]
    assert rep["optimal"] == rep["arrived"] and rep["max_in_flight"] <= 5
    kinds = set()
    for k in range(rep["arrived"]):
        assert same(res[k], ref[idx[k]]), (k, idx[k])
        kinds.add(probs[idx[k]].nnz <= cut)
    assert kinds == {True, False}
    for k, p in enumerate(probs):
        if p.nnz <= cut:
            check_oracle(orc, p, ref[k])


def test_persistent_state_survives_every_launch_boundary(dp):
    """The controller state lives in LDS during a launch and in the LP's
    descriptor between launches: one iteration per launch (budget 1) gives
    the same bits as 250 or all of them per launch."""
    lps = uneven_lps(dp, SMALL[:3], seed0=4100)
    res = residents(dp, [p for p, _ in lps])
    params = dp.default_params(presolve=False)
    runs = [dp.solve_batch(res, params, spmv_cap=CAP, persist_max_nnz=BIG, persist_budget=b)
            for b in (1, 7, 250, 1 << 30)]
    for k in range(len(res)):
        assert runs[0][k].status == "OPTIMAL"
        for r in runs[1:]:
            assert same(runs[0][k], r[k]), k


def test_persistent_time_limit(dp):
    """The loop's time is read from the device clock: a zero time limit stops
    every persistent LP at its first evaluation, as the windowed batch does."""
    lps = uneven_lps(dp, SMALL[:2], seed0=4200)
    res = residents(dp, [p for p, _ in lps])
    params = dp.default_params(presolve=False, termination_criteria__time_sec_limit=0.0)
    got = dp.solve_batch(res, params, spmv_cap=CAP, persist_max_nnz=BIG)
    win = dp.solve_batch(res, params, spmv_cap=CAP)
    for g, w in zip(got, win):
        assert g.status == w.status == "TIME_LIMIT"
        assert g.iterations == w.iterations == 0
