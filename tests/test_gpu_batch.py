"""Batched solving: every LP of a batch gives the same result alone and in a
mix (bit for bit), the same as the single-LP path on the same row-block
layout, and matches the oracle; LPs joining a running batch (continuous
batching) keep that property."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu

CAP = 1024
SIZES = [(300, 700), (1200, 2500), (2500, 6000), (500, 1400), (4000, 7000), (150, 400)]


def uneven_lps(dp, sizes=SIZES, seed0=3000):
    """Poisson(8)+1 nonzeros per row, as the batch workload generator."""
    out = []
    for k, (m, n) in enumerate(sizes):
        prob, opt = dp.generate(m, n, 0, row_rule=2, poisson_lambda=8.0, seed=seed0 + k)
        out.append((prob, opt))
    return out


def same(a, b):
    return (a.status == b.status and a.iterations == b.iterations
            and np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
            and np.array_equal(a.reduced_cost, b.reduced_cost)
            and a.primal_obj == b.primal_obj and a.dual_obj == b.dual_obj)


def residents(dp, lps):
    return [dp.Resident(p, spmv_cap=CAP, panels=(1, 1)) for p, _ in lps]


def alone(dp, res, params, **kw):
    return [dp.solve_batch([r], params, spmv_cap=CAP, **kw)[0] for r in res]


def test_batch_alone_equals_mix_equals_single(dp):
    lps = uneven_lps(dp)
    res = residents(dp, lps)
    params = dp.default_params(presolve=False)
    mix = dp.solve_batch(res, params, spmv_cap=CAP)
    solo_batch = alone(dp, res, params)
    for k, r in enumerate(res):
        single, _ = r.solve(params)
        assert same(mix[k], solo_batch[k]), k
        assert same(mix[k], single), k
        assert mix[k].status == "OPTIMAL"


def test_batch_matches_oracle(dp, orc):
    lps = uneven_lps(dp, SIZES[:4], seed0=3100)
    res = residents(dp, lps)
    params = dp.default_params(presolve=False)
    mix = dp.solve_batch(res, params, spmv_cap=CAP)
    for k, (prob, opt) in enumerate(lps):
        ref, _ = orc.solve(prob.ptr, orc.default_params(presolve=False))
        assert mix[k].status == ref["status"]
        assert abs(mix[k].primal_obj - ref["primal_objective_value"]) <= 1e-6 * max(
            1.0, abs(ref["primal_objective_value"]))
        assert abs(mix[k].iterations - ref["total_count"]) <= 10
        assert abs(mix[k].primal_obj - opt) <= 1e-3 * max(1.0, abs(opt))


def test_continuous_batching_join_while_running(dp):
    """LPs added while others are mid-solve, and into slots freed by
    finished ones, give the same result as alone."""
    lps = uneven_lps(dp, SIZES + [(800, 1600), (2000, 3000)], seed0=3200)
    res = residents(dp, lps)
    params = dp.default_params(presolve=False)
    ref = alone(dp, res, params)
    b = dp.Batch(params, max_lps=3, spmv_cap=CAP)
    pending = list(range(len(res)))
    got = {}
    steps = 0
    while pending or b.in_flight() > 0:
        # join at staggered windows: one new LP every 7 windows while a slot is free
        if pending and b.in_flight() < 3 and steps % 7 == 0:
            k = pending.pop(0)
            b.add(res[k], k)
        if b.in_flight() > 0:
            b.step()
        steps += 1
        for tag, r, lat in b.take():
            assert lat > 0
            got[tag] = r
    assert sorted(got) == list(range(len(res)))
    for k in range(len(res)):
        assert same(got[k], ref[k]), k
    assert b.counters()["windows"] > 0
    b.free()


def test_batch_mixed_precision_alone_equals_mix(dp):
    lps = uneven_lps(dp, SIZES[:5], seed0=3300)
    res = residents(dp, lps)
    params = dp.default_params(presolve=False)
    mix = dp.solve_batch(res, params, spmv_cap=CAP, precision=dp.PREC_MIXED)
    solo = alone(dp, res, params, precision=dp.PREC_MIXED)
    fp64 = dp.solve_batch(res, params, spmv_cap=CAP)
    for k in range(len(res)):
        assert same(mix[k], solo[k]), k
        assert mix[k].status == "OPTIMAL"
        # same LP, same accuracy target: objectives agree to the tolerance
        assert abs(mix[k].primal_obj - fp64[k].primal_obj) <= 1e-3 * max(
            1.0, abs(fp64[k].primal_obj))


def test_batch_edge_cases_match_single(dp):
    """Infeasible LPs, an iteration limit that is not a window multiple, an
    eval frequency that is not a multiple of 10, and polishing."""
    A = np.array([[1.0, 1.0], [1.0, 1.0]])
    infeas = dp.Problem.from_dense(A, c=np.zeros(2), con_lb=np.array([2.0, -np.inf]),
                                   con_ub=np.array([np.inf, 1.0]), var_lb=np.zeros(2),
                                   var_ub=np.ones(2))
    A2 = np.array([[1.0, -1.0]])
    unbdd = dp.Problem.from_dense(A2, c=np.array([-1.0, 0.0]), con_lb=np.array([0.0]),
                                  con_ub=np.array([np.inf]), var_lb=np.zeros(2),
                                  var_ub=np.full(2, np.inf))
    lps = uneven_lps(dp, SIZES[:3], seed0=3400)
    probs = [infeas, unbdd] + [p for p, _ in lps]
    res = [dp.Resident(p, spmv_cap=CAP, panels=(1, 1)) for p in probs]
    cases = [
        dict(presolve=False, termination_criteria__iteration_limit=5000),
        dict(presolve=False, termination_criteria__iteration_limit=37),
        dict(presolve=False, termination_evaluation_frequency=25,
             termination_criteria__iteration_limit=5000),
        dict(presolve=False, feasibility_polishing=True,
             termination_criteria__iteration_limit=5000),
    ]
    for kw in cases:
        params = dp.default_params(**kw)
        mix = dp.solve_batch(res, params, spmv_cap=CAP)
        for k, r in enumerate(res):
            single, _ = r.solve(params)
            assert same(mix[k], single), (kw, k)
    params = dp.default_params(**cases[0])
    mix = dp.solve_batch(res, params, spmv_cap=CAP)
    assert mix[0].status == "PRIMAL_INFEASIBLE"
    assert mix[1].status == "DUAL_INFEASIBLE"


def test_batch_empty_and_zero_limit(dp):
    params = dp.default_params(presolve=False, termination_criteria__iteration_limit=0)
    lps = uneven_lps(dp, SIZES[:2], seed0=3500)
    res = residents(dp, lps)
    out = dp.solve_batch(res, params, spmv_cap=CAP)
    assert all(r.status == "ITERATION_LIMIT" and r.iterations == 0 for r in out)
    assert dp.solve_batch([], params, spmv_cap=CAP) == []


def test_batch_rejects_tiled_layout(dp):
    prob, _ = dp.generate(300, 700, 0, row_rule=2, poisson_lambda=8.0, seed=1)
    r = dp.Resident(prob, spmv_cap=CAP, panels=(2, 1))
    b = dp.Batch(dp.default_params(presolve=False), spmv_cap=CAP)
    with pytest.raises(RuntimeError):
        b.add(r, 0)
    b.free()


def test_window_graphs_replay_identically(dp):
    """Windows replayed from captured hipGraphs give the same bits as direct
    launches, in a batch whose membership changes (graphs re-captured) and for
    a single LP (one graph replayed for most windows)."""
    lps = uneven_lps(dp, SIZES[:4], seed0=3700)
    res = residents(dp, lps)
    for prec in (dp.PREC_FP64, dp.PREC_MIXED):
        params = dp.default_params(presolve=False)
        st = {}
        g = dp.solve_batch(res, params, stats=st, spmv_cap=CAP, precision=prec, use_graph=1,
                           solo_lps=0)
        d = dp.solve_batch(res, params, spmv_cap=CAP, precision=prec, use_graph=0)
        assert st["graph_replays"] > 0
        for k in range(len(res)):
            assert same(g[k], d[k]), (prec, k)
    # single-LP solves through the graph-replayed windowed engine
    params = dp.default_params(presolve=False)
    for r in res[:2]:
        a, _ = r.solve(params, use_graph=1)
        b, _ = r.solve(params)
        assert same(a, b)
    # a batch stepped by hand with eager host work between replays
    b = dp.Batch(params, max_lps=4, spmv_cap=CAP, use_graph=1, solo_lps=0)
    for k, r in enumerate(res):
        b.add(r, k)
    got = {}
    while b.in_flight() > 0:
        b.step()
        _ = [r.m for r in res]  # host work between replays
        for tag, r, _lat in b.take():
            got[tag] = r
    ref = dp.solve_batch(res, params, spmv_cap=CAP)
    for k in range(len(res)):
        assert same(got[k], ref[k]), k
    assert b.counters()["graph_replays"] > 0
    b.free()


def test_tail_single_lp_launches_identical(dp):
    """Windows that issue single-LP launches for the last few LPs give the
    same bits as batched launches."""
    lps = uneven_lps(dp, SIZES[:5], seed0=3800)
    res = residents(dp, lps)
    params = dp.default_params(presolve=False)
    for prec in (dp.PREC_FP64, dp.PREC_MIXED):
        a = dp.solve_batch(res, params, spmv_cap=CAP, precision=prec, solo_lps=0)
        b = dp.solve_batch(res, params, spmv_cap=CAP, precision=prec, solo_lps=64)
        c = dp.solve_batch(res, params, spmv_cap=CAP, precision=prec, solo_lps=2)
        for k in range(len(res)):
            assert same(a[k], b[k]) and same(a[k], c[k]), (prec, k)


def degenerate_lps(dp):
    """Edge shapes: no constraints, empty rows and columns, a free variable
    with a zero objective, fixed variables."""
    out = []
    # m = 0: min c'x over the box
    out.append(dp.Problem.from_csr(0, 3, [0], [], [], c=np.array([1.0, -1.0, 0.0]),
                                   var_lb=np.zeros(3), var_ub=np.array([2.0, 3.0, 1.0])))
    # empty row 1 and empty column 2 (x2 free with c = 0)
    out.append(dp.Problem.from_csr(3, 3, [0, 2, 2, 3], [0, 1, 0], [1.0, 1.0, 1.0],
                                   c=np.array([1.0, 2.0, 0.0]),
                                   con_lb=np.array([1.0, -np.inf, 0.5]),
                                   con_ub=np.array([np.inf, np.inf, 4.0]),
                                   var_lb=np.array([0.0, 0.0, -np.inf]),
                                   var_ub=np.array([np.inf, np.inf, np.inf])))
    # fixed variables (l = u) and an equality row
    out.append(dp.Problem.from_dense(np.array([[1.0, 1.0, 1.0]]), c=np.array([1.0, 2.0, 3.0]),
                                     con_lb=np.array([3.0]), con_ub=np.array([3.0]),
                                     var_lb=np.array([1.0, 0.0, 0.0]),
                                     var_ub=np.array([1.0, 5.0, 5.0])))
    return out


def test_degenerate_shapes_batch_equals_single_and_oracle(dp, orc):
    probs = degenerate_lps(dp)
    params = dp.default_params(presolve=False, termination_criteria__iteration_limit=20000)
    res = [dp.Resident(p, spmv_cap=CAP, panels=(1, 1)) for p in probs]
    mix = dp.solve_batch(res, params, spmv_cap=CAP)
    for k, (p, r) in enumerate(zip(probs, res)):
        single, _ = r.solve(params)
        assert same(mix[k], single), k
        ref, _ = orc.solve(p.ptr, orc.default_params(presolve=False,
                                                     termination_criteria__iteration_limit=20000))
        assert mix[k].status == ref["status"], (k, mix[k].status, ref["status"])
        assert abs(mix[k].primal_obj - ref["primal_objective_value"]) <= 1e-6 * max(
            1.0, abs(ref["primal_objective_value"])), k


def boxed_lp_with_dense_column(dp, m, n, seed):
    """Poisson(8)+1 random entries per row plus column 0 in every row (an A'
    row longer than a 1024-nonzero block); every variable in [0, 10], rows
    feasible at a random box point, so the LP is feasible and bounded."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(m):
        k = min(n - 1, int(rng.poisson(8.0)) + 1)
        cs = np.concatenate(([0], 1 + rng.choice(n - 1, size=k, replace=False)))
        rows.extend([i] * len(cs))
        cols.extend(np.sort(cs).tolist())
    rows, cols = np.array(rows), np.array(cols)
    vals = rng.standard_normal(len(rows))
    rp = np.concatenate(([0], np.cumsum(np.bincount(rows, minlength=m))))
    x0 = rng.uniform(0.0, 10.0, n)
    ax = np.bincount(rows, weights=vals * x0[cols], minlength=m)
    eq = rng.random(m) < 0.3
    lb = np.where(eq, ax, ax - rng.uniform(0.0, 1.0, m))
    ub = np.where(eq, ax, np.inf)
    return dp.Problem.from_csr(m, n, rp, cols, vals, c=rng.standard_normal(n), con_lb=lb,
                               con_ub=ub, var_lb=np.zeros(n), var_ub=np.full(n, 10.0))


@pytest.mark.parametrize("caps", [(4096, 2048), (2048, 1024)])
def test_separate_At_capacity_alone_equals_mix_equals_single(dp, caps):
    """A' planned and launched with a smaller row-block capacity than A:
    batch alone == batch mix == the single-LP path on the same layout, bit
    for bit, including an A' row longer than its blocks."""
    cap, cap_at = caps
    probs = [p for p, _ in uneven_lps(dp, SIZES[:4], seed0=3300)]
    probs.append(boxed_lp_with_dense_column(dp, 1500, 2600, seed=7))
    res = [dp.Resident(p, spmv_cap=cap, panels=(1, 1), spmv_cap_At=cap_at) for p in probs]
    params = dp.default_params(presolve=False)
    for prec in (dp.PREC_FP64, dp.PREC_MIXED):
        mix = dp.solve_batch(res, params, spmv_cap=cap, spmv_cap_At=cap_at, precision=prec)
        for k, r in enumerate(res):
            one = dp.solve_batch([r], params, spmv_cap=cap, spmv_cap_At=cap_at,
                                 precision=prec)[0]
            assert same(mix[k], one), (caps, prec, k)
            assert mix[k].status == "OPTIMAL", (caps, prec, k, mix[k].status)
            if prec == dp.PREC_FP64:
                single, _ = r.solve(params)
                assert same(mix[k], single), (caps, k)
    # a batch refuses LPs whose row blocks are larger than its kernels' capacity
    with pytest.raises(RuntimeError):
        dp.solve_batch(res[:1], params, spmv_cap=cap // 2, spmv_cap_At=cap_at)


def test_graph_option_on_renumbered_upload(dp, monkeypatch):
    """use_graph on a problem whose upload renumbered its rows (the windowed
    engine takes the caller's row order only) runs the direct loop and gives
    its result."""
    monkeypatch.setenv("DPDLP_ROW_ORDER", "1")
    prob, _ = dp.generate(1500, 3500, 30000, row_rule=1, seed=11)
    r = dp.Resident(prob)
    params = dp.default_params(presolve=False)
    a, _ = r.solve(params, use_graph=1)
    b, _ = r.solve(params)
    assert a.status == "OPTIMAL"
    assert same(a, b)


def test_autotuned_layouts_mix_equals_alone(dp, orc, tmp_path):
    """The batch autotuner picks a row-block layout per size class (timing the
    batched phase kernels on a sample of each class), caches it by device and
    class, and a batch of LPs with different per-class layouts gives each LP
    the bits of that LP solved alone with its layout, and the oracle's
    objective."""
    probs = [dp.batch_member(k, 3200, 300, 6000)[0] for k in range(10)]
    params = dp.default_params(presolve=False)
    cache = str(tmp_path / "layouts.txt")
    lays, rep = dp.batch_autotune(probs, params, precision=dp.PREC_FP64, cache_path=cache)
    print("autotune", rep, sorted(set(lays)), flush=True)
    assert rep["classes"] >= 2 and rep["tuned"] == rep["classes"] and rep["cached"] == 0
    for L in lays:
        assert L[0] in (1024, 2048, 4096) and L[2] in (1024, 2048, 4096)
        assert L[1] in (64, 128, 256) and L[3] in (64, 128, 256)
    lays2, rep2 = dp.batch_autotune(probs, params, precision=dp.PREC_FP64, cache_path=cache)
    assert lays2 == lays and rep2["cached"] == rep["classes"] and rep2["tuned"] == 0
    # plus an LP whose dense column (an A' row of 1500 entries) is longer than
    # its 1024-nonzero blocks but fits the batch's 4096: it is summed as a long
    # row alone and in the mix
    probs.append(boxed_lp_with_dense_column(dp, 1500, 2600, seed=9))
    lays.append((1024, 256, 1024, 64))
    res = [dp.Resident(p, layout=L) for p, L in zip(probs, lays)]
    mix = dp.solve_batch(res, params, spmv_cap=4096, spmv_cap_At=4096)
    for k, r in enumerate(res):
        alone = dp.solve_batch([r], params, spmv_cap=4096, spmv_cap_At=4096)[0]
        assert mix[k].status == alone.status == "OPTIMAL"
        assert np.array_equal(mix[k].x, alone.x) and np.array_equal(mix[k].y, alone.y), k
        ref, _ = orc.solve(probs[k].ptr, orc.default_params(presolve=False))
        obj = ref["primal_objective_value"]
        assert abs(mix[k].primal_obj - obj) <= 1e-6 * max(1.0, abs(obj)), k
    for r in res:
        r.free()
