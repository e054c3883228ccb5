"""GPU parity against the oracle across parameter variants, feasibility
polishing, warm starts, and the CLI end to end."""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

VARIANTS = {
    "linf": dict(optimality_norm=1),
    "no_bound_obj": dict(bound_objective_rescaling=False),
    "no_ruiz_no_pc": dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False),
    "pc_alpha_half": dict(pock_chambolle_alpha=0.5),
    "eval_freq_50": dict(termination_evaluation_frequency=50),
    "reflection_0.8": dict(reflection_coefficient=0.8),
    "tight_eps": dict(termination_criteria__eps_optimal_relative=1e-6,
                      termination_criteria__eps_feasible_relative=1e-6),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_variant_parity(dp, orc, name):
    kw = dict(VARIANTS[name], presolve=False)
    prob, opt = dp.generate(800, 2000, 16000, row_rule=0, seed=11)
    iters = [0, 10, 100, 300, 600]
    res, o, (tx, ty) = dp.solve(prob, dp.default_params(**kw), trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(**kw), trace_iters=iters,
                                 n=prob.n, m=prob.m)
    assert res.status == ref["status"]
    assert len(tx) == len(ox)
    tol = 1e-7 if name == "pc_alpha_half" else 1e-8  # device pow vs libm pow: ulp level
    for i in range(len(ox)):
        assert rel(tx[i], ox[i]) < tol and rel(ty[i], oy[i]) < tol, (name, iters[i])
    assert abs(res.primal_obj - ref["primal_objective_value"]) <= 1e-6 * max(
        1.0, abs(ref["primal_objective_value"]))
    assert abs(res.iterations - ref["total_count"]) <= 10


def test_warm_start_parity(dp, orc, row_order):
    prob, opt = dp.generate(600, 1500, 12000, row_rule=0, seed=21)
    rng = np.random.default_rng(0)
    x0 = rng.uniform(0, 5, prob.n)
    y0 = rng.standard_normal(prob.m)
    prob.set_start(x0, y0)
    iters = [0, 1, 50, 400]
    res, o, (tx, ty) = dp.solve(prob, dp.default_params(presolve=False), trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(presolve=False), trace_iters=iters,
                                 n=prob.n, m=prob.m)
    assert res.status == ref["status"]
    for i in range(len(ox)):
        assert rel(tx[i], ox[i]) < 1e-8 and rel(ty[i], oy[i]) < 1e-8


def test_feasibility_polish_parity(dp, orc, row_order):
    prob, opt = dp.generate(700, 1800, 14000, row_rule=0, seed=31)
    kw = dict(presolve=False, feasibility_polishing=True)
    res = dp.solve_plain(prob, dp.default_params(**kw))
    ref, _ = orc.solve(prob.ptr, orc.default_params(**kw))
    assert res.status == ref["status"] == "OPTIMAL"
    assert res.raw["feasibility_polishing_time"] > 0
    assert res.iterations == ref["total_count"]
    assert res.raw["feasibility_iteration"] == ref["feasibility_iteration"]
    assert res.rel_primal <= 1e-6 and res.rel_dual <= 1e-6
    assert abs(res.primal_obj - ref["primal_objective_value"]) <= 1e-8 * max(
        1.0, abs(ref["primal_objective_value"]))
    assert rel(res.x, ref["x"]) < 1e-8 and rel(res.y, ref["y"]) < 1e-8


def test_feasibility_polish_hand_count(dp):
    """The hand-worked polish LP (its count is derived step by step in
    tests/test_oracle_handworked.py): OPTIMAL at T = 0, then a primal polish
    that succeeds at once (+0) and a dual polish that succeeds at T = 10
    (+10), through the C API."""
    from test_oracle_handworked import POLISH_ITERATIONS, polish_lp
    for scaled in (False, True):
        kw = dict(presolve=False, feasibility_polishing=True)
        if not scaled:
            kw.update(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False,
                      bound_objective_rescaling=False)
        prob = polish_lp(lambda m, n, rp, ci, v, c, lc, uc, lv, uv: dp.Problem.from_csr(
            m, n, rp, ci, v, c=c, con_lb=lc, con_ub=uc, var_lb=lv, var_ub=uv))
        res = dp.solve_plain(prob, dp.default_params(**kw))
        assert res.status == "OPTIMAL" and res.iterations == 0
        assert res.raw["feasibility_iteration"] == POLISH_ITERATIONS, scaled
        assert res.rel_primal == 0.0 and res.rel_dual <= 1e-6
        assert np.all(res.x == 0.0) and np.all(res.y == 0.0)


def test_result_vectors_match_oracle(dp, orc, row_order):
    prob, opt = dp.generate(900, 2100, 17000, row_rule=2, seed=41, poisson_lambda=8.0)
    params = dict(presolve=False, termination_criteria__iteration_limit=1000)
    res = dp.solve_plain(prob, dp.default_params(**params))
    ref, _ = orc.solve(prob.ptr, orc.default_params(**params))
    # both converge inside the 1000-iteration cap (the oracle at T = 550), so
    # the comparison is at the same converged point
    assert res.status == ref["status"] == "OPTIMAL"
    assert res.iterations == ref["total_count"]
    assert rel(res.x, ref["x"]) < 1e-8 and rel(res.y, ref["y"]) < 1e-8
    assert rel(res.reduced_cost, ref["rc"]) < 1e-8
    for k in ("relative_primal_residual", "relative_dual_residual", "relative_objective_gap"):
        assert abs(res.raw[k] - ref[k]) <= 1e-8 * max(1e-12, abs(ref[k])) + 1e-14


def test_cli_end_to_end(dp, tmp_path):
    mps = os.path.join(ROOT, "tests", "golden", "capi_example.mps")
    r = subprocess.run([dp.CLI_PATH, "--no_presolve", mps, str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Solution Summary" in r.stdout and "OPTIMAL" in r.stdout
    summ = (tmp_path / "capi_example_summary.txt").read_text().splitlines()
    keys = [l.split(":")[0] for l in summ]
    assert keys == ["Termination Reason", "Runtime (sec)", "Iterations Count",
                    "Primal Objective Value", "Dual Objective Value", "Relative Primal Residual",
                    "Relative Dual Residual", "Absolute Objective Gap", "Relative Objective Gap",
                    "Rows", "Columns", "Nonzeros"]
    assert summ[0] == "Termination Reason: OPTIMAL"
    assert summ[9:] == ["Rows: 3", "Columns: 2", "Nonzeros: 5"]
    x = [float(v) for v in (tmp_path / "capi_example_primal_solution.txt").read_text().split()]
    y = [float(v) for v in (tmp_path / "capi_example_dual_solution.txt").read_text().split()]
    assert np.allclose(x, [1, 2], atol=1e-3) and np.allclose(y, [1, -1, 0], atol=1e-3)
    r = subprocess.run([dp.CLI_PATH, "-f", "--eps_feas_polish", "1e-7", mps, str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0


def test_cli_reads_gzipped_mps(dp, tmp_path):
    """The command line on a gzipped file (inflated on a producer thread while
    the sections are parsed, here in 64-byte windows): the output files take
    the name up to the first '.', and the summary and solutions equal those
    of the plain file's run bit for bit apart from the runtime line."""
    import gzip
    mps = os.path.join(ROOT, "tests", "golden", "capi_example.mps")
    gz = tmp_path / "capi_example.mps.gz"
    with open(mps, "rb") as f, gzip.open(gz, "wb") as g:
        g.write(f.read())
    outs = {}
    for name, path in (("plain", mps), ("gz", str(gz))):
        d = tmp_path / name
        d.mkdir()
        env = dict(os.environ, DPDLP_MPS_WINDOW="64")
        r = subprocess.run([dp.CLI_PATH, "--no_presolve", path, str(d)], capture_output=True,
                           text=True, timeout=120, env=env)
        assert r.returncode == 0, r.stderr
        summ = (d / "capi_example_summary.txt").read_text().splitlines()
        outs[name] = ([l for l in summ if not l.startswith("Runtime")],
                      (d / "capi_example_primal_solution.txt").read_text(),
                      (d / "capi_example_dual_solution.txt").read_text())
    assert outs["gz"] == outs["plain"]
    assert outs["gz"][0][0] == "Termination Reason: OPTIMAL"


@pytest.mark.parametrize("panels", ["1,1", "8,4"])
def test_mixed_precision_reaches_fp64_accuracy(dp, orc, panels, monkeypatch):
    """FP32 matrix / gathers with FP64 state: the FP64 finish reaches the same
    tolerance, and the solution matches the FP64 oracle to the tolerance."""
    monkeypatch.setenv("DPDLP_PANELS", panels)
    prob, opt = dp.generate(1500, 3500, 30000, row_rule=0, seed=51)
    kw = dict(presolve=False, termination_criteria__eps_optimal_relative=1e-6,
              termination_criteria__eps_feasible_relative=1e-6)
    res, o, _ = dp.solve(prob, dp.default_params(**kw), precision=dp.PREC_MIXED)
    ref, _ = orc.solve(prob.ptr, orc.default_params(**kw))
    assert res.status == ref["status"] == "OPTIMAL"
    assert o.iterations_fp32 > 0
    assert res.rel_primal < 1e-6 and res.rel_dual < 1e-6 and res.rel_gap < 1e-6
    assert abs(res.primal_obj - ref["primal_objective_value"]) <= 1e-5 * max(
        1.0, abs(ref["primal_objective_value"]))
    assert rel(res.x, ref["x"]) < 1e-3



@pytest.mark.parametrize("panels", ["1,1", "4,2"])
def test_anchored_mixed_precision_has_no_fp32_floor(dp, panels, monkeypatch):
    """Single-LP mixed precision gathers x_ref - x_init and y_cur - y_init in
    FP32 and adds the FP64 products of the anchors: to 1e-7, with the FP64
    switch tolerance set to 0, every iteration stays in FP32, the iteration
    count is within 5 % of the FP64 run's, and the result is the FP64 one to
    well below the tolerance."""
    monkeypatch.setenv("DPDLP_PANELS", panels)
    prob, opt = dp.generate(20000, 50000, 400000, row_rule=0, seed=52)
    kw = dict(presolve=False, termination_criteria__eps_optimal_relative=1e-7,
              termination_criteria__eps_feasible_relative=1e-7)
    R = dp.Resident(prob)
    ref, _ = R.solve(dp.default_params(**kw))
    res, o = R.solve(dp.default_params(**kw), precision=dp.PREC_MIXED, mixed_switch_tol=0.0)
    print(f"fp64 T={ref.iterations}; mixed T={res.iterations} fp32 {o.iterations_fp32}; "
          f"objective {abs(res.primal_obj - ref.primal_obj) / abs(ref.primal_obj):.2e}, "
          f"x {rel(res.x, ref.x):.2e}, y {rel(res.y, ref.y):.2e}", flush=True)
    assert res.status == ref.status == "OPTIMAL"
    assert o.iterations_fp32 == res.iterations
    assert abs(res.iterations - ref.iterations) <= 0.05 * ref.iterations
    assert res.rel_primal < 1e-7 and res.rel_dual < 1e-7 and res.rel_gap < 1e-7
    assert abs(res.primal_obj - ref.primal_obj) <= 1e-8 * abs(ref.primal_obj)
    assert rel(res.x, ref.x) < 1e-4 and rel(res.y, ref.y) < 1e-4
    R.free()


def test_anchored_mixed_precision_warm_start_and_polish(dp):
    """Anchored FP32 shadows start from a warm start (the first anchors are
    the start point, so their FP64 products are taken before the first
    iteration) and hand over to FP64 feasibility polishing: the mixed solve
    begins at the same iterate as the FP64 one, reaches the tolerance, and
    polishes to the same polish tolerance."""
    prob, opt = dp.generate(3000, 7000, 60000, row_rule=0, seed=53)
    rng = np.random.default_rng(3)
    prob.set_start(rng.uniform(0, 5, prob.n), rng.standard_normal(prob.m))
    kw = dict(presolve=False, termination_criteria__eps_optimal_relative=1e-6,
              termination_criteria__eps_feasible_relative=1e-6)
    iters = [0, 1]
    r64, o64, (x64, y64) = dp.solve(prob, dp.default_params(**kw), trace_iters=iters)
    r32, o32, (x32, y32) = dp.solve(prob, dp.default_params(**kw), trace_iters=iters,
                                    precision=dp.PREC_MIXED)
    assert np.array_equal(x64[0], x32[0]) and np.array_equal(y64[0], y32[0])
    assert rel(x32[1], x64[1]) < 1e-6 and rel(y32[1], y64[1]) < 1e-6
    assert r64.status == r32.status == "OPTIMAL" and o32.iterations_fp32 > 0
    assert r32.rel_primal < 1e-6 and r32.rel_dual < 1e-6 and r32.rel_gap < 1e-6
    assert abs(r32.primal_obj - r64.primal_obj) <= 1e-6 * max(1.0, abs(r64.primal_obj))
    # polishing after a mixed solve to the default 1e-4 (the 1e-6 polish
    # tolerance is then not met yet, so both polish loops run, in FP64)
    kwp = dict(presolve=False, feasibility_polishing=True)
    p32, _, _ = dp.solve(prob, dp.default_params(**kwp), precision=dp.PREC_MIXED)
    assert p32.status == "OPTIMAL" and p32.raw["feasibility_polishing_time"] > 0
    assert p32.rel_primal <= 1e-6 and p32.rel_dual <= 1e-6


def test_skewed_rows_renumbered_parity(dp, orc):
    """Zipf row lengths trigger the row renumbering of the upload (rows by
    decreasing length): iterates (in the caller's order), duals, status and
    objective still match the oracle, and so does the standalone SpMV."""
    from helpers import csr_matvec
    prob, opt = dp.generate(6000, 12000, 120000, row_rule=1, pm1_frac=0.1, seed=41)
    a = prob.arrays()
    lens = np.diff(a["row_ptr"])
    assert lens.max() * len(lens) >= 16 * a["row_ptr"][-1]  # skewed enough to renumber
    iters = [0, 10, 100, 300, 600]
    params = dp.default_params(presolve=False)
    res, o, (tx, ty) = dp.solve(prob, params, trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(presolve=False),
                                 trace_iters=iters, n=prob.n, m=prob.m)
    assert res.status == ref["status"] == "OPTIMAL"
    assert min(len(tx), len(ox)) == len(iters)
    for i in range(len(iters)):
        assert rel(tx[i], ox[i]) < 1e-8 and rel(ty[i], oy[i]) < 1e-8, iters[i]
    assert abs(res.primal_obj - ref["primal_objective_value"]) <= 1e-6 * max(
        1.0, abs(ref["primal_objective_value"]))
    if res.iterations == ref["total_count"]:
        assert rel(res.y, ref["y"]) < 1e-6 and rel(res.x, ref["x"]) < 1e-6
    # the standalone SpMV of a renumbered resident, in the caller's order
    R = dp.Resident(prob)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(prob.n)
    y = rng.standard_normal(prob.m)
    ax = csr_matvec(a["row_ptr"], a["col_ind"], a["vals"], x)
    assert np.allclose(R.spmv(x), ax, rtol=1e-12, atol=1e-12)
    rows = np.repeat(np.arange(prob.m), lens)
    aty = np.zeros(prob.n)
    np.add.at(aty, a["col_ind"], a["vals"] * y[rows])
    assert np.allclose(R.spmv(y, transpose=True), aty, rtol=1e-12, atol=1e-12)
    R.free()


def test_projection_invariants_every_row_and_bound_type(dp, orc):
    """The returned x (the projected PDHG iterate) lies inside its box
    (finite, one-sided and free bounds, fixed variables) and y has the sign its
    row type allows (>= rows y >= 0, <= rows y <= 0, equality and ranged rows
    free); the iterates match the oracle's at the trace points."""
    rng = np.random.default_rng(9)
    m, n = 40, 60
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.3)
    x0 = rng.uniform(0.0, 2.0, n)
    ax = A @ x0
    kinds = np.arange(m) % 4          # 0: >=, 1: <=, 2: =, 3: ranged
    con_lb = np.where(kinds == 1, -np.inf, ax - np.where(kinds == 2, 0.0, 0.5))
    con_ub = np.where(kinds == 0, np.inf, ax + np.where(kinds == 2, 0.0, 0.5))
    con_lb[kinds == 2] = con_ub[kinds == 2] = ax[kinds == 2]
    vk = np.arange(n) % 4             # 0: [0, inf), 1: (-inf, 3], 2: free, 3: fixed
    var_lb = np.select([vk == 0, vk == 1, vk == 2, vk == 3], [0.0, -np.inf, -np.inf, x0])
    var_ub = np.select([vk == 0, vk == 1, vk == 2, vk == 3], [np.inf, 3.0, np.inf, x0])
    c = rng.standard_normal(n) * (vk != 2)
    prob = dp.Problem.from_dense(A, c=c, con_lb=con_lb, con_ub=con_ub, var_lb=var_lb,
                                 var_ub=var_ub)
    iters = [1, 10, 100, 400]
    params = dict(presolve=False, termination_criteria__iteration_limit=2000)
    res, o, (tx, ty) = dp.solve(prob, dp.default_params(**params), trace_iters=iters)
    ref, (ox, oy, _) = orc.solve(prob.ptr, orc.default_params(**params), trace_iters=iters,
                                 n=prob.n, m=prob.m)
    assert res.status == ref["status"]
    assert len(tx) == len(ox) > 0
    for i in range(len(ox)):
        assert rel(tx[i], ox[i]) < 1e-8 and rel(ty[i], oy[i]) < 1e-8, iters[i]
    x, y = res.x, res.y
    tol = 1e-12 * (1 + np.abs(x0))
    assert np.all(x >= np.where(np.isfinite(var_lb), var_lb, -np.inf) - tol)
    assert np.all(x <= np.where(np.isfinite(var_ub), var_ub, np.inf) + tol)
    # fixed variables come back at their value up to the unscaling's rounding
    # (x / D_c / beta of a scaled value that the projection pinned)
    assert np.allclose(x[vk == 3], x0[vk == 3], rtol=1e-15, atol=0)
    assert np.all(y[kinds == 0] >= 0) and np.all(y[kinds == 1] <= 0)


def _pin_problem(dp):
    from test_oracle_handworked import INF
    return dp.Problem.from_csr(2, 2, [0, 2, 4], [0, 1, 0, 1], [3.0, 4.0, 4.0, -3.0],
                               c=[3.0, 4.0], con_lb=[2.0, -INF], con_ub=[2.0, 0.0],
                               var_lb=[0.0, -1.0], var_ub=[INF, 1.0])


def _unscaled(dp, **kw):
    return dp.default_params(presolve=False, l_inf_ruiz_iterations=0,
                             has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                             **kw)


@pytest.mark.gpu
def test_hand_worked_iterations_on_device(dp):
    """The device path reproduces the hand-worked iterates of the 2 x 2 pin LP
    (derivation in tests/test_oracle_handworked.py): step 0.998 / 5, x_cur and
    y_cur at the tops of iterations 1-3."""
    from test_oracle_handworked import HAND_TRACE, _frac
    prob = _pin_problem(dp)
    res, opts, (tx, ty) = dp.solve(prob, _unscaled(dp), trace_iters=[0, 1, 2, 3])
    assert abs(opts.step_size - 0.1996) <= 1e-15 * 0.1996
    assert len(tx) == 4
    for k in (1, 2, 3):
        xs, ys = HAND_TRACE[k]
        assert np.allclose(tx[k], [_frac(s) for s in xs], rtol=1e-13, atol=1e-15), k
        assert np.allclose(ty[k], [_frac(s) for s in ys], rtol=1e-13, atol=1e-15), k


@pytest.mark.gpu
def test_major_variant_plus_two_rule_on_device(dp):
    """(T+2) % f(T+2) == 0 selects the major variant at T = 8: a solve capped
    at 8 iterations reports x = y = 0, capped at 9 or 10 the hand-worked
    T = 8 pdhg iterate."""
    prob = _pin_problem(dp)
    r8 = dp.solve_plain(prob, _unscaled(dp, termination_criteria__iteration_limit=8))
    assert r8.status == "ITERATION_LIMIT" and r8.iterations == 8
    assert np.all(r8.x == 0.0) and np.all(r8.y == 0.0)
    want_x = np.array([0.17813274459085818, 0.28974413803225535])
    want_y = np.array([1.1196146839559613, 0.0])
    for lim in (9, 10):
        r = dp.solve_plain(prob, _unscaled(dp, termination_criteria__iteration_limit=lim))
        assert r.iterations == lim
        assert np.allclose(r.x, want_x, rtol=1e-12, atol=1e-15), (lim, r.x)
        assert np.allclose(r.y, want_y, rtol=1e-12, atol=1e-15), (lim, r.y)


@pytest.mark.gpu
def test_repeated_solves_in_one_process_are_identical(dp):
    """Each solve reseeds the power iteration's start vector, so solving the
    same LP twice in one process gives the same power-iteration count, step
    size and result bits (the reference's second solve in a process would
    continue its generator and may differ)."""
    prob, _ = dp.generate(600, 1500, 12000, row_rule=0, seed=77)
    params = dp.default_params(presolve=False)
    r1, o1, _ = dp.solve(prob, params)
    r2, o2, _ = dp.solve(prob, params)
    assert o1.power_iterations == o2.power_iterations and o1.step_size == o2.step_size
    assert r1.iterations == r2.iterations
    assert np.array_equal(r1.x, r2.x) and np.array_equal(r1.y, r2.y)


def _write_mps_gz(path, a, m, n):
    """Free-format MPS of an LP whose rows are equalities or >= rows (the
    generator's), numbers as %.17g so they read back exactly."""
    import gzip
    lc, uc = a["con_lb"], a["con_ub"]
    rp, ci, v = a["row_ptr"], a["col_ind"], a["vals"]
    cols = [[] for _ in range(n)]
    for i in range(m):
        for k in range(rp[i], rp[i + 1]):
            cols[ci[k]].append((i, v[k]))
    out = ["NAME GEN", "ROWS", " N OBJ"]
    for i in range(m):
        out.append((" E R%d" if lc[i] == uc[i] else " G R%d") % i)
    out.append("COLUMNS")
    for j in range(n):
        out.append("    C%d OBJ %.17g" % (j, a["c"][j]))
        for i, val in cols[j]:
            out.append("    C%d R%d %.17g" % (j, i, val))
    out.append("RHS")
    for i in range(m):
        out.append("    RHS R%d %.17g" % (i, lc[i]))
    out.append("BOUNDS")
    for j in range(n):
        if a["var_lb"][j] != 0.0:
            out.append(" LO BND C%d %.17g" % (j, a["var_lb"][j]))
        if np.isfinite(a["var_ub"][j]):
            out.append(" UP BND C%d %.17g" % (j, a["var_ub"][j]))
    out.append("ENDATA")
    with gzip.open(path, "wt") as f:
        f.write("\n".join(out) + "\n")


@pytest.mark.gpu
def test_mps_gz_file_to_device_solve(dp, orc, tmp_path):
    """Ingestion end to end: a generated LP written as a gzipped MPS file is
    read back (read_mps_file: threaded section parsing, parallel counting
    sort to CSR) with the same matrix, bounds and objective, and its device
    solve through the C API matches the oracle's solve of the read problem
    (same status and iteration count, objective 1e-6, x and y 1e-8)."""
    m, n = 500, 1200
    prob, _ = dp.generate(m, n, 9000, row_rule=0, seed=55)
    a = prob.arrays()
    path = str(tmp_path / "gen.mps.gz")
    _write_mps_gz(path, a, m, n)
    rd = dp.Problem.from_mps(path)
    b = rd.arrays()
    assert (rd.m, rd.n, rd.nnz) == (m, n, prob.nnz)
    assert np.array_equal(b["row_ptr"], a["row_ptr"])
    for i in range(m):
        s0, s1 = a["row_ptr"][i], a["row_ptr"][i + 1]
        o1, o2 = np.argsort(a["col_ind"][s0:s1]), np.argsort(b["col_ind"][s0:s1])
        assert np.array_equal(a["col_ind"][s0:s1][o1], b["col_ind"][s0:s1][o2])
        assert np.array_equal(a["vals"][s0:s1][o1], b["vals"][s0:s1][o2])
    for k in ("c", "con_lb", "con_ub", "var_lb", "var_ub"):
        assert np.array_equal(a[k], b[k]), k
    params = dp.default_params(presolve=False)
    res = dp.solve_plain(rd, params)
    ref, _ = orc.solve(rd.ptr, orc.default_params(presolve=False))
    assert res.status == ref["status"] == "OPTIMAL"
    assert res.iterations == ref["total_count"]
    obj = ref["primal_objective_value"]
    assert abs(res.primal_obj - obj) <= 1e-6 * max(1.0, abs(obj))
    assert rel(res.x, ref["x"]) < 1e-8 and rel(res.y, ref["y"]) < 1e-8
