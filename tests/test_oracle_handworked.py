"""Hand-worked pins of the oracle's iteration, fixed-point error, major-variant
cadence, restart / PID weight and feasibility polish (CPU).

Every expected value below was worked out by hand, in exact rational
arithmetic, from the reference's definitions; none is produced by running the
oracle's own code.  Each check is chosen so that a plausible slip in the
oracle (a Halpern weight of k/(k+1), reflecting about x instead of the
projection, dropping or flipping the 2*eta*<A'dy, dx> term of the fixed-point
error, the "+2" major rule off by one, swapped PID gains, a wrong polish
iteration count) fails it.

The pin LP (2 x 2, scaling switched off, so scaled = unscaled):

    A = [[3, 4], [4, -3]]      A A' = 25 I  ->  sigma_max = 5 exactly
    c = (3, 4)                 ||c|| = 5
    row 0:  3x0 + 4x1 = 2      (l = u = 2: only u enters ||b||)
    row 1:  4x0 - 3x1 <= 0     (l = -inf, u = 0)
    x0 in [0, +inf),  x1 in [-1, 1]

    power iteration: A A' q = 25 q for every q, so it stops after one step
    eta   = 0.998 / 5 = 0.1996                   (pdlp_core_op.cu:646)
    omega = (||c|| + 1) / (||b|| + 1) = 6 / 3 = 2  (pdlp_core_op.cu:649-652)
    tau   = eta / omega = 0.0998,  sigma = eta * omega = 0.3992

Iteration 0 by hand (pdlp_core_op.cu:393-462), from x = y = 0:
    A'y = 0;  t = x - tau c = (-0.2994, -0.3992)
    p = proj(t) = (0, -0.3992);  x_ref = 2p - x = (0, -0.7984)
    A x_ref = (-3.1936, 2.3952);  t_y = y/sigma - A x_ref = (3.1936, -2.3952)
    proj onto [-u, -l] = ([-2, -2], [0, +inf)) -> (-2, 0)
    y_ref = 2 (t_y - proj) sigma - y = (4.14656..., -1.91232...)
          = (1619754/390625, -747003/390625)
    Halpern, inner 0: w = 1/2 -> x_cur = (0, -0.3992),
          y_cur = (809877/390625, -747003/781250)
"""
import ctypes as C
import math

import numpy as np
import pytest

INF = math.inf


def _pin_lp(orc):
    return orc.build_problem(2, 2, [0, 2, 4], [0, 1, 0, 1], [3.0, 4.0, 4.0, -3.0],
                             [3.0, 4.0], [2.0, -INF], [2.0, 0.0], [0.0, -1.0], [INF, 1.0])


def _params(orc, **kw):
    base = dict(presolve=False, l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False,
                bound_objective_rescaling=False)
    base.update(kw)
    return orc.default_params(**base)


# x_cur, y_cur at the top of iterations 1, 2, 3 (exact fractions, by hand)
HAND_TRACE = {
    1: (["0", "-499/1250"], ["809877/390625", "-747003/781250"]),
    2: (["0", "26829733/39062500"], ["4636733449/18310546875", "249001/390625"]),
    3: (["1413728681889/30517578125000", "-3331176321449/15258789062500"],
        ["1505862370881213/762939453125000", "-747003/1562500"]),
}


def _frac(s):
    from fractions import Fraction
    return float(Fraction(s))


def test_step_size_and_first_iterations_by_hand(orc):
    """pdlp_core_op.cu:393-462 (primal / dual steps, reflection, Halpern with
    w = (k+1)/(k+2)) and :638-656 (eta, omega), against the hand values."""
    prob = _pin_lp(orc)
    iters = [0, 1, 2, 3]
    r, (tx, ty, sc) = orc.solve(C.pointer(prob), _params(orc), trace_iters=iters, n=2, m=2)
    assert sc.shape[0] == 4
    assert abs(sc[0, 0] - 0.1996) <= 1e-15 * 0.1996     # eta = 0.998 / 5
    assert sc[0, 1] == 2.0                               # omega = 6 / 3
    assert np.all(tx[0] == 0.0) and np.all(ty[0] == 0.0)
    for k in (1, 2, 3):
        xs, ys = HAND_TRACE[k]
        want_x = np.array([_frac(s) for s in xs])
        want_y = np.array([_frac(s) for s in ys])
        assert np.allclose(tx[k], want_x, rtol=1e-13, atol=1e-15), (k, tx[k], want_x)
        assert np.allclose(ty[k], want_y, rtol=1e-13, atol=1e-15), (k, ty[k], want_y)
        assert sc[k, 2] == k                             # inner count
    # a Halpern weight of k/(k+1) would have given x_cur(1) = 0 (w = 0 at k = 0)
    assert tx[1][1] != 0.0


def test_major_variant_plus_two_rule_by_hand(orc):
    """pdlp_core_op.cu:508-510,541-543: the major variant (the only one that
    writes x_pdhg / y_pdhg) runs when (T+2) % f(T+2) == 0, i.e. at T = 8 for
    the default cadence.  A solve stopped by the iteration limit reports the
    last x_pdhg: still 0 after 8 iterations, the T = 8 value after 9 and 10."""
    prob = _pin_lp(orc)
    r8, _ = orc.solve(C.pointer(prob), _params(orc, termination_criteria__iteration_limit=8))
    assert r8["status"] == "ITERATION_LIMIT" and r8["total_count"] == 8
    assert np.all(r8["x"] == 0.0) and np.all(r8["y"] == 0.0)
    # by hand: x_pdhg, y_pdhg written at T = 8 from x_cur(8), y_cur(8)
    want_x = np.array([_frac("486032362252967288878766870119398965084713/"
                             "2728484105318784713745117187500000000000000"),
                       _frac("592921706422725549128858422853773965084713/"
                             "2046363078989088535308837890625000000000000")])
    want_y = np.array([1.1196146839559613, 0.0])
    for lim in (9, 10):
        r, _ = orc.solve(C.pointer(prob),
                         _params(orc, termination_criteria__iteration_limit=lim))
        assert r["total_count"] == lim
        assert np.allclose(r["x"], want_x, rtol=1e-12, atol=1e-15), (lim, r["x"])
        assert np.allclose(r["y"], want_y, rtol=1e-12, atol=1e-15), (lim, r["y"])


def test_fixed_point_error_with_cross_term_by_hand(orc):
    """pdlp_core_op.cu:658-696 at the major step T = 9 (eval_freq 10):
    FPE^2 = omega |dx|^2 + |dy|^2 / omega + 2 eta <A'dy, dx>, with dx, dy the
    reflected minus the current iterate.  By hand: movement 2.1337553951...,
    interaction -2.1177627653..., FPE 0.12646196974107457.  Without the cross
    term the FPE would be 1.46; with its sign flipped, 2.06."""
    prob = _pin_lp(orc)
    r, (tx, ty, sc) = orc.solve(C.pointer(prob), _params(orc, termination_evaluation_frequency=10),
                                trace_iters=[9, 10], n=2, m=2)
    assert sc.shape[0] == 2
    fpe = sc[1, 3]            # the T = 9 value, seen at the top of T = 10
    fpe2 = (1420430865443518938474650899342705179100970656559301 /
            88817841970012523233890533447265625000000000000000000)
    # the radicand cancels 2.13 against 2.12, so allow a few ulps of 2.13
    assert abs(fpe * fpe - fpe2) <= 1e-14, (fpe, math.sqrt(fpe2))
    assert abs(fpe - 0.12646196974107457) <= 1e-12


def test_forced_restart_and_pid_weight_closed_form(orc):
    """utils.cu:265-293 and pdlp_core_op.cu:577-636.  With eval_freq 10 the
    restart at T = 10 is forced; x_init = y_init = 0 before it, so with the
    default gains (k_p + k_i = 1, k_d = 0) the error sum equals the error e
    and omega_1 = omega_0 exp(e) = |y_pdhg(9)| / |x_pdhg(9)|, whatever
    omega_0.  By hand: x_pdhg(9) = (0.19484221560931897, 0.3068000515324253),
    y_pdhg(9) = (1.0781304961589708, 0)."""
    prob = _pin_lp(orc)
    prm = _params(orc, termination_evaluation_frequency=10)
    r, (tx, ty, sc) = orc.solve(C.pointer(prob), prm, trace_iters=[10, 11], n=2, m=2)
    assert sc[0, 1] == 2.0 and sc[0, 2] == 10          # before: omega_0, inner 10
    assert sc[1, 2] == 1                               # restarted at T = 10
    xp = np.array([0.19484221560931897, 0.3068000515324253])
    yp = np.array([1.0781304961589708, 0.0])
    omega1 = np.linalg.norm(yp) / np.linalg.norm(xp)
    assert abs(sc[1, 1] - omega1) <= 1e-12 * omega1, (sc[1, 1], omega1)
    # the restart moved the anchor: the current iterate restarts from pdhg(9)
    r10, _ = orc.solve(C.pointer(prob), _params(orc, termination_evaluation_frequency=10,
                                                 termination_criteria__iteration_limit=10))
    assert np.allclose(r10["x"], xp, rtol=1e-12) and np.allclose(r10["y"], yp, rtol=1e-12)


def test_pid_gains_second_restart_closed_form(orc):
    """PID primal weight over two restarts (pdlp_core_op.cu:593-613) with
    k_p = 1/2, k_i = 1/4, k_d = 1/8, i_smooth = 1/2, chosen so that the
    exponents collapse:
        restart 1 (T = 10, forced):      S1 = e1,
            omega_1 = omega_0 exp(7/8 e1) = omega_0^(1/8) (d_d1 / d_p1)^(7/8)
        restart 2 (T = 20, inner 10 >= 0.36 T):  S2 = e1/2 + e2,
            exponent = e2/2 + (e1/2 + e2)/4 + (e2 - e1)/8 = 7/8 e2
            omega_2 = omega_1^(1/8) (d_d2 / d_p2)^(7/8)
    Swapping k_p and k_i adds e1/8 to the second exponent; swapping k_p and
    k_d subtracts 3/8 e1; S = i_smooth (S + e) gives 3/4 e2."""
    prob = _pin_lp(orc)
    kw = dict(termination_evaluation_frequency=10, restart_params__k_p=0.5,
              restart_params__k_i=0.25, restart_params__k_d=0.125,
              restart_params__i_smooth=0.5)
    r, (tx, ty, sc) = orc.solve(C.pointer(prob), _params(orc, **kw), trace_iters=[11, 21],
                                n=2, m=2)
    assert sc.shape[0] == 2 and sc[0, 2] == 1 and sc[1, 2] == 1   # restarts at 10 and 20
    p9, _ = orc.solve(C.pointer(prob), _params(orc, termination_criteria__iteration_limit=10, **kw))
    p19, _ = orc.solve(C.pointer(prob), _params(orc, termination_criteria__iteration_limit=20, **kw))
    dp1, dd1 = np.linalg.norm(p9["x"]), np.linalg.norm(p9["y"])
    dp2, dd2 = np.linalg.norm(p19["x"] - p9["x"]), np.linalg.norm(p19["y"] - p9["y"])
    assert min(dp1, dd1, dp2, dd2) > 1e-6                      # guard branch not taken
    omega1 = 2.0 ** 0.125 * (dd1 / dp1) ** 0.875
    omega2 = omega1 ** 0.125 * (dd2 / dp2) ** 0.875
    assert abs(sc[0, 1] - omega1) <= 1e-12 * omega1, (sc[0, 1], omega1)
    assert abs(sc[1, 1] - omega2) <= 1e-12 * omega2, (sc[1, 1], omega2)


def test_pid_guard_keeps_best_weight(orc):
    """pdlp_core_op.cu:608-612: when the primal distance is below 1e-16 (x is
    fixed at 0 by its bounds, so x_pdhg = x_init = 0) the weight falls back to
    the best weight, omega_0 = (|c| + 1) / (|b| + 1) = (5 + 1) / (0 + 1) = 6,
    at every restart (T = 10 forced, then 20 and 30 by the artificial rule).
    Without the guard, log(0) would drive the weight to +inf.  eps = 0 keeps
    the solve from stopping before the restarts."""
    prob = orc.build_problem(2, 2, [0, 2, 4], [0, 1, 0, 1], [3.0, 4.0, 4.0, -3.0],
                             [3.0, 4.0], [0.0, -INF], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0])
    prm = _params(orc, termination_evaluation_frequency=10,
                  termination_criteria__eps_optimal_relative=0.0,
                  termination_criteria__eps_feasible_relative=0.0,
                  termination_criteria__iteration_limit=40)
    r, (tx, ty, sc) = orc.solve(C.pointer(prob), prm, trace_iters=[10, 11, 21, 31], n=2, m=2)
    assert r["status"] == "ITERATION_LIMIT" and r["total_count"] == 40
    assert sc.shape[0] == 4
    assert list(sc[:, 2]) == [10, 1, 1, 1]           # restarted at 10, 20, 30
    assert np.all(sc[:, 1] == 6.0)
    assert np.all(tx == 0.0)


POLISH_C = [3e-5, -4e-5]   # |c| = 5e-5


def polish_lp(build):
    """The LP of the hand-worked polish count: A as the pin LP, rows
    -5 <= 3x0 + 4x1 <= 5 and 4x0 - 3x1 >= -2, x0 in [0, 10], x1 in [-1, 1],
    c = (3e-5, -4e-5).  build(m, n, rp, ci, v, c, lc, uc, lv, uv)."""
    return build(2, 2, [0, 2, 4], [0, 1, 0, 1], [3.0, 4.0, 4.0, -3.0], POLISH_C,
                 [-5.0, -2.0], [5.0, INF], [0.0, -1.0], [10.0, 1.0])


POLISH_ITERATIONS = 10


def test_polish_iteration_count_by_hand(orc):
    """Feasibility polish (pdlp_core_op.cu:862-1030) worked by hand.
      * main loop: at T = 0 (x = y = 0) rel_p = 0 (x = 0 satisfies every row),
        the gap is 0 and rel_d = |c| / (1 + |c|) = 5e-5 < 1e-4: OPTIMAL before
        any iteration; rel_d is not below eps_polish = 1e-6, so polishing runs
        (:866-872);
      * primal polish (c = 0, y = 0): its T = 0 evaluation has rel_p = 0 ->
        FEAS_POLISH_SUCCESS.  The loop body still finishes that iteration
        (the reference's polish loops test the status only in the while
        condition, :938-977), so it ends at total 1 and adds 1 - 1 = 0 (:895);
      * dual polish: every variable bound is finite, so it becomes 0 and
        x_pdhg = x_ref = 0 throughout; every row bound becomes 0 or stays
        infinite, so y stays 0.  The first step that writes the dual slack is
        the "+2" major step T = 8, where z = (0 - (0 - tau c)) / tau = c; the
        next evaluation, T = 10, sees r_d = c - A'0 - z = 0 -> success, the
        body finishes iteration 10, total 11, adding 10 (:920).
    feasibility_iteration = 0 + 10 = 10; the main status stays OPTIMAL."""
    prob = polish_lp(orc.build_problem)
    for scaled in (False, True):
        kw = dict(presolve=False, feasibility_polishing=True)
        prm = orc.default_params(**kw) if scaled else _params(orc, **kw)
        r, _ = orc.solve(C.pointer(prob), prm)
        assert r["status"] == "OPTIMAL" and r["total_count"] == 0
        assert r["feasibility_iteration"] == POLISH_ITERATIONS, (scaled, r["feasibility_iteration"])
        assert r["relative_primal_residual"] == 0.0
        assert r["relative_dual_residual"] <= 1e-6
        assert np.all(r["x"] == 0.0) and np.all(r["y"] == 0.0)
        # without polishing the same solve reports rel_d = 5e-5 / (1 + 5e-5)
        r0, _ = orc.solve(C.pointer(prob), orc.default_params(presolve=False) if scaled
                          else _params(orc))
        assert r0["feasibility_iteration"] == 0
        assert abs(r0["relative_dual_residual"] - 5e-5 / (1 + 5e-5)) <= 1e-15


def test_print_frequency_cadence(orc):
    """utils.cu:492-501: 10 below 1e3, then x10 per decade."""
    f = orc.lib().orc_get_print_frequency
    assert [f(t) for t in (0, 9, 999, 1000, 9999, 10000, 123456)] == \
        [10, 10, 10, 100, 100, 1000, 10000]
