"""The library's host decisions, called directly (CPU, no device): the
termination truth table (utils.cu:214-263), the adaptive restart test
(utils.cu:265-293) and the PID primal-weight update of a restart
(pdlp_core_op.cu:593-619), against expectations worked out by hand from
those rules (strict '<' for optimality, the OPTIMAL > PRIMAL_INF > DUAL_INF >
ITERATION > TIME order, the 0.2 / 0.5 / 0.36 restart thresholds, the PID
guard branches)."""
import ctypes as C
import math

import pytest


class HState(C.Structure):
    """Mirror of hstate_t (d-pdlp_amd/csrc/host/control.h)."""
    _fields_ = ([(n, C.c_double) for n in ("step", "omega", "best_omega", "pw_err_sum",
                                            "pw_last_err", "best_gap")]
                + [(n, C.c_int) for n in ("total", "inner", "is_major")]
                + [(n, C.c_double) for n in ("fpe", "fpe0", "fpe_last", "abs_p", "rel_p",
                                              "abs_d", "rel_d", "pobj", "dobj", "gap", "rel_gap",
                                              "max_p_ray_inf", "max_d_ray_inf", "p_ray_lin_obj",
                                              "d_ray_obj")]
                + [("reason", C.c_int), ("t_cum", C.c_double), ("pend_rst", C.c_int),
                   ("pend_maj", C.c_int), ("m32_best", C.c_double), ("m32_best_T", C.c_int)])


OPTIMAL, PRIMAL_INF, DUAL_INF, TIME, ITER = 1, 2, 3, 5, 6


@pytest.fixture(scope="module")
def ctl(dp):
    L = dp.lib
    L.ctl_check_termination.argtypes = [C.POINTER(HState), C.POINTER(dp.TerminationCriteria)]
    L.ctl_should_restart.argtypes = [C.POINTER(HState), C.POINTER(dp.RestartParams), C.c_int]
    L.ctl_should_restart.restype = C.c_int
    L.ctl_pid_update.argtypes = [C.POINTER(HState), C.POINTER(dp.RestartParams), C.c_double,
                                 C.c_double]
    return L


def _criteria(dp, **kw):
    tc = dp.default_params().termination_criteria
    for k, v in kw.items():
        setattr(tc, k, v)
    return tc


def _term(ctl, dp, tc, **st):
    h = HState()
    h.rel_p = h.rel_d = h.rel_gap = 1.0
    h.d_ray_obj = 0.0
    h.p_ray_lin_obj = 0.0
    h.total = 0
    h.t_cum = 0.0
    for k, v in st.items():
        setattr(h, k, v)
    ctl.ctl_check_termination(C.byref(h), C.byref(tc))
    return h.reason


def test_termination_truth_table(dp, ctl):
    tc = _criteria(dp, eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                   eps_infeasible=1e-10, iteration_limit=1000, time_sec_limit=60.0)
    # optimality: all three strictly below their tolerances
    assert _term(ctl, dp, tc, rel_p=9e-5, rel_d=9e-5, rel_gap=9e-5) == OPTIMAL
    assert _term(ctl, dp, tc, rel_p=1e-4, rel_d=9e-5, rel_gap=9e-5) == 0     # '<', not '<='
    assert _term(ctl, dp, tc, rel_p=9e-5, rel_d=9e-5, rel_gap=2e-4) == 0
    # primal infeasibility: dual ray objective > 0 and max infeasibility / objective <= eps
    assert _term(ctl, dp, tc, d_ray_obj=2.0, max_d_ray_inf=2e-10) == PRIMAL_INF
    assert _term(ctl, dp, tc, d_ray_obj=2.0, max_d_ray_inf=3e-10) == 0
    assert _term(ctl, dp, tc, d_ray_obj=0.0, max_d_ray_inf=0.0) == 0       # needs > 0
    # dual infeasibility: primal ray objective < 0, ratio against its magnitude
    assert _term(ctl, dp, tc, p_ray_lin_obj=-4.0, max_p_ray_inf=4e-10) == DUAL_INF
    assert _term(ctl, dp, tc, p_ray_lin_obj=-4.0, max_p_ray_inf=5e-10) == 0
    assert _term(ctl, dp, tc, p_ray_lin_obj=0.0, max_p_ray_inf=0.0) == 0
    # limits: '>='
    assert _term(ctl, dp, tc, total=1000) == ITER
    assert _term(ctl, dp, tc, total=999) == 0
    assert _term(ctl, dp, tc, t_cum=60.0) == TIME
    # order: optimal over infeasible over limits; iteration limit over time
    assert _term(ctl, dp, tc, rel_p=0, rel_d=0, rel_gap=0, d_ray_obj=1.0, total=5000,
                 t_cum=99.0) == OPTIMAL
    assert _term(ctl, dp, tc, d_ray_obj=1.0, p_ray_lin_obj=-1.0) == PRIMAL_INF
    assert _term(ctl, dp, tc, p_ray_lin_obj=-1.0, total=5000) == DUAL_INF
    assert _term(ctl, dp, tc, total=5000, t_cum=99.0) == ITER


def test_restart_sequence(dp, ctl):
    rp = dp.default_params().restart_params      # 0.36 artificial, 0.2 sufficient, 0.5 necessary
    freq = 200
    h = HState()
    h.fpe0 = 1.0
    h.fpe_last = math.inf
    steps = [  # (total, inner, fpe, expected)
        (100, 100, 0.9, 0),     # before the first evaluation frequency: never
        (200, 200, 0.9, 1),     # at it: forced
        (400, 50, 0.6, 0),      # no criterion (0.6 > 0.5, inner < 0.36 T)
        (600, 60, 0.45, 0),     # necessary reduction but still improving (0.45 < 0.6)
        (800, 80, 0.48, 1),     # necessary reduction and no progress (0.48 > 0.45)
        (1000, 100, 0.19, 1),   # sufficient reduction
        (1200, 432, 0.9, 1),    # artificial: inner >= 0.36 * 1200 = 432
        (1400, 503, 0.9, 0),    # 503 < 0.36 * 1400 = 504
    ]
    for total, inner, fpe, want in steps:
        h.total, h.inner, h.fpe = total, inner, fpe
        assert ctl.ctl_should_restart(C.byref(h), C.byref(rp), freq) == want, total
        assert h.fpe_last == fpe      # the trial value is remembered every time


def test_pid_primal_weight_sequence(dp, ctl):
    """pdlp_core_op.cu:593-619 with the default gains (k_p 0.99, k_i 0.01,
    k_d 0, i_smooth 0.3), every expected value worked out by hand:
      1. omega 1, d_p 2, d_d 8: e = ln 8 - ln 2 - ln 1 = ln 4, S = ln 4,
         omega = exp(0.99 ln 4 + 0.01 ln 4) = 4; gap |log10(1e-3 / 1e-2)| = 1
         beats +inf, so best = (1, 4)
      2. omega 4, d_p 1, d_d 4: e = 0, S = 0.3 ln 4, omega = 4 exp(0.003 ln 4)
         = 4^1.003; ratio 1, gap 0 beats 1: best = (0, 4^1.003)
      3. ratio 1e-9 (outside (1e-8, 1e8)): guard, omega = best = 4^1.003,
         S = e_prev = 0; gap 9 does not beat 0
      4. d_p = 1e-17 (<= 1e-16): guard again, nothing changes
      5. d_p 1, d_d = 2 omega: e = ln 2, S = ln 2, omega = 2 * 4^1.003;
         gap log10(2) does not beat 0"""
    rp = dp.default_params().restart_params
    assert (rp.k_p, rp.k_i, rp.k_d, rp.i_smooth) == (0.99, 0.01, 0.0, 0.3)
    h = HState(omega=1.0, best_omega=1.0, pw_err_sum=0.0, pw_last_err=0.0, best_gap=math.inf)
    w1003 = 4.0 ** 1.003
    seq = [  # (rel_p, rel_d, d_p, d_d) -> (omega, best_omega, S, e_prev, best_gap)
        ((1e-2, 1e-3, 2.0, 8.0), (4.0, 4.0, math.log(4.0), math.log(4.0), 1.0)),
        ((1e-3, 1e-3, 1.0, 4.0), (w1003, w1003, 0.3 * math.log(4.0), 0.0, 0.0)),
        ((1e-3, 1e-12, 1.0, 1.0), (w1003, w1003, 0.0, 0.0, 0.0)),
        ((1e-4, 1e-4, 1e-17, 1.0), (w1003, w1003, 0.0, 0.0, 0.0)),
        ((1e-4, 2e-4, 1.0, 2.0 * w1003), (2.0 * w1003, w1003, math.log(2.0), math.log(2.0), 0.0)),
    ]
    for (rel_p, rel_d, pd, dd), want in seq:
        h.rel_p, h.rel_d = rel_p, rel_d
        ctl.ctl_pid_update(C.byref(h), C.byref(rp), pd, dd)
        got = (h.omega, h.best_omega, h.pw_err_sum, h.pw_last_err, h.best_gap)
        for g, w in zip(got, want):
            assert math.isclose(g, w, rel_tol=1e-14, abs_tol=1e-15), (rel_p, got, want)


def test_pid_gains_are_not_interchangeable(dp, ctl):
    """Gains k_p 1/2, k_i 1/4, k_d 1/8, i_smooth 1/2 from omega 1: restart 1
    with e1 = 2 (d_p 1, d_d e^2) gives S = 2 and omega = exp(1 + 0.5 + 0.25)
    = e^1.75; restart 2 with e2 = -1 (d_d = omega / e) gives S = 1 - 1 = 0,
    derivative -3 and omega = e^1.75 exp(-0.5 + 0 - 0.375) = e^0.875.
    Swapping k_p and k_i would give e^1.125, k_p and k_d e^0.125."""
    prm = dp.default_params()
    rp = prm.restart_params
    rp.k_p, rp.k_i, rp.k_d, rp.i_smooth = 0.5, 0.25, 0.125, 0.5
    h = HState(omega=1.0, best_omega=1.0, pw_err_sum=0.0, pw_last_err=0.0, best_gap=math.inf)
    h.rel_p, h.rel_d = 1e-3, 1e-3
    ctl.ctl_pid_update(C.byref(h), C.byref(rp), 1.0, math.exp(2.0))
    assert math.isclose(h.omega, math.exp(1.75), rel_tol=1e-14)
    assert math.isclose(h.pw_err_sum, 2.0, rel_tol=1e-14)
    ctl.ctl_pid_update(C.byref(h), C.byref(rp), 1.0, h.omega / math.e)
    assert math.isclose(h.omega, math.exp(0.875), rel_tol=1e-14)
    assert abs(h.pw_err_sum) <= 1e-15 and math.isclose(h.pw_last_err, -1.0, rel_tol=1e-14)


def _keep(ctl, h, T, kkt, switch_tol, eps, anchored=0):
    h.total = T
    return ctl.ctl_mixed_keep_fp32(C.byref(h), kkt, switch_tol, eps, anchored)


def test_mixed_precision_switch_rule(ctl):
    """When a mixed-precision solve leaves FP32 for good, by hand: below the
    switch tolerance at once; a 10% improvement of the best FP32 value (kkt
    <= 0.9 best) restarts the stall clock; a stall of W = max(200, T/4)
    iterations ends FP32 when the best is within 100x of the target, and a
    stall of 2W wherever it is (so no LP stays in FP32 above its floor);
    anchored shadows make W four times longer."""
    ctl.ctl_mixed_keep_fp32.argtypes = [C.POINTER(HState), C.c_double, C.c_double, C.c_double,
                                        C.c_int]
    ctl.ctl_mixed_keep_fp32.restype = C.c_int
    h = HState(m32_best=math.inf, m32_best_T=0)
    assert _keep(ctl, h, 10, 5e-7, 1e-6, 1e-8) == 0            # below switch_tol
    # far from the target (eps = 1e-6: 100 eps = 1e-4)
    h = HState(m32_best=math.inf, m32_best_T=0)
    assert _keep(ctl, h, 100, 1e-3, 1e-6, 1e-6) == 1           # first value: the best
    assert (h.m32_best, h.m32_best_T) == (1e-3, 100)
    assert _keep(ctl, h, 200, 0.95e-3, 1e-6, 1e-6) == 1        # not 10% better; 100 < W = 200
    assert h.m32_best_T == 100
    assert _keep(ctl, h, 300, 1e-3, 1e-6, 1e-6) == 1           # 200 >= W but best > 100 eps
    assert _keep(ctl, h, 499, 1e-3, 1e-6, 1e-6) == 1           # 399 < 2W = 400
    assert _keep(ctl, h, 500, 1e-3, 1e-6, 1e-6) == 0           # 400 >= 2W: floor far above
    # near the target (eps = 1e-4: 100 eps = 1e-2)
    h = HState(m32_best=math.inf, m32_best_T=0)
    assert _keep(ctl, h, 1000, 5e-3, 1e-6, 1e-4) == 1
    assert _keep(ctl, h, 1100, 4.5e-3, 1e-6, 1e-4) == 1        # exactly 0.9 best: improved
    assert (h.m32_best, h.m32_best_T) == (4.5e-3, 1100)
    assert _keep(ctl, h, 1400, 4.2e-3, 1e-6, 1e-4) == 1        # W = max(200, 350) = 350 > 300
    assert _keep(ctl, h, 1465, 4.2e-3, 1e-6, 1e-4) == 1        # W = 366 > 365
    assert _keep(ctl, h, 1466, 4.2e-3, 1e-6, 1e-4) == 0        # W = 366 = 366: near, stalled
    # anchored shadows: W = max(800, T/2), far stall 3W/2
    h = HState(m32_best=math.inf, m32_best_T=0)
    assert _keep(ctl, h, 1000, 5e-3, 1e-6, 1e-4, 1) == 1       # near: 5e-3 < 100 eps
    assert _keep(ctl, h, 1998, 4.9e-3, 1e-6, 1e-4, 1) == 1     # 998 < W = 999
    assert _keep(ctl, h, 1999, 4.9e-3, 1e-6, 1e-4, 1) == 0     # 999 >= W = 999
    h = HState(m32_best=math.inf, m32_best_T=0)
    assert _keep(ctl, h, 1000, 1e-3, 1e-6, 1e-6, 1) == 1       # far: 1e-3 > 100 eps
    assert _keep(ctl, h, 3994, 1e-3, 1e-6, 1e-6, 1) == 1       # 2994 < 3W/2 = 1997 + 998
    assert _keep(ctl, h, 3995, 1e-3, 1e-6, 1e-6, 1) == 0       # 2995 >= 1997 + 998
