"""Pins for the FP64 oracle (CPU).  Each check is independent of the oracle's
own formula: brute force, closed forms, library routines, invariants."""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import capi_example, csr_matvec, random_csr, rel


def P(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


# ---------------------------------------------------------------- RNG

STD_PROG = r"""
#include <random>
#include <cstdio>
int main(int argc, char** argv) {
  std::mt19937 gen(1);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (int i = 0; i < 2001; ++i) std::printf("%a\n", dist(gen));
  return 0;
}
"""


def test_rng_matches_std_library(orc):
    """utils.cu:29-30,80-82: the oracle's own mt19937 + polar normal must
    reproduce std::mt19937(1) + std::normal_distribution<double> bit for bit."""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "g.cc")
        exe = os.path.join(d, "g")
        open(src, "w").write(STD_PROG)
        subprocess.run(["g++", "-O1", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    ref = np.array([float.fromhex(s) for s in out])
    got = np.zeros(len(ref))
    orc.lib().orc_mt_normal_fill(1, len(ref), P(got))
    assert np.array_equal(got, ref)


# ---------------------------------------------------------------- sparse

def _csr(seed=3, m=60, n=40):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, m)
    return random_csr(m, n, lens, seed)


def test_transpose_brute_force_and_involution(orc):
    rp, ci, v = _csr()
    m, n = len(rp) - 1, 40
    nnz = len(v)
    trp = np.zeros(n + 1, np.int32)
    tci = np.zeros(nnz, np.int32)
    tv = np.zeros(nnz)
    orc.lib().orc_transpose(m, n, P(rp, C.c_int), P(ci, C.c_int), P(v), P(trp, C.c_int),
                            P(tci, C.c_int), P(tv))
    dense = np.zeros((m, n))
    for i in range(m):
        dense[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    dt = np.zeros((n, m))
    for j in range(n):
        seg = tci[trp[j]:trp[j + 1]]
        assert np.all(np.diff(seg) > 0)  # stable: increasing source rows
        dt[j, seg] = tv[trp[j]:trp[j + 1]]
    assert np.array_equal(dt, dense.T)
    rp2 = np.zeros(m + 1, np.int32)
    ci2 = np.zeros(nnz, np.int32)
    v2 = np.zeros(nnz)
    orc.lib().orc_transpose(n, m, P(trp, C.c_int), P(tci, C.c_int), P(tv), P(rp2, C.c_int),
                            P(ci2, C.c_int), P(v2))
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci) and np.array_equal(v2, v)


def test_spmv_brute_force_and_adjoint(orc):
    rp, ci, v = _csr(seed=9)
    m, n = len(rp) - 1, 40
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n)
    y = np.zeros(m)
    orc.lib().orc_spmv(m, P(rp, C.c_int), P(ci, C.c_int), P(v), P(x), P(y))
    dense = np.zeros((m, n))
    for i in range(m):
        dense[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    assert np.allclose(y, dense @ x, rtol=1e-14, atol=1e-14)
    assert np.allclose(y, csr_matvec(rp, ci, v, x), rtol=1e-14, atol=1e-14)


# ---------------------------------------------------------------- reductions

@pytest.mark.parametrize("scale", [1.0, 1e200, 1e-200, 1e-160])
def test_nrm2_safe_scaling(orc, scale):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(1000) * scale
    got = orc.lib().orc_nrm2(len(x), P(x))
    ref = scale * math.sqrt(math.fsum((x / scale) ** 2))
    assert math.isfinite(got)
    assert abs(got - ref) <= 1e-14 * ref


def test_nrm2_mixed_magnitudes_and_empty(orc):
    x = np.array([1e300, 1.0, 1e-300, -3e299])
    got = orc.lib().orc_nrm2(4, P(x))
    assert abs(got - math.hypot(1e300, 3e299)) <= 1e-15 * got
    assert orc.lib().orc_nrm2(0, P(x)) == 0.0


def test_dot_amax_sum(orc):
    rng = np.random.default_rng(4)
    x = rng.standard_normal(777)
    y = rng.standard_normal(777)
    L = orc.lib()
    assert abs(L.orc_dot(777, P(x), P(y)) - math.fsum(x * y)) < 1e-12
    assert L.orc_amax(777, P(x)) == np.max(np.abs(x))
    assert abs(L.orc_sum(777, P(x)) - math.fsum(x)) < 1e-12
    assert L.orc_amax(0, P(x)) == 0.0 and L.orc_sum(0, P(x)) == 0.0


def test_print_frequency(orc):
    f = orc.lib().orc_get_print_frequency
    assert [f(0), f(999), f(1000), f(9999), f(10000), f(123456)] == [10, 10, 100, 100, 1000,
                                                                      10000]


# ---------------------------------------------------------------- scaling

def _rescale(orc, prob, **kw):
    prm = orc.default_params(**kw)
    m, n, nnz = prob.num_constraints, prob.num_variables, prob.constraint_matrix_num_nonzeros
    out = dict(val=np.zeros(nnz), c=np.zeros(n), lv=np.zeros(n), uv=np.zeros(n), lc=np.zeros(m),
               uc=np.zeros(m), Dr=np.zeros(m), Dc=np.zeros(n))
    beta, gamma = C.c_double(), C.c_double()
    orc.lib().orc_rescale(C.byref(prm), C.byref(prob), P(out["val"]), P(out["c"]),
                          P(out["lv"]), P(out["uv"]), P(out["lc"]), P(out["uc"]),
                          P(out["Dr"]), P(out["Dc"]), C.byref(beta), C.byref(gamma))
    out["beta"], out["gamma"] = beta.value, gamma.value
    return out


def test_ruiz_diagonal_closed_form(orc):
    d = np.array([4.0, 0.25, 9.0, 1e-6])
    n = 4
    prob = orc.build_problem(n, n, np.arange(n + 1), np.arange(n), d, c=np.ones(n),
                             con_lb=np.ones(n), con_ub=np.ones(n), var_lb=np.zeros(n),
                             var_ub=np.full(n, np.inf))
    out = _rescale(orc, prob, l_inf_ruiz_iterations=1, has_pock_chambolle_alpha=False,
                   bound_objective_rescaling=False)
    assert np.allclose(out["Dr"], np.sqrt(d)) and np.allclose(out["Dc"], np.sqrt(d))
    assert np.allclose(out["val"], 1.0)
    assert np.allclose(out["c"], 1.0 / np.sqrt(d))
    assert np.allclose(out["lc"], 1.0 / np.sqrt(d))
    assert out["beta"] == 1.0 and out["gamma"] == 1.0


def test_rank_one_and_ruiz_limit(orc):
    u = np.array([1.0, 2.0, 3.0])
    w = np.array([0.5, 4.0])
    m, n = 3, 2
    vals = np.outer(u, w).ravel()
    prob = orc.build_problem(m, n, np.arange(0, 7, 2), np.tile([0, 1], 3), vals, c=np.ones(n),
                             con_lb=np.zeros(m), con_ub=np.zeros(m), var_lb=np.zeros(n),
                             var_ub=np.ones(n))
    out = _rescale(orc, prob, l_inf_ruiz_iterations=40, has_pock_chambolle_alpha=False,
                   bound_objective_rescaling=False)
    A = out["val"].reshape(m, n)
    # Ruiz drives every row and column infinity norm to 1
    assert np.allclose(np.abs(A).max(axis=1), 1.0, atol=1e-6)
    assert np.allclose(np.abs(A).max(axis=0), 1.0, atol=1e-6)
    # scaled = D_r^-1 A D_c^-1
    assert np.allclose(A, np.outer(u, w) / np.outer(out["Dr"], out["Dc"]))


def test_pock_chambolle_and_bound_obj(orc):
    A = np.array([[1.0, -2.0, 0.0], [0.0, 3.0, 4.0]])
    m, n = A.shape
    rp = np.array([0, 2, 4])
    ci = np.array([0, 1, 1, 2])
    vals = np.array([1.0, -2.0, 3.0, 4.0])
    c = np.array([1.0, 2.0, -2.0])
    lc = np.array([1.0, -np.inf])
    uc = np.array([1.0, 5.0])
    prob = orc.build_problem(m, n, rp, ci, vals, c=c, con_lb=lc, con_ub=uc,
                             var_lb=np.zeros(n), var_ub=np.full(n, 10.0))
    out = _rescale(orc, prob, l_inf_ruiz_iterations=0, pock_chambolle_alpha=1.0)
    r = np.sqrt(np.abs(A).sum(axis=1))
    cc = np.sqrt(np.abs(A).sum(axis=0))
    assert np.allclose(out["Dr"], r) and np.allclose(out["Dc"], cc)
    # bound norm counts l only if finite and l != u, u if finite
    lcs, ucs = lc / r, uc / r
    bsq = ucs[0] ** 2 + ucs[1] ** 2
    beta = 1.0 / (math.sqrt(bsq) + 1.0)
    cs = c / cc
    gamma = 1.0 / (np.linalg.norm(cs) + 1.0)
    assert abs(out["beta"] - beta) < 1e-15 and abs(out["gamma"] - gamma) < 1e-15
    assert np.allclose(out["c"], cs * gamma)
    assert np.allclose(out["uv"], 10.0 * cc * beta)


# ---------------------------------------------------------------- power iteration

def _power(orc, rp, ci, v, m, n, tol=1e-10, iters=20000):
    nnz = len(v)
    trp = np.zeros(n + 1, np.int32)
    tci = np.zeros(nnz, np.int32)
    tv = np.zeros(nnz)
    L = orc.lib()
    L.orc_transpose(m, n, P(rp, C.c_int), P(ci, C.c_int), P(v), P(trp, C.c_int),
                    P(tci, C.c_int), P(tv))
    it = C.c_int()
    s = L.orc_power_iteration(m, n, P(rp, C.c_int), P(ci, C.c_int), P(v), P(trp, C.c_int),
                              P(tci, C.c_int), P(tv), iters, tol, C.byref(it))
    return s, it.value


def test_power_iteration_closed_forms(orc):
    d = np.array([3.0, -7.0, 2.0, 0.5])
    rp = np.arange(5, dtype=np.int32)
    s, _ = _power(orc, rp, np.arange(4, dtype=np.int32), d, 4, 4)
    assert abs(s - 7.0) < 1e-8
    u = np.array([1.0, 2.0, 2.0])
    w = np.array([3.0, 4.0])
    s, it = _power(orc, np.arange(0, 7, 2, dtype=np.int32), np.tile([0, 1], 3).astype(np.int32),
                   np.outer(u, w).ravel(), 3, 2)
    assert abs(s - np.linalg.norm(u) * np.linalg.norm(w)) < 1e-9 and it <= 3


def test_power_iteration_vs_svd(orc):
    rp, ci, v = random_csr(30, 20, np.full(30, 6), seed=8)
    dense = np.zeros((30, 20))
    for i in range(30):
        dense[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    s, _ = _power(orc, rp, ci, v, 30, 20, tol=1e-12, iters=200000)
    assert abs(s - np.linalg.svd(dense, compute_uv=False)[0]) < 1e-6 * s


# ---------------------------------------------------------------- whole solves

def test_solve_capi_example_closed_form(orc):
    A, c, l, u = capi_example()
    m, n = A.shape
    rp = np.array([0, 2, 3, 5])
    ci = np.array([0, 1, 1, 0, 1])
    vals = np.array([1.0, 2.0, 1.0, 3.0, 2.0])
    prob = orc.build_problem(m, n, rp, ci, vals, c, l, u, np.full(n, -np.inf),
                             np.full(n, np.inf))
    r, _ = orc.solve(C.pointer(prob), orc.default_params(presolve=False))
    assert r["status"] == "OPTIMAL"
    assert np.allclose(r["x"], [1, 2], atol=1e-3)
    assert np.allclose(r["y"], [1, -1, 0], atol=1e-3)
    assert abs(r["primal_objective_value"] - 3.0) < 1e-3


def _gen_arrays(dp, m, n, nnz, seed, rule=0):
    prob, opt = dp.generate(m, n, nnz, row_rule=rule, seed=seed)
    return prob, opt


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_solve_matches_known_optimum_and_highs(dp, orc, seed):
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    prob, opt = _gen_arrays(dp, 60, 150, 900, seed)
    r, _ = orc.solve(prob.ptr, orc.default_params(
        presolve=False, termination_criteria__eps_optimal_relative=1e-8,
        termination_criteria__eps_feasible_relative=1e-8))
    assert r["status"] == "OPTIMAL"
    a = prob.arrays()
    A = csr_matrix((a["vals"], a["col_ind"], a["row_ptr"]), shape=(a["m"], a["n"])).toarray()
    lc, uc = a["con_lb"], a["con_ub"]
    eq = lc == uc
    ub_rows = ~eq & np.isfinite(uc)
    lb_rows = ~eq & np.isfinite(lc)
    A_ub = np.vstack([A[ub_rows], -A[lb_rows]])
    b_ub = np.concatenate([uc[ub_rows], -lc[lb_rows]])
    h = linprog(a["c"], A_ub=A_ub, b_ub=b_ub, A_eq=A[eq], b_eq=lc[eq],
                bounds=list(zip(a["var_lb"], np.where(np.isinf(a["var_ub"]), None, a["var_ub"]))),
                method="highs")
    assert h.status == 0
    assert abs(h.fun - opt) <= 1e-7 * max(1, abs(opt))
    assert abs(r["primal_objective_value"] - h.fun) <= 1e-6 * max(1, abs(h.fun))


def test_infeasible_unbounded_and_limits(orc):
    inf = np.inf
    cap = dict(presolve=False, termination_criteria__iteration_limit=5000)
    # x + y >= 2 and x + y <= 1 over the unit box: the dual ray is certified at T = 200
    p1 = orc.build_problem(2, 2, [0, 2, 4], [0, 1, 0, 1], [1.0] * 4, [0.0, 0.0], [2.0, -inf],
                           [inf, 1.0], [0.0, 0.0], [1.0, 1.0])
    r, _ = orc.solve(C.pointer(p1), orc.default_params(**cap))
    assert r["status"] == "PRIMAL_INFEASIBLE" and r["total_count"] == 200
    p3 = orc.build_problem(1, 1, [0, 1], [0], [1.0], [1.0], [-inf], [-1.0], [0.0], [inf])
    r, _ = orc.solve(C.pointer(p3), orc.default_params(**cap))
    assert r["status"] == "PRIMAL_INFEASIBLE"
    # min -x s.t. x - y >= 0, x, y >= 0: unbounded along x = y
    p2 = orc.build_problem(1, 2, [0, 2], [0, 1], [1.0, -1.0], [-1.0, 0.0], [0.0], [inf],
                           [0.0, 0.0], [inf, inf])
    r, _ = orc.solve(C.pointer(p2), orc.default_params(**cap))
    assert r["status"] == "DUAL_INFEASIBLE" and r["total_count"] == 200
    r, _ = orc.solve(C.pointer(p2), orc.default_params(
        presolve=False, termination_criteria__iteration_limit=0))
    assert r["status"] == "ITERATION_LIMIT" and r["total_count"] == 0


def test_polish_reaches_tolerance(dp, orc):
    prob, opt = _gen_arrays(dp, 80, 200, 1500, 5)
    r, _ = orc.solve(prob.ptr, orc.default_params(presolve=False, feasibility_polishing=True))
    r0, _ = orc.solve(prob.ptr, orc.default_params(presolve=False))
    assert r["status"] == r0["status"] == "OPTIMAL" and r["total_count"] == r0["total_count"]
    # unpolished: feasible only to the 1e-4 termination tolerance
    assert r0["relative_primal_residual"] > 1e-6 and r0["relative_dual_residual"] > 1e-6
    # both polishes succeed: each residual reaches eps_polish = 1e-6
    assert r["relative_primal_residual"] <= 1e-6 and r["relative_dual_residual"] <= 1e-6
    assert r["feasibility_iteration"] > 0
    # the gap is recomputed from the polished objectives (pdlp_core_op.cu:923-928)
    pobj, dobj = r["primal_objective_value"], r["dual_objective_value"]
    assert r["objective_gap"] == abs(pobj - dobj)
    assert r["relative_objective_gap"] == abs(pobj - dobj) / (1 + abs(pobj) + abs(dobj))
    assert r["feasibility_polishing_time"] >= 0.0


def test_omp_build_agrees(dp, orc):
    prob, opt = _gen_arrays(dp, 200, 500, 4000, 11)
    r1, _ = orc.solve(prob.ptr, orc.default_params(presolve=False))
    r2, _ = orc.solve(prob.ptr, orc.default_params(presolve=False), omp=True)
    assert r1["status"] == r2["status"] == "OPTIMAL"
    assert abs(r1["primal_objective_value"] - r2["primal_objective_value"]) <= 1e-4 * max(
        1, abs(r1["primal_objective_value"]))
