"""ORACLE binding — test infrastructure only.

ctypes access to oracle/lib/liborcpdlp.so (single-thread FP64 restatement of
the reference solve path) and liborcpdlp_omp.so (its OpenMP build, the CPU
baseline).  Imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only; never by the product library or its binding.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "d-pdlp_amd"))
from dpdlp._types import (CupdlpxResult, LpProblem, PdhgParameters, TERMINATION,  # noqa: E402
                          _dp, _ip)

LIB_SEQ = os.path.join(_HERE, "lib", "liborcpdlp.so")
LIB_OMP = os.path.join(_HERE, "lib", "liborcpdlp_omp.so")


class OrcTrace(C.Structure):
    _fields_ = [("num_points", C.c_int), ("iters", _ip), ("x_cur", _dp), ("y_cur", _dp),
                ("scalars", _dp), ("recorded", C.c_int)]


def _load(path):
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run make -C oracle")
    lib = C.CDLL(path)
    P, R = C.POINTER(LpProblem), C.POINTER(CupdlpxResult)
    sig = {
        "orc_solve": (R, [P, C.POINTER(PdhgParameters), C.POINTER(OrcTrace)]),
        "orc_result_free": (None, [R]),
        "orc_last_timings": (None, [_dp]),
        "orc_default_parameters": (None, [C.POINTER(PdhgParameters)]),
        "orc_num_threads": (C.c_int, []),
        "orc_get_print_frequency": (C.c_int, [C.c_int]),
        "orc_transpose": (None, [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip, _dp]),
        "orc_spmv": (None, [C.c_int, _ip, _ip, _dp, _dp, _dp]),
        "orc_nrm2": (C.c_double, [C.c_int, _dp]),
        "orc_dot": (C.c_double, [C.c_int, _dp, _dp]),
        "orc_amax": (C.c_double, [C.c_int, _dp]),
        "orc_sum": (C.c_double, [C.c_int, _dp]),
        "orc_rescale": (None, [C.POINTER(PdhgParameters), P, _dp, _dp, _dp, _dp, _dp, _dp,
                               _dp, _dp, _dp, _dp]),
        "orc_mt_normal_fill": (None, [C.c_uint32, C.c_int, _dp]),
        "orc_power_iteration": (C.c_double, [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip, _dp,
                                             C.c_int, C.c_double, _ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_libs = {}


def lib(omp=False):
    key = LIB_OMP if omp else LIB_SEQ
    if key not in _libs:
        _libs[key] = _load(key)
    return _libs[key]


def p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def default_params(**kw):
    prm = PdhgParameters()
    lib().orc_default_parameters(C.byref(prm))
    for k, v in kw.items():
        obj, name = prm, k
        if "__" in k:
            outer, name = k.split("__", 1)
            obj = getattr(prm, outer)
        setattr(obj, name, v)
    # the oracle has no presolver either; the field is informational
    return prm


def solve(problem_ptr, params, trace_iters=None, n=0, m=0, omp=False):
    """orc_solve on an lp_problem_t* (from the product's create_lp_problem or a
    LpProblem built here); returns (dict result, (trace_x, trace_y) or None)."""
    L = lib(omp)
    tr = None
    keep = []
    if trace_iters is not None:
        it = np.ascontiguousarray(trace_iters, dtype=np.int32)
        tx = np.zeros((len(it), max(n, 1)))
        ty = np.zeros((len(it), max(m, 1)))
        sc = np.zeros((len(it), 4))
        tr = OrcTrace(num_points=len(it), iters=p(it, C.c_int), x_cur=p(tx), y_cur=p(ty),
                      scalars=p(sc), recorded=0)
        keep = [it, tx, ty, sc]
    rp = L.orc_solve(problem_ptr, C.byref(params), C.byref(tr) if tr is not None else None)
    if not rp:
        raise RuntimeError("oracle solve failed")
    r = rp.contents
    nn, mm = r.num_variables, r.num_constraints
    out = {name: getattr(r, name) for name, t in CupdlpxResult._fields_ if t is not _dp}
    out["status"] = TERMINATION.get(r.termination_reason, "UNKNOWN")
    out["x"] = np.ctypeslib.as_array(r.primal_solution, shape=(nn,)).copy() if nn else np.zeros(0)
    out["y"] = np.ctypeslib.as_array(r.dual_solution, shape=(mm,)).copy() if mm else np.zeros(0)
    out["rc"] = np.ctypeslib.as_array(r.reduced_cost, shape=(nn,)).copy() if nn else np.zeros(0)
    L.orc_result_free(rp)
    trace = None
    if tr is not None:
        k = tr.recorded
        trace = (keep[1][:k, :n], keep[2][:k, :m], keep[3][:k])
    return out, trace


def build_problem(m, n, row_ptr, col_ind, vals, c, con_lb, con_ub, var_lb, var_ub,
                  obj_const=0.0):
    """An LpProblem owned by Python (arrays kept alive on the object)."""
    arrs = dict(rp=np.ascontiguousarray(row_ptr, dtype=np.int32),
                ci=np.ascontiguousarray(col_ind, dtype=np.int32),
                v=np.ascontiguousarray(vals, dtype=np.float64),
                c=np.ascontiguousarray(c, dtype=np.float64),
                lc=np.ascontiguousarray(con_lb, dtype=np.float64),
                uc=np.ascontiguousarray(con_ub, dtype=np.float64),
                lv=np.ascontiguousarray(var_lb, dtype=np.float64),
                uv=np.ascontiguousarray(var_ub, dtype=np.float64))
    prob = LpProblem(num_variables=n, num_constraints=m,
                     variable_lower_bound=p(arrs["lv"]), variable_upper_bound=p(arrs["uv"]),
                     objective_vector=p(arrs["c"]), objective_constant=obj_const,
                     constraint_matrix_row_pointers=p(arrs["rp"], C.c_int),
                     constraint_matrix_col_indices=p(arrs["ci"], C.c_int),
                     constraint_matrix_values=p(arrs["v"]),
                     constraint_matrix_num_nonzeros=len(arrs["v"]),
                     constraint_lower_bound=p(arrs["lc"]), constraint_upper_bound=p(arrs["uc"]),
                     primal_start=None, dual_start=None)
    prob._keep = arrs
    return prob


def last_timings(omp=False):
    """(rescale, transpose, power, loop, polish) wall seconds of the last solve."""
    out = np.zeros(5)
    lib(omp).orc_last_timings(p(out))
    return dict(zip(("rescale", "transpose", "power", "loop", "polish"), out.tolist()))
