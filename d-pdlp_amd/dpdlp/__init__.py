"""Python (ctypes) binding of the MI355X PDLP library.

Mirrors the reference's C API (include/cupdlpx.h: create_lp_problem,
set_start_values, solve_lp_problem, set_default_parameters,
cupdlpx_result_free, lp_problem_free; include/mps_parser.h: read_mps_file)
over ``lib/libcupdlpx-dist.so`` plus the extension entry points of
include/dpdlp_ext.h.  The binding only moves arguments; every computation runs
in the native library on the GPU.  Importing fails loudly when the library has
not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libcupdlpx-dist.so")
CLI_PATH = os.path.join(PKG_ROOT, "lib", "cupdlpx-dist")

from ._types import *  # noqa: F401,F403
from ._types import (_dp, _ip, _Csr, _Dense, AutotuneReport, BatchLayout, BatchOpts,
                     CupdlpxResult, GenSpec, LpProblem,
                     ServeReport, ServeSpec,
                     MatrixDesc, PdhgParameters, SolveOpts, TERMINATION, MATRIX_COO,
                     MATRIX_CSC, MATRIX_CSR, MATRIX_DENSE, PREC_FP64)

# ---------------------------------------------------------------- library

def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()' or make -C d-pdlp_amd)")
    lib = C.CDLL(LIB_PATH)
    P, R, O = C.POINTER(LpProblem), C.POINTER(CupdlpxResult), C.POINTER(SolveOpts)
    sig = {
        "create_lp_problem": (P, [_dp, C.POINTER(MatrixDesc), _dp, _dp, _dp, _dp, _dp]),
        "set_start_values": (None, [P, _dp, _dp]),
        "solve_lp_problem": (R, [P, C.POINTER(PdhgParameters)]),
        "set_default_parameters": (None, [C.POINTER(PdhgParameters)]),
        "cupdlpx_result_free": (None, [R]),
        "lp_problem_free": (None, [P]),
        "read_mps_file": (P, [C.c_char_p]),
        "dpdlp_default_opts": (None, [O]),
        "dpdlp_solve_ex": (R, [P, C.POINTER(PdhgParameters), O]),
        "dpdlp_upload": (C.c_void_p, [P, C.c_int]),
        "dpdlp_upload_tuned": (C.c_void_p, [P, C.c_int, _ip]),
        "dpdlp_solve_resident": (R, [C.c_void_p, C.POINTER(PdhgParameters), O]),
        "dpdlp_resident_free": (None, [C.c_void_p]),
        "dpdlp_resident_info": (None, [C.c_void_p, _ip, _ip, _ip, _ip, _ip]),
        "dpdlp_resident_layout": (None, [C.c_void_p, _ip, _ip, _ip, _ip, _ip]),
        "dpdlp_last_error": (C.c_char_p, []),
        "dpdlp_write_vector": (C.c_int, [C.c_char_p, _dp, C.c_int]),
        "dpdlp_time_spmv": (C.c_double, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
        "dpdlp_generate": (P, [C.POINTER(GenSpec), _dp]),
        "dpdlp_copy_calibration": (C.c_double, [C.c_longlong, C.c_int]),
        "dpdlp_gather_calibration": (C.c_double, [C.c_longlong, C.c_longlong, C.c_int]),
        "dpdlp_gather_calibration_ex": (C.c_double, [C.c_longlong, C.c_longlong, C.c_int,
                                                     C.c_int, C.c_int]),
        "dpdlp_gather_calibration_values": (C.c_double, [C.c_longlong, C.c_longlong, C.c_int,
                                                         C.c_int, C.c_int, C.c_int]),
        "dpdlp_reduce_host": (C.c_int, [C.c_int, _dp, _dp, C.c_longlong, _dp]),
        "dpdlp_scaling_host": (C.c_int, [P, C.POINTER(PdhgParameters)] + [_dp] * 9),
        "dpdlp_plan_blocks": (C.c_int, [C.c_int, _ip, C.c_int, _ip]),
        "dpdlp_resident_spmv": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.c_int]),
        "dpdlp_resident_spmv_ex": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                             C.c_int]),
        "dpdlp_resident_transpose": (C.c_int, [C.c_void_p, _ip, _ip, _dp]),
        "dpdlp_upload_ex": (C.c_void_p, [P, C.c_int, C.c_int, C.c_int]),
        "dpdlp_upload_caps": (C.c_void_p, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
        "dpdlp_upload_batch_layout": (C.c_void_p, [P, C.POINTER(BatchLayout)]),
        "dpdlp_batch_size_class": (C.c_int, [P]),
        "dpdlp_batch_autotune": (C.c_int, [C.POINTER(P), C.c_int, C.POINTER(PdhgParameters),
                                          C.c_int, C.c_char_p, C.POINTER(BatchLayout),
                                          C.POINTER(AutotuneReport)]),
        "dpdlp_resident_bytes": (C.c_size_t, [C.c_void_p]),
        "dpdlp_batch_member_spec": (None, [C.c_int, C.c_uint64, C.c_int, C.c_int,
                                           C.POINTER(GenSpec)]),
        "dpdlp_serve_default_spec": (None, [C.POINTER(ServeSpec)]),
        "dpdlp_serve_run": (C.c_int, [C.POINTER(ServeSpec), C.POINTER(PdhgParameters),
                                      C.POINTER(ServeReport), _dp, _ip, C.POINTER(R),
                                      C.c_int]),
        "dpdlp_batch_default_opts": (None, [C.POINTER(BatchOpts)]),
        "dpdlp_batch_create": (C.c_void_p, [C.POINTER(PdhgParameters), C.POINTER(BatchOpts)]),
        "dpdlp_batch_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_longlong]),
        "dpdlp_batch_step": (C.c_int, [C.c_void_p]),
        "dpdlp_batch_take": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(R),
                                       C.POINTER(C.c_double)]),
        "dpdlp_batch_in_flight": (C.c_int, [C.c_void_p]),
        "dpdlp_batch_bytes": (C.c_size_t, [C.c_void_p]),
        "dpdlp_batch_counters": (None, [C.c_void_p, C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "dpdlp_batch_free": (None, [C.c_void_p]),
        "dpdlp_batch_phase_stats": (None, [C.c_void_p, _ip, _dp, _dp, _dp, _dp]),
        "dpdlp_batch_graph_replays": (C.c_longlong, [C.c_void_p]),
        "dpdlp_solve_batch": (C.c_int, [C.POINTER(C.c_void_p), C.c_int,
                                        C.POINTER(PdhgParameters), C.POINTER(BatchOpts),
                                        C.POINTER(R)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _ptr(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def default_params(**kw) -> PdhgParameters:
    """set_default_parameters() plus keyword overrides (nested names use '__',
    e.g. termination_criteria__iteration_limit=100)."""
    p = PdhgParameters()
    lib.set_default_parameters(C.byref(p))
    for k, v in kw.items():
        obj, name = p, k
        if "__" in k:
            outer, name = k.split("__", 1)
            obj = getattr(p, outer)
        if not hasattr(obj, name):
            raise AttributeError(k)
        setattr(obj, name, v)
    return p


class Problem:
    """Owns an lp_problem_t* from create_lp_problem / read_mps_file / dpdlp_generate."""

    def __init__(self, ptr):
        if not ptr:
            raise ValueError("problem construction failed: " + last_error())
        self.ptr = ptr

    @classmethod
    def from_csr(cls, m, n, row_ptr, col_ind, vals, c=None, con_lb=None, con_ub=None,
                 var_lb=None, var_ub=None, obj_const=None):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_ind, dtype=np.int32)
        v = _f64(vals)
        d = MatrixDesc(m=m, n=n, fmt=MATRIX_CSR, zero_tolerance=0.0)
        d.data.csr = _Csr(nnz=len(v), row_ptr=_ptr(rp, C.c_int), col_ind=_ptr(ci, C.c_int),
                          vals=_ptr(v))
        return cls._make(d, c, con_lb, con_ub, var_lb, var_ub, obj_const, keep=(rp, ci, v))

    @classmethod
    def from_dense(cls, A, c=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None,
                   obj_const=None, zero_tolerance=0.0):
        A = _f64(np.atleast_2d(A))
        m, n = A.shape
        d = MatrixDesc(m=m, n=n, fmt=MATRIX_DENSE, zero_tolerance=zero_tolerance)
        d.data.dense = _Dense(A=_ptr(A))
        return cls._make(d, c, con_lb, con_ub, var_lb, var_ub, obj_const, keep=(A,))

    @classmethod
    def from_coo(cls, m, n, rows, cols, vals, fmt=MATRIX_COO, zero_tolerance=0.0, **kw):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        cc = np.ascontiguousarray(cols, dtype=np.int32)
        v = _f64(vals)
        d = MatrixDesc(m=m, n=n, fmt=fmt, zero_tolerance=zero_tolerance)
        d.data.coo = _Csr(nnz=len(v), row_ptr=_ptr(r, C.c_int), col_ind=_ptr(cc, C.c_int),
                          vals=_ptr(v))
        return cls._make(d, kw.get("c"), kw.get("con_lb"), kw.get("con_ub"), kw.get("var_lb"),
                         kw.get("var_ub"), kw.get("obj_const"), keep=(r, cc, v))

    @classmethod
    def from_csc(cls, m, n, col_ptr, row_ind, vals, zero_tolerance=0.0, **kw):
        cp = np.ascontiguousarray(col_ptr, dtype=np.int32)
        ri = np.ascontiguousarray(row_ind, dtype=np.int32)
        v = _f64(vals)
        d = MatrixDesc(m=m, n=n, fmt=MATRIX_CSC, zero_tolerance=zero_tolerance)
        d.data.csc = _Csr(nnz=len(v), row_ptr=_ptr(cp, C.c_int), col_ind=_ptr(ri, C.c_int),
                          vals=_ptr(v))
        return cls._make(d, kw.get("c"), kw.get("con_lb"), kw.get("con_ub"), kw.get("var_lb"),
                         kw.get("var_ub"), kw.get("obj_const"), keep=(cp, ri, v))

    @classmethod
    def _make(cls, d, c, con_lb, con_ub, var_lb, var_ub, obj_const, keep):
        c, con_lb, con_ub, var_lb, var_ub = map(_f64, (c, con_lb, con_ub, var_lb, var_ub))
        oc = None if obj_const is None else C.byref(C.c_double(obj_const))
        ptr = lib.create_lp_problem(_ptr(c), C.byref(d), _ptr(con_lb), _ptr(con_ub),
                                    _ptr(var_lb), _ptr(var_ub), oc)
        return cls(ptr)

    @classmethod
    def from_mps(cls, path):
        return cls(lib.read_mps_file(os.fsencode(path)))

    @property
    def m(self):
        return self.ptr.contents.num_constraints

    @property
    def n(self):
        return self.ptr.contents.num_variables

    @property
    def nnz(self):
        return self.ptr.contents.constraint_matrix_num_nonzeros

    def arrays(self):
        """Copies of the problem arrays (CSR, bounds, objective)."""
        p = self.ptr.contents
        m, n, nnz = p.num_constraints, p.num_variables, p.constraint_matrix_num_nonzeros

        def g(ptr, k, dt=np.float64):
            if k == 0 or not ptr:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt, copy=True)

        return dict(m=m, n=n, row_ptr=g(p.constraint_matrix_row_pointers, m + 1, np.int32),
                    col_ind=g(p.constraint_matrix_col_indices, nnz, np.int32),
                    vals=g(p.constraint_matrix_values, nnz), c=g(p.objective_vector, n),
                    var_lb=g(p.variable_lower_bound, n), var_ub=g(p.variable_upper_bound, n),
                    con_lb=g(p.constraint_lower_bound, m), con_ub=g(p.constraint_upper_bound, m),
                    obj_const=p.objective_constant)

    def set_start(self, primal=None, dual=None):
        primal, dual = _f64(primal), _f64(dual)
        lib.set_start_values(self.ptr, _ptr(primal), _ptr(dual))

    def free(self):
        if self.ptr:
            lib.lp_problem_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Result:
    status: str
    x: np.ndarray
    y: np.ndarray
    reduced_cost: np.ndarray
    iterations: int
    primal_obj: float
    dual_obj: float
    rel_primal: float
    rel_dual: float
    rel_gap: float
    abs_primal: float
    abs_dual: float
    gap: float
    solve_time: float
    raw: dict


def result_from_ptr(rp, free_fn=None) -> Result:
    if not rp:
        raise RuntimeError("solve failed: " + last_error())
    r = rp.contents
    n, m = r.num_variables, r.num_constraints

    def g(ptr, k):
        return np.ctypeslib.as_array(ptr, shape=(k,)).copy() if k > 0 else np.zeros(0)

    raw = {name: getattr(r, name) for name, t in CupdlpxResult._fields_ if t not in (_dp,)}
    out = Result(status=TERMINATION.get(r.termination_reason, "UNKNOWN"),
                 x=g(r.primal_solution, n), y=g(r.dual_solution, m),
                 reduced_cost=g(r.reduced_cost, n), iterations=r.total_count,
                 primal_obj=r.primal_objective_value, dual_obj=r.dual_objective_value,
                 rel_primal=r.relative_primal_residual, rel_dual=r.relative_dual_residual,
                 rel_gap=r.relative_objective_gap, abs_primal=r.absolute_primal_residual,
                 abs_dual=r.absolute_dual_residual, gap=r.objective_gap,
                 solve_time=r.cumulative_time_sec, raw=raw)
    (free_fn or lib.cupdlpx_result_free)(rp)
    return out


def make_opts(precision=PREC_FP64, quiet=True, spmv_cap=0, profile_phases=0,
              trace_iters=None, n=0, m=0, mixed_switch_tol=None, use_graph=0):
    o = SolveOpts()
    lib.dpdlp_default_opts(C.byref(o))
    o.precision = precision
    o.quiet = int(quiet)
    o.spmv_cap = spmv_cap
    o.profile_phases = profile_phases
    o.use_graph = use_graph
    if mixed_switch_tol is not None:
        o.mixed_switch_tol = mixed_switch_tol
    keep = []
    if trace_iters is not None:
        it = np.ascontiguousarray(trace_iters, dtype=np.int32)
        tx = np.zeros((len(it), max(n, 1)))
        ty = np.zeros((len(it), max(m, 1)))
        o.trace_points = len(it)
        o.trace_iters = _ptr(it, C.c_int)
        o.trace_x = _ptr(tx)
        o.trace_y = _ptr(ty)
        keep = [it, tx, ty]
    return o, keep


def solve(problem: Problem, params: PdhgParameters | None = None, **opt_kw):
    """solve_lp_problem with options; returns (Result, SolveOpts, trace arrays)."""
    o, keep = make_opts(n=problem.n, m=problem.m, **opt_kw)
    pp = None if params is None else C.byref(params)
    rp = lib.dpdlp_solve_ex(problem.ptr, pp, C.byref(o))
    res = result_from_ptr(rp)
    trace = None
    if keep:
        trace = (keep[1][: o.trace_recorded, : problem.n], keep[2][: o.trace_recorded, : problem.m])
    return res, o, trace


def solve_plain(problem: Problem, params: PdhgParameters | None = None) -> Result:
    """Exactly solve_lp_problem() of the C API."""
    pp = None if params is None else C.byref(params)
    return result_from_ptr(lib.solve_lp_problem(problem.ptr, pp))


class Resident:
    """A problem uploaded to HBM once (dpdlp_upload) and solved many times."""

    def __init__(self, problem: Problem, spmv_cap=0, tune=False, panels=None, spmv_cap_At=0,
                 layout=None):
        self.panels = None
        if layout is not None:
            lay = layout if isinstance(layout, BatchLayout) else BatchLayout(*layout)
            self.h = lib.dpdlp_upload_batch_layout(problem.ptr, C.byref(lay))
        elif spmv_cap_At:
            pa, pt = panels if panels is not None else (0, 0)
            self.h = lib.dpdlp_upload_caps(problem.ptr, spmv_cap, spmv_cap_At, pa, pt)
            self.panels = None if panels is None else tuple(panels)
        elif panels is not None:
            self.h = lib.dpdlp_upload_ex(problem.ptr, spmv_cap, panels[0], panels[1])
            self.panels = tuple(panels)
        elif tune:
            ch = (C.c_int * 2)()
            self.h = lib.dpdlp_upload_tuned(problem.ptr, spmv_cap, ch)
            self.panels = (ch[0], ch[1])
        else:
            self.h = lib.dpdlp_upload(problem.ptr, spmv_cap)
        if not self.h:
            raise RuntimeError("upload failed: " + last_error())
        m, n, nnz, ba, bt = (C.c_int() for _ in range(5))
        lib.dpdlp_resident_info(self.h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(ba),
                                C.byref(bt))
        self.m, self.n, self.nnz, self.nblk_A, self.nblk_At = (
            m.value, n.value, nnz.value, ba.value, bt.value)
        pa, pt, ca, ct, rn = (C.c_int() for _ in range(5))
        lib.dpdlp_resident_layout(self.h, C.byref(pa), C.byref(pt), C.byref(ca), C.byref(ct),
                                  C.byref(rn))
        self.panels = (pa.value, pt.value)
        self.caps = (ca.value, ct.value)
        self.renumbered = bool(rn.value)

    def solve(self, params=None, **opt_kw):
        o, _ = make_opts(n=self.n, m=self.m, **opt_kw)
        pp = None if params is None else C.byref(params)
        rp = lib.dpdlp_solve_resident(self.h, pp or C.byref(default_params()), C.byref(o))
        return result_from_ptr(rp), o

    def spmv(self, x, transpose=False, use32=False, reps=1, two_kernel=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n if transpose else self.m)
        rc = lib.dpdlp_resident_spmv_ex(self.h, int(transpose), _ptr(x), _ptr(y), int(use32),
                                        int(reps), int(two_kernel))
        if rc != 0:
            raise RuntimeError("spmv failed: " + last_error())
        return y

    def transpose(self):
        """The stored unscaled A' as (row_ptr, col_ind, vals) host arrays."""
        rp = np.zeros(self.n + 1, dtype=np.int32)
        ci = np.zeros(max(self.nnz, 1), dtype=np.int32)
        v = np.zeros(max(self.nnz, 1))
        rc = lib.dpdlp_resident_transpose(self.h, _ptr(rp, C.c_int), _ptr(ci, C.c_int), _ptr(v))
        if rc != 0:
            raise RuntimeError("A' is stored as tiles only or the copy failed")
        return rp, ci[: self.nnz], v[: self.nnz]

    def time_spmv(self, transpose=False, reps=20, use32=False):
        return lib.dpdlp_time_spmv(self.h, int(transpose), reps, int(use32))

    def free(self):
        if self.h:
            lib.dpdlp_resident_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def batch_opts(max_lps=1024, precision=PREC_FP64, spmv_cap=4096, mixed_switch_tol=None,
               profile_phases=0, use_graph=0, solo_lps=None, spmv_cap_At=None,
               persist_max_nnz=0, persist_budget=None):
    o = BatchOpts()
    lib.dpdlp_batch_default_opts(C.byref(o))
    o.max_lps = max_lps
    o.persist_max_nnz = persist_max_nnz
    if persist_budget is not None:
        o.persist_budget = persist_budget
    o.precision = precision
    o.spmv_cap = spmv_cap
    if spmv_cap_At is not None:
        o.spmv_cap_At = spmv_cap_At
    o.profile_phases = profile_phases
    o.use_graph = use_graph
    if solo_lps is not None:
        o.solo_lps = solo_lps
    if mixed_switch_tol is not None:
        o.mixed_switch_tol = mixed_switch_tol
    return o


class Batch:
    """dpdlp_batch_*: LPs added at any time advance together, one window of
    iterations per step(); finished LPs come out of take()."""

    def __init__(self, params=None, **opt_kw):
        self.opts = batch_opts(**opt_kw)
        self.params = params if params is not None else default_params()
        self.h = lib.dpdlp_batch_create(C.byref(self.params), C.byref(self.opts))
        if not self.h:
            raise RuntimeError("batch create failed: " + last_error())
        self._keep = {}

    def add(self, resident: "Resident", tag: int):
        """Adds an LP uploaded with Resident(problem, spmv_cap=cap, panels=(1, 1))."""
        if lib.dpdlp_batch_add(self.h, resident.h, 0, tag) != 0:
            raise RuntimeError("batch add failed: " + last_error())
        self._keep[tag] = resident

    def step(self) -> int:
        k = lib.dpdlp_batch_step(self.h)
        if k < 0:
            raise RuntimeError("batch step failed: " + last_error())
        return k

    def take(self):
        """[(tag, Result, latency_sec)] of the LPs finished so far."""
        out = []
        tag, lat = C.c_longlong(), C.c_double()
        rp = C.POINTER(CupdlpxResult)()
        while lib.dpdlp_batch_take(self.h, C.byref(tag), C.byref(rp), C.byref(lat)) == 1:
            out.append((tag.value, result_from_ptr(rp), lat.value))
            self._keep.pop(tag.value, None)
        return out

    def in_flight(self) -> int:
        return lib.dpdlp_batch_in_flight(self.h)

    def counters(self):
        w, s_, l_ = C.c_longlong(), C.c_longlong(), C.c_longlong()
        lib.dpdlp_batch_counters(self.h, C.byref(w), C.byref(s_), C.byref(l_))
        return {"windows": w.value, "syncs": s_.value, "launches": l_.value,
                "graph_replays": lib.dpdlp_batch_graph_replays(self.h)}

    def phase_stats(self):
        """Sampled phase-kernel timing: samples, ms and algorithmic bytes of each phase."""
        n = C.c_int()
        a, b, c, d = (C.c_double() for _ in range(4))
        lib.dpdlp_batch_phase_stats(self.h, C.byref(n), C.byref(a), C.byref(b), C.byref(c),
                                    C.byref(d))
        return {"samples": n.value, "primal_ms": a.value, "primal_bytes": b.value,
                "dual_ms": c.value, "dual_bytes": d.value}

    def run(self):
        """Steps until every LP finished; returns {tag: Result}."""
        res = {}
        while self.in_flight() > 0:
            self.step()
            for tag, r, _ in self.take():
                res[tag] = r
        for tag, r, _ in self.take():
            res[tag] = r
        return res

    def free(self):
        if self.h:
            lib.dpdlp_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def solve_batch(residents, params=None, stats=None, **opt_kw):
    """dpdlp_solve_batch over Resident objects; list of Results in input order.
    stats (a dict) receives the batch counters and phase timing."""
    o = batch_opts(**opt_kw)
    pp = params if params is not None else default_params()
    arr = (C.c_void_p * len(residents))(*[r.h for r in residents])
    out = (C.POINTER(CupdlpxResult) * len(residents))()
    k = lib.dpdlp_solve_batch(arr, len(residents), C.byref(pp), C.byref(o), out)
    if k != len(residents):
        raise RuntimeError("batch solve failed: " + last_error())
    if stats is not None:
        for name in ("windows", "syncs", "launches", "graph_replays", "phase_samples", "primal_ms",
                     "primal_bytes", "dual_ms", "dual_bytes"):
            stats[name] = getattr(o, name)
    return [result_from_ptr(out[i]) for i in range(len(residents))]


def generate(m, n, nnz=0, row_rule=0, seed=0, zipf_s=1.5, zipf_max=50000,
             poisson_lambda=8.0, pm1_frac=0.0):
    """dpdlp_generate: (Problem, known optimal objective)."""
    spec = GenSpec(m=m, n=n, nnz=nnz, row_rule=row_rule, zipf_s=zipf_s, zipf_max=zipf_max,
                   poisson_lambda=poisson_lambda, pm1_frac=pm1_frac, seed=seed)
    opt = C.c_double(0.0)
    ptr = lib.dpdlp_generate(C.byref(spec), C.byref(opt))
    return Problem(ptr), opt.value


def serve(params=None, max_requests=0, want_results=False, **spec_kw):
    """dpdlp_serve_run: returns (report dict, latencies, pool indices, results or None)."""
    sp = ServeSpec()
    lib.dpdlp_serve_default_spec(C.byref(sp))
    for k, v in spec_kw.items():
        if not hasattr(sp, k):
            raise KeyError(k)
        setattr(sp, k, v)
    rep = ServeReport()
    cap = max_requests if max_requests > 0 else int(sp.rate * sp.duration_sec * 1.5 + 64)
    lat = np.full(cap, np.nan)
    idx = np.full(cap, -1, dtype=np.int32)
    res = (C.POINTER(CupdlpxResult) * cap)() if want_results else None
    pp = params if params is not None else default_params()
    rc = lib.dpdlp_serve_run(C.byref(sp), C.byref(pp), C.byref(rep), _ptr(lat),
                             _ptr(idx, C.c_int), res, cap)
    if rc != 0:
        raise RuntimeError("serve run failed: " + last_error())
    n = rep.arrived
    report = {name: getattr(rep, name) for name, _ in ServeReport._fields_}
    results = None
    if want_results:
        results = [result_from_ptr(res[i]) if res[i] else None for i in range(n)]
    return report, lat[:n], idx[:n], results


def batch_autotune(problems, params=None, precision=PREC_FP64, cache_path=None):
    """dpdlp_batch_autotune: per size class, the row-block layout of A and A'
    whose batched phase kernels run fastest; returns ([(cap_A, rows_A,
    cap_At, rows_At) per problem], report dict)."""
    n = len(problems)
    arr = (C.POINTER(LpProblem) * max(n, 1))(*[p.ptr for p in problems])
    out = (BatchLayout * max(n, 1))()
    rep = AutotuneReport()
    pp = params if params is not None else default_params()
    path = cache_path.encode() if cache_path else None
    if lib.dpdlp_batch_autotune(arr, n, C.byref(pp), precision, path, out, C.byref(rep)) != 0:
        raise RuntimeError("autotune failed: " + last_error())
    lays = [(o.cap_A, o.rows_A, o.cap_At, o.rows_At) for o in out[:n]]
    return lays, {k: getattr(rep, k) for k, _ in AutotuneReport._fields_}


def batch_member(k, base_seed=3000, m_lo=500, m_hi=200_000):
    """Member k of the uneven batch workload: (Problem, known optimum)."""
    spec = GenSpec()
    lib.dpdlp_batch_member_spec(k, base_seed, m_lo, m_hi, C.byref(spec))
    opt = C.c_double(0.0)
    ptr = lib.dpdlp_generate(C.byref(spec), C.byref(opt))
    return Problem(ptr), opt.value


def serve_pool_member(k, pool_small=64, member_seed=3000, m_lo=500, m_hi=200_000,
                      large_m=1_000_000, large_nnz=10_000_000):
    """Entry k of the serving request pool (dpdlp_serve_run builds the same
    pool): members 0..pool_small-1 of the uneven batch workload, then the
    large requests (batch-generator draws at m = large_m with
    large_nnz / large_m - 1 as the Poisson row-length mean)."""
    spec = GenSpec()
    if k < pool_small:
        lib.dpdlp_batch_member_spec(k, member_seed, m_lo, m_hi, C.byref(spec))
    else:
        lib.dpdlp_batch_member_spec(k - pool_small, member_seed + 100000, large_m, large_m,
                                    C.byref(spec))
        spec.poisson_lambda = max(large_nnz / large_m - 1.0, 0.5)
    opt = C.c_double(0.0)
    ptr = lib.dpdlp_generate(C.byref(spec), C.byref(opt))
    return Problem(ptr), opt.value


def device_scaling(problem: Problem, params: PdhgParameters):
    """The device scaling of a solve, copied back (dpdlp_scaling_host): dict of
    scaled values (CSR order), c, lv, uv, lc, uc, Dr, Dc, beta, gamma."""
    m, n, nnz = problem.m, problem.n, problem.nnz
    out = dict(val=np.zeros(max(nnz, 1)), c=np.zeros(max(n, 1)), lv=np.zeros(max(n, 1)),
               uv=np.zeros(max(n, 1)), lc=np.zeros(max(m, 1)), uc=np.zeros(max(m, 1)),
               Dr=np.zeros(max(m, 1)), Dc=np.zeros(max(n, 1)))
    bg = np.zeros(2)
    rc = lib.dpdlp_scaling_host(problem.ptr, C.byref(params), _ptr(out["val"]), _ptr(out["c"]),
                                _ptr(out["lv"]), _ptr(out["uv"]), _ptr(out["lc"]),
                                _ptr(out["uc"]), _ptr(out["Dr"]), _ptr(out["Dc"]), _ptr(bg))
    if rc != 0:
        raise RuntimeError("device scaling failed: " + last_error())
    for k, length in (("val", nnz), ("c", n), ("lv", n), ("uv", n), ("lc", m), ("uc", m),
                      ("Dr", m), ("Dc", n)):
        out[k] = out[k][:length]
    out["beta"], out["gamma"] = float(bg[0]), float(bg[1])
    return out


RED_NRM2, RED_DOT, RED_AMAX, RED_SUM = 0, 1, 2, 3


def reduce(op, x, y=None):
    """One device reduction (dpdlp_dev_reduce) of host vectors."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = C.c_double(0.0)
    if lib.dpdlp_reduce_host(op, _ptr(x), _ptr(y), len(x), C.byref(out)) != 0:
        raise RuntimeError("reduction failed: " + last_error())
    return out.value


def last_error() -> str:
    s = lib.dpdlp_last_error()
    return s.decode() if s else ""


def plan_blocks(row_ptr, cap):
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    rows = len(rp) - 1
    out = np.zeros(rows + 2, dtype=np.int32)
    nb = lib.dpdlp_plan_blocks(rows, _ptr(rp, C.c_int), cap, _ptr(out, C.c_int))
    return out[: nb + 1]


EXPORTED_SYMBOLS = [
    "create_lp_problem", "set_start_values", "solve_lp_problem", "set_default_parameters",
    "cupdlpx_result_free", "lp_problem_free", "read_mps_file",
]
