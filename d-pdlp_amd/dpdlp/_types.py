"""ctypes mirrors of the public C structs (include/cupdlpx_types.h,
include/dpdlp_ext.h).  Loading nothing, so the oracle binding can share them."""
import ctypes as C

# ---------------------------------------------------------------- C structs

TERMINATION = {
    0: "UNSPECIFIED", 1: "OPTIMAL", 2: "PRIMAL_INFEASIBLE", 3: "DUAL_INFEASIBLE",
    4: "INFEASIBLE_OR_UNBOUNDED", 5: "TIME_LIMIT", 6: "ITERATION_LIMIT",
    7: "FEAS_POLISH_SUCCESS",
}
NORM_L2, NORM_LINF = 0, 1
MATRIX_DENSE, MATRIX_CSR, MATRIX_CSC, MATRIX_COO = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class LpProblem(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int), ("num_constraints", C.c_int),
        ("variable_lower_bound", _dp), ("variable_upper_bound", _dp),
        ("objective_vector", _dp), ("objective_constant", C.c_double),
        ("constraint_matrix_row_pointers", _ip), ("constraint_matrix_col_indices", _ip),
        ("constraint_matrix_values", _dp), ("constraint_matrix_num_nonzeros", C.c_int),
        ("constraint_lower_bound", _dp), ("constraint_upper_bound", _dp),
        ("primal_start", _dp), ("dual_start", _dp),
    ]


class RestartParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "artificial_restart_threshold", "sufficient_reduction_for_restart",
        "necessary_reduction_for_restart", "k_p", "k_i", "k_d", "i_smooth")]


class TerminationCriteria(C.Structure):
    _fields_ = [
        ("eps_optimal_relative", C.c_double), ("eps_feasible_relative", C.c_double),
        ("eps_feas_polish_relative", C.c_double), ("eps_infeasible", C.c_double),
        ("time_sec_limit", C.c_double), ("iteration_limit", C.c_int),
    ]


class GridSize(C.Structure):
    _fields_ = [("row_dims", C.c_int), ("col_dims", C.c_int), ("decided", C.c_bool)]


class PdhgParameters(C.Structure):
    _fields_ = [
        ("l_inf_ruiz_iterations", C.c_int), ("has_pock_chambolle_alpha", C.c_bool),
        ("pock_chambolle_alpha", C.c_double), ("bound_objective_rescaling", C.c_bool),
        ("verbose", C.c_bool), ("termination_evaluation_frequency", C.c_int),
        ("sv_max_iter", C.c_int), ("sv_tol", C.c_double),
        ("termination_criteria", TerminationCriteria), ("restart_params", RestartParams),
        ("reflection_coefficient", C.c_double), ("feasibility_polishing", C.c_bool),
        ("optimality_norm", C.c_int), ("presolve", C.c_bool),
        ("partition_method", C.c_int), ("permute_method", C.c_int),
        ("grid_size", GridSize), ("permute_block_size", C.c_int),
    ]


class CupdlpxResult(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int), ("num_constraints", C.c_int), ("num_nonzeros", C.c_int),
        ("num_reduced_variables", C.c_int), ("num_reduced_constraints", C.c_int),
        ("num_reduced_nonzeros", C.c_int),
        ("primal_solution", _dp), ("dual_solution", _dp), ("reduced_cost", _dp),
        ("total_count", C.c_int), ("rescaling_time_sec", C.c_double),
        ("cumulative_time_sec", C.c_double), ("presolve_time", C.c_double),
        ("presolve_status", C.c_int),
        ("absolute_primal_residual", C.c_double), ("relative_primal_residual", C.c_double),
        ("absolute_dual_residual", C.c_double), ("relative_dual_residual", C.c_double),
        ("primal_objective_value", C.c_double), ("dual_objective_value", C.c_double),
        ("objective_gap", C.c_double), ("relative_objective_gap", C.c_double),
        ("max_primal_ray_infeasibility", C.c_double), ("max_dual_ray_infeasibility", C.c_double),
        ("primal_ray_linear_objective", C.c_double), ("dual_ray_objective", C.c_double),
        ("termination_reason", C.c_int), ("feasibility_polishing_time", C.c_double),
        ("feasibility_iteration", C.c_int),
    ]


class _Dense(C.Structure):
    _fields_ = [("A", _dp)]


class _Csr(C.Structure):
    _fields_ = [("nnz", C.c_int), ("row_ptr", _ip), ("col_ind", _ip), ("vals", _dp)]


class _Union(C.Union):
    _fields_ = [("dense", _Dense), ("csr", _Csr), ("csc", _Csr), ("coo", _Csr)]


class MatrixDesc(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("fmt", C.c_int),
                ("zero_tolerance", C.c_double), ("data", _Union)]


class SolveOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("mixed_switch_tol", C.c_double), ("spmv_cap", C.c_int),
        ("quiet", C.c_int), ("use_graph", C.c_int),
        ("trace_points", C.c_int), ("trace_iters", _ip), ("trace_x", _dp), ("trace_y", _dp),
        ("trace_recorded", C.c_int),
        ("setup_time_sec", C.c_double), ("power_time_sec", C.c_double),
        ("loop_time_sec", C.c_double), ("total_time_sec", C.c_double),
        ("power_iterations", C.c_int), ("step_size", C.c_double),
        ("iterations_fp32", C.c_int), ("time_to_1e4_sec", C.c_double),
        ("time_to_1e6_sec", C.c_double),
        ("profile_phases", C.c_int), ("primal_phase_ms", C.c_double),
        ("dual_phase_ms", C.c_double), ("phase_samples", C.c_int),
    ]


class GenSpec(C.Structure):
    _fields_ = [
        ("m", C.c_int), ("n", C.c_int), ("nnz", C.c_longlong), ("row_rule", C.c_int),
        ("zipf_s", C.c_double), ("zipf_max", C.c_int), ("poisson_lambda", C.c_double),
        ("pm1_frac", C.c_double), ("seed", C.c_uint64),
    ]




class BatchOpts(C.Structure):
    """dpdlp_batch_opts_t (include/dpdlp_ext.h)."""
    _fields_ = [("max_lps", C.c_int), ("precision", C.c_int), ("mixed_switch_tol", C.c_double),
                ("spmv_cap", C.c_int), ("profile_phases", C.c_int), ("use_graph", C.c_int),
                ("solo_lps", C.c_int), ("spmv_cap_At", C.c_int), ("persist_max_nnz", C.c_int),
                ("persist_budget", C.c_int),
                ("windows", C.c_longlong), ("syncs", C.c_longlong), ("launches", C.c_longlong),
                ("graph_replays", C.c_longlong),
                ("phase_samples", C.c_int), ("primal_ms", C.c_double),
                ("primal_bytes", C.c_double), ("dual_ms", C.c_double),
                ("dual_bytes", C.c_double)]


class ServeSpec(C.Structure):
    """dpdlp_serve_spec_t (include/dpdlp_ext.h)."""
    _fields_ = [("rate", C.c_double), ("duration_sec", C.c_double), ("drain_sec", C.c_double),
                ("seed", C.c_uint64), ("pool_small", C.c_int), ("pool_large", C.c_int),
                ("large_frac", C.c_double), ("large_m", C.c_int), ("large_nnz", C.c_longlong),
                ("m_lo", C.c_int), ("m_hi", C.c_int), ("member_seed", C.c_uint64),
                ("hbm_budget_gb", C.c_double), ("max_lps", C.c_int), ("precision", C.c_int),
                ("mode", C.c_int), ("spmv_cap", C.c_int), ("large_lane_nnz", C.c_longlong),
                ("persist_max_nnz", C.c_int)]


class ServeReport(C.Structure):
    """dpdlp_serve_report_t (include/dpdlp_ext.h)."""
    _fields_ = [("arrived", C.c_int), ("completed", C.c_int), ("optimal", C.c_int),
                ("unfinished", C.c_int), ("makespan_sec", C.c_double),
                ("lps_per_sec", C.c_double), ("p50_sec", C.c_double), ("p90_sec", C.c_double),
                ("p99_sec", C.c_double), ("mean_sec", C.c_double), ("max_sec", C.c_double),
                ("max_in_flight", C.c_int), ("peak_gb", C.c_double),
                ("pool_gen_sec", C.c_double), ("windows", C.c_longlong),
                ("syncs", C.c_longlong)]


class BatchLayout(C.Structure):
    """dpdlp_batch_layout_t"""
    _fields_ = [("cap_A", C.c_int), ("rows_A", C.c_int), ("cap_At", C.c_int),
                ("rows_At", C.c_int)]


class AutotuneReport(C.Structure):
    """dpdlp_autotune_report_t"""
    _fields_ = [("classes", C.c_int), ("tuned", C.c_int), ("cached", C.c_int),
                ("seconds", C.c_double)]
