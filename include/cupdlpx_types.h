/*
 * The declarations in this header keep the interface of the cuPDLPx / D-PDLP
 * reference header of the same name, which carries this notice:
 *
 *   Copyright 2025 Haihao Lu
 *
 *   Licensed under the Apache License, Version 2.0 (the "License");
 *   you may not use this file except in compliance with the License.
 *   You may obtain a copy of the License at
 *
 *       http://www.apache.org/licenses/LICENSE-2.0
 *
 *   Unless required by applicable law or agreed to in writing, software
 *   distributed under the License is distributed on an "AS IS" BASIS,
 *   WITHOUT WARRANTIES OR CONDITIONS OF ANY KIND, either express or implied.
 *   See the License for the specific language governing permissions and
 *   limitations under the License.
 */
/*
 * Public data types of the single-GPU PDHG (PDLP) solver.
 *
 * Drop-in replacement for include/cupdlpx_types.h of the reference
 * (cuPDLPx / D-PDLP).  Field names, field order, field types and enum values
 * are kept identical so that a program compiled against the reference header
 * links against this library unchanged (x86-64 sizes: lp_problem_t 104 B,
 * pdhg_parameters_t 192 B, cupdlpx_result_t 208 B, matrix_desc_t 56 B; the
 * library checks every offset with static assertions).
 *
 * The reference includes "PSLP_stats.h" here (cupdlpx_types.h:19) without
 * using any of its types; that include is dropped.
 */
#pragma once

#include <stdbool.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Why a solve stopped.  Values 0..7 as in the reference (cupdlpx_types.h:27-36). */
typedef enum {
  TERMINATION_REASON_UNSPECIFIED,
  TERMINATION_REASON_OPTIMAL,
  TERMINATION_REASON_PRIMAL_INFEASIBLE,
  TERMINATION_REASON_DUAL_INFEASIBLE,
  TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED,
  TERMINATION_REASON_TIME_LIMIT,
  TERMINATION_REASON_ITERATION_LIMIT,
  TERMINATION_REASON_FEAS_POLISH_SUCCESS
} termination_reason_t;

/* Norm used in the optimality residuals (cupdlpx_types.h:38). */
typedef enum { NORM_TYPE_L2 = 0, NORM_TYPE_L_INF = 1 } norm_type_t;

/*
 * min c'x + c0  s.t.  l_c <= A x <= u_c,  l_v <= x <= u_v
 * A is CSR with int32 indices; every pointer is host memory.
 */
typedef struct {
  int num_variables;
  int num_constraints;
  double *variable_lower_bound;
  double *variable_upper_bound;
  double *objective_vector;
  double objective_constant;

  int *constraint_matrix_row_pointers;
  int *constraint_matrix_col_indices;
  double *constraint_matrix_values;
  int constraint_matrix_num_nonzeros;

  double *constraint_lower_bound;
  double *constraint_upper_bound;

  double *primal_start; /* optional warm start, NULL if absent */
  double *dual_start;
} lp_problem_t;

/* Adaptive-restart thresholds and the PID gains of the primal weight. */
typedef struct {
  double artificial_restart_threshold;
  double sufficient_reduction_for_restart;
  double necessary_reduction_for_restart;
  double k_p;
  double k_i;
  double k_d;
  double i_smooth;
} restart_parameters_t;

typedef struct {
  double eps_optimal_relative;
  double eps_feasible_relative;
  double eps_feas_polish_relative;
  double eps_infeasible;
  double time_sec_limit;
  int iteration_limit;
} termination_criteria_t;

/* Multi-GPU fields: kept for ABI compatibility, ignored by this solver. */
typedef enum {
  UNIFORM_PARTITION,
  NNZ_BALANCE_PARTITION,
} partition_method_t;

typedef enum {
  NO_PERMUTATION,
  FULL_RANDOM_PERMUTATION,
  BLOCK_RANDOM_PERMUTATION,
} permute_method_t;

typedef struct {
  int row_dims;
  int col_dims;
  bool decided;
} grid_size_t;

typedef struct {
  int l_inf_ruiz_iterations;
  bool has_pock_chambolle_alpha;
  double pock_chambolle_alpha;
  bool bound_objective_rescaling;
  bool verbose;
  int termination_evaluation_frequency;
  int sv_max_iter;
  double sv_tol;
  termination_criteria_t termination_criteria;
  restart_parameters_t restart_params;
  double reflection_coefficient;
  bool feasibility_polishing;
  norm_type_t optimality_norm;
  bool presolve;                      /* accepted, ignored (no presolver) */
  partition_method_t partition_method; /* ignored (single GPU) */
  permute_method_t permute_method;     /* ignored (single GPU) */
  grid_size_t grid_size;               /* ignored (single GPU) */
  int permute_block_size;              /* ignored (single GPU) */
} pdhg_parameters_t;

typedef struct {
  int num_variables;
  int num_constraints;
  int num_nonzeros;

  int num_reduced_variables;
  int num_reduced_constraints;
  int num_reduced_nonzeros;

  double *primal_solution; /* host, num_variables */
  double *dual_solution;   /* host, num_constraints */
  double *reduced_cost;    /* host, num_variables */

  int total_count;
  double rescaling_time_sec;
  double cumulative_time_sec;
  double presolve_time;
  int presolve_status;

  double absolute_primal_residual;
  double relative_primal_residual;
  double absolute_dual_residual;
  double relative_dual_residual;
  double primal_objective_value;
  double dual_objective_value;
  double objective_gap;
  double relative_objective_gap;
  double max_primal_ray_infeasibility;
  double max_dual_ray_infeasibility;
  double primal_ray_linear_objective;
  double dual_ray_objective;
  termination_reason_t termination_reason;
  double feasibility_polishing_time;
  int feasibility_iteration;
} cupdlpx_result_t;

typedef enum {
  matrix_dense = 0,
  matrix_csr = 1,
  matrix_csc = 2,
  matrix_coo = 3
} matrix_format_t;

/* Input matrix in one of four layouts; converted to CSR by create_lp_problem. */
typedef struct {
  int m; /* rows (constraints) */
  int n; /* columns (variables) */
  matrix_format_t fmt;

  /* entries with |a| <= zero_tolerance are dropped (dense: default 1e-12) */
  double zero_tolerance;

  union MatrixData {
    struct MatrixDense {
      const double *A; /* row-major m*n */
    } dense;

    struct MatrixCSR {
      int nnz;
      const int *row_ptr;
      const int *col_ind;
      const double *vals;
    } csr;

    struct MatrixCSC {
      int nnz;
      const int *col_ptr;
      const int *row_ind;
      const double *vals;
    } csc;

    struct MatrixCOO {
      int nnz;
      const int *row_ind;
      const int *col_ind;
      const double *vals;
    } coo;
  } data;
} matrix_desc_t;

#ifdef __cplusplus
}
#endif
