/*
 * Compile-time proof that the public structs keep the reference's x86-64
 * layout (offsets measured from the reference's include/cupdlpx_types.h with
 * gcc 11), so binaries built against the reference headers can link here.
 */
#include "cupdlpx_types.h"
#include <stddef.h>

#define OFF(T, f, v) _Static_assert(offsetof(T, f) == (v), #T "." #f " offset")
#define SIZE(T, v) _Static_assert(sizeof(T) == (v), #T " size")

SIZE(lp_problem_t, 104);
OFF(lp_problem_t, num_variables, 0);
OFF(lp_problem_t, num_constraints, 4);
OFF(lp_problem_t, variable_lower_bound, 8);
OFF(lp_problem_t, variable_upper_bound, 16);
OFF(lp_problem_t, objective_vector, 24);
OFF(lp_problem_t, objective_constant, 32);
OFF(lp_problem_t, constraint_matrix_row_pointers, 40);
OFF(lp_problem_t, constraint_matrix_col_indices, 48);
OFF(lp_problem_t, constraint_matrix_values, 56);
OFF(lp_problem_t, constraint_matrix_num_nonzeros, 64);
OFF(lp_problem_t, constraint_lower_bound, 72);
OFF(lp_problem_t, constraint_upper_bound, 80);
OFF(lp_problem_t, primal_start, 88);
OFF(lp_problem_t, dual_start, 96);

SIZE(pdhg_parameters_t, 192);
OFF(pdhg_parameters_t, l_inf_ruiz_iterations, 0);
OFF(pdhg_parameters_t, has_pock_chambolle_alpha, 4);
OFF(pdhg_parameters_t, pock_chambolle_alpha, 8);
OFF(pdhg_parameters_t, bound_objective_rescaling, 16);
OFF(pdhg_parameters_t, verbose, 17);
OFF(pdhg_parameters_t, termination_evaluation_frequency, 20);
OFF(pdhg_parameters_t, sv_max_iter, 24);
OFF(pdhg_parameters_t, sv_tol, 32);
OFF(pdhg_parameters_t, termination_criteria, 40);
OFF(pdhg_parameters_t, restart_params, 88);
OFF(pdhg_parameters_t, reflection_coefficient, 144);
OFF(pdhg_parameters_t, feasibility_polishing, 152);
OFF(pdhg_parameters_t, optimality_norm, 156);
OFF(pdhg_parameters_t, presolve, 160);
OFF(pdhg_parameters_t, partition_method, 164);
OFF(pdhg_parameters_t, permute_method, 168);
OFF(pdhg_parameters_t, grid_size, 172);
OFF(pdhg_parameters_t, permute_block_size, 184);
SIZE(termination_criteria_t, 48);
SIZE(restart_parameters_t, 56);
SIZE(grid_size_t, 12);

SIZE(cupdlpx_result_t, 208);
OFF(cupdlpx_result_t, primal_solution, 24);
OFF(cupdlpx_result_t, dual_solution, 32);
OFF(cupdlpx_result_t, reduced_cost, 40);
OFF(cupdlpx_result_t, total_count, 48);
OFF(cupdlpx_result_t, termination_reason, 184);
OFF(cupdlpx_result_t, feasibility_iteration, 200);

SIZE(matrix_desc_t, 56);
OFF(matrix_desc_t, data, 24);
SIZE(termination_reason_t, 4);
SIZE(norm_type_t, 4);
SIZE(matrix_format_t, 4);

int dpdlp_abi_checked(void) { return 1; }
