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
 * Public C API of the MI355X PDHG (PDLP) solver.
 *
 * Drop-in for include/cupdlpx.h of the reference: same six functions with the
 * same signatures, ownership and NULL-on-error behaviour.  Each declaration
 * names the reference declaration and definition it replaces.
 */
#pragma once

#include "cupdlpx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Deep-copies c, A (dense/CSR/CSC/COO -> CSR) and the bounds; NULL vectors
 * become 0 / -inf / +inf.  Returns NULL on an unsupported format or an index
 * out of range.  Replaces cupdlpx.h:26-30 (src/cupdlpx.c:26-117). */
lp_problem_t *create_lp_problem(const double *objective_c,
                                const matrix_desc_t *A_desc,
                                const double *con_lb, const double *con_ub,
                                const double *var_lb, const double *var_ub,
                                const double *objective_constant);

/* Copies a warm start into prob (either pointer may be NULL).
 * Replaces cupdlpx.h:33-34 (src/cupdlpx.c:146-172). */
void set_start_values(lp_problem_t *prob, const double *primal,
                      const double *dual);

/* Solves prob on the current HIP device; params == NULL means defaults.
 * Blocking.  Returns a heap result owned by the caller, or NULL on error
 * (bad arguments or a HIP failure; the library never calls exit()).
 * Replaces cupdlpx.h:37-38 (src/cupdlpx.c:174-198 -> src/solver.cu:32-127). */
cupdlpx_result_t *solve_lp_problem(const lp_problem_t *prob,
                                   const pdhg_parameters_t *params);

/* Replaces cupdlpx.h:41 (src/utils.cu:295-329). */
void set_default_parameters(pdhg_parameters_t *params);

/* Frees the result and all three of its arrays.
 * Replaces cupdlpx.h:43 (src/cupdlpx.c:119-127, which leaks reduced_cost). */
void cupdlpx_result_free(cupdlpx_result_t *results);

/* Replaces cupdlpx.h:45 (src/cupdlpx.c:129-144). */
void lp_problem_free(lp_problem_t *prob);

#ifdef __cplusplus
}
#endif
