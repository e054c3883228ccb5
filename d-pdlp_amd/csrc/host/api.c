/*
 * Public C API (include/cupdlpx.h): problem construction with the dense /
 * CSC / COO -> CSR converters, warm starts, defaults and frees.
 *
 * Follows src/cupdlpx.c:26-198 and the converters of src/utils.cu:816-1013.
 * Differences from the reference (documented behaviour changes):
 *   - allocation failures and bad arguments return NULL instead of exit();
 *   - cupdlpx_result_free also frees reduced_cost (leaked at cupdlpx.c:119-127);
 *   - A_desc == NULL is rejected instead of dereferenced.
 */
#include "cupdlpx.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* utils.cu:295-329 */
void set_default_parameters(pdhg_parameters_t *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->l_inf_ruiz_iterations = 10;
  p->has_pock_chambolle_alpha = true;
  p->pock_chambolle_alpha = 1.0;
  p->bound_objective_rescaling = true;
  p->verbose = false;
  p->termination_evaluation_frequency = 200;
  p->feasibility_polishing = false;
  p->reflection_coefficient = 1.0;
  p->sv_max_iter = 5000;
  p->sv_tol = 1e-4;
  p->termination_criteria.eps_optimal_relative = 1e-4;
  p->termination_criteria.eps_feasible_relative = 1e-4;
  p->termination_criteria.eps_infeasible = 1e-10;
  p->termination_criteria.time_sec_limit = 3600.0;
  p->termination_criteria.iteration_limit = INT32_MAX;
  p->termination_criteria.eps_feas_polish_relative = 1e-6;
  p->restart_params.artificial_restart_threshold = 0.36;
  p->restart_params.sufficient_reduction_for_restart = 0.2;
  p->restart_params.necessary_reduction_for_restart = 0.5;
  p->restart_params.k_p = 0.99;
  p->restart_params.k_i = 0.01;
  p->restart_params.k_d = 0.0;
  p->restart_params.i_smooth = 0.3;
  p->optimality_norm = NORM_TYPE_L2;
  p->presolve = true;
  p->grid_size.decided = false;
  p->partition_method = NNZ_BALANCE_PARTITION;
  p->permute_method = BLOCK_RANDOM_PERMUTATION;
}

static double *vec_or_fill(int n, const double *src, double fill) {
  double *d = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (!d) return NULL;
  if (src) {
    if (n > 0) memcpy(d, src, (size_t)n * sizeof(double));
  } else {
    for (int i = 0; i < n; ++i) d[i] = fill;
  }
  return d;
}

/* Dense row-major -> CSR, dropping |a| <= tol (default 1e-12), utils.cu:825-859 */
static int dense_to_csr(const matrix_desc_t *d, int **rp, int **ci, double **v, int *nnz_out) {
  int m = d->m, n = d->n;
  double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 1e-12;
  const double *A = d->data.dense.A;
  long long cnt = 0;
  if (!A && (long long)m * n > 0) return -1;
  for (long long k = 0; k < (long long)m * n; ++k)
    if (fabs(A[k]) > tol) ++cnt;
  if (cnt > INT32_MAX) return -1;
  *rp = (int *)malloc((size_t)(m + 1) * sizeof(int));
  *ci = (int *)malloc((size_t)(cnt ? cnt : 1) * sizeof(int));
  *v = (double *)malloc((size_t)(cnt ? cnt : 1) * sizeof(double));
  if (!*rp || !*ci || !*v) return -1;
  int nz = 0;
  for (int i = 0; i < m; ++i) {
    (*rp)[i] = nz;
    for (int j = 0; j < n; ++j) {
      double a = A[(size_t)i * n + j];
      if (fabs(a) > tol) {
        (*ci)[nz] = j;
        (*v)[nz] = a;
        ++nz;
      }
    }
  }
  (*rp)[m] = nz;
  *nnz_out = nz;
  return 0;
}

/* CSC -> CSR, dropping |a| <= tol only when tol > 0, utils.cu:861-922 */
static int csc_to_csr(const matrix_desc_t *d, int **rp, int **ci, double **v, int *nnz_out) {
  int m = d->m, n = d->n;
  const int *cp = d->data.csc.col_ptr, *ri = d->data.csc.row_ind;
  const double *vals = d->data.csc.vals;
  double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
  *rp = (int *)calloc((size_t)m + 1, sizeof(int));
  if (!*rp) return -1;
  if (n > 0 && !cp) return -1;
  int eff = 0;
  for (int j = 0; j < n; ++j) {
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
      int r = ri[k];
      if (r < 0 || r >= m) {
        fprintf(stderr, "[interface] CSC: row index out of range\n");
        return -1;
      }
      if (tol > 0 && fabs(vals[k]) <= tol) continue;
      (*rp)[r + 1]++;
      ++eff;
    }
  }
  for (int i = 0; i < m; ++i) (*rp)[i + 1] += (*rp)[i];
  *ci = (int *)malloc((size_t)(eff ? eff : 1) * sizeof(int));
  *v = (double *)malloc((size_t)(eff ? eff : 1) * sizeof(double));
  int *next = (int *)malloc((size_t)(m ? m : 1) * sizeof(int));
  if (!*ci || !*v || !next) {
    free(next);
    return -1;
  }
  for (int i = 0; i < m; ++i) next[i] = (*rp)[i];
  for (int j = 0; j < n; ++j) {
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
      if (tol > 0 && fabs(vals[k]) <= tol) continue;
      int pos = next[ri[k]]++;
      (*ci)[pos] = j;
      (*v)[pos] = vals[k];
    }
  }
  free(next);
  *nnz_out = eff;
  return 0;
}

/* COO -> CSR keeping input order within rows, utils.cu:924-1013 */
static int coo_to_csr(const matrix_desc_t *d, int **rp, int **ci, double **v, int *nnz_out) {
  int m = d->m, n = d->n, nin = d->data.coo.nnz;
  const int *r = d->data.coo.row_ind, *c = d->data.coo.col_ind;
  const double *vals = d->data.coo.vals;
  double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
  int nnz = 0;
  for (int k = 0; k < nin; ++k)
    if (!(tol > 0) || fabs(vals[k]) > tol) ++nnz;
  *rp = (int *)calloc((size_t)m + 1, sizeof(int));
  *ci = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
  *v = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
  if (!*rp || !*ci || !*v) return -1;
  for (int k = 0; k < nin; ++k) {
    if (tol > 0 && !(fabs(vals[k]) > tol)) continue;
    if (r[k] < 0 || r[k] >= m) {
      fprintf(stderr, "[interface] COO: row index out of range\n");
      return -1;
    }
    (*rp)[r[k] + 1]++;
  }
  for (int i = 0; i < m; ++i) (*rp)[i + 1] += (*rp)[i];
  int *next = (int *)malloc((size_t)(m ? m : 1) * sizeof(int));
  if (!next) return -1;
  for (int i = 0; i < m; ++i) next[i] = (*rp)[i];
  for (int k = 0; k < nin; ++k) {
    if (tol > 0 && !(fabs(vals[k]) > tol)) continue;
    if (c[k] < 0 || c[k] >= n) {
      fprintf(stderr, "[interface] COO: col index out of range\n");
      free(next);
      return -1;
    }
    int pos = next[r[k]]++;
    (*ci)[pos] = c[k];
    (*v)[pos] = vals[k];
  }
  free(next);
  *nnz_out = nnz;
  return 0;
}

/* cupdlpx.c:26-117 */
lp_problem_t *create_lp_problem(const double *objective_c, const matrix_desc_t *A_desc,
                                const double *con_lb, const double *con_ub,
                                const double *var_lb, const double *var_ub,
                                const double *objective_constant) {
  if (!A_desc || A_desc->m < 0 || A_desc->n < 0) {
    fprintf(stderr, "[interface] create_lp_problem: invalid matrix descriptor.\n");
    return NULL;
  }
  lp_problem_t *p = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
  if (!p) return NULL;
  p->num_variables = A_desc->n;
  p->num_constraints = A_desc->m;
  int *rp = NULL, *ci = NULL, nnz = 0, rc = 0;
  double *v = NULL;
  switch (A_desc->fmt) {
  case matrix_dense:
    rc = dense_to_csr(A_desc, &rp, &ci, &v, &nnz);
    break;
  case matrix_csc:
    rc = csc_to_csr(A_desc, &rp, &ci, &v, &nnz);
    if (rc) fprintf(stderr, "[interface] CSC->CSR failed.\n");
    break;
  case matrix_coo:
    rc = coo_to_csr(A_desc, &rp, &ci, &v, &nnz);
    if (rc) fprintf(stderr, "[interface] COO->CSR failed.\n");
    break;
  case matrix_csr: {
    nnz = A_desc->data.csr.nnz;
    if (nnz < 0) {
      rc = -1;
      break;
    }
    rp = (int *)malloc((size_t)(A_desc->m + 1) * sizeof(int));
    ci = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
    v = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    if (!rp || !ci || !v || !A_desc->data.csr.row_ptr) {
      rc = -1;
      break;
    }
    memcpy(rp, A_desc->data.csr.row_ptr, (size_t)(A_desc->m + 1) * sizeof(int));
    if (nnz) {
      memcpy(ci, A_desc->data.csr.col_ind, (size_t)nnz * sizeof(int));
      memcpy(v, A_desc->data.csr.vals, (size_t)nnz * sizeof(double));
    }
    break;
  }
  default:
    fprintf(stderr, "[interface] make_problem_from_matrix: unsupported matrix format %d.\n",
            (int)A_desc->fmt);
    rc = -1;
  }
  if (rc) {
    free(rp);
    free(ci);
    free(v);
    free(p);
    return NULL;
  }
  p->constraint_matrix_row_pointers = rp;
  p->constraint_matrix_col_indices = ci;
  p->constraint_matrix_values = v;
  p->constraint_matrix_num_nonzeros = nnz;
  p->objective_constant = objective_constant ? *objective_constant : 0.0;
  p->objective_vector = vec_or_fill(p->num_variables, objective_c, 0.0);
  p->variable_lower_bound = vec_or_fill(p->num_variables, var_lb, -INFINITY);
  p->variable_upper_bound = vec_or_fill(p->num_variables, var_ub, INFINITY);
  p->constraint_lower_bound = vec_or_fill(p->num_constraints, con_lb, -INFINITY);
  p->constraint_upper_bound = vec_or_fill(p->num_constraints, con_ub, INFINITY);
  if (!p->objective_vector || !p->variable_lower_bound || !p->variable_upper_bound ||
      !p->constraint_lower_bound || !p->constraint_upper_bound) {
    lp_problem_free(p);
    return NULL;
  }
  return p;
}

/* cupdlpx.c:119-127 (+ reduced_cost) */
void cupdlpx_result_free(cupdlpx_result_t *r) {
  if (!r) return;
  free(r->primal_solution);
  free(r->dual_solution);
  free(r->reduced_cost);
  free(r);
}

/* cupdlpx.c:129-144 */
void lp_problem_free(lp_problem_t *p) {
  if (!p) return;
  free(p->constraint_matrix_row_pointers);
  free(p->constraint_matrix_col_indices);
  free(p->constraint_matrix_values);
  free(p->variable_lower_bound);
  free(p->variable_upper_bound);
  free(p->objective_vector);
  free(p->constraint_lower_bound);
  free(p->constraint_upper_bound);
  free(p->primal_start);
  free(p->dual_start);
  memset(p, 0, sizeof(*p));
  free(p);
}

/* cupdlpx.c:146-172 */
void set_start_values(lp_problem_t *p, const double *primal, const double *dual) {
  if (!p) return;
  free(p->primal_start);
  free(p->dual_start);
  p->primal_start = NULL;
  p->dual_start = NULL;
  if (primal) {
    p->primal_start = (double *)malloc((size_t)(p->num_variables > 0 ? p->num_variables : 1) *
                                       sizeof(double));
    if (p->primal_start && p->num_variables > 0)
      memcpy(p->primal_start, primal, (size_t)p->num_variables * sizeof(double));
  }
  if (dual) {
    p->dual_start = (double *)malloc(
        (size_t)(p->num_constraints > 0 ? p->num_constraints : 1) * sizeof(double));
    if (p->dual_start && p->num_constraints > 0)
      memcpy(p->dual_start, dual, (size_t)p->num_constraints * sizeof(double));
  }
}

/* cupdlpx.c:174-198 */
cupdlpx_result_t *solve_lp_problem(const lp_problem_t *prob, const pdhg_parameters_t *params) {
  if (!prob) {
    fprintf(stderr, "[interface] solve_lp_problem: invalid arguments.\n");
    return NULL;
  }
  pdhg_parameters_t local;
  if (params)
    local = *params;
  else
    set_default_parameters(&local);
  dpdlp_solve_opts_t opts;
  dpdlp_default_opts(&opts);
  cupdlpx_result_t *r = dpdlp_solve_ex(prob, &local, &opts);
  if (!r) fprintf(stderr, "[interface] optimize returned NULL.\n");
  return r;
}
