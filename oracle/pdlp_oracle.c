/*
 * ORACLE — test infrastructure only (see pdlp_oracle.h for the pinning list).
 *
 * Plain FP64 C restatement of the reference's single-GPU solve path.  Every
 * function cites the reference lines whose semantics it follows.  Reductions
 * are sequential in the single-thread build; with -fopenmp the loops are
 * parallel (that build is the timed CPU baseline, not the parity reference).
 */
#define _POSIX_C_SOURCE 200809L
#include "pdlp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#define OMP_FOR_DYN _Pragma("omp parallel for schedule(dynamic, 256)")

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}
static double *dalloc(int n) { return (double *)xmalloc((size_t)n * sizeof(double)); }
static double *dzero(int n) { return (double *)calloc(n ? (size_t)n : 1, sizeof(double)); }
static double *dcopy(int n, const double *s) {
  double *d = dalloc(n);
  if (n) memcpy(d, s, (size_t)n * sizeof(double));
  return d;
}

static double wall_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* utils.cu:295-329 */
void orc_default_parameters(pdhg_parameters_t *p) {
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
  p->termination_criteria.iteration_limit = 2147483647;
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
  p->partition_method = NNZ_BALANCE_PARTITION;
  p->permute_method = BLOCK_RANDOM_PERMUTATION;
}

/* utils.cu:492-501: 10 below 1e3, then x10 per decade. */
int orc_get_print_frequency(int iter) {
  int step = 10;
  long long threshold = 1000;
  while (iter >= threshold) {
    step *= 10;
    threshold *= 10;
  }
  return step;
}

/* ------------------------------------------------------------------------ */
/* Sparse kernels                                                            */
/* ------------------------------------------------------------------------ */

/* Stable CSR transpose (the CSC of A, pdlp_core_op.cu:111-128): within each
 * output row the source rows appear in increasing order. */
void orc_transpose(int m, int n, const int *rp, const int *ci, const double *v,
                   int *t_rp, int *t_ci, double *t_v) {
  for (int j = 0; j <= n; ++j) t_rp[j] = 0;
  int nnz = rp[m];
  for (int k = 0; k < nnz; ++k) t_rp[ci[k] + 1]++;
  for (int j = 0; j < n; ++j) t_rp[j + 1] += t_rp[j];
  int *next = (int *)xmalloc((size_t)(n + 1) * sizeof(int));
  memcpy(next, t_rp, (size_t)(n + 1) * sizeof(int));
  for (int i = 0; i < m; ++i) {
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int pos = next[ci[k]]++;
      t_ci[pos] = i;
      t_v[pos] = v[k];
    }
  }
  free(next);
}

/* y = A x for CSR A (the cusparseSpMV call sites of pdlp_core_op.cu:501,534);
 * sequential left-to-right sum per row. */
void orc_spmv(int rows, const int *rp, const int *ci, const double *v,
              const double *x, double *y) {
  OMP_FOR_DYN
  for (int i = 0; i < rows; ++i) {
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

/* Euclidean norm with Blue's three-accumulator scaling (the safe-scaling
 * nrm2 of Anderson, ACM TOMS 44(1), Algorithm 978), standing in for
 * cublasDnrm2 at pdlp_core_op.cu:587,590,682,684 and utils.cu:127,152,669,680. */
#define BLUE_TSML 0x1p-511
#define BLUE_TBIG 0x1p+486
#define BLUE_SSML 0x1p+537
#define BLUE_SBIG 0x1p-538
double orc_nrm2(int n, const double *x) {
  if (n <= 0) return 0.0;
  double asml = 0.0, amed = 0.0, abig = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : asml, amed, abig)
#endif
  for (int i = 0; i < n; ++i) {
    double ax = fabs(x[i]);
    if (ax > BLUE_TBIG) {
      double t = ax * BLUE_SBIG;
      abig += t * t;
    } else if (ax < BLUE_TSML) {
      double t = ax * BLUE_SSML;
      asml += t * t;
    } else {
      amed += ax * ax;
    }
  }
  double scl, sumsq;
  if (abig > 0.0) {
    if (amed > 0.0 || isnan(amed)) abig += (amed * BLUE_SBIG) * BLUE_SBIG;
    scl = 1.0 / BLUE_SBIG;
    sumsq = abig;
  } else if (asml > 0.0) {
    if (amed > 0.0 || isnan(amed)) {
      double ymed = sqrt(amed);
      double ysml = sqrt(asml) / BLUE_SSML;
      double ymin = ysml > ymed ? ymed : ysml;
      double ymax = ysml > ymed ? ysml : ymed;
      scl = 1.0;
      double r = ymin / ymax;
      sumsq = ymax * ymax * (1.0 + r * r);
    } else {
      scl = 1.0 / BLUE_SSML;
      sumsq = asml;
    }
  } else {
    scl = 1.0;
    sumsq = amed;
  }
  return scl * sqrt(sumsq);
}

/* cublasDdot (pdlp_core_op.cu:690, utils.cu:142,687,696,760). */
double orc_dot(int n, const double *x, const double *y) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : s)
#endif
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* get_vector_inf_norm (utils.cu:616-627): |x_k| for the first k maximising
 * |x_i|, i.e. max |x_i|; 0 for n <= 0. */
double orc_amax(int n, const double *x) {
  if (n <= 0) return 0.0;
  double mx = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(max : mx)
#endif
  for (int i = 0; i < n; ++i) {
    double a = fabs(x[i]);
    if (a > mx) mx = a;
  }
  return mx;
}

/* get_vector_sum (utils.cu:629-637) = dot(x, ones). */
double orc_sum(int n, const double *x) {
  if (n <= 0) return 0.0;
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : s)
#endif
  for (int i = 0; i < n; ++i) s += x[i] * 1.0;
  return s;
}

/* ------------------------------------------------------------------------ */
/* Random numbers: std::mt19937 (Matsumoto & Nishimura 1998) and libstdc++'s */
/* normal_distribution (Marsaglia polar method), utils.cu:29-30,80-82.       */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t mt[624];
  int idx;
  int has_saved;
  double saved;
} orc_rng_t;

static void mt_seed(orc_rng_t *r, uint32_t seed) {
  r->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
  r->idx = 624;
  r->has_saved = 0;
  r->saved = 0.0;
}

static uint32_t mt_next(orc_rng_t *r) {
  if (r->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = r->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      r->mt[i] = v;
    }
    r->idx = 0;
  }
  uint32_t y = r->mt[r->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* generate_canonical<double, 53> over a 32-bit engine: two draws. */
static double mt_canonical(orc_rng_t *r) {
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)mt_next(r) * tmp;
    tmp *= 4294967296.0;
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

static double mt_normal(orc_rng_t *r) {
  if (r->has_saved) {
    r->has_saved = 0;
    return r->saved;
  }
  double x, y, r2;
  do {
    x = 2.0 * mt_canonical(r) - 1.0;
    y = 2.0 * mt_canonical(r) - 1.0;
    r2 = x * x + y * y;
  } while (r2 > 1.0 || r2 == 0.0);
  double mult = sqrt(-2.0 * log(r2) / r2);
  r->saved = x * mult;
  r->has_saved = 1;
  return y * mult;
}

void orc_mt_normal_fill(uint32_t seed, int count, double *out) {
  orc_rng_t r;
  mt_seed(&r, seed);
  for (int i = 0; i < count; ++i) out[i] = mt_normal(&r) * 1.0 + 0.0;
}

/* utils.cu:66-171: power iteration on A A' from q0 ~ N(0,1)^m. */
double orc_power_iteration(int m, int n, const int *rp, const int *ci,
                           const double *v, const int *t_rp, const int *t_ci,
                           const double *t_v, int max_iter, double tol,
                           int *iters_out) {
  double *eig = dalloc(m), *next = dalloc(m), *dp = dalloc(n);
  orc_rng_t rng;
  mt_seed(&rng, 1u); /* reseeded per solve: the reference's first solve */
  for (int i = 0; i < m; ++i) eig[i] = mt_normal(&rng);
  double sigma_sq = 1.0;
  int it = 0;
  for (int i = 0; i < max_iter; ++i) {
    it = i + 1;
    memcpy(next, eig, (size_t)m * sizeof(double));
    double nrm = orc_nrm2(m, next);
    double inv = 1.0 / nrm;
    OMP_FOR
    for (int k = 0; k < m; ++k) next[k] *= inv;
    orc_spmv(n, t_rp, t_ci, t_v, next, dp);
    orc_spmv(m, rp, ci, v, dp, eig);
    sigma_sq = orc_dot(m, next, eig);
    double neg = -sigma_sq;
    OMP_FOR
    for (int k = 0; k < m; ++k) next[k] = next[k] * neg + eig[k];
    double res = orc_nrm2(m, next);
    if (res < tol) break;
  }
  free(eig);
  free(next);
  free(dp);
  if (iters_out) *iters_out = it;
  return sqrt(sigma_sq);
}

/* ------------------------------------------------------------------------ */
/* Preconditioning (preconditioner.c:89-273)                                 */
/* ------------------------------------------------------------------------ */

#define SCALING_EPSILON 1e-12

/* preconditioner.c:89-110 */
static void scale_in_place(int m, int n, const int *rp, const int *ci,
                           double *val, double *c, double *lv, double *uv,
                           double *lc, double *uc, const double *rs,
                           const double *cs) {
  for (int j = 0; j < n; ++j) {
    c[j] /= cs[j];
    uv[j] *= cs[j];
    lv[j] *= cs[j];
  }
  for (int i = 0; i < m; ++i) {
    lc[i] /= rs[i];
    uc[i] /= rs[i];
  }
  OMP_FOR_DYN
  for (int i = 0; i < m; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) val[k] /= (rs[i] * cs[ci[k]]);
}

void orc_rescale(const pdhg_parameters_t *params, const lp_problem_t *prob,
                 double *val, double *c, double *lv, double *uv, double *lc,
                 double *uc, double *D_r, double *D_c, double *beta,
                 double *gamma) {
  int m = prob->num_constraints, n = prob->num_variables;
  int nnz = prob->constraint_matrix_num_nonzeros;
  const int *rp = prob->constraint_matrix_row_pointers;
  const int *ci = prob->constraint_matrix_col_indices;
  if (nnz) memcpy(val, prob->constraint_matrix_values, (size_t)nnz * sizeof(double));
  if (n) {
    memcpy(c, prob->objective_vector, (size_t)n * sizeof(double));
    memcpy(lv, prob->variable_lower_bound, (size_t)n * sizeof(double));
    memcpy(uv, prob->variable_upper_bound, (size_t)n * sizeof(double));
  }
  if (m) {
    memcpy(lc, prob->constraint_lower_bound, (size_t)m * sizeof(double));
    memcpy(uc, prob->constraint_upper_bound, (size_t)m * sizeof(double));
  }
  for (int i = 0; i < m; ++i) D_r[i] = 1.0;
  for (int j = 0; j < n; ++j) D_c[j] = 1.0;
  double *rs = dalloc(m), *cs = dalloc(n);

  /* Ruiz L-inf, preconditioner.c:112-160 */
  for (int it = 0; it < params->l_inf_ruiz_iterations; ++it) {
    for (int j = 0; j < n; ++j) cs[j] = 0.0;
    for (int i = 0; i < m; ++i) rs[i] = 0.0;
    for (int i = 0; i < m; ++i) {
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        double a = fabs(val[k]);
        if (a > cs[ci[k]]) cs[ci[k]] = a;
        if (a > rs[i]) rs[i] = a;
      }
    }
    for (int j = 0; j < n; ++j) cs[j] = (cs[j] < SCALING_EPSILON) ? 1.0 : sqrt(cs[j]);
    for (int i = 0; i < m; ++i) rs[i] = (rs[i] < SCALING_EPSILON) ? 1.0 : sqrt(rs[i]);
    scale_in_place(m, n, rp, ci, val, c, lv, uv, lc, uc, rs, cs);
    for (int j = 0; j < n; ++j) D_c[j] *= cs[j];
    for (int i = 0; i < m; ++i) D_r[i] *= rs[i];
  }

  /* Pock-Chambolle, preconditioner.c:162-195; column sums accumulate in
   * increasing row order. */
  if (params->has_pock_chambolle_alpha) {
    double alpha = params->pock_chambolle_alpha;
    for (int j = 0; j < n; ++j) cs[j] = 0.0;
    for (int i = 0; i < m; ++i) rs[i] = 0.0;
    for (int i = 0; i < m; ++i) {
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        double a = fabs(val[k]);
        cs[ci[k]] += pow(a, 2.0 - alpha);
        rs[i] += pow(a, alpha);
      }
    }
    for (int j = 0; j < n; ++j) cs[j] = (cs[j] < SCALING_EPSILON) ? 1.0 : sqrt(cs[j]);
    for (int i = 0; i < m; ++i) rs[i] = (rs[i] < SCALING_EPSILON) ? 1.0 : sqrt(rs[i]);
    scale_in_place(m, n, rp, ci, val, c, lv, uv, lc, uc, rs, cs);
    for (int j = 0; j < n; ++j) D_c[j] *= cs[j];
    for (int i = 0; i < m; ++i) D_r[i] *= rs[i];
  }

  /* bound / objective rescaling, preconditioner.c:226-269 */
  if (params->bound_objective_rescaling) {
    double bsq = 0.0;
    for (int i = 0; i < m; ++i) {
      if (isfinite(lc[i]) && lc[i] != uc[i]) bsq += lc[i] * lc[i];
      if (isfinite(uc[i])) bsq += uc[i] * uc[i];
    }
    double csq = 0.0;
    for (int j = 0; j < n; ++j) csq += c[j] * c[j];
    *beta = 1.0 / (sqrt(bsq) + 1.0);
    *gamma = 1.0 / (sqrt(csq) + 1.0);
    for (int i = 0; i < m; ++i) {
      lc[i] *= *beta;
      uc[i] *= *beta;
    }
    for (int j = 0; j < n; ++j) {
      lv[j] *= *beta;
      uv[j] *= *beta;
      c[j] *= *gamma;
    }
  } else {
    *beta = 1.0;
    *gamma = 1.0;
  }
  free(rs);
  free(cs);
}

/* ------------------------------------------------------------------------ */
/* Solver state (internal_types.h:50-145, host arrays)                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  int m, n;
  const int *A_rp, *A_ci;
  const double *A_v;
  const int *At_rp, *At_ci;
  const double *At_v;
  double *lv, *uv, *c, *lc, *uc;
  double *lc_fin, *uc_fin, *lv_fin, *uv_fin;
  const double *D_r, *D_c;
  double beta, gamma, c0;
  double obj_norm, bound_norm;

  double *x_init, *x_cur, *x_pdhg, *x_ref, *dual_product, *dual_slack, *r_d,
      *dx;
  double *y_init, *y_cur, *y_pdhg, *y_ref, *primal_product, *primal_slack,
      *r_p, *dy;

  double step, omega, best_omega, pw_err_sum, pw_last_err, best_gap;
  int total, inner;
  int is_major;
  double fpe, fpe0, fpe_last;

  double abs_p, rel_p, abs_d, rel_d, pobj, dobj, gap, rel_gap;
  double max_p_ray_inf, max_d_ray_inf, p_ray_lin_obj, d_ray_obj;
  termination_reason_t reason;
  double t_cum;
} ost;

static void clampf_bounds(int n, const double *v, double *out) {
  for (int i = 0; i < n; ++i) out[i] = isfinite(v[i]) ? v[i] : 0.0;
}

/* pdlp_core_op.cu:393-418 and :495-526 */
static void pdhg_primal(ost *s) {
  orc_spmv(s->n, s->At_rp, s->At_ci, s->At_v, s->y_cur, s->dual_product);
  double step = s->step / s->omega;
  int major = s->is_major ||
              ((s->total + 2) % orc_get_print_frequency(s->total + 2)) == 0;
  if (major) {
    OMP_FOR
    for (int j = 0; j < s->n; ++j) {
      double t = s->x_cur[j] - step * (s->c[j] - s->dual_product[j]);
      s->x_pdhg[j] = fmax(s->lv[j], fmin(t, s->uv[j]));
      s->dual_slack[j] = (s->x_pdhg[j] - t) / step;
      s->x_ref[j] = 2.0 * s->x_pdhg[j] - s->x_cur[j];
    }
  } else {
    OMP_FOR
    for (int j = 0; j < s->n; ++j) {
      double t = s->x_cur[j] - step * (s->c[j] - s->dual_product[j]);
      double p = fmax(s->lv[j], fmin(t, s->uv[j]));
      s->x_ref[j] = 2.0 * p - s->x_cur[j];
    }
  }
}

/* pdlp_core_op.cu:420-443 and :528-557 */
static void pdhg_dual(ost *s) {
  orc_spmv(s->m, s->A_rp, s->A_ci, s->A_v, s->x_ref, s->primal_product);
  double step = s->step * s->omega;
  int major = s->is_major ||
              ((s->total + 2) % orc_get_print_frequency(s->total + 2)) == 0;
  if (major) {
    OMP_FOR
    for (int i = 0; i < s->m; ++i) {
      double t = s->y_cur[i] / step - s->primal_product[i];
      double p = fmax(-s->uc[i], fmin(t, -s->lc[i]));
      s->y_pdhg[i] = (t - p) * step;
      s->y_ref[i] = 2.0 * s->y_pdhg[i] - s->y_cur[i];
    }
  } else {
    OMP_FOR
    for (int i = 0; i < s->m; ++i) {
      double t = s->y_cur[i] / step - s->primal_product[i];
      double p = fmax(-s->uc[i], fmin(t, -s->lc[i]));
      s->y_ref[i] = 2.0 * (t - p) * step - s->y_cur[i];
    }
  }
}

/* pdlp_core_op.cu:445-462 and :559-567 */
static void halpern(ost *s, double rho) {
  double w = (double)(s->inner + 1) / (s->inner + 2);
  OMP_FOR
  for (int j = 0; j < s->n; ++j) {
    double r = rho * s->x_ref[j] + (1.0 - rho) * s->x_cur[j];
    s->x_cur[j] = w * r + (1.0 - w) * s->x_init[j];
  }
  OMP_FOR
  for (int i = 0; i < s->m; ++i) {
    double r = rho * s->y_ref[i] + (1.0 - rho) * s->y_cur[i];
    s->y_cur[i] = w * r + (1.0 - w) * s->y_init[i];
  }
}

/* pdlp_core_op.cu:658-696 (overwrites dual_product with A' dy). */
static void fixed_point_error(ost *s) {
  OMP_FOR
  for (int j = 0; j < s->n; ++j) s->dx[j] = s->x_ref[j] - s->x_cur[j];
  OMP_FOR
  for (int i = 0; i < s->m; ++i) s->dy[i] = s->y_ref[i] - s->y_cur[i];
  orc_spmv(s->n, s->At_rp, s->At_ci, s->At_v, s->dy, s->dual_product);
  double dn = orc_nrm2(s->m, s->dy);
  double pn = orc_nrm2(s->n, s->dx);
  double movement = pn * pn * s->omega + dn * dn / s->omega;
  double cross = orc_dot(s->n, s->dual_product, s->dx);
  double interaction = 2 * s->step * cross;
  s->fpe = sqrt(movement + interaction);
}

/* utils.cu:639-719 */
static void compute_residual(ost *s, norm_type_t norm) {
  orc_spmv(s->m, s->A_rp, s->A_ci, s->A_v, s->x_pdhg, s->primal_product);
  orc_spmv(s->n, s->At_rp, s->At_ci, s->At_v, s->y_pdhg, s->dual_product);
  OMP_FOR
  for (int i = 0; i < s->m; ++i) {
    double cl = fmax(s->lc[i], fmin(s->primal_product[i], s->uc[i]));
    s->r_p[i] = (s->primal_product[i] - cl) * s->D_r[i];
    s->primal_slack[i] = fmax(s->y_pdhg[i], 0.0) * s->lc_fin[i] +
                         fmin(s->y_pdhg[i], 0.0) * s->uc_fin[i];
  }
  OMP_FOR
  for (int j = 0; j < s->n; ++j)
    s->r_d[j] = (s->c[j] - s->dual_product[j] - s->dual_slack[j]) * s->D_c[j];
  s->abs_p = (norm == NORM_TYPE_L_INF) ? orc_amax(s->m, s->r_p) : orc_nrm2(s->m, s->r_p);
  s->abs_p /= s->beta;
  s->abs_d = (norm == NORM_TYPE_L_INF) ? orc_amax(s->n, s->r_d) : orc_nrm2(s->n, s->r_d);
  s->abs_d /= s->gamma;
  s->pobj = orc_dot(s->n, s->c, s->x_pdhg) / (s->beta * s->gamma) + s->c0;
  double base = orc_dot(s->n, s->dual_slack, s->x_pdhg);
  double ssum = orc_sum(s->m, s->primal_slack);
  s->dobj = (base + ssum) / (s->beta * s->gamma) + s->c0;
  s->rel_p = s->abs_p / (1.0 + s->bound_norm);
  s->rel_d = s->abs_d / (1.0 + s->obj_norm);
  s->gap = fabs(s->pobj - s->dobj);
  s->rel_gap = s->gap / (1.0 + fabs(s->pobj) + fabs(s->dobj));
}

/* utils.cu:721-813 (clobbers primal_product, dual_product, primal_slack,
 * dual_slack exactly as the reference does). */
static void compute_infeasibility(ost *s) {
  OMP_FOR
  for (int j = 0; j < s->n; ++j) {
    if (isfinite(s->lv[j])) s->dx[j] = fmax(s->dx[j], 0.0);
    if (isfinite(s->uv[j])) s->dx[j] = fmin(s->dx[j], 0.0);
  }
  OMP_FOR
  for (int i = 0; i < s->m; ++i) {
    if (!isfinite(s->lc[i])) s->dy[i] = fmin(s->dy[i], 0.0);
    if (!isfinite(s->uc[i])) s->dy[i] = fmax(s->dy[i], 0.0);
  }
  double p_inf = orc_amax(s->n, s->dx);
  if (p_inf > 0.0) {
    double sc = 1.0 / p_inf;
    OMP_FOR
    for (int j = 0; j < s->n; ++j) s->dx[j] *= sc;
  }
  double d_inf = orc_amax(s->m, s->dy);
  orc_spmv(s->m, s->A_rp, s->A_ci, s->A_v, s->dx, s->primal_product);
  orc_spmv(s->n, s->At_rp, s->At_ci, s->At_v, s->dy, s->dual_product);
  s->p_ray_lin_obj = orc_dot(s->n, s->c, s->dx) / (s->beta * s->gamma);
  OMP_FOR
  for (int i = 0; i < s->m; ++i)
    s->primal_slack[i] = fmax(s->dy[i], 0.0) * s->lc_fin[i] +
                         fmin(s->dy[i], 0.0) * s->uc_fin[i];
  OMP_FOR
  for (int j = 0; j < s->n; ++j)
    s->dual_slack[j] = fmax(-s->dual_product[j], 0.0) * s->lv_fin[j] +
                       fmin(-s->dual_product[j], 0.0) * s->uv_fin[j];
  double sp = orc_sum(s->m, s->primal_slack);
  double sd = orc_sum(s->n, s->dual_slack);
  s->d_ray_obj = (sp + sd) / (s->beta * s->gamma);
  OMP_FOR
  for (int i = 0; i < s->m; ++i) {
    double pp = s->primal_product[i];
    s->primal_slack[i] = (fmax(0.0, -pp) * (double)(isfinite(s->lc[i]) != 0) +
                          fmax(0.0, pp) * (double)(isfinite(s->uc[i]) != 0)) *
                         s->D_r[i];
  }
  OMP_FOR
  for (int j = 0; j < s->n; ++j) {
    double dp = -s->dual_product[j];
    s->dual_slack[j] = (fmax(0.0, dp) * (double)(!isfinite(s->lv[j])) -
                        fmin(0.0, dp) * (double)(!isfinite(s->uv[j]))) *
                       s->D_c[j];
  }
  s->max_p_ray_inf = orc_amax(s->m, s->primal_slack);
  double ds_norm = orc_amax(s->n, s->dual_slack);
  s->max_d_ray_inf = ds_norm;
  double sf = fmax(d_inf, ds_norm);
  if (sf > 0.0) {
    s->max_d_ray_inf /= sf;
    s->d_ray_obj /= sf;
  } else {
    s->max_d_ray_inf = 0.0;
    s->d_ray_obj = 0.0;
  }
}

/* utils.cu:214-263 */
static void check_termination(ost *s, const termination_criteria_t *tc) {
  if (s->rel_d < tc->eps_feasible_relative && s->rel_p < tc->eps_feasible_relative &&
      s->rel_gap < tc->eps_optimal_relative) {
    s->reason = TERMINATION_REASON_OPTIMAL;
    return;
  }
  if (s->d_ray_obj > 0.0 && s->max_d_ray_inf / s->d_ray_obj <= tc->eps_infeasible) {
    s->reason = TERMINATION_REASON_PRIMAL_INFEASIBLE;
    return;
  }
  if (s->p_ray_lin_obj < 0.0 &&
      s->max_p_ray_inf / (-s->p_ray_lin_obj) <= tc->eps_infeasible) {
    s->reason = TERMINATION_REASON_DUAL_INFEASIBLE;
    return;
  }
  if (s->total >= tc->iteration_limit) {
    s->reason = TERMINATION_REASON_ITERATION_LIMIT;
    return;
  }
  if (s->t_cum >= tc->time_sec_limit) {
    s->reason = TERMINATION_REASON_TIME_LIMIT;
    return;
  }
}

/* utils.cu:265-293 */
static int should_restart(ost *s, const restart_parameters_t *rp, int freq) {
  int r = 0;
  if (s->total == freq) {
    r = 1;
  } else if (s->total > freq) {
    if (s->fpe <= rp->sufficient_reduction_for_restart * s->fpe0) r = 1;
    if (s->fpe <= rp->necessary_reduction_for_restart * s->fpe0 && s->fpe > s->fpe_last) r = 1;
    if (s->inner >= rp->artificial_restart_threshold * s->total) r = 1;
  }
  s->fpe_last = s->fpe;
  return r;
}

/* pdlp_core_op.cu:577-636 */
static void perform_restart(ost *s, const restart_parameters_t *rp) {
  double pd2 = 0.0, dd2 = 0.0;
  double *tmpx = dalloc(s->n), *tmpy = dalloc(s->m);
  for (int j = 0; j < s->n; ++j) tmpx[j] = s->x_pdhg[j] - s->x_init[j];
  for (int i = 0; i < s->m; ++i) tmpy[i] = s->y_pdhg[i] - s->y_init[i];
  memcpy(s->dx, tmpx, (size_t)s->n * sizeof(double));
  memcpy(s->dy, tmpy, (size_t)s->m * sizeof(double));
  pd2 = orc_nrm2(s->n, tmpx);
  dd2 = orc_nrm2(s->m, tmpy);
  free(tmpx);
  free(tmpy);
  double ratio = s->rel_d / s->rel_p;
  if (pd2 > 1e-16 && dd2 > 1e-16 && pd2 < 1e12 && dd2 < 1e12 && ratio > 1e-8 &&
      ratio < 1e8) {
    double err = log(dd2) - log(pd2) - log(s->omega);
    s->pw_err_sum *= rp->i_smooth;
    s->pw_err_sum += err;
    double derr = err - s->pw_last_err;
    s->omega *= exp(rp->k_p * err + rp->k_i * s->pw_err_sum + rp->k_d * derr);
    s->pw_last_err = err;
  } else {
    s->omega = s->best_omega;
    s->pw_err_sum = 0.0;
    s->pw_last_err = 0.0;
  }
  double g = fabs(log10(s->rel_d / s->rel_p));
  if (g < s->best_gap) {
    s->best_gap = g;
    s->best_omega = s->omega;
  }
  memcpy(s->x_init, s->x_pdhg, (size_t)s->n * sizeof(double));
  memcpy(s->x_cur, s->x_pdhg, (size_t)s->n * sizeof(double));
  memcpy(s->y_init, s->y_pdhg, (size_t)s->m * sizeof(double));
  memcpy(s->y_cur, s->y_pdhg, (size_t)s->m * sizeof(double));
  s->inner = 0;
  s->fpe_last = INFINITY;
}

static void trace_record(ost *s, orc_trace_t *tr) {
  if (!tr || tr->recorded >= tr->num_points) return;
  if (tr->iters[tr->recorded] != s->total) return;
  int k = tr->recorded++;
  if (s->n) memcpy(tr->x_cur + (size_t)k * s->n, s->x_cur, (size_t)s->n * sizeof(double));
  if (s->m) memcpy(tr->y_cur + (size_t)k * s->m, s->y_cur, (size_t)s->m * sizeof(double));
  if (tr->scalars) {
    tr->scalars[4 * k + 0] = s->step;
    tr->scalars[4 * k + 1] = s->omega;
    tr->scalars[4 * k + 2] = s->inner;
    tr->scalars[4 * k + 3] = s->fpe;
  }
}

static void alloc_iterates(ost *s) {
  s->x_init = dzero(s->n); s->x_cur = dzero(s->n); s->x_pdhg = dzero(s->n);
  s->x_ref = dzero(s->n); s->dual_product = dzero(s->n); s->dual_slack = dzero(s->n);
  s->r_d = dzero(s->n); s->dx = dzero(s->n);
  s->y_init = dzero(s->m); s->y_cur = dzero(s->m); s->y_pdhg = dzero(s->m);
  s->y_ref = dzero(s->m); s->primal_product = dzero(s->m); s->primal_slack = dzero(s->m);
  s->r_p = dzero(s->m); s->dy = dzero(s->m);
}

static void free_iterates(ost *s) {
  free(s->x_init); free(s->x_cur); free(s->x_pdhg); free(s->x_ref);
  free(s->dual_product); free(s->dual_slack); free(s->r_d); free(s->dx);
  free(s->y_init); free(s->y_cur); free(s->y_pdhg); free(s->y_ref);
  free(s->primal_product); free(s->primal_slack); free(s->r_p); free(s->dy);
}

/* ------------------------------------------------------------------------ */
/* Feasibility polishing (pdlp_core_op.cu:862-1353, utils.cu:1015-1229)    */
/* ------------------------------------------------------------------------ */

/* utils.cu:1145-1183 */
static void primal_polish_residual(ost *s, const ost *o, norm_type_t norm) {
  orc_spmv(s->m, s->A_rp, s->A_ci, s->A_v, s->x_pdhg, s->primal_product);
  for (int i = 0; i < s->m; ++i) {
    double cl = fmax(s->lc[i], fmin(s->primal_product[i], s->uc[i]));
    s->r_p[i] = (s->primal_product[i] - cl) * s->D_r[i];
  }
  s->abs_p = (norm == NORM_TYPE_L_INF) ? orc_amax(s->m, s->r_p) : orc_nrm2(s->m, s->r_p);
  s->abs_p /= s->beta;
  s->rel_p = s->abs_p / (1.0 + s->bound_norm);
  s->pobj = orc_dot(s->n, o->c, s->x_pdhg) / (s->beta * s->gamma) + s->c0;
}

/* utils.cu:1185-1229 */
static void dual_polish_residual(ost *s, const ost *o, norm_type_t norm) {
  orc_spmv(s->n, s->At_rp, s->At_ci, s->At_v, s->y_pdhg, s->dual_product);
  for (int j = 0; j < s->n; ++j)
    s->r_d[j] = (s->c[j] - s->dual_product[j] - s->dual_slack[j]) * s->D_c[j];
  for (int i = 0; i < s->m; ++i)
    s->primal_slack[i] = fmax(s->y_pdhg[i], 0.0) * o->lc_fin[i] +
                         fmin(s->y_pdhg[i], 0.0) * o->uc_fin[i];
  s->abs_d = (norm == NORM_TYPE_L_INF) ? orc_amax(s->n, s->r_d) : orc_nrm2(s->n, s->r_d);
  s->abs_d /= s->gamma;
  s->rel_d = s->abs_d / (1.0 + s->obj_norm);
  double base = orc_dot(s->n, s->dual_slack, o->x_pdhg);
  double ssum = orc_sum(s->m, s->primal_slack);
  s->dobj = (base + ssum) / (s->beta * s->gamma) + s->c0;
}

/* utils.cu:1015-1039 */
static void polish_check(ost *s, const termination_criteria_t *tc, int primal) {
  if (primal) {
    if (s->rel_p <= tc->eps_feas_polish_relative) {
      s->reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
      return;
    }
  } else {
    if (s->rel_d <= tc->eps_feas_polish_relative) {
      s->reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
      return;
    }
  }
  if (s->total >= tc->iteration_limit) {
    s->reason = TERMINATION_REASON_ITERATION_LIMIT;
    return;
  }
  if (s->t_cum >= tc->time_sec_limit) {
    s->reason = TERMINATION_REASON_TIME_LIMIT;
    return;
  }
}

/* Shared scalar reset of initialize_{primal,dual}_feas_polish_state
 * (pdlp_core_op.cu:1082-1096, :1227-1241). */
static void polish_reset_scalars(ost *p, const ost *o) {
  p->pw_err_sum = 0.0;
  p->pw_last_err = 0.0;
  p->best_omega = 0.0;
  p->fpe = INFINITY;
  p->fpe0 = INFINITY;
  p->fpe_last = INFINITY;
  p->step = o->step;
  p->omega = o->omega;
  p->is_major = 0;
  p->total = 0;
  p->inner = 0;
  p->reason = TERMINATION_REASON_UNSPECIFIED;
  p->t_cum = 0.0;
  p->best_gap = INFINITY;
}

/* One polishing loop (pdlp_core_op.cu:932-1030). */
static void polish_loop(const pdhg_parameters_t *params, ost *s, const ost *o,
                        int primal) {
  double t0 = wall_now();
  int do_restart = 0;
  int freq = params->termination_evaluation_frequency;
  while (s->reason == TERMINATION_REASON_UNSPECIFIED) {
    if (s->is_major || s->total == 0 ||
        s->total % orc_get_print_frequency(s->total) == 0) {
      if (primal)
        primal_polish_residual(s, o, params->optimality_norm);
      else
        dual_polish_residual(s, o, params->optimality_norm);
      s->t_cum = wall_now() - t0;
      polish_check(s, &params->termination_criteria, primal);
    }
    if (s->is_major || s->total == 0) {
      do_restart = should_restart(s, &params->restart_params, freq);
      if (do_restart) {
        if (primal) {
          memcpy(s->x_init, s->x_pdhg, (size_t)s->n * sizeof(double));
          memcpy(s->x_cur, s->x_pdhg, (size_t)s->n * sizeof(double));
        } else {
          memcpy(s->y_init, s->y_pdhg, (size_t)s->m * sizeof(double));
          memcpy(s->y_cur, s->y_pdhg, (size_t)s->m * sizeof(double));
        }
        s->inner = 0;
        s->fpe_last = INFINITY;
      }
    }
    s->is_major = ((s->total + 1) % freq) == 0;
    pdhg_primal(s);
    pdhg_dual(s);
    if (s->is_major || do_restart) {
      if (primal) {
        for (int j = 0; j < s->n; ++j) s->dx[j] = s->x_ref[j] - s->x_cur[j];
        double pn = orc_nrm2(s->n, s->dx);
        s->fpe = pn * pn * s->omega;
      } else {
        for (int i = 0; i < s->m; ++i) s->dy[i] = s->y_ref[i] - s->y_cur[i];
        double dn = orc_nrm2(s->m, s->dy);
        s->fpe = dn * dn / s->omega;
      }
      if (do_restart) {
        s->fpe0 = s->fpe;
        do_restart = 0;
      }
    }
    halpern(s, params->reflection_coefficient);
    s->inner++;
    s->total++;
  }
}

/* pdlp_core_op.cu:862-930 */
static void feasibility_polish(const pdhg_parameters_t *params, ost *s,
                               double *polish_time, int *polish_iters) {
  double t0 = wall_now();
  const termination_criteria_t *tc = &params->termination_criteria;
  if (s->rel_p < tc->eps_feas_polish_relative && s->rel_d < tc->eps_feas_polish_relative)
    return;
  double omega0 = params->bound_objective_rescaling
                      ? 1.0
                      : (s->obj_norm + 1.0) / (s->bound_norm + 1.0);

  /* primal polish: c = 0, c0 = 0, x copied, y zeroed (:1032-1105) */
  ost p = *s;
  p.c = dzero(s->n);
  p.c0 = 0.0;
  p.x_init = dcopy(s->n, s->x_init);
  p.x_cur = dcopy(s->n, s->x_cur);
  p.x_pdhg = dcopy(s->n, s->x_pdhg);
  p.x_ref = dcopy(s->n, s->x_ref);
  p.primal_product = dcopy(s->m, s->primal_product);
  p.y_init = dzero(s->m); p.y_cur = dzero(s->m); p.y_pdhg = dzero(s->m);
  p.y_ref = dzero(s->m); p.dual_product = dzero(s->n); p.dual_slack = dzero(s->n);
  p.primal_slack = dzero(s->m); p.r_d = dzero(s->n); p.r_p = dzero(s->m);
  p.dx = dzero(s->n); p.dy = dzero(s->m);
  polish_reset_scalars(&p, s);
  p.rel_d = 0.0; p.abs_d = 0.0; p.rel_gap = 0.0; p.gap = 0.0;
  p.omega = omega0;
  p.best_omega = omega0;
  polish_loop(params, &p, s, 1);
  if (p.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
    memcpy(s->x_pdhg, p.x_pdhg, (size_t)s->n * sizeof(double));
    s->abs_p = p.abs_p;
    s->rel_p = p.rel_p;
    s->pobj = p.pobj;
  }
  *polish_iters += p.total - 1;

  /* dual polish: finite bounds -> 0, y copied, x zeroed (:1145-1249) */
  ost d = *s;
  d.lc = dalloc(s->m); d.uc = dalloc(s->m); d.lv = dalloc(s->n); d.uv = dalloc(s->n);
  for (int i = 0; i < s->m; ++i) {
    d.lc[i] = isfinite(s->lc[i]) ? 0.0 : s->lc[i];
    d.uc[i] = isfinite(s->uc[i]) ? 0.0 : s->uc[i];
  }
  for (int j = 0; j < s->n; ++j) {
    d.lv[j] = isfinite(s->lv[j]) ? 0.0 : s->lv[j];
    d.uv[j] = isfinite(s->uv[j]) ? 0.0 : s->uv[j];
  }
  d.lc_fin = dzero(s->m); d.uc_fin = dzero(s->m); d.lv_fin = dzero(s->n); d.uv_fin = dzero(s->n);
  d.y_init = dcopy(s->m, s->y_init);
  d.y_cur = dcopy(s->m, s->y_cur);
  d.y_pdhg = dcopy(s->m, s->y_pdhg);
  d.y_ref = dcopy(s->m, s->y_ref);
  d.dual_product = dcopy(s->n, s->dual_product);
  d.dual_slack = dcopy(s->n, s->dual_slack);
  d.x_init = dzero(s->n); d.x_cur = dzero(s->n); d.x_pdhg = dzero(s->n);
  d.x_ref = dzero(s->n); d.primal_product = dzero(s->m); d.primal_slack = dzero(s->m);
  d.r_d = dzero(s->n); d.r_p = dzero(s->m); d.dx = dzero(s->n); d.dy = dzero(s->m);
  polish_reset_scalars(&d, s);
  d.rel_p = 0.0; d.abs_p = 0.0; d.rel_gap = 0.0; d.gap = 0.0;
  d.omega = omega0;
  d.best_omega = omega0;
  polish_loop(params, &d, s, 0);
  if (d.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
    memcpy(s->y_pdhg, d.y_pdhg, (size_t)s->m * sizeof(double));
    s->abs_d = d.abs_d;
    s->rel_d = d.rel_d;
    s->dobj = d.dobj;
  }
  *polish_iters += d.total - 1;
  s->gap = fabs(s->pobj - s->dobj);
  s->rel_gap = s->gap / (1.0 + fabs(s->pobj) + fabs(s->dobj));

  free(p.c);
  free(p.x_init); free(p.x_cur); free(p.x_pdhg); free(p.x_ref); free(p.primal_product);
  free(p.y_init); free(p.y_cur); free(p.y_pdhg); free(p.y_ref); free(p.dual_product);
  free(p.dual_slack); free(p.primal_slack); free(p.r_d); free(p.r_p); free(p.dx); free(p.dy);
  free(d.lc); free(d.uc); free(d.lv); free(d.uv);
  free(d.lc_fin); free(d.uc_fin); free(d.lv_fin); free(d.uv_fin);
  free(d.y_init); free(d.y_cur); free(d.y_pdhg); free(d.y_ref); free(d.dual_product);
  free(d.dual_slack); free(d.x_init); free(d.x_cur); free(d.x_pdhg); free(d.x_ref);
  free(d.primal_product); free(d.primal_slack); free(d.r_d); free(d.r_p); free(d.dx); free(d.dy);
  *polish_time = wall_now() - t0;
}

/* ------------------------------------------------------------------------ */
/* Driver (solver.cu:32-127)                                                 */
/* ------------------------------------------------------------------------ */

/* wall times of the last orc_solve: rescale, transpose, power, loop, polish */
static double g_timings[5];
void orc_last_timings(double *out) { memcpy(out, g_timings, sizeof(g_timings)); }

cupdlpx_result_t *orc_solve(const lp_problem_t *prob,
                            const pdhg_parameters_t *params,
                            orc_trace_t *trace) {
  if (!prob || !params) return NULL;
  int m = prob->num_constraints, n = prob->num_variables;
  int nnz = prob->constraint_matrix_num_nonzeros;
  double t_scale0 = wall_now();

  ost s;
  memset(&s, 0, sizeof(s));
  s.m = m;
  s.n = n;
  double *A_v = dalloc(nnz);
  s.c = dalloc(n); s.lv = dalloc(n); s.uv = dalloc(n);
  s.lc = dalloc(m); s.uc = dalloc(m);
  double *D_r = dalloc(m), *D_c = dalloc(n);
  orc_rescale(params, prob, A_v, s.c, s.lv, s.uv, s.lc, s.uc, D_r, D_c, &s.beta, &s.gamma);
  double t_rescale = wall_now() - t_scale0;
  s.A_rp = prob->constraint_matrix_row_pointers;
  s.A_ci = prob->constraint_matrix_col_indices;
  s.A_v = A_v;
  s.D_r = D_r;
  s.D_c = D_c;
  s.c0 = prob->objective_constant;

  /* pdlp_core_op.cu:93-130 */
  double t_tr0 = wall_now();
  int *At_rp = (int *)xmalloc((size_t)(n + 1) * sizeof(int));
  int *At_ci = (int *)xmalloc((size_t)(nnz ? nnz : 1) * sizeof(int));
  double *At_v = dalloc(nnz);
  orc_transpose(m, n, s.A_rp, s.A_ci, A_v, At_rp, At_ci, At_v);
  s.At_rp = At_rp;
  s.At_ci = At_ci;
  s.At_v = At_v;

  alloc_iterates(&s);
  /* warm start, pdlp_core_op.cu:174-201 */
  if (prob->primal_start) {
    for (int j = 0; j < n; ++j) {
      double v = prob->primal_start[j] * D_c[j] * s.beta;
      s.x_init[j] = s.x_cur[j] = s.x_pdhg[j] = v;
    }
  }
  if (prob->dual_start) {
    for (int i = 0; i < m; ++i) {
      double v = prob->dual_start[i] * D_r[i] * s.gamma;
      s.y_init[i] = s.y_cur[i] = s.y_pdhg[i] = v;
    }
  }
  s.lc_fin = dalloc(m); s.uc_fin = dalloc(m); s.lv_fin = dalloc(n); s.uv_fin = dalloc(n);
  clampf_bounds(m, s.lc, s.lc_fin);
  clampf_bounds(m, s.uc, s.uc_fin);
  clampf_bounds(n, s.lv, s.lv_fin);
  clampf_bounds(n, s.uv, s.uv_fin);

  /* norms of the unscaled problem, pdlp_core_op.cu:232-286 */
  {
    double ssq = 0.0, mx = 0.0;
    for (int j = 0; j < n; ++j) {
      double v = prob->objective_vector[j];
      if (params->optimality_norm == NORM_TYPE_L_INF) {
        if (fabs(v) > mx) mx = fabs(v);
      } else {
        ssq += v * v;
      }
    }
    s.obj_norm = (params->optimality_norm == NORM_TYPE_L_INF) ? mx : sqrt(ssq);
    ssq = 0.0;
    mx = 0.0;
    for (int i = 0; i < m; ++i) {
      double lo = prob->constraint_lower_bound[i], up = prob->constraint_upper_bound[i];
      if (params->optimality_norm == NORM_TYPE_L_INF) {
        if (isfinite(lo) && lo != up && fabs(lo) > mx) mx = fabs(lo);
        if (isfinite(up) && fabs(up) > mx) mx = fabs(up);
      } else {
        if (isfinite(lo) && lo != up) ssq += lo * lo;
        if (isfinite(up)) ssq += up * up;
      }
    }
    s.bound_norm = (params->optimality_norm == NORM_TYPE_L_INF) ? mx : sqrt(ssq);
  }
  s.best_gap = INFINITY;
  s.fpe_last = INFINITY;
  s.reason = TERMINATION_REASON_UNSPECIFIED;

  double t_pi0 = wall_now();
  g_timings[0] = t_rescale;
  g_timings[1] = t_pi0 - t_tr0;
  /* pdlp_core_op.cu:638-656 */
  if (nnz == 0) {
    s.step = 1.0;
  } else {
    double sv = orc_power_iteration(m, n, s.A_rp, s.A_ci, A_v, At_rp, At_ci, At_v,
                                    params->sv_max_iter, params->sv_tol, NULL);
    s.step = 0.998 / sv;
  }
  s.omega = params->bound_objective_rescaling ? 1.0 : (s.obj_norm + 1.0) / (s.bound_norm + 1.0);
  s.best_omega = s.omega;

  /* main loop, solver.cu:56-102 */
  double t0 = wall_now();
  g_timings[2] = t0 - t_pi0;
  int do_restart = 0;
  int freq = params->termination_evaluation_frequency;
  while (s.total < params->termination_criteria.iteration_limit) {
    trace_record(&s, trace);
    if (s.is_major || s.total == 0 || s.total % orc_get_print_frequency(s.total) == 0) {
      compute_residual(&s, params->optimality_norm);
      if (s.is_major && s.total < 3 * freq) compute_infeasibility(&s);
      s.t_cum = wall_now() - t0;
      check_termination(&s, &params->termination_criteria);
      if (s.reason != TERMINATION_REASON_UNSPECIFIED) break;
    }
    if (s.is_major || s.total == 0) {
      do_restart = should_restart(&s, &params->restart_params, freq);
      if (do_restart) perform_restart(&s, &params->restart_params);
    }
    s.is_major = ((s.total + 1) % freq) == 0;
    pdhg_primal(&s);
    pdhg_dual(&s);
    if (s.is_major || do_restart) {
      fixed_point_error(&s);
      if (do_restart) {
        s.fpe0 = s.fpe;
        do_restart = 0;
      }
    }
    halpern(&s, params->reflection_coefficient);
    s.inner++;
    s.total++;
  }
  if (s.reason == TERMINATION_REASON_UNSPECIFIED) {
    s.reason = TERMINATION_REASON_ITERATION_LIMIT;
    compute_residual(&s, params->optimality_norm);
  }
  g_timings[3] = wall_now() - t0;
  double polish_time = 0.0;
  int polish_iters = 0;
  if (params->feasibility_polishing && s.reason != TERMINATION_REASON_DUAL_INFEASIBLE &&
      s.reason != TERMINATION_REASON_PRIMAL_INFEASIBLE)
    feasibility_polish(params, &s, &polish_time, &polish_iters);
  g_timings[4] = polish_time;

  /* create_result_from_state, pdlp_core_op.cu:789-859 */
  cupdlpx_result_t *r = (cupdlpx_result_t *)calloc(1, sizeof(cupdlpx_result_t));
  orc_spmv(n, At_rp, At_ci, At_v, s.y_pdhg, s.dual_product);
  r->primal_solution = dalloc(n);
  r->dual_solution = dalloc(m);
  r->reduced_cost = dalloc(n);
  for (int j = 0; j < n; ++j) {
    r->reduced_cost[j] = (s.c[j] - s.dual_product[j]) * D_c[j] / s.gamma;
    r->primal_solution[j] = s.x_pdhg[j] / D_c[j] / s.beta;
  }
  for (int i = 0; i < m; ++i) r->dual_solution[i] = s.y_pdhg[i] / D_r[i] / s.gamma;
  r->num_variables = n;
  r->num_constraints = m;
  r->num_nonzeros = nnz;
  r->total_count = s.total;
  r->rescaling_time_sec = t_rescale;
  r->cumulative_time_sec = s.t_cum;
  r->absolute_primal_residual = s.abs_p;
  r->relative_primal_residual = s.rel_p;
  r->absolute_dual_residual = s.abs_d;
  r->relative_dual_residual = s.rel_d;
  r->primal_objective_value = s.pobj;
  r->dual_objective_value = s.dobj;
  r->objective_gap = s.gap;
  r->relative_objective_gap = s.rel_gap;
  r->max_primal_ray_infeasibility = s.max_p_ray_inf;
  r->max_dual_ray_infeasibility = s.max_d_ray_inf;
  r->primal_ray_linear_objective = s.p_ray_lin_obj;
  r->dual_ray_objective = s.d_ray_obj;
  r->termination_reason = s.reason;
  r->feasibility_polishing_time = polish_time;
  r->feasibility_iteration = polish_iters;

  free_iterates(&s);
  free(s.lc_fin); free(s.uc_fin); free(s.lv_fin); free(s.uv_fin);
  free(s.c); free(s.lv); free(s.uv); free(s.lc); free(s.uc);
  free(A_v); free(D_r); free(D_c); free(At_rp); free(At_ci); free(At_v);
  return r;
}

void orc_result_free(cupdlpx_result_t *r) {
  if (!r) return;
  free(r->primal_solution);
  free(r->dual_solution);
  free(r->reduced_cost);
  free(r);
}
