/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement, in plain FP64 C, of the reference's single-GPU PDHG solve
 * path (src/solver.cu:32-127 and everything it calls).  It is the checker the
 * GPU library is compared with, and (built with -fopenmp) the CPU baseline
 * that bench.py times.  It is never linked into or called by the product
 * library; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.
 *
 * Built twice by oracle/Makefile:
 *   liborcpdlp.so      single thread, -ffp-contract=off  (parity reference)
 *   liborcpdlp_omp.so  OpenMP over rows/columns          (CPU baseline)
 *
 * Pinning (what each function is checked against; see tests/test_oracle_*.py):
 *   orc_mt_normal_fill   std::mt19937(1) + std::normal_distribution<double>
 *                        output produced by a g++-compiled program
 *   orc_transpose        dense brute force, double transpose = identity
 *   orc_spmv             dense brute force, adjointness y'(Ax) = x'(A'y)
 *   orc_nrm2 / dot /
 *   orc_amax / sum       long-double / math.fsum references, 1e200 overflow
 *   orc_rescale          closed forms (diagonal, rank one), Ruiz limit
 *   orc_power_iteration  numpy SVD on small matrices, diagonal closed form
 *   orc_solve            docs/C_API.md example (x=(1,2), y=(1,-1,0), obj 3),
 *                        generated LPs with a known optimum, scipy HiGHS,
 *                        infeasible / unbounded LPs, iteration/time limits
 */
#pragma once

#include "cupdlpx_types.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Iterate trace: x_cur / y_cur (scaled space) recorded at the top of the main
 * loop when total_count equals one of iters[] (ascending). */
typedef struct {
  int num_points;
  const int *iters;
  double *x_cur;   /* num_points * n */
  double *y_cur;   /* num_points * m */
  double *scalars; /* num_points * 4: step_size, primal_weight, inner, fpe */
  int recorded;    /* out: how many points were reached */
} orc_trace_t;

/* Full solve: rescale, transpose, power iteration, restarted Halpern PDHG,
 * optional feasibility polish, unscaled result.  Follows solver.cu:32-127. */
cupdlpx_result_t *orc_solve(const lp_problem_t *prob,
                            const pdhg_parameters_t *params,
                            orc_trace_t *trace);
void orc_result_free(cupdlpx_result_t *r);
/* Wall times (s) of the last orc_solve in this process: rescale, transpose,
 * power iteration, main loop, polish. */
void orc_last_timings(double *out5);
void orc_default_parameters(pdhg_parameters_t *p);
int orc_num_threads(void);

/* ---- components, exposed for unit tests ---- */
int orc_get_print_frequency(int iter);
void orc_transpose(int m, int n, const int *rp, const int *ci, const double *v,
                   int *t_rp, int *t_ci, double *t_v);
void orc_spmv(int rows, const int *rp, const int *ci, const double *v,
              const double *x, double *y);
double orc_nrm2(int n, const double *x);
double orc_dot(int n, const double *x, const double *y);
double orc_amax(int n, const double *x);
double orc_sum(int n, const double *x);

/* Rescales a copy: scaled CSR values and bounds/objective written into the
 * caller's arrays; D_r[m], D_c[n], beta, gamma as in preconditioner.c. */
void orc_rescale(const pdhg_parameters_t *params, const lp_problem_t *prob,
                 double *A_val_s, double *c_s, double *lv_s, double *uv_s,
                 double *lc_s, double *uc_s, double *D_r, double *D_c,
                 double *beta, double *gamma);

/* mt19937 seeded with `seed` followed by `count` draws of the polar normal. */
void orc_mt_normal_fill(uint32_t seed, int count, double *out);

/* Power iteration on A A' (utils.cu:66-171); returns sigma_max and the number
 * of iterations run in *iters_out. */
double orc_power_iteration(int m, int n, const int *rp, const int *ci,
                           const double *v, const int *t_rp, const int *t_ci,
                           const double *t_v, int max_iter, double tol,
                           int *iters_out);

#ifdef __cplusplus
}
#endif
