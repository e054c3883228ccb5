/*
 * Extensions beyond the reference C API (no counterpart in the reference):
 * solve options (precision mode, SpMV block size, iterate trace), problems kept
 * resident in HBM across solves, the seeded LP generators of the benchmark
 * workloads, and the batched / served solving entry points.
 *
 * These never change the behaviour of the six cupdlpx.h functions.
 */
#pragma once

#include "cupdlpx_types.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DPDLP_PREC_FP64 = 0, DPDLP_PREC_MIXED = 1 };

typedef struct {
  int precision;           /* DPDLP_PREC_FP64 (default) or DPDLP_PREC_MIXED */
  double mixed_switch_tol; /* mixed: iterate in FP64 once max(rel_p, rel_d,
                              rel_gap) < this (default 1e-5 * 10) */
  int spmv_cap;            /* 0 = automatic, else 1024 / 2048 / 4096 */
  int quiet;               /* 1: no "Solution Summary" print */
  int use_graph;           /* 1: replay the inner loop from a hipGraph */
  /* iterate trace: scaled x_cur / y_cur at the loop top of these iterations */
  int trace_points;
  const int *trace_iters;
  double *trace_x; /* trace_points * n */
  double *trace_y; /* trace_points * m */
  int trace_recorded;
  /* outputs */
  double setup_time_sec;   /* upload + transpose + plans + scaling */
  double power_time_sec;   /* power iteration */
  double loop_time_sec;    /* main PDHG loop (the reference's solve time) */
  double total_time_sec;   /* whole call */
  int power_iterations;
  double step_size;        /* eta after the power iteration */
  int iterations_fp32;     /* mixed mode: iterations run with FP32 SpMV */
  double time_to_1e4_sec;  /* loop time when max rel KKT first < 1e-4 (or -1) */
  double time_to_1e6_sec;  /* same for 1e-6 (or -1) */
  /* live kernel timing: HIP events around the two fused phase kernels of
   * every 8th iteration of the main loop, on the solve's stream */
  int profile_phases;
  double primal_phase_ms;  /* mean duration of phase 1 (A'y + primal update) */
  double dual_phase_ms;    /* mean duration of phase 2 (A x + dual update) */
  int phase_samples;
} dpdlp_solve_opts_t;

void dpdlp_default_opts(dpdlp_solve_opts_t *o);

/* Solve with options; same result contract as solve_lp_problem. */
cupdlpx_result_t *dpdlp_solve_ex(const lp_problem_t *prob,
                                 const pdhg_parameters_t *params,
                                 dpdlp_solve_opts_t *opts);

/* A problem uploaded once (unscaled A, its transpose, bounds, objective) and
 * solved many times without host-to-device traffic. */
typedef struct dpdlp_resident dpdlp_resident_t;
dpdlp_resident_t *dpdlp_upload(const lp_problem_t *prob, int spmv_cap);
/* Same, choosing the XCD column-panel count of A and A' (1, 2, 4 or 8) by
 * timing each candidate SpMV on the device; chosen[0..1] receives them. */
dpdlp_resident_t *dpdlp_upload_tuned(const lp_problem_t *prob, int spmv_cap, int *chosen);
cupdlpx_result_t *dpdlp_solve_resident(dpdlp_resident_t *res,
                                       const pdhg_parameters_t *params,
                                       dpdlp_solve_opts_t *opts);
void dpdlp_resident_free(dpdlp_resident_t *res);

/* Last error message of this thread's most recent failing call ("" if none). */
const char *dpdlp_last_error(void);

/* Time one standalone FP64 SpMV (A x or A' y) on a resident problem with HIP
 * events: returns the mean kernel time in ms over `reps` launches. */
double dpdlp_time_spmv(dpdlp_resident_t *res, int transpose, int reps, int use32);
/* y = A x (transpose = 0) or y = A' x (transpose = 1) on a resident problem's
 * unscaled matrix, with host input and output (for parity tests). */
int dpdlp_resident_spmv(dpdlp_resident_t *res, int transpose, const double *x, double *y,
                        int use32);

/* Copies `bytes` device-to-device with a plain kernel `reps` times; returns the
 * mean kernel time in ms (reference kernel for counter calibration). */
double dpdlp_copy_calibration(long long bytes, int reps);

/* Problem sizes of a resident problem (m, n, nnz, row blocks of A and A'). */
void dpdlp_resident_info(const dpdlp_resident_t *res, int *m, int *n, int *nnz,
                         int *nblk_A, int *nblk_At);

/* ---- seeded synthetic LPs (feasible by construction) ---- */
typedef struct {
  int m, n;
  long long nnz;     /* exact nonzero count (row-length rule) */
  int row_rule;      /* 0 uniform positions, 1 Zipf(1.5) rescaled, 2 Poisson(lambda)+1 */
  double zipf_s;     /* 1.5 */
  int zipf_max;      /* 50000 */
  double poisson_lambda;
  double pm1_frac;   /* fraction of values replaced by +-1 */
  uint64_t seed;
} dpdlp_gen_spec_t;

/* Generates an lp_problem_t (free with lp_problem_free) and, optionally, the
 * known optimal objective c'x*. */
lp_problem_t *dpdlp_generate(const dpdlp_gen_spec_t *spec, double *opt_obj);

#ifdef __cplusplus
}
#endif
