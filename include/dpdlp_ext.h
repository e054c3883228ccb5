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
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DPDLP_PREC_FP64 = 0, DPDLP_PREC_MIXED = 1 };

typedef struct {
  int precision;           /* DPDLP_PREC_FP64 (default) or DPDLP_PREC_MIXED */
  double mixed_switch_tol; /* mixed: iterate in FP64 once max(rel_p, rel_d,
                              rel_gap) < this (default 1e-6), or once FP32
                              progress stalls */
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

/* Upload with a fixed row-block capacity (1024, 2048, 4096; 0 = by size) and
 * column-panel counts for A and A' (1, 2, 4, 8; 0 = by size).  Batches need
 * panels 1, 1 and the batch's capacity. */
dpdlp_resident_t *dpdlp_upload_ex(const lp_problem_t *prob, int spmv_cap, int panels_A,
                                  int panels_At);
/* dpdlp_upload_ex with a row-block capacity of its own for A' (1024, 2048,
 * 4096; 0 = the rule dpdlp_upload_ex applies).  A batch whose spmv_cap_At
 * differs from its spmv_cap takes LPs uploaded this way. */
dpdlp_resident_t *dpdlp_upload_caps(const lp_problem_t *prob, int spmv_cap, int spmv_cap_At,
                                    int panels_A, int panels_At);
/* Device bytes a resident problem holds (matrix, transpose, plans, vectors). */
size_t dpdlp_resident_bytes(const dpdlp_resident_t *res);

/* Last error message of this thread's most recent failing call ("" if none). */
const char *dpdlp_last_error(void);

/* Writes `size` values to `path`, one "%.10g" per line: the solution files of
 * the command line (replaces save_vector, src/cli.c:59-79).  The text is
 * formatted by host worker threads and written in order.  0 or -1. */
int dpdlp_write_vector(const char *path, const double *data, int size);

/* Time one standalone FP64 SpMV (A x or A' y) on a resident problem with HIP
 * events: returns the mean kernel time in ms over `reps` launches. */
double dpdlp_time_spmv(dpdlp_resident_t *res, int transpose, int reps, int use32);
/* y = A x (transpose = 0) or y = A' x (transpose = 1) on a resident problem's
 * unscaled matrix, with host input and output (for parity tests). */
int dpdlp_resident_spmv(dpdlp_resident_t *res, int transpose, const double *x, double *y,
                        int use32);
/* The device descriptor (a dpdlp_csr_t of include/dpdlp_dev.h) of a
 * resident problem's unscaled A (transpose = 0) or A' (1), with no scratch
 * (xpart and tick NULL), for callers of the dpdlp_dev_* SpMVs.  0 or -1. */
int dpdlp_resident_device_csr(const dpdlp_resident_t *res, int transpose, void *csr_out);
/* dpdlp_resident_spmv run `reps` times on the same scratch (the tiled SpMV's
 * tickets advance across the launches), the last result returned;
 * two_kernel = 1 runs a tiled matrix with the two-kernel form instead of the
 * one-pass kernel. */
int dpdlp_resident_spmv_ex(dpdlp_resident_t *res, int transpose, const double *x, double *y,
                           int use32, int reps, int two_kernel);

/* The device scaling of a solve (Ruiz, Pock-Chambolle, bound/objective as
 * params ask) of p in its given row order, copied back: scaled A values (CSR
 * order), c, l_v, u_v, l_c, u_c, D_r, D_c and {beta, gamma}; for tests. */
int dpdlp_scaling_host(const lp_problem_t *p, const pdhg_parameters_t *params, double *val,
                       double *c, double *lv, double *uv, double *lc, double *uc, double *D_r,
                       double *D_c, double *beta_gamma);

/* One dpdlp_dev_reduce (op DPDLP_RED_*) of host vectors of length n into
 * *out (y is read for the dot product only); for tests. */
int dpdlp_reduce_host(int op, const double *x, const double *y, long long n, double *out);

/* Rate (G gathers/s) of `gathers` random 8-byte reads with the SpMV's access
 * pattern (coalesced index stream, `chunk` indices per workgroup, eight
 * gathers in flight per thread).  xcd_local: every workgroup reads only the
 * slice_bytes slice of its own XCD (an 8 * slice_bytes vector, each XCD's L2
 * holding its slice when it fits); else all workgroups read one slice_bytes
 * vector.  The ceiling of an SpMV's gathers at that working-set size. */
double dpdlp_gather_calibration_ex(long long slice_bytes, long long gathers, int reps,
                                   int xcd_local, int chunk);
/* The same with the FP64 value of each gather streamed beside its index
 * when with_values (the full 12-byte-per-nonzero stream of an FP64 SpMV, so
 * the rate is the ceiling of a row-block pass that also reads its values). */
double dpdlp_gather_calibration_values(long long slice_bytes, long long gathers, int reps,
                                       int xcd_local, int chunk, int with_values);
/* dpdlp_gather_calibration_ex(slice_bytes, gathers, reps, 1, 2048) */
double dpdlp_gather_calibration(long long slice_bytes, long long gathers, int reps);

/* Copies `bytes` device-to-device with a plain kernel `reps` times; returns the
 * mean kernel time in ms (reference kernel for counter calibration). */
double dpdlp_copy_calibration(long long bytes, int reps);

/* Copies a resident problem's unscaled A' (CSR: n+1 row pointers, nnz row
 * indices of A, nnz values) to the host, in its stored row order; -1 when
 * A' is stored as column-panel tiles only (for transpose tests). */
int dpdlp_resident_transpose(const dpdlp_resident_t *res, int *t_rp, int *t_ci, double *t_v);

/* Problem sizes of a resident problem (m, n, nnz, row blocks of A and A'). */
void dpdlp_resident_info(const dpdlp_resident_t *res, int *m, int *n, int *nnz,
                         int *nblk_A, int *nblk_At);
/* The layout a resident problem was uploaded with: column-panel counts and
 * row-block capacities of A and A', and whether its rows were renumbered by
 * length (any output may be NULL). */
void dpdlp_resident_layout(const dpdlp_resident_t *res, int *panels_A, int *panels_At, int *cap_A,
                           int *cap_At, int *renumbered);

/* ---- batched solving: many LPs advancing together ---- */

/* Every LP of a batch runs the single-LP algorithm unchanged (same scaling,
 * step size, restart, termination and polishing; same iterates bit for bit
 * whichever other LPs share the batch); the batch launches each kernel kind
 * once per iteration for all LPs in flight.  LPs join and leave at window
 * boundaries (every gcd(10, termination_evaluation_frequency) iterations). */
typedef struct {
  int max_lps;             /* slot capacity (default 1024) */
  int precision;           /* DPDLP_PREC_FP64 (default) or DPDLP_PREC_MIXED */
  double mixed_switch_tol; /* mixed: FP64 from this max relative KKT on
                              (default 1e-6), or once FP32 progress stalls */
  int spmv_cap;            /* row-block capacity of every LP (default 4096) */
  int profile_phases;      /* 1: HIP events around the two phase kernels of
                              the first iteration of every 4th batched
                              window, power iteration beside it included
                              (dpdlp_batch_phase_stats) */
  int use_graph;           /* 1: capture each window's launches in a hipGraph
                              and replay it while the batch keeps its shape */
  int solo_lps;            /* windows with at most this many iterating LPs use
                              their single-LP launches (default 4; 0: never) */
  int spmv_cap_At;         /* row-block capacity of every LP's A' (0: the
                              same as spmv_cap) */
  int persist_max_nnz;     /* LPs with 1..persist_max_nnz nonzeros run their
                              whole solve (power iteration and PDHG loop, FP64)
                              in the persistent kernel, one workgroup each,
                              beside the windows of the others (0: none; the
                              default) */
  int persist_budget;      /* iterations per LP and persistent launch: an LP
                              that finished retires when its launch ends
                              (default 250) */
  /* outputs of dpdlp_solve_batch */
  long long windows, syncs, launches, graph_replays;
  int phase_samples;
  double primal_ms, primal_bytes, dual_ms, dual_bytes;
} dpdlp_batch_opts_t;

void dpdlp_batch_default_opts(dpdlp_batch_opts_t *o);

typedef struct dpdlp_batch dpdlp_batch_t;

/* params: shared by every LP of the batch (copied). */
dpdlp_batch_t *dpdlp_batch_create(const pdhg_parameters_t *params, const dpdlp_batch_opts_t *o);
/* Sets up an LP (uploaded by dpdlp_upload_ex(prob, cap, 1, 1)) in a free slot;
 * it starts at the next dpdlp_batch_step.  take_ownership: the batch frees
 * res when the LP finishes.  Returns 0, or -1 (batch full, wrong layout, no
 * memory). */
int dpdlp_batch_add(dpdlp_batch_t *b, dpdlp_resident_t *res, int take_ownership, long long tag);
/* Runs one window of iterations for every LP in flight; finished LPs' results
 * become available to dpdlp_batch_take.  Returns the LPs still in flight, or
 * -1 on a device error. */
int dpdlp_batch_step(dpdlp_batch_t *b);
/* Pops one finished LP: returns 1 and sets *tag, *res (caller frees with
 * cupdlpx_result_free) and *latency_sec (add to result), 0 when none waits. */
int dpdlp_batch_take(dpdlp_batch_t *b, long long *tag, cupdlpx_result_t **res,
                     double *latency_sec);
int dpdlp_batch_in_flight(const dpdlp_batch_t *b);
/* Device bytes the batch holds for LP state beyond the resident problems. */
size_t dpdlp_batch_bytes(const dpdlp_batch_t *b);
/* Counters: windows run, host-device synchronisations, kernel launches
 * (captured ones included), windows replayed from a graph. */
void dpdlp_batch_counters(const dpdlp_batch_t *b, long long *windows, long long *syncs,
                          long long *launches);
long long dpdlp_batch_graph_replays(const dpdlp_batch_t *b);
/* Live phase timing (profile_phases): sampled launches, their summed kernel
 * time (ms) and summed algorithmic bytes, for the primal (A'y) and dual (Ax)
 * phase kernels. */
void dpdlp_batch_phase_stats(const dpdlp_batch_t *b, int *samples, double *primal_ms,
                             double *primal_bytes, double *dual_ms, double *dual_bytes);
void dpdlp_batch_free(dpdlp_batch_t *b);

/* Solves `count` resident problems in one batch; results[i] belongs to
 * res[i].  The LPs are set up on a second host thread and stream while the
 * batch already steps.  o (optional) also receives the counters and phase
 * timing.  Returns the number of results (count) or -1. */
int dpdlp_solve_batch(dpdlp_resident_t *const *res, int count, const pdhg_parameters_t *params,
                      dpdlp_batch_opts_t *o, cupdlpx_result_t **results);

/* ---- batch layout autotuner ---- */

/* SpMV row-block layout of one batched LP (untiled A and A'): row-block
 * capacity (1024, 2048 or 4096 nonzeros) and most rows per block (1..256;
 * the row reduction gives each row of a block of R rows a group of 256 / R
 * lanes, rounded down to a power of two, at most 64), for A and for A'. */
typedef struct {
  int cap_A, rows_A, cap_At, rows_At;
} dpdlp_batch_layout_t;

/* Uploads an LP for a batch with this layout.  A batch takes LPs whose
 * capacities are at most its spmv_cap (A) and spmv_cap_At (A'). */
dpdlp_resident_t *dpdlp_upload_batch_layout(const lp_problem_t *prob,
                                            const dpdlp_batch_layout_t *layout);

/* The autotuner's size class of an LP: its nonzeros in factors of 4 and the
 * mean row lengths of A and A' in factors of 2, as one integer. */
int dpdlp_batch_size_class(const lp_problem_t *prob);

typedef struct {
  int classes;    /* size classes among the LPs */
  int tuned;      /* classes timed by this call */
  int cached;     /* classes read from the cache */
  double seconds; /* time the call took */
} dpdlp_autotune_report_t;

/* Picks a layout for every LP: per size class, the capacity and rows per
 * block (1024/2048/4096 x 64/128/256) for which the batch's own dual (A x)
 * and primal (A' y) phase kernels run fastest on up to 16 members of the
 * class solved as one batch for 200 iterations (params otherwise; precision
 * as the batch will run).  Classes found in cache_path (a text file; NULL:
 * no cache) for this device are not timed again; timed ones are appended.
 * layouts[i] receives LP i's layout.  Returns 0 or -1. */
int dpdlp_batch_autotune(const lp_problem_t *const *probs, int count,
                         const pdhg_parameters_t *params, int precision, const char *cache_path,
                         dpdlp_batch_layout_t *layouts, dpdlp_autotune_report_t *report);

/* ---- serving: a Poisson request stream with continuous batching ---- */

typedef struct {
  double rate;          /* arrivals per second (Poisson process) */
  double duration_sec;  /* requests arrive during [0, duration) */
  double drain_sec;     /* the run stops at duration + drain at the latest */
  uint64_t seed;        /* arrival gaps and request draws */
  int pool_small;       /* distinct uneven LPs: batch workload members 0.. */
  int pool_large;       /* distinct large LPs */
  double large_frac;    /* share of requests drawn from the large pool */
  int large_m;          /* large LP rows (columns m * U[1.5, 3]) */
  long long large_nnz;  /* large LP nonzeros (Poisson rows of that mean) */
  int m_lo, m_hi;       /* uneven LPs: m log-uniform in [m_lo, m_hi] */
  uint64_t member_seed; /* LP seeds */
  double hbm_budget_gb; /* admission: device bytes of the LPs in flight */
  int max_lps;          /* admission: LPs in flight */
  int precision;        /* batch precision (DPDLP_PREC_MIXED default) */
  int mode;             /* 0 continuous batching, 1 plain loop (one at a time, FP64) */
  int spmv_cap;         /* row-block capacity (default 4096) */
  long long large_lane_nnz; /* requests with at least this many nonzeros are
                               solved by the single-LP path on a second
                               thread and stream, beside the batch (0: off) */
  int persist_max_nnz;  /* requests with at most this many nonzeros run whole
                           in the batch's persistent kernel (the batch option
                           of the same name; 0: off, the default) */
} dpdlp_serve_spec_t;

typedef struct {
  int arrived, completed, optimal, unfinished;
  double makespan_sec;  /* clock start to last completion */
  double lps_per_sec;   /* completed / makespan */
  double p50_sec, p90_sec, p99_sec, mean_sec, max_sec; /* completion - arrival;
                           unfinished requests count as +inf in the percentiles */
  int max_in_flight;
  double peak_gb;       /* admitted device bytes, peak */
  double pool_gen_sec;  /* host generation of the request pool (not timed) */
  long long windows, syncs;
} dpdlp_serve_report_t;

void dpdlp_serve_default_spec(dpdlp_serve_spec_t *s);
/* Runs one request stream.  Optional outputs, one entry per arrived request
 * (at most max_requests when > 0): latency, pool index, result (caller frees
 * each with cupdlpx_result_free; NULL when unfinished).  Returns 0 or -1. */
int dpdlp_serve_run(const dpdlp_serve_spec_t *spec, const pdhg_parameters_t *params,
                    dpdlp_serve_report_t *rep, double *latency_out, int *pool_index_out,
                    cupdlpx_result_t **results_out, int max_requests);

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
/* Member k of the uneven batch workload: m log-uniform in [m_lo, m_hi],
 * n = m * U[1.5, 3], Poisson(8)+1 nonzeros per row, LP seed base_seed + k. */
void dpdlp_batch_member_spec(int k, uint64_t base_seed, int m_lo, int m_hi,
                             dpdlp_gen_spec_t *out);

#ifdef __cplusplus
}
#endif
