/*
 * Thin C-ABI between the host solver (C) and the hand-written gfx950 kernels.
 *
 * Every function takes device pointers held in plain C structs, sizes, and a
 * HIP stream passed as void*; none allocates; each returns 0 on success or a
 * hipError_t value.  Reductions write raw partial results into device scalar
 * slots (dpdlp_state_t.scal) so that a whole evaluation needs one
 * device-to-host copy.  The comment on each entry point names the reference
 * call sites (cuSPARSE / cuBLAS / CUDA kernels) it replaces.
 */
#pragma once

#include <stddef.h>

#include "cupdlpx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CSR matrix resident in HBM, with its row-block plan for the stream SpMV.
 *
 * When P > 1 the nonzeros are also stored as XCD column-panel tiles: the
 * columns are cut into P equal panels, and the entries of row block b that
 * fall in panel k form tile b*P + k (entries in row order, columns global).
 * Workgroup w of the tiled SpMV handles tile w; the hardware dispatches
 * workgroups round-robin over the 8 XCDs, so every XCD gathers from only one
 * panel of the vector, which stays resident in its 4 MB L2.  Per-panel row
 * sums go to xpart[k*rows + row]; the last of a row block's P workgroups to
 * finish (counted on tick[b]) adds the P partials in panel order and applies
 * the epilogue.  Without tickets a second kernel does that. */
typedef struct {
  int rows, cols, nnz;
  const int *rp;     /* [rows+1] */
  const int *ci;     /* [nnz] */
  const double *v;   /* [nnz] FP64 values */
  const float *v32;  /* [nnz] FP32 copy or NULL */
  const int *blk;    /* [nblk+1] first row of each row block */
  int nblk;
  int cap;           /* nonzeros a multi-row block may hold (1024/2048/4096) */
  /* column-panel tiles (P == 1: not used) */
  int P;
  const int *toff;            /* [nblk*P+1] first entry of each tile */
  const unsigned short *rofs; /* tile-local row offsets, P*(R+1) per block */
  const int *pci;             /* [nnz] column of each tiled entry */
  const double *pv;           /* [nnz] tiled FP64 values */
  const float *pv32;          /* [nnz] tiled FP32 values or NULL */
  double *xpart;              /* [P*rows] partial row sums (scratch) */
  unsigned *tick;             /* [nblk] last-arriver tickets of the one-pass
                                 tiled SpMV (zeroed once, never reset), or NULL
                                 for the two-kernel form */
} dpdlp_csr_t;

/* One LP's device state (scaled space).  Field meaning follows the
 * reference's pdhg_solver_state_t (internal/internal_types.h:50-145). */
typedef struct {
  int m, n;
  dpdlp_csr_t A, At;
  double *lv, *uv, *c, *lc, *uc; /* scaled bounds and objective */
  double *D_r, *D_c;             /* cumulative row / column scaling */
  double *x_init, *x_cur, *x_pdhg, *x_ref, *dual_product, *dual_slack, *dx;
  double *y_init, *y_cur, *y_pdhg, *y_ref, *primal_product, *primal_slack, *dy;
  float *x_ref32, *y_cur32; /* FP32 gather shadows (mixed precision) or NULL */
  /* Anchored shadows (mixed precision, single-LP solves): A x_init and
   * A' y_init in FP64, refreshed whenever the anchors move; the shadows then
   * hold x_ref - x_init and y_cur - y_init, and each FP32 product adds its
   * anchor's.  NULL: the shadows hold x_ref and y_cur. */
  double *ax0, *aty0;
  double *part;   /* partial sums: DPDLP_SLOTS x part_stride */
  int part_stride;
  double *scal;   /* DPDLP_SCAL_LEN doubles of raw reduction results */
} dpdlp_state_t;

/* Workgroups (of 256 lanes) of every grid-stride element-wise launch
 * over `count` items, at most 2048; also the number of partial sums such a
 * launch's reduction writes. */
static inline int dpdlp_ew_grid(long long count) {
  long long g = (count + 255) / 256;
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  return (int)g;
}

/* ---- batched solving: many LPs per launch ---- */

/* One LP of a batch (device pointers; scal is that LP's slot array). */
typedef struct {
  int m, n;
  const int *A_rp, *A_ci, *At_rp, *At_ci;
  const double *A_v, *At_v;
  const float *A_v32, *At_v32;
  double *lv, *uv, *c, *lc, *uc, *D_r, *D_c;
  double *x_init, *x_cur, *x_pdhg, *x_ref, *dual_product, *dual_slack, *dx;
  double *y_init, *y_cur, *y_pdhg, *y_ref, *primal_product, *primal_slack, *dy;
  float *x_ref32, *y_cur32;
  double *pi_eig, *pi_next, *pi_dual;
  double *scal;
  int cap_A, cap_At; /* the LP's row-block capacities: a single-row block
                        longer than this takes the whole-workgroup sum */
} dpdlp_blp_t;

/* Per-LP control of one batched step, written by the host. */
typedef struct {
  double tau, sigma, w, rho;
  int phase;        /* run the PDHG phases */
  int major;        /* major variant: write x_pdhg, dual_slack, y_pdhg */
  int fpe;          /* write dx, dy and finish the fixed-point error */
  int fpe_to;       /* bit 0: copy FPE slots to the restart group, bit 1: major group */
  int eval;         /* KKT residual */
  int restart_dist; /* add restart distances to the residual */
  int inf;          /* infeasibility rays */
  int restart;      /* restart copy mask: 1 x, 2 y */
  int use32;        /* FP32 matrix and gathers */
  int pi;           /* power-iteration step */
  int pi_slot;      /* history index of this power-iteration step */
  int pad;
} dpdlp_bctl_t;

/* Work item: a row block (lp, r0, r1) of one LP's matrix, or an element
 * block (lp, j, G): block j of the G grid-stride blocks of that LP. */
typedef struct {
  int lp, a, b, pad;
} dpdlp_bblk_t;

typedef struct {
  int nlp;
  const dpdlp_blp_t *lps;
  const dpdlp_bctl_t *ctl;
  const dpdlp_bblk_t *bA, *bAt, *eN, *eM, *eNM;
  int nbA, nbAt, neN, neM, neNM;
  const int *rA, *rAt, *rN, *rM, *rNM; /* per LP [lo, hi) into the lists */
  double *part;
  int stride;
  int cap;    /* row-block capacity of the A lists */
  int cap_At; /* of the A' lists */
} dpdlp_batch_dev_t;

/* Batched steps: every launch covers all LPs whose control flag is set. */
int dpdlp_bdev_restart_copy(const dpdlp_batch_dev_t *b, void *stream);
int dpdlp_bdev_phases(const dpdlp_batch_dev_t *b, void *stream);
/* the two halves of dpdlp_bdev_phases, for timing them apart */
int dpdlp_bdev_primal(const dpdlp_batch_dev_t *b, void *stream);
int dpdlp_bdev_dual(const dpdlp_batch_dev_t *b, void *stream);
int dpdlp_bdev_fpe(const dpdlp_batch_dev_t *b, void *stream);
int dpdlp_bdev_residual(const dpdlp_batch_dev_t *b, void *stream);
int dpdlp_bdev_infeasibility(const dpdlp_batch_dev_t *b, void *stream);
int dpdlp_bdev_power_step(const dpdlp_batch_dev_t *b, void *stream);

/* ---- per-LP controller state (the reference's scalar loop state) ---- */

/* The decisions of the reference loop (solver.cu:56-102) read and write
 * this: step size and primal weight with the PID terms
 * (pdlp_core_op.cu:593-619), iteration counters, fixed-point errors of the
 * restart test (utils.cu:265-293), the last evaluation's residuals and
 * objectives (utils.cu:639-719) and infeasibility rays (utils.cu:721-813),
 * and the termination reason.  The host controller (batch and single-LP
 * drivers) and the persistent kernel keep the same struct. */
typedef struct {
  double step, omega, best_omega, pw_err_sum, pw_last_err, best_gap;
  int total, inner, is_major;
  double fpe, fpe0, fpe_last;
  double abs_p, rel_p, abs_d, rel_d, pobj, dobj, gap, rel_gap;
  double max_p_ray_inf, max_d_ray_inf, p_ray_lin_obj, d_ray_obj;
  termination_reason_t reason;
  double t_cum;
  int pend_rst, pend_maj;
  double m32_best; /* mixed precision: best max relative KKT seen in FP32 */
  int m32_best_T;  /* iteration at which it was reached */
} hstate_t;

/* per-LP constants: scaling factors, objective constant, unscaled norms */
typedef struct {
  double beta, gamma, c0, obj_norm, bound_norm;
} lpk_t;

/* ---- persistent solving: one workgroup runs one LP's whole loop ---- */

/* Stages of an LP in the persistent kernel. */
enum { DPDLP_PS_PI = 0, DPDLP_PS_PDHG = 1, DPDLP_PS_DONE = 2 };

/* One LP of a persistent launch (device memory; the kernel reads and writes
 * it).  L: the LP's device arrays (untiled A and A', row-block capacity at
 * most DPDLP_PERSIST_CAP); A_blk/At_blk: its row-block plans (device);
 * pi_eig/pi_next (in L) hold the power iteration's normalised start vector
 * (dpdlp_dev_power_init) when stage is DPDLP_PS_PI.  The raw reduction slots
 * live in L.scal as in the batch.  The loop's time (h.t_cum, checked against
 * the time limit at every evaluation) is read from the device's constant-rate
 * wall clock, clk_hz ticks per second. */
typedef struct {
  dpdlp_blp_t L;
  const int *A_blk, *At_blk;
  int A_nblk, At_nblk;
  lpk_t k;
  hstate_t h;
  double sigma_sq;
  long long clk0; /* device wall clock when the PDHG loop started */
  int stage;
  int pi_iters;
  int do_restart;
  int iters_run; /* PDHG iterations and power-iteration steps of the last launch */
} dpdlp_plp_t;

/* the largest row-block capacity the persistent kernel stages in LDS */
#define DPDLP_PERSIST_CAP 4096

/* Runs up to `budget` iterations (power-iteration steps and PDHG
 * iterations together) of each of the `count` LPs in lps (device array;
 * params: the solve parameters, also in device memory),
 * one workgroup per LP, with the reference's per-LP decisions (evaluation
 * cadence, residuals, infeasibility rays, termination, restart test, PID
 * weight) made on the device: the loop of solver.cu:56-102 after the power
 * iteration of utils.cu:66-171, FP64 throughout.  An LP whose stage is
 * DPDLP_PS_DONE is skipped; one that terminates leaves h.reason set and its
 * stage DPDLP_PS_DONE.  Stream-ordered, no allocation. */
int dpdlp_pdev_run(dpdlp_plp_t *lps, int count, const pdhg_parameters_t *params,
                   double clk_hz, int budget, void *stream);

/* Scalar / partial slot map (raw values; the host combines them). */
enum {
  DPDLP_SL_RESP = 0,   /* rows: rp blue[3], psum, ydist blue[3], rp max  (8) */
  DPDLP_SL_RESD = 8,   /* cols: rd blue[3], c.x, z.x, xdist blue[3], rd max (9) */
  DPDLP_SL_FPE = 17,   /* dx blue[3], dy blue[3], <A'dy,dx>               (7) */
  DPDLP_SL_INF = 24,   /* max|dx|, max|dy|, psum, pinf, dsum, c.dx, dinf   (7) */
  DPDLP_SL_PI = 31,    /* <q,AA'q>, |AA'q| blue[3], residual blue[3]       (7) */
  DPDLP_SL_SCAL = 38,  /* scaling: bound blue[3], obj blue[3]              (6) */
  DPDLP_SL_PI2 = 44,   /* second set of residual blue[3] (alternate steps) (3) */
  DPDLP_SLOTS = 47,
  DPDLP_SC_FPE_RESTART = 48, /* copy of the 7 FPE slots from a restart step */
  DPDLP_SC_FPE_MAJOR = 56,   /* copy of the 7 FPE slots from a major step */
  DPDLP_SC_PI_HIST = 64,     /* power iteration history, 4 per iteration */
  DPDLP_PI_BATCH = 256,
  DPDLP_SCAL_LEN = 64 + 4 * 256
};

/* Row-block plan on the host: rows with more than `cap` nonzeros get a block
 * of their own, other blocks hold <= 256 rows and <= cap nonzeros.  Writes
 * blk[0..nblk] and returns nblk. */
int dpdlp_plan_blocks(int rows, const int *rp_host, int cap, int *blk_out);
/* The same with at most rmax (1..DPDLP_BLOCK_ROWS) rows per block: fewer rows
 * per block give each row a wider lane group in the row reduction
 * (256 / rows, rounded down to a power of two, at most 64 lanes). */
int dpdlp_plan_blocks_rows(int rows, const int *rp_host, int cap, int rmax, int *blk_out);
/* rows per row block of dpdlp_plan_blocks: one per thread of the SpMV's
 * 256-thread workgroup */
#define DPDLP_BLOCK_ROWS 256

/* Host construction of the column-panel tiles of a CSR matrix with a given
 * block plan: toff[nblk*P+1], rofs[P*(rows+nblk)], perm[nnz] (tiled position
 * -> CSR position) and pci[nnz].  Returns 0, or -1 if a multi-row tile would
 * exceed 65535 entries. */
int dpdlp_build_panels(int rows, int cols, const int *rp, const int *ci, const int *blk,
                       int nblk, int P, int *toff, unsigned short *rofs, int *perm, int *pci);
/* Row blocks for the tiled SpMV (<= rmax rows, every tile <= cap entries,
 * rows that do not fit alone).  rmax <= DPDLP_PANEL_RMAX.  Returns nblk. */
int dpdlp_plan_blocks_panels(int rows, int cols, const int *rp, const int *ci, int cap, int P,
                             int rmax, int *blk);
#define DPDLP_PANEL_RMAX 1024
/* The row blocks of the tiled transpose A' (n rows, columns = A's m rows)
 * planned from A itself (m x n CSR on the host), without transposing: the
 * same plan dpdlp_plan_blocks_panels gives on A''s host CSR.  Returns nblk,
 * or -1 if P is not 1..8 or host memory runs out. */
int dpdlp_plan_blocks_panels_T(int m, int n, const int *rp, const int *ci, int cap, int P,
                               int rmax, int *blk);
/* dpdlp_build_panels on the device: rp, ci (device CSR of the matrix being
 * tiled, rows x cols) and blk (a plan of dpdlp_plan_blocks_panels, blocks
 * of at most DPDLP_PANEL_RMAX rows) in device memory; the outputs likewise.
 * Stream-ordered, no allocation. */
int dpdlp_dev_build_panels(int rows, int cols, const int *rp, const int *ci, const int *blk,
                           int nblk, int P, int *toff, unsigned short *rofs, int *perm, int *pci,
                           void *stream);
/* dst[t] = src[perm[t]] (tiled copies of values after scaling). */
int dpdlp_dev_permute(const double *src, const int *perm, double *dst, long long count,
                      void *stream);
int dpdlp_dev_permute_f32(const float *src, const int *perm, float *dst, long long count,
                          void *stream);

/* y = A x (FP64), replaces cusparseSpMV CSR_ALG2 at pdlp_core_op.cu:501,534
 * and utils.cu:134,138,645,650,750,755. */
int dpdlp_dev_spmv(const dpdlp_csr_t *A, const double *x, double *y, void *stream);
/* y = A x with FP32 values and FP32 x, FP64 accumulation. */
int dpdlp_dev_spmv_f32(const dpdlp_csr_t *A, const float *x, double *y, void *stream);

/* y = A x (FP64) without allocating or taking scratch from the caller.  No
 * counterpart in the reference (whose cusparseSpMV call sites,
 * pdlp_core_op.cu:501,534, run on a buffer allocated per solve); it is for
 * callers that cannot allocate device memory (graph capture, serving loops).
 * A tiled matrix (P > 1) uses a fixed per-GPU scratch instead of A->xpart and
 * A->tick: P * rows <= DPDLP_NOALLOC_SUMS and nblk <= DPDLP_NOALLOC_TICKETS,
 * else hipErrorInvalidValue with a message in dpdlp_last_error().  Calls that
 * overlap on one GPU race on that scratch. */
#define DPDLP_NOALLOC_SUMS (4 << 20)
#define DPDLP_NOALLOC_TICKETS (1 << 16)
int dpdlp_dev_spmv_noalloc(const dpdlp_csr_t *A, const double *x, double *y, void *stream);

/* Phase 1 of one PDHG step: A'y_cur, primal step with projection and
 * reflection, Halpern update of x_cur.  Replaces the SpMV at
 * pdlp_core_op.cu:501 + kernels :393-418 + the x half of :445-462.
 * major: also write x_pdhg and dual_slack; fpe: also write dx = x_ref - x_cur
 * and its Blue partials.  use32: gather the FP32 shadow of y with FP32 values. */
int dpdlp_dev_primal_phase(const dpdlp_state_t *st, double tau, double w,
                           double rho, int major, int fpe, int use32,
                           void *stream);
/* Phase 2: A x_ref, dual step, Halpern update of y_cur.  Replaces
 * pdlp_core_op.cu:534 + kernels :420-443 + the y half of :445-462. */
int dpdlp_dev_dual_phase(const dpdlp_state_t *st, double sigma, double w,
                         double rho, int major, int fpe, int use32,
                         void *stream);
/* Finishes the fixed-point error of pdlp_core_op.cu:658-696: A'dy into
 * dual_product with <A'dy, dx> partials, then reduces the 7 FPE slots and
 * copies them to DPDLP_SC_FPE_RESTART and/or DPDLP_SC_FPE_MAJOR. */
int dpdlp_dev_fpe_finish(const dpdlp_state_t *st, int to_restart, int to_major,
                         void *stream);
/* Primal-only / dual-only FPE finish for feasibility polishing
 * (pdlp_core_op.cu:1332-1353): reduces the dx (primal) or dy slots. */
int dpdlp_dev_fpe_finish_polish(const dpdlp_state_t *st, int primal,
                                int to_restart, int to_major, void *stream);
/* KKT residual of utils.cu:639-719 (2 SpMV + residual kernel + norms/dots)
 * into DPDLP_SL_RESP / DPDLP_SL_RESD; restart_dist adds the restart distances
 * of pdlp_core_op.cu:579-591.  which: 3 = both, 1 = primal rows only
 * (utils.cu:1145-1183), 2 = dual columns only (utils.cu:1185-1229).  For
 * which==2, lc_fin/uc_fin/x_obj come from the original state `orig`. */
int dpdlp_dev_residual(const dpdlp_state_t *st, const dpdlp_state_t *orig,
                       int which, int restart_dist, void *stream);
/* Infeasibility rays of utils.cu:721-813 into DPDLP_SL_INF. */
int dpdlp_dev_infeasibility(const dpdlp_state_t *st, void *stream);
/* Restart copies of pdlp_core_op.cu:621-632 (part: 3 both, 1 x, 2 y). */
int dpdlp_dev_restart_copy(const dpdlp_state_t *st, int part, void *stream);
/* ax0 = A x_init and aty0 = A' y_init (FP64 SpMVs), and the shadow y_cur32 =
 * y_cur - y_init, for a state with anchored shadows. */
int dpdlp_dev_anchor_products(const dpdlp_state_t *st, void *stream);
/* Power iteration (utils.cu:66-171), `iters` steps starting at history index
 * 0; q0 must already be in `eig` (m).  Workspace: next (m), dual (n). */
int dpdlp_dev_power_init(const dpdlp_state_t *st, double *eig, double *next,
                         void *stream);
int dpdlp_dev_power_steps(const dpdlp_state_t *st, double *eig, double *next,
                          double *dual, int iters, void *stream);
/* Result of pdlp_core_op.cu:789-859: reduced cost into rc, unscaled x, y. */
int dpdlp_dev_result(const dpdlp_state_t *st, double beta, double gamma,
                     double *x_out, double *y_out, double *rc_out, void *stream);

/* Stable transpose on the device (csr2csc, pdlp_core_op.cu:93-130): fills
 * t_ci/t_v of A' (row pointers from the host's column counts) from A's CSR.
 * t_v's storage doubles as the sort's position buffers; scratch needs
 * dpdlp_dev_transpose_scratch_bytes(nnz) bytes. */
size_t dpdlp_dev_transpose_scratch_bytes(long long nnz);
int dpdlp_dev_transpose(int m, int n, long long nnz, const int *rp, const int *ci,
                        const double *v, int *t_ci, double *t_v, void *scratch, void *stream);

/* Scaling on device (preconditioner.c:89-273).  rs/cs are work vectors. */
int dpdlp_dev_ruiz_step(dpdlp_state_t *st, double *rs, double *cs, void *stream);
int dpdlp_dev_pock_chambolle(dpdlp_state_t *st, double alpha, double *rs,
                             double *cs, void *stream);
int dpdlp_dev_bound_obj_sums(const dpdlp_state_t *st, void *stream);
/* Bound/objective rescaling with beta and gamma formed on the device from
 * the sums dpdlp_dev_bound_obj_sums left in scal (and stored to
 * scal[DPDLP_SL_SCAL + 2], scal[DPDLP_SL_SCAL + 3]): no host round trip. */
int dpdlp_dev_bound_obj_apply_scal(dpdlp_state_t *st, void *stream);
/* FP32 copies of the scaled values (mixed precision). */
int dpdlp_dev_make_f32(const double *src, float *dst, long long count, void *stream);
/* Polishing helpers: v <- 0 where finite (pdlp_core_op.cu:1136-1143). */
int dpdlp_dev_zero_finite(double *v, int count, void *stream);
int dpdlp_dev_fill(double *v, double value, long long count, void *stream);
/* Standalone reductions into out[0] (device memory): ||x||_2 (Blue's scaled
 * sums: no overflow or underflow), x.y, max |x_i|, sum x_i; 0 for n = 0.
 * Partials of a fixed grid combined in a fixed order (deterministic).  part
 * needs 4 * dpdlp_dev_reduce_blocks(n) doubles.  Replace cublasDnrm2 /
 * cublasDdot / get_vector_inf_norm / get_vector_sum (utils.cu:616-637). */
enum { DPDLP_RED_NRM2 = 0, DPDLP_RED_DOT = 1, DPDLP_RED_AMAX = 2, DPDLP_RED_SUM = 3 };
int dpdlp_dev_reduce_blocks(long long n);
int dpdlp_dev_reduce(int op, const double *x, const double *y, long long n, double *part,
                     double *out, void *stream);

/* Gather-rate calibration (the SpMVs' gather ceiling): workgroup w streams
 * idx[w*chunk .. (w+1)*chunk) coalesced and gathers x[base + idx[k]], base =
 * (w mod 8) * slice when xcd_local (each XCD reads only its own slice), else
 * 0; y gets one partial sum per thread (ceil(count/chunk) * 256 doubles). */
int dpdlp_dev_gather_calib(const int *idx, const double *x, double *y, long long count,
                           int chunk, long long slice, int xcd_local, void *stream);
/* The same with the FP64 values val[count] streamed beside the indices (the
 * SpMV's 12 bytes per nonzero) and multiplied into the gathers; val NULL is
 * dpdlp_dev_gather_calib. */
int dpdlp_dev_gather_calib_values(const int *idx, const double *val, const double *x,
                                  double *y, long long count, int chunk, long long slice,
                                  int xcd_local, void *stream);
/* Plain device copy of `bytes` (multiple of 32): the reference point that
 * scales the FETCH_SIZE / WRITE_SIZE counters to bytes. */
int dpdlp_dev_copy_calib(const void *src, void *dst, long long bytes, void *stream);
/* Warm start (pdlp_core_op.cu:174-201): a = b = c = start * D * s. */
int dpdlp_dev_warm_start(double *a, double *b, double *c, const double *start,
                         const double *D, double s, int count, void *stream);

#ifdef __cplusplus
}
#endif
