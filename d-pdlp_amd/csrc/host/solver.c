/*
 * Host driver of the single-GPU solve (the reference's optimize(),
 * src/solver.cu:32-127, with initialize_solver_state, the step size, restart,
 * residual, infeasibility, polishing and result logic of pdlp_core_op.cu and
 * utils.cu).
 *
 * Control stays on the host, as in the reference, but the host never waits
 * inside an iteration: each PDHG step is two fused kernels launched
 * asynchronously, every reduction lands in device scalar slots, and the host
 * reads all of them with ONE device-to-host copy per evaluation point (the
 * reference makes 5-12 blocking cuBLAS calls per evaluation).  Scalar
 * decisions (termination, restart, PID primal weight) are the reference's,
 * made from those values.
 */
#define _POSIX_C_SOURCE 200809L
#include "cupdlpx.h"
#include "dpdlp_dev.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"
#include "control.h"
#include "solver_internal.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void free_perm_copies(dpdlp_resident_t *r);

#define HIPCHK(call)                                                           \
  do {                                                                         \
    hipError_t _e = (hipError_t)(call);                                        \
    if (_e != hipSuccess) {                                                    \
      dpdlp_set_error("HIP error %d (%s) at %s:%d", (int)_e,                   \
                      hipGetErrorString(_e), __FILE__, __LINE__);              \
      goto fail;                                                               \
    }                                                                          \
  } while (0)

/* bytes an array takes in an arena or a resident's blob (256-byte steps) */
static size_t pad256(size_t bytes) { return ((bytes ? bytes : 8) + 255) & ~(size_t)255; }

static void *dmalloc(ctx_t *c, size_t bytes) {
  void *p = NULL;
  size_t need = pad256(bytes);
  if (c->arena && c->arena_off + need <= c->arena_len) {
    p = c->arena + c->arena_off;
    c->arena_off += need;
    return p;
  }
  if (c->nalloc >= MAX_ALLOCS) return NULL;
  if (hipMalloc(&p, need) != hipSuccess) return NULL;
  c->allocs[c->nalloc++] = p;
  return p;
}

/* zero-filled: arena memory is cleared once when the arena is made */
static double *dzero(ctx_t *c, long long count) {
  size_t bytes = (size_t)(count > 0 ? count : 1) * sizeof(double);
  int from_arena = c->arena && c->arena_off + pad256(bytes) <= c->arena_len;
  double *p = (double *)dmalloc(c, bytes);
  if (p && !from_arena && hipMemsetAsync(p, 0, bytes, c->s) != hipSuccess) return NULL;
  return p;
}

static double *dcopy(ctx_t *c, const double *src, long long count) {
  double *p = (double *)dmalloc(c, (size_t)(count > 0 ? count : 1) * sizeof(double));
  if (p && count > 0 &&
      hipMemcpyAsync(p, src, (size_t)count * sizeof(double), hipMemcpyDeviceToDevice, c->s) !=
          hipSuccess)
    return NULL;
  return p;
}

static void harvest_events(ctx_t *c) {
  for (int k = 0; k < c->nev_used; ++k) {
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, c->ev[k][0], c->ev[k][1]) == hipSuccess &&
        hipEventElapsedTime(&b, c->ev[k][1], c->ev[k][2]) == hipSuccess) {
      c->p1_ms += a;
      c->p2_ms += b;
      c->nsamp++;
    }
  }
  c->nev_used = 0;
}

void dpdlp_i_ctx_free(ctx_t *c) {
  if (c->s) hipStreamSynchronize(c->s);
  for (int k = 0; k < c->nev_created; ++k)
    for (int q = 0; q < 3; ++q) hipEventDestroy(c->ev[k][q]);
  c->nev_created = 0;
  for (int i = 0; i < c->nalloc; ++i) hipFree(c->allocs[i]);
  c->nalloc = 0;
  if (c->arena) hipFree(c->arena);
  c->arena = NULL;
  c->arena_off = c->arena_len = c->arena_cap = 0;
  if (c->h_scal && c->owns_scal) hipHostFree(c->h_scal);
  c->h_scal = NULL;
  if (c->s && c->owns_stream) hipStreamDestroy(c->s);
  c->s = NULL;
}

static int sync_scal(ctx_t *c, int count) {
  if (hipMemcpyAsync(c->h_scal, c->st.scal, (size_t)count * sizeof(double),
                     hipMemcpyDeviceToHost, c->s) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(c->s) != hipSuccess) return -1;
  if (c->nev_used) harvest_events(c);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Upload                                                                    */
/* ------------------------------------------------------------------------ */

/* Column panels so that each XCD's slice of the gathered FP64 vector is at
 * most 2 MB (half the 4 MB L2 of an XCD); 1, 2, 4 or 8. */
static int pick_panels(int cols) {
  long long bytes = (long long)cols * 8;
  int P = 1;
  while (P < 8 && bytes / P > (2LL << 20)) P *= 2;
  return P;
}

/* Partial row sums of the tiled SpMV: P vectors of the tiled matrix's rows,
 * shared by A and A' (one SpMV at a time). */
static long long xpart_sums_len(const dpdlp_resident_t *r) {
  long long a = r->A_P > 1 ? (long long)r->A_P * r->m : 0;
  long long b = r->At_P > 1 ? (long long)r->At_P * r->n : 0;
  return a > b ? a : b;
}

void dpdlp_resident_csr(const dpdlp_resident_t *r, int transpose, void *out, double *xpart) {
  dpdlp_csr_t *A = (dpdlp_csr_t *)out;
  memset(A, 0, sizeof(*A));
  /* the tickets follow the partial sums: A's row blocks first, then A''s */
  unsigned *tick = xpart ? (unsigned *)(xpart + xpart_sums_len(r)) : NULL;
  if (!transpose) {
    *A = (dpdlp_csr_t){r->m, r->n, r->nnz, r->A_rp, r->A_ci, r->A_v, NULL, r->A_blk, r->A_nblk,
                       r->cap, r->A_P, r->A_toff, r->A_rofs, r->A_pci, r->A_pv, NULL, xpart,
                       NULL};
    if (r->A_P > 1) A->tick = tick;
  } else {
    *A = (dpdlp_csr_t){r->n, r->m, r->nnz, r->At_rp, r->At_ci, r->At_v, NULL, r->At_blk,
                       r->At_nblk, r->At_cap, r->At_P, r->At_toff, r->At_rofs, r->At_pci, r->At_pv,
                       NULL, xpart, NULL};
    if (r->At_P > 1 && tick) A->tick = tick + (r->A_P > 1 ? r->A_nblk : 0);
  }
}

long long dpdlp_resident_xpart_len(const dpdlp_resident_t *r) {
  long long sums = xpart_sums_len(r);
  if (sums == 0) return 0;
  long long tickets = (r->A_P > 1 ? (long long)r->A_nblk : 0) +
                      (r->At_P > 1 ? (long long)r->At_nblk : 0);
  return sums + (tickets + 1) / 2; /* two 4-byte tickets per double */
}

static int pick_cap(long long nnz, int rows) {
  (void)rows;
  if (nnz >= 4000000) return 2048;
  return 1024;
}

void dpdlp_default_opts(dpdlp_solve_opts_t *o) {
  memset(o, 0, sizeof(*o));
  o->precision = DPDLP_PREC_FP64;
  o->mixed_switch_tol = 1e-6;
  o->time_to_1e4_sec = -1.0;
  o->time_to_1e6_sec = -1.0;
}

static dpdlp_resident_t *upload_with(const lp_problem_t *p, int spmv_cap, int cap_At,
                                     int rows_A, int rows_At, int PA,
                                     int PAt, int order, const dpdlp_uploader_t *u);

/* frees the host side of r and hands back its device blob */
char *dpdlp_i_resident_release(dpdlp_resident_t *r, size_t *cap) {
  if (!r) return NULL;
  char *b = r->blob;
  if (cap) *cap = r->blob_cap;
  free(r->h_pstart);
  free(r->h_dstart);
  free(r->h_A_blk);
  free(r->h_At_blk);
  free(r->h_row_perm);
  free_perm_copies(r);
  free(r);
  return b;
}

dpdlp_resident_t *dpdlp_i_upload_into(const lp_problem_t *p, int spmv_cap, int spmv_cap_At,
                                      int panels_A, int panels_At, const dpdlp_uploader_t *u) {
  return upload_with(p, spmv_cap, spmv_cap_At, 0, 0, panels_A, panels_At,
                     (panels_A > 0 || panels_At > 0) ? 0 : -1, u);
}

void dpdlp_resident_free(dpdlp_resident_t *r) {
  if (!r) return;
  if (r->blob) hipFree(r->blob);
  free(r->h_pstart);
  free(r->h_dstart);
  free(r->h_A_blk);
  free(r->h_At_blk);
  free(r->h_row_perm);
  free_perm_copies(r);
  free(r);
}

#define UPCHK(call)                                                            \
  do {                                                                         \
    hipError_t _e = (hipError_t)(call);                                        \
    if (_e != hipSuccess) {                                                    \
      dpdlp_set_error("HIP error %d (%s) at %s:%d", (int)_e,                   \
                      hipGetErrorString(_e), __FILE__, __LINE__);              \
      goto ufail;                                                              \
    }                                                                          \
  } while (0)


/* Carves a resident's device arrays out of one blob.  Without an uploader
 * the blob is allocated here and every array is copied directly; with one,
 * the blob comes from the uploader (a recycled allocation) and the arrays
 * are gathered in its pinned staging buffer and sent in one asynchronous
 * copy on its stream, so an upload under load neither frees device memory
 * nor waits behind the device's kernels for each array. */
typedef struct {
  char *blob, *stage;
  size_t cap, off;
  hipStream_t s;
} usink_t;

static int sink_open(usink_t *k, const dpdlp_uploader_t *u, size_t need) {
  memset(k, 0, sizeof(*k));
  if (u) {
    k->blob = u->get_blob(u->ctx, need, &k->cap);
    k->stage = u->get_stage(u->ctx, need);
    k->s = (hipStream_t)u->stream;
    if (!k->blob || !k->stage) {
      dpdlp_set_error("no device memory or staging buffer for %zu bytes", need);
      if (k->blob) u->put_blob(u->ctx, k->blob, k->cap);
      k->blob = NULL;
      return -1;
    }
    return 0;
  }
  if (hipMalloc((void **)&k->blob, need) != hipSuccess) {
    dpdlp_set_error("device allocation failed (%zu bytes of problem data)", need);
    k->blob = NULL;
    return -1;
  }
  k->cap = need;
  return 0;
}

static hipError_t up(usink_t *k, void **dst, const void *src, size_t bytes) {
  size_t step = pad256(bytes);
  if (k->off + step > k->cap) return hipErrorOutOfMemory;
  *dst = k->blob + k->off;
  if (bytes && src) {
    if (k->stage)
      memcpy(k->stage + k->off, src, bytes);
    else {
      hipError_t e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) return e;
    }
  }
  k->off += step;
  return hipSuccess;
}

static hipError_t sink_close(usink_t *k) {
  if (!k->stage || k->off == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(k->blob, k->stage, k->off, hipMemcpyHostToDevice, k->s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(k->s);
}


dpdlp_resident_t *dpdlp_upload(const lp_problem_t *p, int spmv_cap) {
  return upload_with(p, spmv_cap, 0, 0, 0, 0, 0, -1, NULL);
}

dpdlp_resident_t *dpdlp_upload_ex(const lp_problem_t *p, int spmv_cap, int panels_A,
                                  int panels_At) {
  if ((panels_A && panels_A != 1 && panels_A != 2 && panels_A != 4 && panels_A != 8) ||
      (panels_At && panels_At != 1 && panels_At != 2 && panels_At != 4 && panels_At != 8)) {
    dpdlp_set_error("panel counts must be 0 (automatic), 1, 2, 4 or 8");
    return NULL;
  }
  /* explicit layouts (the batch's plain one) keep the row order */
  return upload_with(p, spmv_cap, 0, 0, 0, panels_A, panels_At,
                     (panels_A > 0 || panels_At > 0) ? 0 : -1, NULL);
}

dpdlp_resident_t *dpdlp_upload_caps(const lp_problem_t *p, int spmv_cap, int spmv_cap_At,
                                    int panels_A, int panels_At) {
  if ((panels_A && panels_A != 1 && panels_A != 2 && panels_A != 4 && panels_A != 8) ||
      (panels_At && panels_At != 1 && panels_At != 2 && panels_At != 4 && panels_At != 8)) {
    dpdlp_set_error("panel counts must be 0 (automatic), 1, 2, 4 or 8");
    return NULL;
  }
  if (spmv_cap_At && spmv_cap_At != 1024 && spmv_cap_At != 2048 && spmv_cap_At != 4096) {
    dpdlp_set_error("row-block capacities must be 0 (automatic), 1024, 2048 or 4096");
    return NULL;
  }
  return upload_with(p, spmv_cap, spmv_cap_At, 0, 0, panels_A, panels_At,
                     (panels_A > 0 || panels_At > 0) ? 0 : -1, NULL);
}

dpdlp_resident_t *dpdlp_upload_batch_layout(const lp_problem_t *p, const dpdlp_batch_layout_t *L) {
  if (!L) {
    dpdlp_set_error("no layout");
    return NULL;
  }
  const int caps[2] = {L->cap_A, L->cap_At}, rows[2] = {L->rows_A, L->rows_At};
  for (int k = 0; k < 2; ++k) {
    if (caps[k] != 1024 && caps[k] != 2048 && caps[k] != 4096) {
      dpdlp_set_error("row-block capacities must be 1024, 2048 or 4096");
      return NULL;
    }
    if (rows[k] < 1 || rows[k] > DPDLP_BLOCK_ROWS) {
      dpdlp_set_error("rows per block must be 1..%d", DPDLP_BLOCK_ROWS);
      return NULL;
    }
  }
  return upload_with(p, L->cap_A, L->cap_At, L->rows_A, L->rows_At, 1, 1, 0, NULL);
}

size_t dpdlp_resident_bytes(const dpdlp_resident_t *r) {
  if (!r) return 0;
  size_t nnz = (size_t)r->nnz, m = (size_t)r->m, n = (size_t)r->n;
  size_t b = nnz * 24 + (m + n + 2) * 4 + (size_t)(r->A_nblk + r->At_nblk + 2) * 4 +
             (3 * n + 2 * m) * 8;
  if (r->A_P > 1) b += nnz * 16 + (size_t)r->A_P * (m + (size_t)r->A_nblk) * 2;
  if (r->At_P > 1) b += nnz * 16 + (size_t)r->At_P * (n + (size_t)r->At_nblk) * 2;
  return b;
}

/* Times both SpMVs of a resident problem (mean of `reps`, ms). */
static void time_both(dpdlp_resident_t *r, double *ta, double *tt) {
  *ta = dpdlp_time_spmv(r, 0, 10, 0);
  *tt = dpdlp_time_spmv(r, 1, 10, 0);
}

dpdlp_resident_t *dpdlp_upload_tuned(const lp_problem_t *p, int spmv_cap, int *chosen) {
  /* candidates P in {1,2,4,8} for A and A' independently; each candidate
   * upload is timed with the plain SpMV and freed */
  static const int cand[4] = {1, 2, 4, 8};
  double bestA = 1e30, bestT = 1e30;
  int PA = 1, PT = 1;
  for (int i = 0; i < 4; ++i) {
    dpdlp_resident_t *r = upload_with(p, spmv_cap, 0, 0, 0, cand[i], cand[i], -1, NULL);
    if (!r) continue;
    double ta, tt;
    time_both(r, &ta, &tt);
    if (ta > 0 && ta < bestA) {
      bestA = ta;
      PA = cand[i];
    }
    if (tt > 0 && tt < bestT) {
      bestT = tt;
      PT = cand[i];
    }
    dpdlp_resident_free(r);
  }
  if (chosen) {
    chosen[0] = PA;
    chosen[1] = PT;
  }
  return upload_with(p, spmv_cap, 0, 0, 0, PA, PT, -1, NULL);
}

/* row lengths skewed enough for the row renumbering to pay: the longest row
 * holds at least 16 times the mean (DPDLP_ROW_ORDER=0/1 overrides) */
static int rows_skewed(int m, const int *rp) {
  const char *env = getenv("DPDLP_ROW_ORDER");
  if (env) return atoi(env) != 0;
  if (m < 4096) return 0;
  long long mx = 0;
  for (int i = 0; i < m; ++i)
    if (rp[i + 1] - rp[i] > mx) mx = rp[i + 1] - rp[i];
  return mx * (long long)m >= 16LL * rp[m];
}

/* the entries of renumbered rows [m t / ntask, m (t + 1) / ntask) */
typedef struct {
  int m, ntask;
  const int *perm, *rp, *ci;
  const double *v;
  const int *rp2;
  int *ci2;
  double *v2;
} rowcopy_job_t;

static void rowcopy_task(void *arg, int t) {
  rowcopy_job_t *J = (rowcopy_job_t *)arg;
  int i0 = (int)((long long)J->m * t / J->ntask), i1 = (int)((long long)J->m * (t + 1) / J->ntask);
  for (int i = i0; i < i1; ++i) {
    int o = J->perm[i], a = J->rp[o], L = J->rp[o + 1] - a;
    memcpy(J->ci2 + J->rp2[i], J->ci + a, (size_t)L * sizeof(int));
    memcpy(J->v2 + J->rp2[i], J->v + a, (size_t)L * sizeof(double));
  }
}

/* perm[new] = old: rows by decreasing length, ties in row order (a stable
 * counting sort) */
static int row_order_by_length(int m, const int *rp, int *perm) {
  int mx = 0;
  for (int i = 0; i < m; ++i)
    if (rp[i + 1] - rp[i] > mx) mx = rp[i + 1] - rp[i];
  int *cnt = (int *)calloc((size_t)mx + 2, sizeof(int));
  if (!cnt) {
    dpdlp_set_error("host allocation failed");
    return -1;
  }
  for (int i = 0; i < m; ++i) cnt[mx - (rp[i + 1] - rp[i])]++;
  int run = 0;
  for (int L = 0; L <= mx; ++L) {
    int c = cnt[L];
    cnt[L] = run;
    run += c;
  }
  for (int i = 0; i < m; ++i) perm[cnt[mx - (rp[i + 1] - rp[i])]++] = i;
  free(cnt);
  return 0;
}

/* Deals the rows of a length order (perm[rank] = row) round-robin into P
 * contiguous ranges of width W = ceil(m / P), the column panels of A'
 * (panels.c), a full range being skipped: perm[new] = row afterwards.  Every
 * panel then holds an equal share of the long rows, and within a panel the
 * rows still run from longest to shortest, so the dual entries that each
 * XCD gathers most often share cache lines of its own slice. */
static int interleave_panels(int m, int P, int *perm) {
  if (P <= 1 || m <= 0) return 0;
  const long long W = ((long long)m + P - 1) / P;
  int *out = (int *)malloc((size_t)m * sizeof(int));
  long long fill[8] = {0};
  if (!out) {
    dpdlp_set_error("host allocation failed");
    return -1;
  }
  int k = 0;
  for (int r = 0; r < m; ++r) {
    for (;;) {
      const long long capk = k < P - 1 ? W : (long long)m - (long long)(P - 1) * W;
      if (fill[k] < capk) break;
      k = (k + 1) % P;
    }
    out[k * W + fill[k]++] = perm[r];
    k = (k + 1) % P;
  }
  memcpy(perm, out, (size_t)m * sizeof(int));
  free(out);
  return 0;
}

static void free_perm_copies(dpdlp_resident_t *r) {
  free(r->h_perm_rp);
  free(r->h_perm_ci);
  free(r->h_perm_v);
  free(r->h_perm_lc);
  free(r->h_perm_uc);
  r->h_perm_rp = r->h_perm_ci = NULL;
  r->h_perm_v = r->h_perm_lc = r->h_perm_uc = NULL;
}

static dpdlp_resident_t *upload_with(const lp_problem_t *p, int spmv_cap, int cap_At,
                                     int rows_A, int rows_At, int PA,
                                     int PAt, int order, const dpdlp_uploader_t *u) {
  double t0 = dpdlp_wall();
  usink_t K;
  memset(&K, 0, sizeof(K));
  if (!p || p->num_variables < 0 || p->num_constraints < 0 ||
      p->constraint_matrix_num_nonzeros < 0) {
    dpdlp_set_error("invalid problem");
    return NULL;
  }
  int m = p->num_constraints, n = p->num_variables, nnz = p->constraint_matrix_num_nonzeros;
  const int *rp = p->constraint_matrix_row_pointers, *ci = p->constraint_matrix_col_indices;
  if (!rp || (nnz > 0 && (!ci || !p->constraint_matrix_values)) ||
      (n > 0 && (!p->objective_vector || !p->variable_lower_bound ||
                 !p->variable_upper_bound)) ||
      (m > 0 && (!p->constraint_lower_bound || !p->constraint_upper_bound))) {
    dpdlp_set_error("problem has NULL arrays");
    return NULL;
  }
  if (rp[0] != 0 || rp[m] != nnz) {
    dpdlp_set_error("row pointers inconsistent with nnz");
    return NULL;
  }
  for (int i = 0; i < m; ++i)
    if (rp[i + 1] < rp[i]) {
      dpdlp_set_error("row pointers not monotone at row %d", i);
      return NULL;
    }
  for (int k = 0; k < nnz; ++k)
    if (ci[k] < 0 || ci[k] >= n) {
      dpdlp_set_error("column index %d out of range at nonzero %d", ci[k], k);
      return NULL;
    }
  dpdlp_resident_t *r = (dpdlp_resident_t *)calloc(1, sizeof(*r));
  int *t_rp = NULL, *blk = NULL, *t_ci = NULL;
  double *t_v = NULL;
  if (!r) return NULL;
  r->m = m;
  r->n = n;
  r->nnz = nnz;
  r->cap = spmv_cap > 0 ? spmv_cap : pick_cap(nnz, m);
  if (r->cap != 1024 && r->cap != 2048 && r->cap != 4096) r->cap = 2048;
  r->c0 = p->objective_constant;

  /* norms of the unscaled problem, pdlp_core_op.cu:232-286 */
  {
    double ssq = 0.0, mx = 0.0;
    for (int j = 0; j < n; ++j) {
      double v = p->objective_vector[j];
      ssq += v * v;
      if (fabs(v) > mx) mx = fabs(v);
    }
    r->obj_l2 = sqrt(ssq);
    r->obj_linf = mx;
    ssq = 0.0;
    mx = 0.0;
    for (int i = 0; i < m; ++i) {
      double lo = p->constraint_lower_bound[i], upb = p->constraint_upper_bound[i];
      if (isfinite(lo) && lo != upb) {
        ssq += lo * lo;
        if (fabs(lo) > mx) mx = fabs(lo);
      }
      if (isfinite(upb)) {
        ssq += upb * upb;
        if (fabs(upb) > mx) mx = fabs(upb);
      }
    }
    r->bnd_l2 = sqrt(ssq);
    r->bnd_linf = mx;
  }
  if (p->primal_start) {
    r->h_pstart = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    if (!r->h_pstart) goto ufail;
    memcpy(r->h_pstart, p->primal_start, (size_t)n * sizeof(double));
  }
  if (p->dual_start) {
    r->h_dstart = (double *)malloc((size_t)(m ? m : 1) * sizeof(double));
    if (!r->h_dstart) goto ufail;
    memcpy(r->h_dstart, p->dual_start, (size_t)m * sizeof(double));
  }

  /* Row order: with strongly skewed row lengths (automatic layouts only),
   * rows are renumbered by decreasing length, so the dual entries that A'
   * gathers most often share cache lines; vectors indexed by rows are
   * permuted at upload and the dual result back at the end. */
  const int renumber = order > 0 || (order < 0 && rows_skewed(m, rp));

  /* XCD column-panel tiles.  Without a panel count from the caller, A'
   * stays untiled when the rows are renumbered (A' 128 us untiled at
   * m = 1e6, nnz = 2e7 Zipf rows).  When A' is tiled over renumbered rows,
   * the length order is dealt round-robin over its column panels
   * (interleave_panels): in plain length order, panels of equal width would
   * hand the first panel most of the nonzeros (256 / 487 us on 2 / 4
   * panels). */
  r->A_P = PA > 0 ? PA : pick_panels(n);
  r->At_P = PAt > 0 ? PAt : (renumber ? 1 : pick_panels(m));
  if (PA <= 0) {
    const char *env = getenv("DPDLP_PANELS");
    if (env) {
      int a = 0, b = 0;
      if (sscanf(env, "%d,%d", &a, &b) == 2) {
        if (a == 1 || a == 2 || a == 4 || a == 8) r->A_P = a;
        if (b == 1 || b == 2 || b == 4 || b == 8) r->At_P = b;
      }
    }
  }

  const int *rp_u = rp, *ci_u = ci;
  const double *v_u = p->constraint_matrix_values, *lc_u = p->constraint_lower_bound,
               *uc_u = p->constraint_upper_bound;
  if (renumber) {
    int *perm = (int *)malloc((size_t)(m ? m : 1) * sizeof(int));
    int *rp2 = (int *)malloc((size_t)(m + 1) * sizeof(int));
    int *ci2 = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
    double *v2 = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    double *lc2 = (double *)malloc((size_t)(m ? m : 1) * sizeof(double));
    double *uc2 = (double *)malloc((size_t)(m ? m : 1) * sizeof(double));
    r->h_row_perm = perm;
    r->h_perm_rp = rp2;
    r->h_perm_ci = ci2;
    r->h_perm_v = v2;
    r->h_perm_lc = lc2;
    r->h_perm_uc = uc2;
    if (!perm || !rp2 || !ci2 || !v2 || !lc2 || !uc2) {
      dpdlp_set_error("host allocation failed");
      goto ufail;
    }
    if (row_order_by_length(m, rp, perm)) goto ufail;
    if (interleave_panels(m, r->At_P, perm)) goto ufail;
    rp2[0] = 0;
    for (int i = 0; i < m; ++i) {
      int o = perm[i];
      rp2[i + 1] = rp2[i] + (rp[o + 1] - rp[o]);
      lc2[i] = p->constraint_lower_bound[o];
      uc2[i] = p->constraint_upper_bound[o];
    }
    rowcopy_job_t J = {m, 4 * dpdlp_host_threads(), perm, rp, ci, p->constraint_matrix_values,
                       rp2, ci2, v2};
    dpdlp_par_for(J.ntask, rowcopy_task, &J);
    if (r->h_dstart) {
      double *d2 = (double *)malloc((size_t)(m ? m : 1) * sizeof(double));
      if (!d2) goto ufail;
      for (int i = 0; i < m; ++i) d2[i] = r->h_dstart[perm[i]];
      free(r->h_dstart);
      r->h_dstart = d2;
    }
    rp_u = rp2;
    ci_u = ci2;
    v_u = v2;
    lc_u = lc2;
    uc_u = uc2;
  }


  /* an untiled A' of a large problem with an automatic block size takes
   * 4096-nonzero blocks (its rows, the columns of A, are short and many) */
  r->At_cap = r->cap;
  if (spmv_cap <= 0 && r->At_P == 1 && nnz >= 4000000) r->At_cap = 4096;
  if (cap_At > 0) r->At_cap = cap_At;
  /* most rows per untiled row block (the lane-group width of the reduction) */
  r->rows_A = rows_A > 0 && rows_A <= DPDLP_BLOCK_ROWS ? rows_A : DPDLP_BLOCK_ROWS;
  r->rows_At = rows_At > 0 && rows_At <= DPDLP_BLOCK_ROWS ? rows_At : DPDLP_BLOCK_ROWS;

  /* A' is transposed on the device after the upload (the host only counts
   * columns for its row pointers); its column-panel tiles, when it has
   * them, are planned from A's host CSR and cut on the device from the
   * device transpose.  Uploads through an uploader (serving: the device is
   * busy solving other LPs, and the device transpose and tile cutting
   * delay their launches) transpose and cut tiled matrices on the host
   * instead. */
  const int host_At = u != NULL && r->At_P > 1, host_A = u != NULL && r->A_P > 1;
  t_rp = (int *)malloc((size_t)(n + 1) * sizeof(int));
  blk = (int *)malloc((size_t)((m > n ? m : n) + 2) * sizeof(int));
  if (host_At) {
    t_ci = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
    t_v = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
  }
  if (!t_rp || !blk || (host_At && (!t_ci || !t_v))) {
    dpdlp_set_error("host allocation failed");
    goto ufail;
  }
  if (host_At) {
    if (dpdlp_host_transpose(m, n, rp_u, ci_u, v_u, t_rp, t_ci, t_v)) goto ufail;
  } else {
    dpdlp_host_col_ptr(m, n, rp_u, ci_u, t_rp);
  }
  /* row-block plans (host copies kept for the batch's work lists) */
  for (int which = 0; which < 2; ++which) {
    int P = which ? r->At_P : r->A_P, rows = which ? n : m;
    const int *hrp = which ? t_rp : rp_u;
    int cap = which ? r->At_cap : r->cap;
    int nb = P <= 1 ? dpdlp_plan_blocks_rows(rows, hrp, cap, which ? r->rows_At : r->rows_A, blk)
             : !which ? dpdlp_plan_blocks_panels(m, n, rp_u, ci_u, cap, P, DPDLP_PANEL_RMAX, blk)
             : host_At ? dpdlp_plan_blocks_panels(n, m, t_rp, t_ci, cap, P, DPDLP_PANEL_RMAX, blk)
                       : dpdlp_plan_blocks_panels_T(m, n, rp_u, ci_u, cap, P, DPDLP_PANEL_RMAX, blk);
    if (nb < 0) {
      dpdlp_set_error("host allocation failed (panel plan)");
      goto ufail;
    }
    int *h = (int *)malloc((size_t)(nb + 1) * sizeof(int));
    if (!h) {
      dpdlp_set_error("host allocation failed");
      goto ufail;
    }
    memcpy(h, blk, (size_t)(nb + 1) * sizeof(int));
    if (which) {
      r->At_nblk = nb;
      r->h_At_blk = h;
    } else {
      r->A_nblk = nb;
      r->h_A_blk = h;
    }
  }

  /* every device array lives in one blob, carved in 256-byte steps in the
   * order below */
  {
    size_t need = pad256(4 * ((size_t)m + 1)) + pad256(4 * (size_t)nnz) + pad256(8 * (size_t)nnz) +
                  pad256(4 * ((size_t)n + 1)) + pad256(4 * (size_t)nnz) + pad256(8 * (size_t)nnz) +
                  pad256(4 * ((size_t)r->A_nblk + 1)) + pad256(4 * ((size_t)r->At_nblk + 1)) +
                  3 * pad256(8 * (size_t)n) + 2 * pad256(8 * (size_t)m);
    for (int which = 0; which < 2; ++which) {
      int P = which ? r->At_P : r->A_P, rows = which ? n : m;
      size_t nb = (size_t)(which ? r->At_nblk : r->A_nblk);
      if (P > 1)
        need += pad256(4 * (nb * P + 1)) + pad256(2 * (size_t)P * ((size_t)rows + nb)) +
                2 * pad256(4 * (size_t)nnz) + pad256(8 * (size_t)nnz);
    }
    if (sink_open(&K, u, need)) goto ufail;
    r->blob = K.blob;
    r->blob_cap = K.cap;
    r->pooled = u != NULL;
  }
  UPCHK(up(&K, (void **)&r->A_rp, rp_u, (size_t)(m + 1) * sizeof(int)));
  UPCHK(up(&K, (void **)&r->A_ci, ci_u, (size_t)nnz * sizeof(int)));
  UPCHK(up(&K, (void **)&r->A_v, v_u, (size_t)nnz * sizeof(double)));
  UPCHK(up(&K, (void **)&r->At_rp, t_rp, (size_t)(n + 1) * sizeof(int)));
  UPCHK(up(&K, (void **)&r->At_ci, t_ci, (size_t)nnz * sizeof(int)));
  UPCHK(up(&K, (void **)&r->At_v, t_v, (size_t)nnz * sizeof(double)));
  UPCHK(up(&K, (void **)&r->A_blk, r->h_A_blk, (size_t)(r->A_nblk + 1) * sizeof(int)));
  UPCHK(up(&K, (void **)&r->At_blk, r->h_At_blk, (size_t)(r->At_nblk + 1) * sizeof(int)));

  /* column-panel tiles: device arrays carved here and cut on the device
   * (k_build_panels) once A is there, A''s after the transpose */
  for (int which = 0; which < 2; ++which) {
    size_t P = (size_t)(which ? r->At_P : r->A_P), nb = (size_t)(which ? r->At_nblk : r->A_nblk);
    size_t rows = (size_t)(which ? n : m);
    if (P <= 1) continue;
    int *toff = NULL, *perm = NULL, *pci = NULL;
    unsigned short *rofs = NULL;
    if (which ? host_At : host_A) {
      /* cut on the host (dpdlp_build_panels) and copied with the rest */
      toff = (int *)malloc((nb * P + 1) * sizeof(int));
      rofs = (unsigned short *)malloc(P * (rows + nb) * sizeof(unsigned short));
      perm = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
      pci = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
      int ok = toff && rofs && perm && pci &&
               dpdlp_build_panels((int)rows, which ? m : n, which ? t_rp : rp_u,
                                  which ? t_ci : ci_u, which ? r->h_At_blk : r->h_A_blk,
                                  (int)nb, (int)P, toff, rofs, perm, pci) == 0;
      if (!ok) {
        free(toff);
        free(rofs);
        free(perm);
        free(pci);
        dpdlp_set_error("building column-panel tiles failed");
        goto ufail;
      }
    }
    hipError_t e =
        up(&K, (void **)(which ? &r->At_toff : &r->A_toff), toff, (nb * P + 1) * sizeof(int));
    if (e == hipSuccess)
      e = up(&K, (void **)(which ? &r->At_rofs : &r->A_rofs), rofs,
             P * (rows + nb) * sizeof(unsigned short));
    if (e == hipSuccess)
      e = up(&K, (void **)(which ? &r->At_perm : &r->A_perm), perm, (size_t)nnz * sizeof(int));
    if (e == hipSuccess)
      e = up(&K, (void **)(which ? &r->At_pci : &r->A_pci), pci, (size_t)nnz * sizeof(int));
    if (e == hipSuccess)
      e = up(&K, (void **)(which ? &r->At_pv : &r->A_pv), NULL, (size_t)nnz * sizeof(double));
    free(toff);
    free(rofs);
    free(perm);
    free(pci);
    UPCHK(e);
  }
  UPCHK(up(&K, (void **)&r->c, p->objective_vector, (size_t)n * sizeof(double)));
  UPCHK(up(&K, (void **)&r->lv, p->variable_lower_bound, (size_t)n * sizeof(double)));
  UPCHK(up(&K, (void **)&r->uv, p->variable_upper_bound, (size_t)n * sizeof(double)));
  UPCHK(up(&K, (void **)&r->lc, lc_u, (size_t)m * sizeof(double)));
  UPCHK(up(&K, (void **)&r->uc, uc_u, (size_t)m * sizeof(double)));
  UPCHK(sink_close(&K));
  if (r->A_P > 1) {
    if (!host_A)
      UPCHK((hipError_t)dpdlp_dev_build_panels(m, n, r->A_rp, r->A_ci, r->A_blk, r->A_nblk,
                                               r->A_P, r->A_toff, r->A_rofs, r->A_perm, r->A_pci,
                                               K.s));
    if (nnz > 0) UPCHK((hipError_t)dpdlp_dev_permute(r->A_v, r->A_perm, r->A_pv, nnz, K.s));
  }
  if (host_At) {
    /* A' and its tiles came from the host: only the tiled values are gathered */
    if (nnz > 0) UPCHK((hipError_t)dpdlp_dev_permute(r->At_v, r->At_perm, r->At_pv, nnz, K.s));
    UPCHK(hipStreamSynchronize(K.s));
  } else if (nnz == 0) {
    if (r->At_P > 1)
      UPCHK((hipError_t)dpdlp_dev_build_panels(n, m, r->At_rp, r->At_ci, r->At_blk, r->At_nblk,
                                               r->At_P, r->At_toff, r->At_rofs, r->At_perm,
                                               r->At_pci, K.s));
    UPCHK(hipStreamSynchronize(K.s));
  } else {
    size_t sb = dpdlp_dev_transpose_scratch_bytes(nnz), scap = sb;
    char *sc = NULL;
    if (u)
      sc = u->get_blob(u->ctx, sb, &scap);
    else if (hipMalloc((void **)&sc, sb) != hipSuccess)
      sc = NULL;
    if (!sc) {
      dpdlp_set_error("device allocation failed (transpose scratch)");
      goto ufail;
    }
    hipError_t e = (hipError_t)dpdlp_dev_transpose(m, n, nnz, r->A_rp, r->A_ci, r->A_v,
                                                   r->At_ci, r->At_v, sc, K.s);
    if (e == hipSuccess && r->At_P > 1)
      e = (hipError_t)dpdlp_dev_build_panels(n, m, r->At_rp, r->At_ci, r->At_blk, r->At_nblk,
                                             r->At_P, r->At_toff, r->At_rofs, r->At_perm,
                                             r->At_pci, K.s);
    if (e == hipSuccess && r->At_P > 1)
      e = (hipError_t)dpdlp_dev_permute(r->At_v, r->At_perm, r->At_pv, nnz, K.s);
    if (e == hipSuccess) e = hipStreamSynchronize(K.s);
    if (u)
      u->put_blob(u->ctx, sc, scap);
    else
      hipFree(sc);
    UPCHK(e);
  }
  free(t_rp);
  free(t_ci);
  free(t_v);
  free(blk);
  free_perm_copies(r);
  r->upload_time = dpdlp_wall() - t0;
  return r;
ufail:
  free(t_rp);
  free(t_ci);
  free(t_v);
  free(blk);
  dpdlp_resident_free(r);
  return NULL;
}

/* ------------------------------------------------------------------------ */
/* One PDHG step (solver.cu:85-101)                                          */
/* ------------------------------------------------------------------------ */

static int pdhg_step(ctx_t *c, dpdlp_state_t *st, hstate_t *h, double rho, int do_restart,
                     int use32, int polish, int timed) {
  int major_variant =
      h->is_major || ((h->total + 2) % dpdlp_print_frequency(h->total + 2)) == 0;
  int fpe = h->is_major || do_restart;
  double w = (double)(h->inner + 1) / (h->inner + 2);
  int e;
  hipEvent_t *ev = NULL;
  if (timed && c->nev_used < c->nev_created) ev = c->ev[c->nev_used++];
  if (ev && hipEventRecord(ev[0], c->s) != hipSuccess) return -1;
  e = dpdlp_dev_primal_phase(st, h->step / h->omega, w, rho, major_variant, fpe, use32, c->s);
  if (e) return e;
  if (ev && hipEventRecord(ev[1], c->s) != hipSuccess) return -1;
  e = dpdlp_dev_dual_phase(st, h->step * h->omega, w, rho, major_variant, fpe, use32, c->s);
  if (e) return e;
  if (ev && hipEventRecord(ev[2], c->s) != hipSuccess) return -1;
  if (fpe) {
    if (polish)
      e = dpdlp_dev_fpe_finish_polish(st, polish == 1, do_restart, h->is_major, c->s);
    else
      e = dpdlp_dev_fpe_finish(st, do_restart, h->is_major, c->s);
    if (e) return e;
    if (do_restart) h->pend_rst = 1;
    if (h->is_major) h->pend_maj = 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Feasibility polishing (pdlp_core_op.cu:862-1030)                          */
/* ------------------------------------------------------------------------ */

static void polish_reset(hstate_t *p, const hstate_t *o, double omega0) {
  memset(p, 0, sizeof(*p));
  p->best_omega = 0.0;
  p->fpe = INFINITY;
  p->fpe0 = INFINITY;
  p->fpe_last = INFINITY;
  p->step = o->step;
  p->omega = omega0;
  p->best_omega = omega0;
  p->reason = TERMINATION_REASON_UNSPECIFIED;
  p->best_gap = INFINITY;
}

static int polish_loop(ctx_t *c, dpdlp_state_t *ps, hstate_t *h, const pdhg_parameters_t *params,
                       int primal, double c0_state) {
  double t0 = dpdlp_wall();
  int do_restart = 0, freq = params->termination_evaluation_frequency;
  const termination_criteria_t *tc = &params->termination_criteria;
  dpdlp_print_polish_header(primal, params);
  while (h->reason == TERMINATION_REASON_UNSPECIFIED) {
    if (h->is_major || h->total == 0 || h->total % dpdlp_print_frequency(h->total) == 0) {
      if (dpdlp_dev_residual(ps, &c->st, primal ? 1 : 2, 0, c->s)) return -1;
      if (sync_scal(c, 64)) return -1;
      ctl_pending_fpe(h, c->h_scal, primal ? 1 : 2);
      const double *S = c->h_scal;
      if (primal) {
        h->abs_p = (params->optimality_norm == NORM_TYPE_L_INF ? S[DPDLP_SL_RESP + 7]
                                                               : dpdlp_blue(S + DPDLP_SL_RESP)) /
                   c->k.beta;
        h->rel_p = h->abs_p / (1.0 + c->k.bound_norm);
        h->pobj = S[DPDLP_SL_RESD + 3] / (c->k.beta * c->k.gamma) + c0_state;
      } else {
        h->abs_d = (params->optimality_norm == NORM_TYPE_L_INF ? S[DPDLP_SL_RESD + 8]
                                                               : dpdlp_blue(S + DPDLP_SL_RESD)) /
                   c->k.gamma;
        h->rel_d = h->abs_d / (1.0 + c->k.obj_norm);
        h->dobj = (S[DPDLP_SL_RESD + 4] + S[DPDLP_SL_RESP + 3]) / (c->k.beta * c->k.gamma) + c0_state;
      }
      h->t_cum = dpdlp_wall() - t0;
      ctl_check_polish_termination(h, tc, primal);
      if (params->verbose)
        dpdlp_print_polish_iteration(primal, h->total, h->t_cum, primal ? h->pobj : h->dobj,
                                     primal ? h->abs_p : h->abs_d, primal ? h->rel_p : h->rel_d);
    }
    if (h->is_major || h->total == 0) {
      do_restart = ctl_should_restart(h, &params->restart_params, freq);
      if (do_restart) {
        if (dpdlp_dev_restart_copy(ps, primal ? 1 : 2, c->s)) return -1;
        h->inner = 0;
        h->fpe_last = INFINITY;
      }
    }
    h->is_major = ((h->total + 1) % freq) == 0;
    if (pdhg_step(c, ps, h, params->reflection_coefficient, do_restart, 0, primal ? 1 : 2, 0))
      return -1;
    do_restart = 0;
    h->inner++;
    h->total++;
  }
  return 0;
}

int dpdlp_i_feasibility_polish(ctx_t *c, hstate_t *h, const pdhg_parameters_t *params,
                              double *ptime, int *piters) {
  double t0 = dpdlp_wall();
  const termination_criteria_t *tc = &params->termination_criteria;
  if (h->rel_p < tc->eps_feas_polish_relative && h->rel_d < tc->eps_feas_polish_relative) {
    printf("Skipping feasibility polishing as the solution is already sufficiently feasible.\n");
    return 0;
  }
  double omega0 = params->bound_objective_rescaling
                      ? 1.0
                      : (c->k.obj_norm + 1.0) / (c->k.bound_norm + 1.0);
  dpdlp_state_t *o = &c->st;
  int n = o->n, m = o->m;
  dpdlp_state_t P = *o, D = *o;
  hstate_t hp, hd;

  /* primal polish state (pdlp_core_op.cu:1032-1105) */
  P.c = dzero(c, n);
  P.x_init = dcopy(c, o->x_init, n);
  P.x_cur = dcopy(c, o->x_cur, n);
  P.x_pdhg = dcopy(c, o->x_pdhg, n);
  P.x_ref = dcopy(c, o->x_ref, n);
  P.primal_product = dcopy(c, o->primal_product, m);
  P.y_init = dzero(c, m);
  P.y_cur = dzero(c, m);
  P.y_pdhg = dzero(c, m);
  P.y_ref = dzero(c, m);
  P.dual_product = dzero(c, n);
  P.dual_slack = dzero(c, n);
  P.primal_slack = dzero(c, m);
  P.dx = dzero(c, n);
  P.dy = dzero(c, m);
  P.x_ref32 = NULL;
  P.y_cur32 = NULL;
  P.ax0 = P.aty0 = NULL;
  if (!P.c || !P.x_init || !P.x_cur || !P.x_pdhg || !P.x_ref || !P.primal_product || !P.y_init ||
      !P.y_cur || !P.y_pdhg || !P.y_ref || !P.dual_product || !P.dual_slack || !P.primal_slack ||
      !P.dx || !P.dy)
    return -1;
  polish_reset(&hp, h, omega0);
  if (polish_loop(c, &P, &hp, params, 1, 0.0)) return -1;
  if (hp.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
    if (hipMemcpyAsync(o->x_pdhg, P.x_pdhg, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                       c->s) != hipSuccess)
      return -1;
    h->abs_p = hp.abs_p;
    h->rel_p = hp.rel_p;
    h->pobj = hp.pobj;
  }
  *piters += hp.total - 1;

  /* dual polish state (pdlp_core_op.cu:1145-1249) */
  D.lc = dcopy(c, o->lc, m);
  D.uc = dcopy(c, o->uc, m);
  D.lv = dcopy(c, o->lv, n);
  D.uv = dcopy(c, o->uv, n);
  if (!D.lc || !D.uc || !D.lv || !D.uv) return -1;
  if (dpdlp_dev_zero_finite(D.lc, m, c->s) || dpdlp_dev_zero_finite(D.uc, m, c->s) ||
      dpdlp_dev_zero_finite(D.lv, n, c->s) || dpdlp_dev_zero_finite(D.uv, n, c->s))
    return -1;
  D.y_init = dcopy(c, o->y_init, m);
  D.y_cur = dcopy(c, o->y_cur, m);
  D.y_pdhg = dcopy(c, o->y_pdhg, m);
  D.y_ref = dcopy(c, o->y_ref, m);
  D.dual_product = dcopy(c, o->dual_product, n);
  D.dual_slack = dcopy(c, o->dual_slack, n);
  D.x_init = dzero(c, n);
  D.x_cur = dzero(c, n);
  D.x_pdhg = dzero(c, n);
  D.x_ref = dzero(c, n);
  D.primal_product = dzero(c, m);
  D.primal_slack = dzero(c, m);
  D.dx = dzero(c, n);
  D.dy = dzero(c, m);
  D.x_ref32 = NULL;
  D.y_cur32 = NULL;
  D.ax0 = D.aty0 = NULL;
  if (!D.y_init || !D.y_cur || !D.y_pdhg || !D.y_ref || !D.dual_product || !D.dual_slack ||
      !D.x_init || !D.x_cur || !D.x_pdhg || !D.x_ref || !D.primal_product || !D.primal_slack ||
      !D.dx || !D.dy)
    return -1;
  polish_reset(&hd, h, omega0);
  if (polish_loop(c, &D, &hd, params, 0, c->k.c0)) return -1;
  if (hd.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
    if (hipMemcpyAsync(o->y_pdhg, D.y_pdhg, (size_t)m * sizeof(double), hipMemcpyDeviceToDevice,
                       c->s) != hipSuccess)
      return -1;
    h->abs_d = hd.abs_d;
    h->rel_d = hd.rel_d;
    h->dobj = hd.dobj;
  }
  *piters += hd.total - 1;
  h->gap = fabs(h->pobj - h->dobj);
  h->rel_gap = h->gap / (1.0 + fabs(h->pobj) + fabs(h->dobj));
  if (params->verbose) printf("---------------------------------------------------------------------------------------\n");
  dpdlp_print_polish_final(hp.reason, hp.total - 1, hp.t_cum, hd.reason, hd.total - 1, hd.t_cum,
                           hp.rel_p, hd.rel_d,
                           fabs(hp.pobj - hd.dobj) / (1.0 + fabs(hp.pobj) + fabs(hd.dobj)));
  *ptime = dpdlp_wall() - t0;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Solve                                                                     */
/* ------------------------------------------------------------------------ */

static size_t piece(long long count, size_t elem) {
  size_t b = (size_t)(count > 0 ? count : 1) * elem;
  return (b + 255) & ~(size_t)255;
}

size_t dpdlp_i_ctx_footprint(const dpdlp_resident_t *R, int mixed, int own_scal) {
  long long m = R->m, n = R->n, nnz = R->nnz;
  long long stride = R->A_nblk > R->At_nblk ? R->A_nblk : R->At_nblk;
  if (stride < 2048) stride = 2048;
  long long xl = dpdlp_resident_xpart_len(R);
  size_t b = 2 * piece(nnz, 8);
  if (xl > 0) b += piece(xl, 8);
  if (R->A_P > 1) b += piece(nnz, 8) + (mixed ? piece(nnz, 4) : 0);
  if (R->At_P > 1) b += piece(nnz, 8) + (mixed ? piece(nnz, 4) : 0);
  b += 13 * piece(n, 8) + 12 * piece(m, 8); /* vectors, scaling, work */
  b += piece((long long)DPDLP_SLOTS * stride, 8);
  if (own_scal) b += piece(DPDLP_SCAL_LEN, 8);
  if (mixed) b += 2 * piece(nnz, 4) + piece(n, 4) + piece(m, 4);
  return b;
}

int dpdlp_i_ctx_setup(ctx_t *c, dpdlp_resident_t *R, const pdhg_parameters_t *params, int mixed,
                      hipStream_t stream, double *scal_dev, double *scal_host,
                      double *rescale_time) {
  if (dpdlp_i_ctx_setup_issue(c, R, params, mixed, stream, scal_dev, scal_host)) return -1;
  if (hipStreamSynchronize(c->s) != hipSuccess) {
    dpdlp_set_error("device sync failed");
    return -1;
  }
  return dpdlp_i_ctx_setup_complete(c, R, params, rescale_time);
}

int dpdlp_i_ctx_setup_issue(ctx_t *c, dpdlp_resident_t *R, const pdhg_parameters_t *params,
                            int mixed, hipStream_t stream, double *scal_dev, double *scal_host) {
  int m = R->m, n = R->n, nnz = R->nnz;
  dpdlp_state_t *st = &c->st;
  c->mixed = mixed;
  if (stream) {
    c->s = stream;
  } else {
    HIPCHK(hipStreamCreate(&c->s));
    c->owns_stream = 1;
  }
  if (scal_host) {
    c->h_scal = scal_host;
  } else {
    HIPCHK(hipHostMalloc((void **)&c->h_scal, DPDLP_SCAL_LEN * sizeof(double), 0));
    c->owns_scal = 1;
  }
  /* one device arena for the LP's state, cleared once; a caller may hand in
   * a recycled arena that is large enough */
  {
    size_t need = dpdlp_i_ctx_footprint(R, mixed, scal_dev == NULL);
    if (!c->arena || c->arena_cap < need) {
      if (c->arena) hipFree(c->arena);
      c->arena_cap = need;
      if (hipMalloc((void **)&c->arena, c->arena_cap) != hipSuccess) {
        dpdlp_set_error("device allocation failed (%zu bytes of LP state)", c->arena_cap);
        c->arena = NULL;
        c->arena_cap = 0;
        goto fail;
      }
    }
    c->arena_off = 0;
    HIPCHK(hipMemsetAsync(c->arena, 0, need, c->s));
    c->arena_len = need; /* bump allocation stays within what was cleared */
  }

  /* ---- state (pdlp_core_op.cu:42-391) ---- */
  st->m = m;
  st->n = n;
  c->row_perm = R->h_row_perm;
  c->w_A_v = dcopy(c, R->A_v, nnz);
  c->w_At_v = dcopy(c, R->At_v, nnz);
  {
    long long xl = dpdlp_resident_xpart_len(R);
    double *xpart = xl > 0 ? (double *)dmalloc(c, (size_t)xl * sizeof(double)) : NULL;
    if (xl > 0 && !xpart) {
      dpdlp_set_error("device allocation failed (panel partials)");
      goto fail;
    }
    dpdlp_resident_csr(R, 0, &st->A, xpart);
    dpdlp_resident_csr(R, 1, &st->At, xpart);
    st->A.v = c->w_A_v;
    st->At.v = c->w_At_v;
    st->A.pv = st->A.P > 1 ? (double *)dmalloc(c, (size_t)(nnz ? nnz : 1) * sizeof(double)) : NULL;
    st->At.pv = st->At.P > 1 ? (double *)dmalloc(c, (size_t)(nnz ? nnz : 1) * sizeof(double)) : NULL;
    if ((st->A.P > 1 && !st->A.pv) || (st->At.P > 1 && !st->At.pv)) {
      dpdlp_set_error("device allocation failed (tiled values)");
      goto fail;
    }
  }
  /* c, l_v, u_v, l_c, u_c are laid out as in the resident's blob (same
   * order and padding), so one copy brings all five */
  st->c = (double *)dmalloc(c, (size_t)(n ? n : 1) * sizeof(double));
  st->lv = (double *)dmalloc(c, (size_t)(n ? n : 1) * sizeof(double));
  st->uv = (double *)dmalloc(c, (size_t)(n ? n : 1) * sizeof(double));
  st->lc = (double *)dmalloc(c, (size_t)(m ? m : 1) * sizeof(double));
  st->uc = (double *)dmalloc(c, (size_t)(m ? m : 1) * sizeof(double));
  if (st->c && st->lv && st->uv && st->lc && st->uc) {
    const char *rb = (const char *)R->c, *cb = (const char *)st->c;
    if ((const char *)R->uc - rb == (const char *)st->uc - cb &&
        (const char *)R->lv - rb == (const char *)st->lv - cb &&
        (const char *)R->uv - rb == (const char *)st->uv - cb &&
        (const char *)R->lc - rb == (const char *)st->lc - cb) {
      HIPCHK(hipMemcpyAsync(st->c, R->c, (size_t)((const char *)R->uc - rb) + (size_t)m * 8,
                            hipMemcpyDeviceToDevice, c->s));
    } else {
      HIPCHK(hipMemcpyAsync(st->c, R->c, (size_t)n * 8, hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(st->lv, R->lv, (size_t)n * 8, hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(st->uv, R->uv, (size_t)n * 8, hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(st->lc, R->lc, (size_t)m * 8, hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(st->uc, R->uc, (size_t)m * 8, hipMemcpyDeviceToDevice, c->s));
    }
  }
  st->D_r = (double *)dmalloc(c, (size_t)(m ? m : 1) * sizeof(double));
  st->D_c = (double *)dmalloc(c, (size_t)(n ? n : 1) * sizeof(double));
  st->x_init = dzero(c, n);
  st->x_cur = dzero(c, n);
  st->x_pdhg = dzero(c, n);
  st->x_ref = dzero(c, n);
  st->dual_product = dzero(c, n);
  st->dual_slack = dzero(c, n);
  st->dx = dzero(c, n);
  st->y_init = dzero(c, m);
  st->y_cur = dzero(c, m);
  st->y_pdhg = dzero(c, m);
  st->y_ref = dzero(c, m);
  st->primal_product = dzero(c, m);
  st->primal_slack = dzero(c, m);
  st->dy = dzero(c, m);
  st->part_stride = R->A_nblk;
  if (R->At_nblk > st->part_stride) st->part_stride = R->At_nblk;
  if (st->part_stride < 2048) st->part_stride = 2048;
  st->part = dzero(c, (long long)DPDLP_SLOTS * st->part_stride);
  st->scal = scal_dev ? scal_dev : dzero(c, DPDLP_SCAL_LEN);
  double *rs = dzero(c, m), *cs = dzero(c, n);
  double *pi_eig = dzero(c, m), *pi_next = dzero(c, m), *pi_dual = dzero(c, n);
  c->rs = rs;
  c->cs = cs;
  c->pi_eig = pi_eig;
  c->pi_next = pi_next;
  c->pi_dual = pi_dual;
  if (!c->w_A_v || !c->w_At_v || !st->c || !st->lv || !st->uv || !st->lc || !st->uc ||
      !st->D_r || !st->D_c || !st->x_init || !st->x_cur || !st->x_pdhg || !st->x_ref ||
      !st->dual_product || !st->dual_slack || !st->dx || !st->y_init || !st->y_cur ||
      !st->y_pdhg || !st->y_ref || !st->primal_product || !st->primal_slack || !st->dy ||
      !st->part || !st->scal || !rs || !cs || !pi_eig || !pi_next || !pi_dual) {
    dpdlp_set_error("device allocation failed");
    goto fail;
  }
  c->k.c0 = R->c0;
  c->k.obj_norm = params->optimality_norm == NORM_TYPE_L_INF ? R->obj_linf : R->obj_l2;
  c->k.bound_norm = params->optimality_norm == NORM_TYPE_L_INF ? R->bnd_linf : R->bnd_l2;

  /* ---- scaling on device (preconditioner.c:197-273) ---- */
  double t_sc0 = dpdlp_wall();
  if ((char *)st->D_c - (char *)st->D_r == (ptrdiff_t)pad256((size_t)m * 8)) {
    /* adjacent: one fill covers both (and the padding between them) */
    HIPCHK(dpdlp_dev_fill(st->D_r, 1.0, (int)(st->D_c - st->D_r) + n, c->s));
  } else {
    HIPCHK(dpdlp_dev_fill(st->D_r, 1.0, m, c->s));
    HIPCHK(dpdlp_dev_fill(st->D_c, 1.0, n, c->s));
  }
  for (int it = 0; it < params->l_inf_ruiz_iterations; ++it)
    HIPCHK(dpdlp_dev_ruiz_step(st, rs, cs, c->s));
  if (params->has_pock_chambolle_alpha)
    HIPCHK(dpdlp_dev_pock_chambolle(st, params->pock_chambolle_alpha, rs, cs, c->s));
  /* beta and gamma are formed on the device; the host reads them at the one
   * synchronisation that ends the scaling */
  if (params->bound_objective_rescaling) {
    HIPCHK(dpdlp_dev_bound_obj_sums(st, c->s));
    HIPCHK(dpdlp_dev_bound_obj_apply_scal(st, c->s));
  }
  /* scaled values in tile order */
  if (st->A.P > 1) HIPCHK(dpdlp_dev_permute(c->w_A_v, R->A_perm, (double *)st->A.pv, nnz, c->s));
  if (st->At.P > 1)
    HIPCHK(dpdlp_dev_permute(c->w_At_v, R->At_perm, (double *)st->At.pv, nnz, c->s));
  /* the scalars come back with the stream's next synchronisation */
  HIPCHK(hipMemcpyAsync(c->h_scal, st->scal, 64 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  c->t_issue = t_sc0;
  return 0;
fail:
  return -1;
}

int dpdlp_i_ctx_setup_complete(ctx_t *c, dpdlp_resident_t *R, const pdhg_parameters_t *params,
                               double *rescale_time) {
  int m = R->m, n = R->n, nnz = R->nnz;
  dpdlp_state_t *st = &c->st;
  if (params->bound_objective_rescaling) {
    c->k.beta = c->h_scal[DPDLP_SL_SCAL + 2];
    c->k.gamma = c->h_scal[DPDLP_SL_SCAL + 3];
  } else {
    c->k.beta = 1.0;
    c->k.gamma = 1.0;
  }
  *rescale_time = dpdlp_wall() - c->t_issue;

  /* warm start, pdlp_core_op.cu:174-201: the start vectors are staged in the
   * power-iteration work vectors, which that iteration overwrites before
   * reading them */
  if (R->h_pstart && n > 0) {
    HIPCHK(hipMemcpyAsync(c->pi_dual, R->h_pstart, (size_t)n * sizeof(double),
                          hipMemcpyHostToDevice, c->s));
    HIPCHK(dpdlp_dev_warm_start(st->x_init, st->x_cur, st->x_pdhg, c->pi_dual, st->D_c,
                                c->k.beta, n, c->s));
  }
  if (R->h_dstart && m > 0) {
    HIPCHK(hipMemcpyAsync(c->pi_next, R->h_dstart, (size_t)m * sizeof(double),
                          hipMemcpyHostToDevice, c->s));
    HIPCHK(dpdlp_dev_warm_start(st->y_init, st->y_cur, st->y_pdhg, c->pi_next, st->D_r,
                                c->k.gamma, m, c->s));
  }

  /* mixed precision: FP32 values and gather shadows */
  if (c->mixed) {
    c->w_A_v32 = (float *)dmalloc(c, (size_t)(nnz ? nnz : 1) * sizeof(float));
    c->w_At_v32 = (float *)dmalloc(c, (size_t)(nnz ? nnz : 1) * sizeof(float));
    st->x_ref32 = (float *)dmalloc(c, (size_t)(n ? n : 1) * sizeof(float));
    st->y_cur32 = (float *)dmalloc(c, (size_t)(m ? m : 1) * sizeof(float));
    if (!c->w_A_v32 || !c->w_At_v32 || !st->x_ref32 || !st->y_cur32) {
      dpdlp_set_error("device allocation failed (FP32 copies)");
      goto fail;
    }
    HIPCHK(dpdlp_dev_make_f32(c->w_A_v, c->w_A_v32, nnz, c->s));
    HIPCHK(dpdlp_dev_make_f32(c->w_At_v, c->w_At_v32, nnz, c->s));
    if (st->A.P > 1) {
      float *t = (float *)dmalloc(c, (size_t)(nnz ? nnz : 1) * sizeof(float));
      if (!t) goto fail;
      HIPCHK(dpdlp_dev_make_f32(st->A.pv, t, nnz, c->s));
      st->A.pv32 = t;
    }
    if (st->At.P > 1) {
      float *t = (float *)dmalloc(c, (size_t)(nnz ? nnz : 1) * sizeof(float));
      if (!t) goto fail;
      HIPCHK(dpdlp_dev_make_f32(st->At.pv, t, nnz, c->s));
      st->At.pv32 = t;
    }
    HIPCHK(dpdlp_dev_make_f32(st->y_cur, st->y_cur32, m, c->s));
    st->A.v32 = c->w_A_v32;
    st->At.v32 = c->w_At_v32;
    c->use32 = 1;
  }

  return 0;
fail:
  return -1;
}

void dpdlp_i_init_hstate(hstate_t *h, const ctx_t *c, const pdhg_parameters_t *params,
                         double step) {
  ctl_init_state(h, &c->k, params, step);
}

void dpdlp_i_result_scalars(cupdlpx_result_t *res, const ctx_t *c, const hstate_t *h,
                            double rescale_time, double polish_time, int polish_iters) {
  res->num_variables = c->st.n;
  res->num_constraints = c->st.m;
  res->num_nonzeros = c->st.A.nnz;
  res->total_count = h->total;
  res->rescaling_time_sec = rescale_time;
  res->cumulative_time_sec = h->t_cum;
  res->absolute_primal_residual = h->abs_p;
  res->relative_primal_residual = h->rel_p;
  res->absolute_dual_residual = h->abs_d;
  res->relative_dual_residual = h->rel_d;
  res->primal_objective_value = h->pobj;
  res->dual_objective_value = h->dobj;
  res->objective_gap = h->gap;
  res->relative_objective_gap = h->rel_gap;
  res->max_primal_ray_infeasibility = h->max_p_ray_inf;
  res->max_dual_ray_infeasibility = h->max_d_ray_inf;
  res->primal_ray_linear_objective = h->p_ray_lin_obj;
  res->dual_ray_objective = h->d_ray_obj;
  res->termination_reason = h->reason;
  res->feasibility_polishing_time = polish_time;
  res->feasibility_iteration = polish_iters;
}

int dpdlp_i_result_launch(ctx_t *c, double *x_host, double *y_host, double *rc_host) {
  dpdlp_state_t *st = &c->st;
  int m = st->m, n = st->n;
  /* unscaled vectors (pdlp_core_op.cu:789-859) into the LP's work vectors */
  double *xo = c->pi_dual, *rco = st->dual_slack, *yo = c->pi_next;
  if (n > 0 && st->At.nblk == 0) HIPCHK(hipMemsetAsync(rco, 0, (size_t)n * 8, c->s));
  HIPCHK(dpdlp_dev_result(st, c->k.beta, c->k.gamma, xo, yo, rco, c->s));
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(x_host, xo, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(rc_host, rco, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                          c->s));
  }
  if (m > 0)
    HIPCHK(hipMemcpyAsync(y_host, yo, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, c->s));
  return 0;
fail:
  return -1;
}

cupdlpx_result_t *dpdlp_i_ctx_result(ctx_t *c, const hstate_t *h, double rescale_time,
                                     double polish_time, int polish_iters) {
  dpdlp_state_t *st = &c->st;
  int m = st->m, n = st->n;
  cupdlpx_result_t *res = (cupdlpx_result_t *)calloc(1, sizeof(*res));
  if (!res) goto fail;
  res->primal_solution = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
  res->dual_solution = (double *)malloc((size_t)(m ? m : 1) * sizeof(double));
  res->reduced_cost = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
  if (!res->primal_solution || !res->dual_solution || !res->reduced_cost) goto fail;
  if (dpdlp_i_result_launch(c, res->primal_solution, res->dual_solution, res->reduced_cost))
    goto fail;
  HIPCHK(hipStreamSynchronize(c->s));
  if (c->row_perm && m > 0) { /* dual back to the caller's row order */
    double *t = (double *)malloc((size_t)m * sizeof(double));
    if (!t) goto fail;
    for (int i = 0; i < m; ++i) t[c->row_perm[i]] = res->dual_solution[i];
    memcpy(res->dual_solution, t, (size_t)m * sizeof(double));
    free(t);
  }
  dpdlp_i_result_scalars(res, c, h, rescale_time, polish_time, polish_iters);
  return res;
fail:
  cupdlpx_result_free(res);
  return NULL;
}

cupdlpx_result_t *dpdlp_solve_resident(dpdlp_resident_t *R, const pdhg_parameters_t *params,
                                       dpdlp_solve_opts_t *opts) {
  double t_begin = dpdlp_wall();
  ctx_t C;
  memset(&C, 0, sizeof(C));
  ctx_t *c = &C;
  cupdlpx_result_t *res = NULL;
  double *h_q = NULL, *d_start = NULL;
  dpdlp_solve_opts_t local_opts;
  if (!opts) {
    dpdlp_default_opts(&local_opts);
    opts = &local_opts;
  }
  opts->time_to_1e4_sec = -1.0;
  opts->time_to_1e6_sec = -1.0;
  opts->trace_recorded = 0;
  opts->iterations_fp32 = 0;
  if (!R || !params) {
    dpdlp_set_error("NULL problem or parameters");
    return NULL;
  }
  if (params->termination_evaluation_frequency <= 0) {
    dpdlp_set_error("termination_evaluation_frequency must be positive");
    return NULL;
  }
  if (opts->use_graph && R->A_P == 1 && R->At_P == 1 && !R->h_row_perm &&
      opts->trace_points <= 0 &&
      !opts->profile_phases) {
    /* the windowed engine with a single LP: per-iteration parameters live in
     * device memory, so each window of iterations replays one hipGraph */
    dpdlp_batch_opts_t bo;
    dpdlp_batch_default_opts(&bo);
    bo.max_lps = 1;
    bo.precision = opts->precision;
    bo.mixed_switch_tol = opts->mixed_switch_tol;
    bo.spmv_cap = R->cap;
    bo.spmv_cap_At = R->At_cap;
    bo.use_graph = 1;
    cupdlpx_result_t *r1 = NULL;
    if (dpdlp_solve_batch(&R, 1, params, &bo, &r1) != 1) return NULL;
    opts->total_time_sec = dpdlp_wall() - t_begin;
    opts->loop_time_sec = r1->cumulative_time_sec;
    if (!opts->quiet) dpdlp_print_final(r1, params);
    return r1;
  }
  int m = R->m, n = R->n, nnz = R->nnz;
  dpdlp_state_t *st = &c->st;
  hstate_t H;
  memset(&H, 0, sizeof(H));
  hstate_t *h = &H;
  c->mixed = (opts->precision == DPDLP_PREC_MIXED);

  if (opts->profile_phases) {
    HIPCHK(hipStreamCreate(&c->s));
    c->owns_stream = 1;
    for (int k = 0; k < 64; ++k) {
      for (int q = 0; q < 3; ++q) HIPCHK(hipEventCreate(&c->ev[k][q]));
      c->nev_created = k + 1;
    }
  }
  double rescale_time = 0.0;
  if (dpdlp_i_ctx_setup_issue(c, R, params, opts->precision == DPDLP_PREC_MIXED, c->s, NULL,
                              NULL))
    goto fail;
  /* the power iteration's start vector (utils.cu:80-82) is drawn on the host
   * while the device scales the problem */
  if (nnz != 0) {
    h_q = (double *)malloc((size_t)m * sizeof(double));
    if (!h_q) {
      dpdlp_set_error("host allocation failed");
      goto fail;
    }
    dpdlp_normal_fill(1u, m, h_q);
    if (R->h_row_perm) { /* draw i belongs to original row i */
      double *t = (double *)malloc((size_t)m * sizeof(double));
      if (!t) {
        dpdlp_set_error("host allocation failed");
        goto fail;
      }
      for (int i = 0; i < m; ++i) t[i] = h_q[R->h_row_perm[i]];
      free(h_q);
      h_q = t;
    }
  }
  if (hipStreamSynchronize(c->s) != hipSuccess) {
    dpdlp_set_error("device sync failed");
    goto fail;
  }
  if (dpdlp_i_ctx_setup_complete(c, R, params, &rescale_time)) goto fail;
  double *pi_eig = c->pi_eig, *pi_next = c->pi_next, *pi_dual = c->pi_dual;
  /* mixed precision: anchored FP32 shadows.  The FP32 SpMVs multiply
   * x_ref - x_init and y_cur - y_init, and the FP64 products of the anchors
   * (refreshed at every restart) are added back, so the FP32 rounding scales
   * with the distance from the last restart point instead of with the
   * iterates themselves. */
  if (c->mixed && n > 0 && m > 0) {
    st->ax0 = (double *)dmalloc(c, (size_t)m * sizeof(double));
    st->aty0 = (double *)dmalloc(c, (size_t)n * sizeof(double));
    if (!st->ax0 || !st->aty0) {
      dpdlp_set_error("device allocation failed (anchor products)");
      goto fail;
    }
    HIPCHK(dpdlp_dev_anchor_products(st, c->s));
  }

  /* ---- step size: power iteration (pdlp_core_op.cu:638-656) ---- */
  double t_pi0 = dpdlp_wall();
  int pi_iters = 0;
  if (nnz == 0) {
    h->step = 1.0;
  } else {
    HIPCHK(hipMemcpy(pi_eig, h_q, (size_t)m * sizeof(double), hipMemcpyHostToDevice));
    double sigma_sq = 1.0;
    int done = 0, batch = 8;
    if (params->sv_max_iter > 0) HIPCHK(dpdlp_dev_power_init(st, pi_eig, pi_next, c->s));
    while (!done && pi_iters < params->sv_max_iter) {
      int k = params->sv_max_iter - pi_iters;
      if (k > batch) k = batch;
      HIPCHK(dpdlp_dev_power_steps(st, pi_eig, pi_next, pi_dual, k, c->s));
      if (sync_scal(c, DPDLP_SC_PI_HIST + 4 * k)) goto fail;
      double r_first = 0.0, r_last = 0.0;
      for (int i = 0; i < k; ++i) {
        const double *hh = c->h_scal + DPDLP_SC_PI_HIST + 4 * i;
        sigma_sq = hh[0];
        pi_iters++;
        const double r = dpdlp_blue(hh + 1);
        if (i == 0) r_first = r;
        r_last = r;
        if (r < params->sv_tol) {
          done = 1;
          break;
        }
      }
      /* the next batch of speculative steps: doubling up to 32, or, once the
       * residuals fall geometrically, the steps they predict to the
       * tolerance plus two (steps past the first converged one are wasted;
       * sigma is always read at that step) */
      batch = batch * 2 > 32 ? 32 : batch * 2;
      if (!done && k >= 4 && r_last > 0.0 && r_first > r_last) {
        const double rate = log(r_last / r_first) / (k - 1); /* < 0 */
        const double need = log(params->sv_tol / r_last) / rate;
        if (need >= 0.0 && need < 30.0) batch = (int)need + 2;
      }
    }
    h->step = 0.998 / sqrt(sigma_sq);
  }
  dpdlp_i_init_hstate(h, c, params, h->step);
  double power_time = dpdlp_wall() - t_pi0;

  /* ---- main loop (solver.cu:56-102) ---- */
  if (params->verbose) dpdlp_print_table_header();
  double t0 = dpdlp_wall();
  int do_restart = 0, freq = params->termination_evaluation_frequency;
  const termination_criteria_t *tc = &params->termination_criteria;
  int trace_k = 0;
  while (h->total < tc->iteration_limit) {
    if (opts->trace_points > 0 && trace_k < opts->trace_points &&
        opts->trace_iters[trace_k] == h->total) {
      if (n > 0)
        HIPCHK(hipMemcpyAsync(opts->trace_x + (size_t)trace_k * n, st->x_cur,
                              (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->s));
      if (m > 0)
        HIPCHK(hipMemcpyAsync(opts->trace_y + (size_t)trace_k * m, st->y_cur,
                              (size_t)m * sizeof(double), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      if (R->h_row_perm && m > 0) {
        double *ty = opts->trace_y + (size_t)trace_k * m;
        double *t = (double *)malloc((size_t)m * sizeof(double));
        if (!t) goto fail;
        for (int i = 0; i < m; ++i) t[R->h_row_perm[i]] = ty[i];
        memcpy(ty, t, (size_t)m * sizeof(double));
        free(t);
      }
      trace_k++;
      opts->trace_recorded = trace_k;
    }
    int restart_check = h->is_major || h->total == 0;
    if (restart_check || h->total % dpdlp_print_frequency(h->total) == 0) {
      HIPCHK(dpdlp_dev_residual(st, NULL, 3, restart_check, c->s));
      int inf = h->is_major && h->total < 3 * freq;
      if (inf) HIPCHK(dpdlp_dev_infeasibility(st, c->s));
      if (sync_scal(c, 64)) {
        dpdlp_set_error("device sync failed");
        goto fail;
      }
      ctl_pending_fpe(h, c->h_scal, 0);
      ctl_residual(h, &c->k, c->h_scal, params->optimality_norm);
      if (inf) ctl_infeasibility(h, &c->k, c->h_scal);
      h->t_cum = dpdlp_wall() - t0;
      double kkt = fmax(h->rel_p, fmax(h->rel_d, h->rel_gap));
      if (opts->time_to_1e4_sec < 0 && kkt < 1e-4) opts->time_to_1e4_sec = h->t_cum;
      if (opts->time_to_1e6_sec < 0 && kkt < 1e-6) opts->time_to_1e6_sec = h->t_cum;
      if (c->use32 && !ctl_mixed_keep_fp32(h, kkt, opts->mixed_switch_tol,
                                           fmax(tc->eps_optimal_relative,
                                                tc->eps_feasible_relative),
                                           st->ax0 != NULL))
        c->use32 = 0;
      ctl_check_termination(h, tc);
      if (params->verbose)
        dpdlp_print_iteration(h->total, h->t_cum, h->pobj, h->dobj, h->abs_p, h->abs_d, h->gap,
                              h->rel_p, h->rel_d, h->rel_gap);
      if (h->reason != TERMINATION_REASON_UNSPECIFIED) break;
    }
    if (restart_check) {
      do_restart = ctl_should_restart(h, &params->restart_params, freq);
      if (do_restart) {
        const double *S = c->h_scal;
        ctl_pid_update(h, &params->restart_params, dpdlp_blue(S + DPDLP_SL_RESD + 5),
                       dpdlp_blue(S + DPDLP_SL_RESP + 4));
        HIPCHK(dpdlp_dev_restart_copy(st, 3, c->s));
        if (c->use32 && st->ax0) HIPCHK(dpdlp_dev_anchor_products(st, c->s));
        h->inner = 0;
        h->fpe_last = INFINITY;
      }
    }
    h->is_major = ((h->total + 1) % freq) == 0;
    HIPCHK(pdhg_step(c, st, h, params->reflection_coefficient, do_restart, c->use32, 0,
                     opts->profile_phases && (h->total & 7) == 0));
    if (c->use32) opts->iterations_fp32++;
    do_restart = 0;
    h->inner++;
    h->total++;
  }
  if (h->reason == TERMINATION_REASON_UNSPECIFIED) {
    h->reason = TERMINATION_REASON_ITERATION_LIMIT;
    HIPCHK(dpdlp_dev_residual(st, NULL, 3, 0, c->s));
    if (sync_scal(c, 64)) goto fail;
    ctl_pending_fpe(h, c->h_scal, 0);
    ctl_residual(h, &c->k, c->h_scal, params->optimality_norm);
    if (params->verbose)
      dpdlp_print_iteration(h->total, h->t_cum, h->pobj, h->dobj, h->abs_p, h->abs_d, h->gap,
                            h->rel_p, h->rel_d, h->rel_gap);
  }
  double loop_time = dpdlp_wall() - t0;

  double polish_time = 0.0;
  int polish_iters = 0;
  if (params->feasibility_polishing && h->reason != TERMINATION_REASON_DUAL_INFEASIBLE &&
      h->reason != TERMINATION_REASON_PRIMAL_INFEASIBLE) {
    if (dpdlp_i_feasibility_polish(c, h, params, &polish_time, &polish_iters)) {
      dpdlp_set_error("feasibility polishing failed on the device");
      goto fail;
    }
  }

  res = dpdlp_i_ctx_result(c, h, rescale_time, polish_time, polish_iters);
  if (!res) goto fail;

  opts->setup_time_sec = rescale_time;
  opts->power_time_sec = power_time;
  opts->loop_time_sec = loop_time;
  opts->power_iterations = pi_iters;
  opts->step_size = h->step;
  opts->phase_samples = c->nsamp;
  opts->primal_phase_ms = c->nsamp ? c->p1_ms / c->nsamp : 0.0;
  opts->dual_phase_ms = c->nsamp ? c->p2_ms / c->nsamp : 0.0;
  free(h_q);
  dpdlp_i_ctx_free(c);
  opts->total_time_sec = dpdlp_wall() - t_begin;
  if (!opts->quiet) dpdlp_print_final(res, params);
  return res;

fail:
  free(h_q);
  if (d_start) hipFree(d_start);
  cupdlpx_result_free(res);
  dpdlp_i_ctx_free(c);
  return NULL;
}

cupdlpx_result_t *dpdlp_solve_ex(const lp_problem_t *prob, const pdhg_parameters_t *params,
                                 dpdlp_solve_opts_t *opts) {
  if (!prob) return NULL;
  pdhg_parameters_t local;
  if (params)
    local = *params;
  else
    set_default_parameters(&local);
  dpdlp_print_initial_info(&local, prob);
  double t0 = dpdlp_wall();
  dpdlp_resident_t *R = dpdlp_upload(prob, opts ? opts->spmv_cap : 0);
  if (!R) return NULL;
  cupdlpx_result_t *r = dpdlp_solve_resident(R, &local, opts);
  dpdlp_resident_free(R);
  if (r && opts) opts->total_time_sec = dpdlp_wall() - t0;
  return r;
}
