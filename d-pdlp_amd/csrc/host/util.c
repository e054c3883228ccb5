/*
 * Host utilities: thread-local error text, wall clock, worker threads,
 * stable transpose.
 */
#define _GNU_SOURCE
#include "dpdlp_ext.h"
#include "dpdlp_host.h"

#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

static _Thread_local char g_err[512];

void dpdlp_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  fprintf(stderr, "[dpdlp] %s\n", g_err);
}

const char *dpdlp_last_error(void) { return g_err; }

double dpdlp_wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static __thread int tl_cap;

void dpdlp_host_threads_cap(int max_threads) { tl_cap = max_threads > 0 ? max_threads : 0; }

int dpdlp_host_threads(void) {
  if (tl_cap == 1) return 1;
  const char *env = getenv("DPDLP_HOST_THREADS");
  if (!env || !*env) env = getenv("OMP_NUM_THREADS");
  long t = env && *env ? atol(env) : 0;
  if (t < 1) {
    cpu_set_t set;
    t = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set)
                                                      : sysconf(_SC_NPROCESSORS_ONLN);
  }
  if (t < 1) t = 1;
  if (t > 32) t = 32;
  if (tl_cap > 0 && t > tl_cap) t = tl_cap;
  return (int)t;
}

typedef struct {
  void (*fn)(void *, int);
  void *ctx;
  int ntask;
  int next; /* next task to take (atomic) */
} par_job_t;

static void *par_worker(void *arg) {
  par_job_t *j = (par_job_t *)arg;
  for (;;) {
    int t = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
    if (t >= j->ntask) break;
    j->fn(j->ctx, t);
  }
  return NULL;
}

void dpdlp_par_for(int ntask, void (*fn)(void *ctx, int task), void *ctx) {
  if (ntask <= 0) return;
  par_job_t job = {fn, ctx, ntask, 0};
  int nth = dpdlp_host_threads();
  if (nth > ntask) nth = ntask;
  pthread_t th[32];
  int started = 0;
  for (int i = 1; i < nth; ++i) {
    if (pthread_create(&th[started], NULL, par_worker, &job) != 0) break;
    ++started;
  }
  par_worker(&job); /* the caller works too; a thread that failed to start is not needed */
  for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
}

/* ---- solution files: "%.10g\n" per value, formatted by worker threads
 * into one buffer per slice, written in order ---- */
typedef struct {
  const double *data;
  int size, nsl;
  char **buf;
  size_t *len;
} fmt_job_t;

static void fmt_slice(void *arg, int t) {
  fmt_job_t *J = (fmt_job_t *)arg;
  int i0 = (int)((long long)J->size * t / J->nsl), i1 = (int)((long long)J->size * (t + 1) / J->nsl);
  size_t cap = (size_t)(i1 - i0) * 24 + 1; /* "%.10g" is at most 17 characters */
  char *b = (char *)malloc(cap);
  if (!b) return;
  size_t o = 0;
  for (int i = i0; i < i1; ++i) o += (size_t)snprintf(b + o, cap - o, "%.10g\n", J->data[i]);
  J->buf[t] = b;
  J->len[t] = o;
}

int dpdlp_write_vector(const char *path, const double *data, int size) {
  if (!path || (size > 0 && !data) || size < 0) {
    dpdlp_set_error("dpdlp_write_vector: bad arguments");
    return -1;
  }
  FILE *f = fopen(path, "w");
  if (!f) {
    dpdlp_set_error("cannot open %s for writing", path);
    return -1;
  }
  int nsl = size / 65536 + 1, nth = dpdlp_host_threads();
  if (nsl > 4 * nth) nsl = 4 * nth;
  char **buf = (char **)calloc((size_t)nsl, sizeof(char *));
  size_t *len = (size_t *)calloc((size_t)nsl, sizeof(size_t));
  int rc = 0;
  if (!buf || !len) {
    rc = -1;
  } else {
    fmt_job_t J = {data, size, nsl, buf, len};
    dpdlp_par_for(nsl, fmt_slice, &J);
    for (int t = 0; t < nsl && rc == 0; ++t)
      if (!buf[t] || fwrite(buf[t], 1, len[t], f) != len[t]) rc = -1;
  }
  for (int t = 0; buf && t < nsl; ++t) free(buf[t]);
  free(buf);
  free(len);
  if (fclose(f) != 0) rc = -1;
  if (rc) dpdlp_set_error("writing %s failed", path);
  return rc;
}

/* Row pointers of A' (column counts of A, exclusive prefix).  Each task
 * owns a range of columns and counts the entries that fall in it. */
typedef struct {
  int n, ntask;
  long long nnz;
  const int *ci;
  int *t_rp;
} colcnt_job_t;

static void colcnt_task(void *arg, int t) {
  colcnt_job_t *J = (colcnt_job_t *)arg;
  int c0 = (int)((long long)J->n * t / J->ntask), c1 = (int)((long long)J->n * (t + 1) / J->ntask);
  unsigned w = (unsigned)(c1 - c0);
  int *cnt = J->t_rp + 1 + c0;
  for (long long k = 0; k < J->nnz; ++k) {
    unsigned d = (unsigned)(J->ci[k] - c0);
    if (d < w) cnt[d]++;
  }
}

void dpdlp_host_col_ptr(int m, int n, const int *rp, const int *ci, int *t_rp) {
  memset(t_rp, 0, (size_t)(n + 1) * sizeof(int));
  long long nnz = rp[m];
  int ntask = nnz < (1 << 20) ? 1 : dpdlp_host_threads();
  if (ntask > n) ntask = n > 0 ? n : 1;
  colcnt_job_t J = {n, ntask, nnz, ci, t_rp};
  dpdlp_par_for(ntask, colcnt_task, &J);
  for (int j = 0; j < n; ++j) t_rp[j + 1] += t_rp[j];
}

/* Counting-sort transpose; entries of each output row keep increasing source
 * row order (the layout csr2csc produces at pdlp_core_op.cu:111-128). */
/* the entries of columns [n t / ntask, n (t + 1) / ntask): a scan of A in
 * row order that keeps its own columns, so each A' row lists its entries in
 * increasing row order */
typedef struct {
  int m, n, ntask;
  const int *rp, *ci;
  const double *v;
  int *cursor, *t_ci;
  double *t_v;
} tr_job_t;

static void tr_task(void *arg, int t) {
  tr_job_t *J = (tr_job_t *)arg;
  int c0 = (int)((long long)J->n * t / J->ntask), c1 = (int)((long long)J->n * (t + 1) / J->ntask);
  unsigned w = (unsigned)(c1 - c0);
  int *cur = J->cursor + c0;
  for (int i = 0; i < J->m; ++i)
    for (int k = J->rp[i]; k < J->rp[i + 1]; ++k) {
      unsigned d = (unsigned)(J->ci[k] - c0);
      if (d < w) {
        int dst = cur[d]++;
        J->t_ci[dst] = i;
        J->t_v[dst] = J->v[k];
      }
    }
}

int dpdlp_host_transpose(int m, int n, const int *rp, const int *ci, const double *v,
                         int *t_rp, int *t_ci, double *t_v) {
  int *cursor = (int *)malloc((size_t)(n + 1) * sizeof(int));
  if (!cursor) return -1;
  dpdlp_host_col_ptr(m, n, rp, ci, t_rp);
  memcpy(cursor, t_rp, (size_t)(n + 1) * sizeof(int));
  int ntask = rp[m] < (1 << 20) ? 1 : dpdlp_host_threads();
  if (ntask > n) ntask = n > 0 ? n : 1;
  tr_job_t J = {m, n, ntask, rp, ci, v, cursor, t_ci, t_v};
  dpdlp_par_for(ntask, tr_task, &J);
  free(cursor);
  return 0;
}
