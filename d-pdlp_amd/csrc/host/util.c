/*
 * Host utilities: thread-local error text, wall clock, stable transpose.
 */
#define _POSIX_C_SOURCE 200809L
#include "dpdlp_ext.h"
#include "dpdlp_host.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

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

/* Row pointers of A' (column counts of A, exclusive prefix). */
void dpdlp_host_col_ptr(int m, int n, const int *rp, const int *ci, int *t_rp) {
  memset(t_rp, 0, (size_t)(n + 1) * sizeof(int));
  for (int k = 0; k < rp[m]; ++k) t_rp[ci[k] + 1]++;
  for (int j = 0; j < n; ++j) t_rp[j + 1] += t_rp[j];
}

/* Counting-sort transpose; entries of each output row keep increasing source
 * row order (the layout csr2csc produces at pdlp_core_op.cu:111-128). */
int dpdlp_host_transpose(int m, int n, const int *rp, const int *ci, const double *v,
                         int *t_rp, int *t_ci, double *t_v) {
  int *cursor = (int *)malloc((size_t)(n + 1) * sizeof(int));
  if (!cursor) return -1;
  memset(t_rp, 0, (size_t)(n + 1) * sizeof(int));
  for (int i = 0; i < m; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) t_rp[ci[k] + 1]++;
  for (int j = 0; j < n; ++j) t_rp[j + 1] += t_rp[j];
  memcpy(cursor, t_rp, (size_t)(n + 1) * sizeof(int));
  for (int i = 0; i < m; ++i) {
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int dst = cursor[ci[k]]++;
      t_ci[dst] = i;
      t_v[dst] = v[k];
    }
  }
  free(cursor);
  return 0;
}
