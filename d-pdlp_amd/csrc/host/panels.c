/*
 * Host planning of the SpMV layouts: row blocks and the XCD column-panel
 * tiles (layout documented at dpdlp_csr_t in include/dpdlp_dev.h).
 */
#include "dpdlp_dev.h"
#include "dpdlp_host.h"

#include <stdlib.h>

/* Consecutive rows per block: at most DPDLP_BLOCK_ROWS rows and cap nonzeros; a row
 * longer than cap gets a block of its own. */
int dpdlp_plan_blocks(int rows, const int *rp, int cap, int *blk) {
  return dpdlp_plan_blocks_rows(rows, rp, cap, DPDLP_BLOCK_ROWS, blk);
}

int dpdlp_plan_blocks_rows(int rows, const int *rp, int cap, int rmax, int *blk) {
  if (rmax < 1 || rmax > DPDLP_BLOCK_ROWS) rmax = DPDLP_BLOCK_ROWS;
  int nb = 0, r = 0;
  while (r < rows) {
    int start = r;
    long long len = (long long)rp[r + 1] - rp[r];
    if (len > cap) {
      r++;
    } else {
      long long nz = 0;
      while (r < rows && r - start < rmax) {
        long long l = (long long)rp[r + 1] - rp[r];
        if (l > cap || nz + l > cap) break;
        nz += l;
        r++;
      }
    }
    blk[nb++] = start;
  }
  blk[nb] = rows;
  return nb;
}

typedef struct {
  const int *rp, *ci, *blk;
  int nblk, P, ntask;
  long long W;
  int *toff, *perm, *pci;
  unsigned short *rofs;
  int bad; /* a multi-row tile above 65535 entries */
} bp_job_t;

/* the tiles of blocks [nblk * t / ntask, nblk * (t + 1) / ntask): block b's
 * tiles fill its own nonzero range [rp[blk[b]], rp[blk[b + 1]]) */
static void bp_task(void *arg, int t) {
  bp_job_t *J = (bp_job_t *)arg;
  const int P = J->P;
  const long long W = J->W;
  int b0 = (int)((long long)J->nblk * t / J->ntask), b1 = (int)((long long)J->nblk * (t + 1) / J->ntask);
  for (int b = b0; b < b1; ++b) {
    int r0 = J->blk[b], r1 = J->blk[b + 1], R = r1 - r0;
    int cnt[8] = {0}, base[9], fill[8];
    for (int e = J->rp[r0]; e < J->rp[r1]; ++e) cnt[(int)(J->ci[e] / W)]++;
    base[0] = J->rp[r0];
    for (int k = 0; k < P; ++k) {
      base[k + 1] = base[k] + cnt[k];
      J->toff[(size_t)b * P + k] = base[k];
      if (R > 1 && cnt[k] > 65535) __atomic_store_n(&J->bad, 1, __ATOMIC_RELAXED);
      fill[k] = base[k];
    }
    unsigned short *ro = J->rofs + (size_t)P * ((size_t)r0 + b);
    for (int i = r0; i < r1; ++i) {
      for (int k = 0; k < P; ++k) ro[(size_t)k * (R + 1) + (i - r0)] = (unsigned short)(fill[k] - base[k]);
      for (int e = J->rp[i]; e < J->rp[i + 1]; ++e) {
        int k = (int)(J->ci[e] / W);
        J->perm[fill[k]] = e;
        J->pci[fill[k]] = J->ci[e];
        fill[k]++;
      }
    }
    for (int k = 0; k < P; ++k) ro[(size_t)k * (R + 1) + R] = (unsigned short)(fill[k] - base[k]);
  }
}

int dpdlp_build_panels(int rows, int cols, const int *rp, const int *ci, const int *blk,
                       int nblk, int P, int *toff, unsigned short *rofs, int *perm, int *pci) {
  (void)rows;
  if (P < 1 || P > 8) return -1;
  long long W = ((long long)cols + P - 1) / P;
  if (W < 1) W = 1;
  int ntask = nblk < 1 ? 1 : (nblk < 4 * dpdlp_host_threads() ? nblk : 4 * dpdlp_host_threads());
  bp_job_t J = {rp, ci, blk, nblk, P, ntask, W, toff, perm, pci, rofs, 0};
  if (nblk > 0) dpdlp_par_for(ntask, bp_task, &J);
  toff[(size_t)nblk * P] = rp[blk[nblk]];
  return J.bad ? -1 : 0;
}

/* Row blocks for the panel-tiled SpMV: consecutive rows while the block has at
 * most rmax rows and every one of its P tiles at most cap entries; a row that
 * cannot share a tile (a panel count above cap) gets a block of its own. */
/* per-row panel counts, saturated at 65535 (above any cap) */
typedef struct {
  const int *rp, *ci;
  long long W;
  int P, rows, ntask;
  unsigned short *rc;
} pc_job_t;

static void row_panel_counts(const int *rp, const int *ci, long long W, int P, int r, int *c) {
  for (int k = 0; k < P; ++k) c[k] = 0;
  for (int e = rp[r]; e < rp[r + 1]; ++e) c[(int)(ci[e] / W)]++;
}

static void pc_task(void *arg, int t) {
  pc_job_t *J = (pc_job_t *)arg;
  int r0 = (int)((long long)J->rows * t / J->ntask), r1 = (int)((long long)J->rows * (t + 1) / J->ntask);
  int c[8];
  for (int r = r0; r < r1; ++r) {
    row_panel_counts(J->rp, J->ci, J->W, J->P, r, c);
    for (int k = 0; k < J->P; ++k) J->rc[(size_t)r * J->P + k] = (unsigned short)(c[k] > 65535 ? 65535 : c[k]);
  }
}

/* greedy row blocks over the rows' panel counts rc[r * P + k] (or counted
 * from rp/ci while planning when rc is NULL) */
static int plan_from_counts(int rows, const int *rp, const int *ci, long long W,
                            const unsigned short *rc, int cap, int P, int rmax, int *blk) {
  int tile[8], rowc[8];
  int nb = 0, r = 0;
  while (r < rows) {
    int start = r;
    for (int k = 0; k < P; ++k) tile[k] = 0;
    while (r < rows && r - start < rmax) {
      if (rc)
        for (int k = 0; k < P; ++k) rowc[k] = rc[(size_t)r * P + k];
      else
        row_panel_counts(rp, ci, W, P, r, rowc);
      int fits = 1, alone = 0;
      for (int k = 0; k < P; ++k) {
        if (rowc[k] > cap) alone = 1;
        if (tile[k] + rowc[k] > cap) fits = 0;
      }
      if (alone) {
        if (r == start) r++; /* a block of its own */
        break;
      }
      if (!fits) break;
      for (int k = 0; k < P; ++k) tile[k] += rowc[k];
      r++;
    }
    blk[nb++] = start;
  }
  blk[nb] = rows;
  return nb;
}

int dpdlp_plan_blocks_panels(int rows, int cols, const int *rp, const int *ci, int cap, int P,
                             int rmax, int *blk) {
  long long W = ((long long)cols + P - 1) / P;
  if (W < 1) W = 1;
  /* the rows' panel counts, counted by the worker threads (counted while
   * planning if there is no memory for them) */
  unsigned short *rc = rows > 0 ? (unsigned short *)malloc((size_t)rows * P * sizeof(unsigned short)) : NULL;
  if (rc) {
    pc_job_t J = {rp, ci, W, P, rows, 4 * dpdlp_host_threads(), rc};
    dpdlp_par_for(J.ntask, pc_task, &J);
  }
  int nb = plan_from_counts(rows, rp, ci, W, rc, cap, P, rmax, blk);
  free(rc);
  return nb;
}

/* Panel counts of the transpose A' (n rows) read from A (m x n, CSR)
 * without transposing: A' row j, panel k of A''s columns (A's rows
 * [k W, (k + 1) W)) counts the entries of A's column j in those rows.  The
 * worker threads add into 32-bit counters (relaxed atomics: the sums do not
 * depend on the order), saturated to 16 bits afterwards as the planner
 * wants. */
typedef struct {
  const int *rp, *ci;
  int m, n, P, ntask;
  long long W;
  unsigned *c32;
  unsigned short *rc;
} tc_job_t;

static void tc_count(void *arg, int t) {
  tc_job_t *J = (tc_job_t *)arg;
  int r0 = (int)((long long)J->m * t / J->ntask), r1 = (int)((long long)J->m * (t + 1) / J->ntask);
  for (int i = r0; i < r1; ++i) {
    int k = (int)(i / J->W);
    for (int e = J->rp[i]; e < J->rp[i + 1]; ++e)
      __atomic_fetch_add(&J->c32[(size_t)J->ci[e] * J->P + k], 1u, __ATOMIC_RELAXED);
  }
}

static void tc_saturate(void *arg, int t) {
  tc_job_t *J = (tc_job_t *)arg;
  size_t len = (size_t)J->n * J->P;
  size_t a = len * (size_t)t / (size_t)J->ntask, b = len * (size_t)(t + 1) / (size_t)J->ntask;
  for (size_t q = a; q < b; ++q) J->rc[q] = (unsigned short)(J->c32[q] > 65535u ? 65535u : J->c32[q]);
}

int dpdlp_plan_blocks_panels_T(int m, int n, const int *rp, const int *ci, int cap, int P,
                               int rmax, int *blk) {
  if (P < 1 || P > 8) return -1;
  long long W = ((long long)m + P - 1) / P;
  if (W < 1) W = 1;
  unsigned short *rc = (unsigned short *)malloc((size_t)(n > 0 ? n : 1) * P * sizeof(unsigned short));
  unsigned *c32 = (unsigned *)calloc((size_t)(n > 0 ? n : 1) * P, sizeof(unsigned));
  if (!rc || !c32) {
    free(rc);
    free(c32);
    return -1;
  }
  tc_job_t J = {rp, ci, m, n, P, 4 * dpdlp_host_threads(), W, c32, rc};
  dpdlp_par_for(J.ntask, tc_count, &J);
  dpdlp_par_for(J.ntask, tc_saturate, &J);
  free(c32);
  int nb = plan_from_counts(n, NULL, NULL, W, rc, cap, P, rmax, blk);
  free(rc);
  return nb;
}
