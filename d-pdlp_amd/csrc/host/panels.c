/*
 * Host planning of the SpMV layouts: row blocks and the XCD column-panel
 * tiles (layout documented at dpdlp_csr_t in include/dpdlp_dev.h).
 */
#include "dpdlp_dev.h"

#include <stdlib.h>

/* Consecutive rows per block: at most DPDLP_BLOCK_ROWS rows and cap nonzeros; a row
 * longer than cap gets a block of its own. */
int dpdlp_plan_blocks(int rows, const int *rp, int cap, int *blk) {
  int nb = 0, r = 0;
  while (r < rows) {
    int start = r;
    long long len = (long long)rp[r + 1] - rp[r];
    if (len > cap) {
      r++;
    } else {
      long long nz = 0;
      while (r < rows && r - start < DPDLP_BLOCK_ROWS) {
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

int dpdlp_build_panels(int rows, int cols, const int *rp, const int *ci, const int *blk,
                       int nblk, int P, int *toff, unsigned short *rofs, int *perm, int *pci) {
  long long W = ((long long)cols + P - 1) / P;
  if (W < 1) W = 1;
  int *cnt = (int *)calloc((size_t)P + 1, sizeof(int));
  if (!cnt) return -1;
  int pos = 0;
  for (int b = 0; b < nblk; ++b) {
    int r0 = blk[b], r1 = blk[b + 1], R = r1 - r0;
    /* tile sizes of this block */
    for (int k = 0; k <= P; ++k) cnt[k] = 0;
    for (int e = rp[r0]; e < rp[r1]; ++e) cnt[(int)(ci[e] / W) + 1]++;
    int base[9];
    base[0] = pos;
    for (int k = 0; k < P; ++k) {
      base[k + 1] = base[k] + cnt[k + 1];
      toff[(size_t)b * P + k] = base[k];
      if (R > 1 && cnt[k + 1] > 65535) {
        free(cnt);
        return -1;
      }
    }
    int fill[8];
    for (int k = 0; k < P; ++k) fill[k] = base[k];
    unsigned short *ro = rofs + (size_t)P * (r0 + b);
    for (int i = r0; i < r1; ++i) {
      for (int k = 0; k < P; ++k) ro[(size_t)k * (R + 1) + (i - r0)] = (unsigned short)(fill[k] - base[k]);
      for (int e = rp[i]; e < rp[i + 1]; ++e) {
        int k = (int)(ci[e] / W);
        perm[fill[k]] = e;
        pci[fill[k]] = ci[e];
        fill[k]++;
      }
    }
    for (int k = 0; k < P; ++k) ro[(size_t)k * (R + 1) + R] = (unsigned short)(fill[k] - base[k]);
    pos = base[P];
  }
  toff[(size_t)nblk * P] = pos;
  free(cnt);
  return 0;
}

/* Row blocks for the panel-tiled SpMV: consecutive rows while the block has at
 * most rmax rows and every one of its P tiles at most cap entries; a row that
 * cannot share a tile (a panel count above cap) gets a block of its own. */
int dpdlp_plan_blocks_panels(int rows, int cols, const int *rp, const int *ci, int cap, int P,
                             int rmax, int *blk) {
  long long W = ((long long)cols + P - 1) / P;
  if (W < 1) W = 1;
  int tile[8], rowc[8];
  int nb = 0, r = 0;
  while (r < rows) {
    int start = r;
    for (int k = 0; k < P; ++k) tile[k] = 0;
    while (r < rows && r - start < rmax) {
      for (int k = 0; k < P; ++k) rowc[k] = 0;
      for (int e = rp[r]; e < rp[r + 1]; ++e) rowc[(int)(ci[e] / W)]++;
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
