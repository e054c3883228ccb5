/*
 * gfx950 kernels of the single-LP PDHG path and their extern "C" launchers
 * (declared in include/dpdlp_dev.h).
 *
 * SpMV: "CSR-stream" row blocks in natural row order.  A workgroup of 256
 * threads (4 wave64) owns a block of <= 256 consecutive rows holding <= CAP
 * nonzeros; it streams the block's values and column indices with coalesced
 * non-temporal loads, multiplies by the gathered vector, stages the products
 * in LDS, and reduces each row with a power-of-two lane group sized to the
 * block (rows/lanes fixed per block, so sums are deterministic).  A row longer
 * than CAP owns a block and is reduced by the whole workgroup.  Because rows
 * stay in natural order, the fused epilogues (projection, reflection, Halpern,
 * residuals) read and write their row vectors coalesced.
 *
 * Every reduction writes per-workgroup partials that k_finalize combines in a
 * fixed order, so results do not depend on scheduling.  Elementwise arithmetic
 * is compiled without FMA contraction so it rounds exactly like the FP64
 * oracle; only SpMV and reduction summation order differ from it.
 */
#include <stdlib.h>

#include "dev_common.h"
#include "dpdlp_dev.h"

/* a row block's rows map one-to-one onto a workgroup's threads */
static_assert(DPDLP_BLOCK_ROWS == DPDLP_BS, "row-block plan and SpMV workgroup disagree");

static inline int ew_grid(long long n) { return dpdlp_ew_grid(n); }

#include "spmv_core.h"

template <typename VT, typename XT, class Epi, int CAP>
__global__ __launch_bounds__(DPDLP_BS) void k_spmv(
    const int *__restrict__ blk, const int *__restrict__ rp,
    const int *__restrict__ ci, const VT *__restrict__ val,
    const XT *__restrict__ x, Epi epi, double *__restrict__ part, int stride,
    int slot0) {
  constexpr int NS = Epi::NS, NM = Epi::NM;
  __shared__ double prod[CAP];
  __shared__ int srp[DPDLP_BS + 1];
  __shared__ double ssum[DPDLP_BS];
  __shared__ double red[(NS + NM > 0 ? NS + NM : 1) * 4];
  double s[NS > 0 ? NS : 1];
  double mx[NM > 0 ? NM : 1];
#pragma unroll
  for (int k = 0; k < (NS > 0 ? NS : 1); ++k) s[k] = 0.0;
#pragma unroll
  for (int k = 0; k < (NM > 0 ? NM : 1); ++k) mx[k] = 0.0;
  spmv_block<VT, XT, CAP>(rp, ci, val, x, blk[blockIdx.x], blk[blockIdx.x + 1], epi, s, mx, prod,
                          srp, ssum);
  if (NS + NM > 0) {
    if (epi.partials) block_partials<NS, NM>(s, mx, red, part, stride, slot0);
  }
}

/* Column-panel tile SpMV, pass 1: workgroup w = b*P + k sums the entries of
 * row block b that lie in column panel k and writes one partial per row to
 * xpart[k*rows + row].  Same staging and row reduction as k_spmv. */
template <typename VT, typename XT, int CAP>
__global__ __launch_bounds__(DPDLP_BS) void k_xp_partial(
    const int *__restrict__ blk, const int *__restrict__ toff,
    const unsigned short *__restrict__ rofs, const int *__restrict__ pci,
    const VT *__restrict__ pv, const XT *__restrict__ x, double *__restrict__ xpart,
    int rows, int P) {
  constexpr int UNR = 8;
  __shared__ double prod[CAP];
  __shared__ int srp[DPDLP_PANEL_RMAX + 1];
  __shared__ double ssum[4];
  const int tid = threadIdx.x;
  const int w = blockIdx.x, b = w / P, k = w - b * P;
  const int r0 = blk[b], r1 = blk[b + 1];
  const int R = r1 - r0;
  const int t0 = toff[w], t1 = toff[w + 1];
  const int nz = t1 - t0;
  double *out = xpart + (size_t)k * rows;
  if (R == 1) {
    double acc = row_sum_wg<VT, XT>(pci, pv, x, t0, t1);
    acc = wave_sum(acc);
    if ((tid & 63) == 0) ssum[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) out[r0] = (ssum[0] + ssum[1]) + (ssum[2] + ssum[3]);
    return;
  }
  const unsigned short *ro = rofs + (size_t)P * ((size_t)r0 + b) + (size_t)k * (R + 1);
  for (int t = tid; t <= R; t += DPDLP_BS) srp[t] = ro[t];
  for (int base = tid; base < nz; base += DPDLP_BS * UNR) {
    int cidx[UNR];
    VT vv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      int q = base + u * DPDLP_BS;
      if (q < nz) {
        cidx[u] = ntload(pci + t0 + q);
        vv[u] = ntload(pv + t0 + q);
      }
    }
    XT xx[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      int q = base + u * DPDLP_BS;
      if (q < nz) xx[u] = x[cidx[u]];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      int q = base + u * DPDLP_BS;
      if (q < nz) prod[q] = (double)vv[u] * (double)xx[u];
    }
  }
  __syncthreads();
  rows_reduce(prod, srp, R, [&](int r, double v) { out[r0 + r] = v; });
}

/* Column-panel tile SpMV, pass 2: row sums in panel order, then the fused
 * epilogue and its per-workgroup partial reductions. */
template <class Epi>
__global__ __launch_bounds__(DPDLP_BS) void k_xp_epi(const double *__restrict__ xpart, int rows,
                                                    int P, Epi epi, double *__restrict__ part,
                                                    int stride, int slot0) {
  constexpr int NS = Epi::NS, NM = Epi::NM;
  __shared__ double red[(NS + NM > 0 ? NS + NM : 1) * 4];
  double s[NS > 0 ? NS : 1];
  double mx[NM > 0 ? NM : 1];
#pragma unroll
  for (int q = 0; q < (NS > 0 ? NS : 1); ++q) s[q] = 0.0;
#pragma unroll
  for (int q = 0; q < (NM > 0 ? NM : 1); ++q) mx[q] = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < rows;
       i += (long long)gridDim.x * blockDim.x) {
    double a = xpart[i];
    for (int k = 1; k < P; ++k) a += xpart[(size_t)k * rows + i];
    epi((int)i, a, s, mx);
  }
  if (NS + NM > 0) {
    if (epi.partials) block_partials<NS, NM>(s, mx, red, part, stride, slot0);
  }
}

/* Column-panel tile SpMV in one pass.  Workgroup w = b*P + k sums the entries
 * of row block b that lie in panel k (as k_xp_partial) and writes its partial
 * row sums to xpart[k*rows + row]; then it takes a ticket on row block b.
 * Each launch adds exactly P tickets per block, so the counter, which is never
 * reset, reads (P-1) mod P for exactly one workgroup of the block: the last
 * to finish.  That workgroup adds the P partials of every row of the block in
 * panel order 0..P-1 (the same sums, in the same order, as k_xp_epi) and runs
 * the epilogue, whose per-workgroup reduction partials go to index b.  No
 * workgroup waits for another, and the result does not depend on which one
 * arrives last.  The agent-scope fences make the partials written on the
 * other XCDs (whose L2s are not coherent with this one's) visible: release
 * before the ticket, acquire after it. */
template <typename VT, typename XT, class Epi, int CAP>
__global__ __launch_bounds__(DPDLP_BS) void k_xp_fused(
    const int *__restrict__ blk, const int *__restrict__ toff,
    const unsigned short *__restrict__ rofs, const int *__restrict__ pci,
    const VT *__restrict__ pv, const XT *__restrict__ x, double *xpart,
    unsigned *__restrict__ tick, int rows, int P, Epi epi, double *__restrict__ part,
    int stride, int slot0) {
  constexpr int UNR = 8;
  constexpr int NS = Epi::NS, NM = Epi::NM;
  __shared__ double prod[CAP];
  __shared__ int srp[DPDLP_PANEL_RMAX + 1];
  __shared__ double ssum[4];
  __shared__ double red[(NS + NM > 0 ? NS + NM : 1) * 4];
  __shared__ int is_last;
  const int tid = threadIdx.x;
  const int w = blockIdx.x, b = w / P, k = w - b * P;
  const int r0 = blk[b], r1 = blk[b + 1];
  const int R = r1 - r0;
  const int t0 = toff[w], t1 = toff[w + 1];
  const int nz = t1 - t0;
  double *out = xpart + (size_t)k * rows;
  if (R == 1) {
    double acc = row_sum_wg<VT, XT>(pci, pv, x, t0, t1);
    acc = wave_sum(acc);
    if ((tid & 63) == 0) ssum[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) out[r0] = (ssum[0] + ssum[1]) + (ssum[2] + ssum[3]);
  } else {
    const unsigned short *ro = rofs + (size_t)P * ((size_t)r0 + b) + (size_t)k * (R + 1);
    for (int t = tid; t <= R; t += DPDLP_BS) srp[t] = ro[t];
    for (int base = tid; base < nz; base += DPDLP_BS * UNR) {
      int cidx[UNR];
      VT vv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        int q = base + u * DPDLP_BS;
        if (q < nz) {
          cidx[u] = ntload(pci + t0 + q);
          vv[u] = ntload(pv + t0 + q);
        }
      }
      XT xx[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        int q = base + u * DPDLP_BS;
        if (q < nz) xx[u] = x[cidx[u]];
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        int q = base + u * DPDLP_BS;
        if (q < nz) prod[q] = (double)vv[u] * (double)xx[u];
      }
    }
    __syncthreads();
    rows_reduce(prod, srp, R, [&](int r, double v) { out[r0 + r] = v; });
  }
  __threadfence(); /* release this workgroup's partial row sums */
  __syncthreads();
  if (tid == 0) {
    const unsigned old = atomicAdd(tick + b, 1u);
    is_last = ((old + 1u) % (unsigned)P) == 0u;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence(); /* acquire the other panels' partials */
  double s[NS > 0 ? NS : 1];
  double mx[NM > 0 ? NM : 1];
#pragma unroll
  for (int q = 0; q < (NS > 0 ? NS : 1); ++q) s[q] = 0.0;
#pragma unroll
  for (int q = 0; q < (NM > 0 ? NM : 1); ++q) mx[q] = 0.0;
  for (int r = tid; r < R; r += DPDLP_BS) {
    const int i = r0 + r;
    double a = xpart[i];
    for (int q = 1; q < P; ++q) a += xpart[(size_t)q * rows + i];
    epi(i, a, s, mx);
  }
  if (NS + NM > 0) {
    if (epi.partials) block_partials_at<NS, NM>(s, mx, red, part, stride, slot0, b);
  }
}

/* dst[t] = src[perm[t]] */
template <typename T>
__global__ void k_permute(const T *__restrict__ src, const int *__restrict__ perm,
                          T *__restrict__ dst, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

/* The one-pass tiled SpMV runs when the matrix has tickets and the
 * environment sets DPDLP_TILE_ONEPASS=1; otherwise the two-kernel form runs.
 * Both give the same row sums.  The two-kernel form is the default: on cfg1's
 * A (4 panels) the one-pass kernel took 282-288 us against 167 us, because
 * the agent-scope release every workgroup needs (buffer_wbl2) writes back
 * its XCD's L2 and costs more than the partial-sum round trip it saves. */
static bool tile_onepass(const dpdlp_csr_t *A) {
  static int enabled = -1;
  if (enabled < 0) {
    const char *e = getenv("DPDLP_TILE_ONEPASS");
    enabled = (e && e[0] == '1') ? 1 : 0;
  }
  return A->P > 1 && A->tick != nullptr && enabled;
}

/* partial-sum count written by an SpMV with a reducing epilogue */
static inline int spmv_nparts(const dpdlp_csr_t *A) {
  if (A->P > 1) return tile_onepass(A) ? A->nblk : ew_grid(A->rows);
  return A->nblk;
}

template <typename VT, typename XT, class Epi>
static hipError_t launch_spmv_t(const dpdlp_csr_t *A, const int *ci, const VT *v, const VT *pv,
                                const XT *x, const Epi &epi, double *part, int stride,
                                int slot0, hipStream_t s) {
  if (A->rows <= 0) return hipSuccess;
  if (tile_onepass(A)) {
    int grid = A->nblk * A->P;
    switch (A->cap) {
    case 1024:
      k_xp_fused<VT, XT, Epi, 1024><<<grid, DPDLP_BS, 0, s>>>(
          A->blk, A->toff, A->rofs, A->pci, pv, x, A->xpart, A->tick, A->rows, A->P, epi, part,
          stride, slot0);
      break;
    case 4096:
      k_xp_fused<VT, XT, Epi, 4096><<<grid, DPDLP_BS, 0, s>>>(
          A->blk, A->toff, A->rofs, A->pci, pv, x, A->xpart, A->tick, A->rows, A->P, epi, part,
          stride, slot0);
      break;
    default:
      k_xp_fused<VT, XT, Epi, 2048><<<grid, DPDLP_BS, 0, s>>>(
          A->blk, A->toff, A->rofs, A->pci, pv, x, A->xpart, A->tick, A->rows, A->P, epi, part,
          stride, slot0);
      break;
    }
    return hipGetLastError();
  }
  if (A->P > 1) {
    int grid = A->nblk * A->P;
    switch (A->cap) {
    case 1024:
      k_xp_partial<VT, XT, 1024><<<grid, DPDLP_BS, 0, s>>>(A->blk, A->toff, A->rofs, A->pci, pv,
                                                          x, A->xpart, A->rows, A->P);
      break;
    case 4096:
      k_xp_partial<VT, XT, 4096><<<grid, DPDLP_BS, 0, s>>>(A->blk, A->toff, A->rofs, A->pci, pv,
                                                          x, A->xpart, A->rows, A->P);
      break;
    default:
      k_xp_partial<VT, XT, 2048><<<grid, DPDLP_BS, 0, s>>>(A->blk, A->toff, A->rofs, A->pci, pv,
                                                          x, A->xpart, A->rows, A->P);
      break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_xp_epi<Epi><<<ew_grid(A->rows), DPDLP_BS, 0, s>>>(A->xpart, A->rows, A->P, epi, part,
                                                        stride, slot0);
    return hipGetLastError();
  }
  switch (A->cap) {
  case 1024:
    k_spmv<VT, XT, Epi, 1024><<<A->nblk, DPDLP_BS, 0, s>>>(A->blk, A->rp, ci, v, x, epi, part,
                                                          stride, slot0);
    break;
  case 4096:
    k_spmv<VT, XT, Epi, 4096><<<A->nblk, DPDLP_BS, 0, s>>>(A->blk, A->rp, ci, v, x, epi, part,
                                                          stride, slot0);
    break;
  default:
    k_spmv<VT, XT, Epi, 2048><<<A->nblk, DPDLP_BS, 0, s>>>(A->blk, A->rp, ci, v, x, epi, part,
                                                          stride, slot0);
    break;
  }
  return hipGetLastError();
}

template <class Epi>
static hipError_t launch_spmv(const dpdlp_csr_t *A, const double *x, const Epi &epi,
                              double *part, int stride, int slot0, hipStream_t s) {
  return launch_spmv_t<double, double, Epi>(A, A->ci, A->v, A->pv, x, epi, part, stride, slot0,
                                            s);
}

template <class Epi>
static hipError_t launch_spmv32(const dpdlp_csr_t *A, const float *x, const Epi &epi,
                                double *part, int stride, int slot0, hipStream_t s) {
  return launch_spmv_t<float, float, Epi>(A, A->ci, A->v32, A->pv32, x, epi, part, stride,
                                          slot0, s);
}

/* ------------------------------------------------------------------------ */
/* Fixed-order finalize: one workgroup per slot                              */
/* ------------------------------------------------------------------------ */

/* The sums of K partial slots slot0.. (nparts partials each), by one
 * workgroup, returned to every thread in tot[].  Thread t adds partials
 * t, t + 256, ... in that order, each wavefront sums its lanes with the
 * fixed reduction (wave_sum) and the four wavefront sums are combined as
 * (w0 + w1) + (w2 + w3): k_finalize's order, so a kernel that totals its
 * inputs itself gets the bits a separate k_finalize launch would give.
 * red: 4 K doubles of LDS. */
template <int K>
__device__ __forceinline__ void wg_slot_sums(const double *__restrict__ part, int stride,
                                             int nparts, int slot0, double *red, double *tot) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < nparts; i += DPDLP_BS) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += part[(size_t)(slot0 + k) * stride + i];
  }
  __syncthreads(); /* red may still be read by an earlier call */
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double w = wave_sum(acc[k]);
    if ((threadIdx.x & 63) == 0) red[4 * k + (threadIdx.x >> 6)] = w;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k)
    tot[k] = (red[4 * k] + red[4 * k + 1]) + (red[4 * k + 2] + red[4 * k + 3]);
}

/* maxmask bit k set: slot k is a maximum, else a sum.  Writes
 * scal[dst0 + k] and, when dst1 >= 0 / dst2 >= 0, copies to those too. */
__global__ __launch_bounds__(DPDLP_BS) void k_finalize(const double *__restrict__ part,
                                                      int stride, int nparts, int slot0,
                                                      unsigned maxmask, double *scal,
                                                      int dst0, int dst1, int dst2) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  const bool ismax = (maxmask >> k) & 1u;
  double v;
  if (ismax) {
    const double *p = part + (size_t)(slot0 + k) * stride;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += DPDLP_BS) acc = fmax(acc, p[i]);
    acc = wave_max(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  } else {
    wg_slot_sums<1>(part, stride, nparts, slot0 + k, red, &v);
  }
  if (threadIdx.x == 0) {
    scal[dst0 + k] = v;
    if (dst1 >= 0) scal[dst1 + k] = v;
    if (dst2 >= 0) scal[dst2 + k] = v;
  }
}

/* Two groups of sums in one launch: workgroups [0, cnt1) total slots
 * slot1.. over n1 partials into scal[slot1 + k] and scal[hist1 + k], the
 * next cnt2 total slots slot2.. over n2 partials into scal[slot2 + k] and
 * scal[hist2 + k]; each in k_finalize's order. */
__global__ __launch_bounds__(DPDLP_BS) void k_finalize_pair(const double *__restrict__ part,
                                                           int stride, double *scal, int n1,
                                                           int slot1, int cnt1, int hist1, int n2,
                                                           int slot2, int hist2,
                                                           unsigned max1 = 0u,
                                                           unsigned max2 = 0u) {
  __shared__ double red[4];
  const bool first = (int)blockIdx.x < cnt1;
  const int k = first ? (int)blockIdx.x : (int)blockIdx.x - cnt1;
  const int slot = (first ? slot1 : slot2) + k;
  const int nparts = first ? n1 : n2;
  double v;
  if (((first ? max1 : max2) >> k) & 1u) {
    /* a maximum, as k_finalize takes it */
    const double *p = part + (size_t)slot * stride;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += DPDLP_BS) acc = fmax(acc, p[i]);
    acc = wave_max(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  } else {
    wg_slot_sums<1>(part, stride, nparts, slot, red, &v);
  }
  if (threadIdx.x == 0) {
    scal[slot] = v;
    const int h = first ? hist1 : hist2;
    if (h >= 0) scal[h + k] = v;
  }
}

static hipError_t finalize(const dpdlp_state_t *st, int nparts, int slot0, int count,
                           unsigned maxmask, int dst0, int dst1, int dst2,
                           hipStream_t s) {
  k_finalize<<<count, DPDLP_BS, 0, s>>>(st->part, st->part_stride, nparts, slot0,
                                        maxmask, st->scal, dst0, dst1, dst2);
  return hipGetLastError();
}

/* ------------------------------------------------------------------------ */
/* Elementwise kernels                                                       */
/* ------------------------------------------------------------------------ */

/* restart copies (pdlp_core_op.cu:621-632 and :1290-1310) */
__global__ void k_restart_copy(double *x_init, double *x_cur, const double *x_pdhg,
                               double *y_init, double *y_cur, const double *y_pdhg,
                               float *y_cur32, int anchored, int n, int m, int part) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)n + m;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < n) {
      if (part & 1) {
        double v = x_pdhg[i];
        x_init[i] = v;
        x_cur[i] = v;
      }
    } else {
      int k = (int)(i - n);
      if (part & 2) {
        double v = y_pdhg[k];
        y_init[k] = v;
        y_cur[k] = v;
        if (y_cur32) y_cur32[k] = anchored ? 0.0f : (float)v; /* y_cur = y_init */
      }
    }
  }
}

/* y_cur32 = y_cur - y_init (anchored shadows) */
__global__ void k_delta_f32(const double *a, const double *b, float *out, int count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (float)(a[i] - b[i]);
}

/* ray projections (utils.cu:533-559) + max|dx|, max|dy| partials */
__global__ __launch_bounds__(DPDLP_BS) void k_inf_project(
    double *dx, const double *lv, const double *uv, int n, double *dy,
    const double *lc, const double *uc, int m, double *part, int stride, int slot0) {
  __shared__ double red[2 * 4];
  double s[1] = {0.0};
  double mx[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)n + m;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < n) {
      double d = dx[i];
      if (isfinite(lv[i])) d = fmax(d, 0.0);
      if (isfinite(uv[i])) d = fmin(d, 0.0);
      dx[i] = d;
      mx[0] = fmax(mx[0], fabs(d));
    } else {
      int k = (int)(i - n);
      double d = dy[k];
      if (!isfinite(lc[k])) d = fmin(d, 0.0);
      if (!isfinite(uc[k])) d = fmax(d, 0.0);
      dy[k] = d;
      mx[1] = fmax(mx[1], fabs(d));
    }
  }
  block_partials<0, 2>(s, mx, red, part, stride, slot0);
}

/* dx *= 1/max|dx| when positive (utils.cu:731-737) */
__global__ void k_scale_by_inv(double *v, int n, const double *scal, int idx) {
  double a = scal[idx];
  if (!(a > 0.0)) return;
  double sc = 1.0 / a;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    v[i] *= sc;
}

/* dot(a, b) partial into one slot */
__global__ __launch_bounds__(DPDLP_BS) void k_dot(const double *a, const double *b, int n,
                                                 double *part, int stride, int slot0) {
  __shared__ double red[4];
  double s[1] = {0.0};
  double mx[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s[0] += a[i] * b[i];
  block_partials<1, 0>(s, mx, red, part, stride, slot0);
}

/* sum of dual-objective contributions max(y,0) l_fin + min(y,0) u_fin */
__global__ __launch_bounds__(DPDLP_BS) void k_dual_contrib_sum(const double *y, const double *lc,
                                                              const double *uc, int m, double *part,
                                                              int stride, int slot0) {
  __shared__ double red[4];
  double s[1] = {0.0};
  double mx[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (long long)gridDim.x * blockDim.x) {
    double yv = y[i];
    s[0] += fmax(yv, 0.0) * lfin(lc[i]) + fmin(yv, 0.0) * lfin(uc[i]);
  }
  block_partials<1, 0>(s, mx, red, part, stride, slot0);
}

/* Blue partials of a vector (power iteration start) */
__global__ __launch_bounds__(DPDLP_BS) void k_nrm_partials(const double *v, int n, double *part,
                                                          int stride, int slot0) {
  __shared__ double red[3 * 4];
  double s[3] = {0.0, 0.0, 0.0};
  double mx[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    blue_acc(v[i], s[0], s[1], s[2]);
  block_partials<3, 0>(s, mx, red, part, stride, slot0);
}

/* next = eig / |eig|, utils.cu:124-132 */
__global__ void k_pow_normalize(const double *eig, double *next, int m, const double *scal,
                                int nrm_slot) {
  double nrm = blue_combine(scal[nrm_slot], scal[nrm_slot + 1], scal[nrm_slot + 2]);
  double inv = 1.0 / nrm;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (long long)gridDim.x * blockDim.x)
    next[i] = eig[i] * inv;
}

/* residual z - sigma^2 q (utils.cu:145-153), then next = z / |z| for the
 * following step (both read the old next first).  The four totals of the
 * A-product's slots (DPDLP_SL_PI: <q, z>, Blue sums of |z|) are read from
 * scal, or, when pow_nparts > 0, summed here from their pow_nparts partials
 * (every workgroup the same bits; workgroup 0 stores them at DPDLP_SL_PI and
 * pow_hist).  When prev_slot >= 0 the previous step's residual partials
 * (one per workgroup of that launch, same grid) are totalled the same way
 * into scal[prev_slot] and scal[prev_hist].  This step's residual partials
 * go to res_slot, which must differ from prev_slot. */
__global__ __launch_bounds__(DPDLP_BS) void k_pow_update(const double *eig, double *next, int m,
                                                        double *scal, double *part, int stride,
                                                        int pow_nparts, int pow_hist,
                                                        int res_slot, int prev_slot,
                                                        int prev_hist) {
  __shared__ double red[4 * 4];
  double s[3] = {0.0, 0.0, 0.0};
  double mx[1] = {0.0};
  double t[4];
  if (pow_nparts > 0) {
    wg_slot_sums<4>(part, stride, pow_nparts, DPDLP_SL_PI, red, t);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) {
        scal[DPDLP_SL_PI + k] = t[k];
        scal[pow_hist + k] = t[k];
      }
  } else {
    for (int k = 0; k < 4; ++k) t[k] = scal[DPDLP_SL_PI + k];
  }
  if (prev_slot >= 0) {
    double r[3];
    wg_slot_sums<3>(part, stride, (int)gridDim.x, prev_slot, red, r);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < 3; ++k) {
        scal[prev_slot + k] = r[k];
        scal[prev_hist + k] = r[k];
      }
  }
  __syncthreads(); /* red is reused by block_partials */
  double neg = -t[0];
  double nrm = blue_combine(t[1], t[2], t[3]);
  double inv = 1.0 / nrm;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (long long)gridDim.x * blockDim.x) {
    double e = eig[i];
    double r = next[i] * neg + e;
    blue_acc(r, s[0], s[1], s[2]);
    next[i] = e * inv;
  }
  block_partials<3, 0>(s, mx, red, part, stride, res_slot);
}

/* A power-iteration step's A' product with the previous step's update
 * folded in (utils.cu:124-153).  Every workgroup first totals the previous
 * A product's partials (<q, z> = sigma^2 and the Blue sums of |z|), or reads
 * them from scal when pow_nparts == 0; workgroup 0 stores them at
 * DPDLP_SL_PI and pow_hist.  When prev_slot >= 0 the residual partials of the
 * update before (prev_nparts of them) are totalled the same way into
 * scal[prev_slot] and scal[prev_hist].  The workgroups then do that update's
 * elementwise work with k_pow_update's decomposition: virtual workgroup w of
 * g = ew_grid(m) covers i = w*256 + tid + k*g*256 and workgroup b takes
 * w = b, b + gridDim.x, ...; residual z - sigma^2 q into Blue partials at
 * res_slot (index w), q <- z / |z|.  Last, each runs its row block of the
 * untiled A' on the new q, gathering z[j] * (1 / |z|): the value q[j] gets,
 * bit for bit, so nothing reads the q written in the same launch. */
template <int CAP>
__global__ __launch_bounds__(DPDLP_BS) void k_pow_at(
    const int *__restrict__ blk, const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ val, const double *__restrict__ eig, double *__restrict__ next,
    double *__restrict__ dual, int m, int g, double *scal, double *part, int stride,
    int pow_nparts, int pow_hist, int res_slot, int prev_slot, int prev_nparts, int prev_hist) {
  __shared__ double prod[CAP];
  __shared__ int srp[DPDLP_BS + 1];
  __shared__ double ssum[DPDLP_BS];
  __shared__ double red[4 * 4];
  double t[4];
  if (pow_nparts > 0) {
    wg_slot_sums<4>(part, stride, pow_nparts, DPDLP_SL_PI, red, t);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) {
        scal[DPDLP_SL_PI + k] = t[k];
        scal[pow_hist + k] = t[k];
      }
  } else {
    for (int k = 0; k < 4; ++k) t[k] = scal[DPDLP_SL_PI + k];
  }
  if (prev_slot >= 0) {
    double r[3];
    wg_slot_sums<3>(part, stride, prev_nparts, prev_slot, red, r);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < 3; ++k) {
        scal[prev_slot + k] = r[k];
        scal[prev_hist + k] = r[k];
      }
  }
  const double neg = -t[0];
  const double inv = 1.0 / blue_combine(t[1], t[2], t[3]);
  for (int w = blockIdx.x; w < g; w += gridDim.x) {
    double s[3] = {0.0, 0.0, 0.0};
    double mx[1] = {0.0};
    for (long long i = (long long)w * DPDLP_BS + threadIdx.x; i < m;
         i += (long long)g * DPDLP_BS) {
      double e = eig[i];
      double r = next[i] * neg + e;
      blue_acc(r, s[0], s[1], s[2]);
      next[i] = e * inv;
    }
    __syncthreads(); /* red is still read by the totals above / the last w */
    block_partials_at<3, 0>(s, mx, red, part, stride, res_slot, w);
  }
  EpiStore st{dual, 0};
  double s0[1] = {0.0}, m0[1] = {0.0};
  spmv_block<double, double, CAP, EpiStore, true>(rp, ci, val, eig, blk[blockIdx.x],
                                                  blk[blockIdx.x + 1], st, s0, m0, prod, srp,
                                                  ssum, CAP, inv);
}

__global__ void k_copy_scal(double *scal, int src, int dst, int count) {
  if (threadIdx.x < count) scal[dst + threadIdx.x] = scal[src + threadIdx.x];
}

/* result unscale (pdlp_core_op.cu:464-480) */
__global__ void k_unscale(const double *x, const double *y, const double *D_c, const double *D_r,
                          double beta, double gamma, double *x_out, double *y_out, int n, int m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)n + m;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < n) {
      x_out[i] = x[i] / D_c[i] / beta;
    } else {
      int k = (int)(i - n);
      y_out[k] = y[k] / D_r[k] / gamma;
    }
  }
}

/* ---- scaling (preconditioner.c) ---- */

/* Per-row reduction of |a| by a group of G lanes per row (G = 4..64, picked
 * from the mean row length), loads coalesced within the group:
 * mode 0 = max (exact in any order); mode 1 = sum of |a|^p accumulated
 * strictly in storage order by the group's first lane (so column sums over
 * the transpose match a row-major host accumulation bit for bit); pow is
 * skipped for p == 1, which is exact.  The result does not depend on G. */
template <int G>
__device__ __forceinline__ double row_abs_reduce(const int *rp, const double *v, long long r,
                                                 int gl, int mode, double p) {
  /* unsigned positions: no overflow stepping past a row ending near 2^31 - 1 */
  const unsigned a = (unsigned)rp[r], e = (unsigned)rp[r + 1];
  double acc = 0.0;
  if (mode == 0) {
    for (unsigned q = a + gl; q < e; q += G) acc = fmax(acc, fabs(v[q]));
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc = fmax(acc, __shfl_xor(acc, off, G));
  } else {
    for (unsigned base = a; base < e; base += G) {
      const unsigned q = base + gl;
      double t = 0.0;
      if (q < e) t = (p == 1.0) ? fabs(v[q]) : pow(fabs(v[q]), p);
      const int cnt = (e - base) < (unsigned)G ? (int)(e - base) : G;
      for (int j = 0; j < cnt; ++j) acc += __shfl(t, j, G);
    }
  }
  return acc;
}

/* Row reductions of A (rows [0, mA), exponent pA, into outA) and of A'
 * (the next mB groups, exponent pB, into outB) in one launch. */
template <int G>
__global__ __launch_bounds__(DPDLP_BS) void k_row_reduce2(const int *rpA, const double *vA, int mA,
                                                         double pA, double *outA,
                                                         const int *rpB, const double *vB,
                                                         int mB, double pB, double *outB,
                                                         int mode) {
  const int gl = threadIdx.x & (G - 1);
  const long long g0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const long long ng = ((long long)gridDim.x * blockDim.x) / G;
  for (long long r = g0; r < (long long)mA + mB; r += ng) {
    const bool a = r < mA;
    const long long rr = a ? r : r - mA;
    const double acc = a ? row_abs_reduce<G>(rpA, vA, rr, gl, mode, pA)
                         : row_abs_reduce<G>(rpB, vB, rr, gl, mode, pB);
    if (gl == 0) (a ? outA : outB)[rr] = acc;
  }
}

/* The factors of one Ruiz / Pock-Chambolle step and the vectors they scale,
 * one index per thread: f <- sqrt(f) (1 below 1e-12), D *= f; then
 * c /= cs, lv,uv *= cs and lc,uc /= rs (preconditioner.c:92-100,126-156). */
__global__ void k_scale_factors_vectors(double *cs, double *D_c, double *c, double *lv,
                                        double *uv, int n, double *rs, double *D_r, double *lc,
                                        double *uc, int m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)n + m;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < n) {
      double f = cs[i];
      f = (f < 1e-12) ? 1.0 : sqrt(f);
      cs[i] = f;
      D_c[i] *= f;
      c[i] /= f;
      uv[i] *= f;
      lv[i] *= f;
    } else {
      int k = (int)(i - n);
      double f = rs[k];
      f = (f < 1e-12) ? 1.0 : sqrt(f);
      rs[k] = f;
      D_r[k] *= f;
      lc[k] /= f;
      uc[k] /= f;
    }
  }
}

/* a_k /= (rs[row] * cs[col]) for A and A' in one launch, G lanes per row;
 * the product is formed as rs[i] * cs[j] in both matrices (for A' the row
 * index is a column of A). */
template <int G>
__global__ __launch_bounds__(DPDLP_BS) void k_scale_values2(const int *rpA, const int *ciA,
                                                           double *vA, int mA, const int *rpB,
                                                           const int *ciB, double *vB, int mB,
                                                           const double *rs, const double *cs) {
  const int gl = threadIdx.x & (G - 1);
  const long long g0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const long long ng = ((long long)gridDim.x * blockDim.x) / G;
  for (long long r = g0; r < (long long)mA + mB; r += ng) {
    if (r < mA) {
      const double fr = rs[r];
      const unsigned e = (unsigned)rpA[r + 1];
      for (unsigned q = (unsigned)rpA[r] + gl; q < e; q += G) vA[q] /= fr * cs[ciA[q]];
    } else {
      const long long rr = r - mA;
      const double fc = cs[rr];
      const unsigned e = (unsigned)rpB[rr + 1];
      for (unsigned q = (unsigned)rpB[rr] + gl; q < e; q += G) vB[q] /= rs[ciB[q]] * fc;
    }
  }
}

/* lanes per row for a mean row length: the next power of two, 4..64 */
static inline int lanes_per_row(long long nnz, long long rows) {
  long long mean = rows > 0 ? (nnz + rows - 1) / rows : 1;
  int g = 4;
  while (g < 64 && g < mean) g <<= 1;
  return g;
}

static inline int group_grid(long long rows, int G) {
  long long per = DPDLP_BS / G;
  long long g = (rows + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (int)g;
}

/* rs <- row reduction of A (exponent pA), cs <- of A' (exponent pB) */
static void row_reduce_both(const dpdlp_state_t *st, int mode, double pA, double pB, double *rs,
                            double *cs, hipStream_t s) {
  const int G = lanes_per_row(2LL * st->A.nnz, (long long)st->m + st->n);
  const int grid = group_grid((long long)st->m + st->n, G);
#define RR(GG)                                                                              \
  k_row_reduce2<GG><<<grid, DPDLP_BS, 0, s>>>(st->A.rp, st->A.v, st->m, pA, rs, st->At.rp, \
                                              st->At.v, st->n, pB, cs, mode)
  switch (G) {
  case 4: RR(4); break;
  case 8: RR(8); break;
  case 16: RR(16); break;
  case 32: RR(32); break;
  default: RR(64); break;
  }
#undef RR
}

static void scale_values_both(dpdlp_state_t *st, const double *rs, const double *cs,
                              hipStream_t s) {
  const int G = lanes_per_row(2LL * st->A.nnz, (long long)st->m + st->n);
  const int grid = group_grid((long long)st->m + st->n, G);
#define SV(GG)                                                                         \
  k_scale_values2<GG><<<grid, DPDLP_BS, 0, s>>>(st->A.rp, st->A.ci, (double *)st->A.v, \
                                                st->m, st->At.rp, st->At.ci,           \
                                                (double *)st->At.v, st->n, rs, cs)
  switch (G) {
  case 4: SV(4); break;
  case 8: SV(8); break;
  case 16: SV(16); break;
  case 32: SV(32); break;
  default: SV(64); break;
  }
#undef SV
}

/* bound norm^2 and objective norm^2 partials (preconditioner.c:227-247) */
__global__ __launch_bounds__(DPDLP_BS) void k_bound_obj_sums(const double *lc, const double *uc,
                                                            int m, const double *c, int n,
                                                            double *part, int stride, int slot0) {
  __shared__ double red[2 * 4];
  double s[2] = {0.0, 0.0};
  double mx[1] = {0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)m + n;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < m) {
      double l = lc[i], u = uc[i];
      if (isfinite(l) && l != u) s[0] += l * l;
      if (isfinite(u)) s[0] += u * u;
    } else {
      double v = c[i - m];
      s[1] += v * v;
    }
  }
  block_partials<2, 0>(s, mx, red, part, stride, slot0);
}

/* l_c, u_c, l_v, u_v *= beta and c *= gamma (preconditioner.c:252-265), with
 * beta = 1/(sqrt(bound sum of squares) + 1) and gamma =
 * 1/(sqrt(objective sum of squares) + 1) formed from the reduced sums in
 * S[DPDLP_SL_SCAL], S[DPDLP_SL_SCAL + 1] (preconditioner.c:249-250); beta
 * and gamma are also stored to S[DPDLP_SL_SCAL + 2], S[DPDLP_SL_SCAL + 3]
 * for the host. */
__global__ void k_bound_obj_apply_scal(double *lc, double *uc, int m, double *lv, double *uv,
                                       double *c, int n, double *S) {
  const double beta = 1.0 / (sqrt(S[DPDLP_SL_SCAL]) + 1.0);
  const double gamma = 1.0 / (sqrt(S[DPDLP_SL_SCAL + 1]) + 1.0);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)m + n;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < m) {
      lc[i] *= beta;
      uc[i] *= beta;
    } else {
      int k = (int)(i - m);
      lv[k] *= beta;
      uv[k] *= beta;
      c[k] *= gamma;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S[DPDLP_SL_SCAL + 2] = beta;
    S[DPDLP_SL_SCAL + 3] = gamma;
  }
}

__global__ void k_to_f32(const double *src, float *dst, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

__global__ void k_zero_finite(double *v, int count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    if (isfinite(v[i])) v[i] = 0.0;
}

__global__ void k_fill(double *v, double value, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    v[i] = value;
}

/* plain copy of known size, used to calibrate the HBM byte counters */
__global__ __launch_bounds__(DPDLP_BS) void k_copy_calib(const double4 *__restrict__ src,
                                                        double4 *__restrict__ dst, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

/* Standalone reductions (get_vector_inf_norm, get_vector_sum and the cuBLAS
 * nrm2 / dot calls, utils.cu:616-637): partials of a fixed grid, combined by
 * k_reduce_final in a fixed order; nrm2 accumulates Blue's three scaled sums,
 * so it neither overflows nor underflows. */
__global__ __launch_bounds__(DPDLP_BS) void k_reduce_partials(int op, const double *__restrict__ x,
                                                             const double *__restrict__ y,
                                                             long long n,
                                                             double *__restrict__ part, int nb) {
  __shared__ double red[4 * 4];
  double s[3] = {0.0, 0.0, 0.0};
  double mx[1] = {0.0};
  for (long long i = (long long)blockIdx.x * DPDLP_BS + threadIdx.x; i < n;
       i += (long long)gridDim.x * DPDLP_BS) {
    const double v = x[i];
    if (op == DPDLP_RED_NRM2)
      blue_acc(v, s[0], s[1], s[2]);
    else if (op == DPDLP_RED_DOT)
      s[0] += v * y[i];
    else if (op == DPDLP_RED_SUM)
      s[0] += v;
    else
      mx[0] = fmax(mx[0], fabs(v));
  }
  block_partials<3, 1>(s, mx, red, part, nb, 0);
}

__global__ __launch_bounds__(DPDLP_BS) void k_reduce_final(int op, const double *__restrict__ part,
                                                          int nb, double *__restrict__ out) {
  __shared__ double red[4];
  double v[4];
  for (int k = 0; k < 4; ++k) {
    const double *p = part + (size_t)k * nb;
    double acc = 0.0;
    if (k == 3) {
      for (int i = threadIdx.x; i < nb; i += DPDLP_BS) acc = fmax(acc, p[i]);
      acc = wave_max(acc);
    } else {
      for (int i = threadIdx.x; i < nb; i += DPDLP_BS) acc += p[i];
      acc = wave_sum(acc);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    v[k] = k == 3 ? fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))
                  : (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (op == DPDLP_RED_NRM2)
      out[0] = blue_combine(v[0], v[1], v[2]);
    else if (op == DPDLP_RED_AMAX)
      out[0] = v[3];
    else
      out[0] = v[0];
  }
}

/* Gather-rate calibration with the SpMV's access pattern: workgroup w streams
 * its chunk of `chunk` column indices coalesced (consecutive threads read
 * consecutive indices, eight loads in flight per thread, non-temporal, as
 * spmv_block does) and gathers x[base + idx] for each, where base is the
 * slice of its own XCD ((w mod 8) * slice; workgroup w runs on XCD w mod 8)
 * when xcd_local is set, else 0.  With WITHV the FP64 values are streamed
 * beside the indices (the SpMV's 12 bytes per nonzero) and each gather is
 * multiplied by its value.  Each thread writes the sum of its products. */
template <bool WITHV>
__global__ __launch_bounds__(DPDLP_BS) void k_gather_calib(const int *__restrict__ idx,
                                                          const double *__restrict__ val,
                                                          const double *__restrict__ x,
                                                          double *__restrict__ y, long long count,
                                                          int chunk, long long slice,
                                                          int xcd_local) {
  constexpr int UNR = 8;
  const long long c0 = (long long)blockIdx.x * chunk;
  const double *xs = x + (xcd_local ? (long long)(blockIdx.x & 7) * slice : 0);
  long long lim = count - c0;
  if (lim > chunk) lim = chunk;
  double acc = 0.0;
  for (int base = threadIdx.x; base < lim; base += DPDLP_BS * UNR) {
    int ci[UNR];
    double vv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int k = base + u * DPDLP_BS;
      ci[u] = k < lim ? __builtin_nontemporal_load(idx + c0 + k) : -1;
      vv[u] = 1.0;
      if (WITHV && k < lim) vv[u] = __builtin_nontemporal_load(val + c0 + k);
    }
    double v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) v[u] = ci[u] >= 0 ? xs[ci[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < UNR; ++u) acc += WITHV ? vv[u] * v[u] : v[u];
  }
  y[(long long)blockIdx.x * DPDLP_BS + threadIdx.x] = acc;
}

/* warm start: v = start * D * s into init, cur and pdhg (pdlp_core_op.cu:174-201) */
__global__ void k_warm_start(double *a, double *b, double *c, const double *start,
                             const double *D, double s, int count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    double v = start[i] * D[i] * s;
    a[i] = v;
    b[i] = v;
    c[i] = v;
  }
}

/* ------------------------------------------------------------------------ */
/* extern "C" launchers                                                      */
/* ------------------------------------------------------------------------ */

#define ST(s) ((hipStream_t)(s))
#define RET(e)                                                                 \
  do {                                                                         \
    hipError_t _e = (e);                                                       \
    if (_e != hipSuccess) return (int)_e;                                      \
  } while (0)

extern "C" {

int dpdlp_dev_spmv(const dpdlp_csr_t *A, const double *x, double *y, void *stream) {
  EpiStore e{y, 0};
  return (int)launch_spmv(A, x, e, nullptr, 0, 0, ST(stream));
}

int dpdlp_dev_spmv_f32(const dpdlp_csr_t *A, const float *x, double *y, void *stream) {
  EpiStore e{y, 0};
  return (int)launch_spmv32(A, x, e, nullptr, 0, 0, ST(stream));
}

int dpdlp_dev_primal_phase(const dpdlp_state_t *st, double tau, double w, double rho, int major,
                           int fpe, int use32, void *stream) {
  EpiPrimal e;
  e.x_cur = st->x_cur;
  e.x_init = st->x_init;
  e.c = st->c;
  e.lv = st->lv;
  e.uv = st->uv;
  /* FP32 iterations write only the FP32 copies the next SpMV gathers */
  e.x_ref = use32 ? nullptr : st->x_ref;
  e.x_pdhg = st->x_pdhg;
  e.z = st->dual_slack;
  e.dx = st->dx;
  e.x_ref32 = use32 ? st->x_ref32 : nullptr;
  const bool anchored = use32 && st->ax0 && st->aty0;
  e.aty0 = anchored ? st->aty0 : nullptr;
  e.delta32 = anchored;
  e.tau = tau;
  e.w = w;
  e.rho = rho;
  e.major = major;
  e.fpe = fpe;
  e.partials = fpe;
  if (use32)
    return (int)launch_spmv32(&st->At, st->y_cur32, e, st->part, st->part_stride,
                              DPDLP_SL_FPE, ST(stream));
  return (int)launch_spmv(&st->At, st->y_cur, e, st->part, st->part_stride, DPDLP_SL_FPE,
                          ST(stream));
}

int dpdlp_dev_dual_phase(const dpdlp_state_t *st, double sigma, double w, double rho, int major,
                         int fpe, int use32, void *stream) {
  EpiDual e;
  e.y_cur = st->y_cur;
  e.y_init = st->y_init;
  e.lc = st->lc;
  e.uc = st->uc;
  e.y_pdhg = st->y_pdhg;
  e.dy = st->dy;
  e.y_cur32 = use32 ? st->y_cur32 : nullptr;
  const bool anchored = use32 && st->ax0 && st->aty0;
  e.ax0 = anchored ? st->ax0 : nullptr;
  e.delta32 = anchored;
  e.sigma = sigma;
  e.w = w;
  e.rho = rho;
  e.major = major;
  e.fpe = fpe;
  e.partials = fpe;
  if (use32)
    return (int)launch_spmv32(&st->A, st->x_ref32, e, st->part, st->part_stride,
                              DPDLP_SL_FPE + 3, ST(stream));
  return (int)launch_spmv(&st->A, st->x_ref, e, st->part, st->part_stride, DPDLP_SL_FPE + 3,
                          ST(stream));
}

int dpdlp_dev_fpe_finish(const dpdlp_state_t *st, int to_restart, int to_major, void *stream) {
  EpiFpe e{st->dual_product, st->dx, 1};
  RET(launch_spmv(&st->At, st->dy, e, st->part, st->part_stride, DPDLP_SL_FPE + 6,
                  ST(stream)));
  int d1 = to_restart ? DPDLP_SC_FPE_RESTART : -1;
  int d2 = to_major ? DPDLP_SC_FPE_MAJOR : -1;
  /* slots come from three kernels with different partial counts */
  RET(finalize(st, spmv_nparts(&st->At), DPDLP_SL_FPE, 3, 0u, DPDLP_SL_FPE,
               d1 >= 0 ? d1 : -1, d2 >= 0 ? d2 : -1, ST(stream)));
  RET(finalize(st, spmv_nparts(&st->A), DPDLP_SL_FPE + 3, 3, 0u, DPDLP_SL_FPE + 3,
               d1 >= 0 ? d1 + 3 : -1, d2 >= 0 ? d2 + 3 : -1, ST(stream)));
  RET(finalize(st, spmv_nparts(&st->At), DPDLP_SL_FPE + 6, 1, 0u, DPDLP_SL_FPE + 6,
               d1 >= 0 ? d1 + 6 : -1, d2 >= 0 ? d2 + 6 : -1, ST(stream)));
  return 0;
}

int dpdlp_dev_fpe_finish_polish(const dpdlp_state_t *st, int primal, int to_restart,
                                int to_major, void *stream) {
  int d1 = to_restart ? DPDLP_SC_FPE_RESTART : -1;
  int d2 = to_major ? DPDLP_SC_FPE_MAJOR : -1;
  if (primal)
    return (int)finalize(st, spmv_nparts(&st->At), DPDLP_SL_FPE, 3, 0u, DPDLP_SL_FPE, d1, d2,
                         ST(stream));
  return (int)finalize(st, spmv_nparts(&st->A), DPDLP_SL_FPE + 3, 3, 0u, DPDLP_SL_FPE + 3,
                       d1 >= 0 ? d1 + 3 : -1, d2 >= 0 ? d2 + 3 : -1, ST(stream));
}

int dpdlp_dev_residual(const dpdlp_state_t *st, const dpdlp_state_t *orig, int which,
                       int restart_dist, void *stream) {
  hipStream_t s = ST(stream);
  if (which & 1) {
    EpiResP e;
    e.pprod = st->primal_product;
    e.lc = st->lc;
    e.uc = st->uc;
    e.D_r = st->D_r;
    e.y = st->y_pdhg;
    e.y_init = st->y_init;
    e.dist = restart_dist;
    e.partials = 1;
    RET(launch_spmv(&st->A, st->x_pdhg, e, st->part, st->part_stride, DPDLP_SL_RESP, s));
    /* with both halves, the two totals share one launch after the A' product */
    if (which != 3)
      RET(finalize(st, spmv_nparts(&st->A), DPDLP_SL_RESP, 8, 1u << 7, DPDLP_SL_RESP, -1, -1, s));
    if (which == 1) {
      /* primal polish: objective uses the original c (utils.cu:1176-1182) */
      int g = ew_grid(st->n);
      k_dot<<<g, DPDLP_BS, 0, s>>>(orig->c, st->x_pdhg, st->n, st->part, st->part_stride,
                                   DPDLP_SL_RESD + 3);
      RET(hipGetLastError());
      RET(finalize(st, g, DPDLP_SL_RESD + 3, 1, 0u, DPDLP_SL_RESD + 3, -1, -1, s));
    }
  }
  if (which & 2) {
    EpiResD e;
    e.dprod = st->dual_product;
    e.c = st->c;
    e.z = st->dual_slack;
    e.D_c = st->D_c;
    e.x_obj = (which == 2) ? orig->x_pdhg : st->x_pdhg;
    e.x_init = st->x_init;
    e.dist = restart_dist;
    e.partials = 1;
    RET(launch_spmv(&st->At, st->y_pdhg, e, st->part, st->part_stride, DPDLP_SL_RESD, s));
    if (which == 3) {
      k_finalize_pair<<<8 + 9, DPDLP_BS, 0, s>>>(st->part, st->part_stride, st->scal,
                                                 spmv_nparts(&st->A), DPDLP_SL_RESP, 8, -1,
                                                 spmv_nparts(&st->At), DPDLP_SL_RESD, -1,
                                                 1u << 7, 1u << 8);
      RET(hipGetLastError());
    } else {
      RET(finalize(st, spmv_nparts(&st->At), DPDLP_SL_RESD, 9, 1u << 8, DPDLP_SL_RESD, -1, -1,
                   s));
    }
    if (which == 2) {
      /* dual polish: contributions use the original finite bounds */
      int g = ew_grid(st->m);
      k_dual_contrib_sum<<<g, DPDLP_BS, 0, s>>>(st->y_pdhg, orig->lc, orig->uc, st->m, st->part,
                                                st->part_stride, DPDLP_SL_RESP + 3);
      RET(hipGetLastError());
      RET(finalize(st, g, DPDLP_SL_RESP + 3, 1, 0u, DPDLP_SL_RESP + 3, -1, -1, s));
    }
  }
  return 0;
}

int dpdlp_dev_infeasibility(const dpdlp_state_t *st, void *stream) {
  hipStream_t s = ST(stream);
  int g = ew_grid((long long)st->n + st->m);
  k_inf_project<<<g, DPDLP_BS, 0, s>>>(st->dx, st->lv, st->uv, st->n, st->dy, st->lc, st->uc,
                                       st->m, st->part, st->part_stride, DPDLP_SL_INF);
  RET(hipGetLastError());
  RET(finalize(st, g, DPDLP_SL_INF, 2, 3u, DPDLP_SL_INF, -1, -1, s));
  k_scale_by_inv<<<ew_grid(st->n), DPDLP_BS, 0, s>>>(st->dx, st->n, st->scal, DPDLP_SL_INF);
  RET(hipGetLastError());
  EpiInfP ep;
  ep.pprod = st->primal_product;
  ep.pslack = st->primal_slack;
  ep.lc = st->lc;
  ep.uc = st->uc;
  ep.D_r = st->D_r;
  ep.dy = st->dy;
  ep.partials = 1;
  RET(launch_spmv(&st->A, st->dx, ep, st->part, st->part_stride, DPDLP_SL_INF + 2, s));
  EpiInfD ed;
  ed.dprod = st->dual_product;
  ed.dslack = st->dual_slack;
  ed.lv = st->lv;
  ed.uv = st->uv;
  ed.D_c = st->D_c;
  ed.c = st->c;
  ed.dx = st->dx;
  ed.partials = 1;
  RET(launch_spmv(&st->At, st->dy, ed, st->part, st->part_stride, DPDLP_SL_INF + 4, s));
  RET(finalize(st, spmv_nparts(&st->A), DPDLP_SL_INF + 2, 2, 2u, DPDLP_SL_INF + 2, -1, -1, s));
  RET(finalize(st, spmv_nparts(&st->At), DPDLP_SL_INF + 4, 3, 4u, DPDLP_SL_INF + 4, -1, -1, s));
  return 0;
}

int dpdlp_dev_restart_copy(const dpdlp_state_t *st, int part, void *stream) {
  k_restart_copy<<<ew_grid((long long)st->n + st->m), DPDLP_BS, 0, ST(stream)>>>(
      st->x_init, st->x_cur, st->x_pdhg, st->y_init, st->y_cur, st->y_pdhg, st->y_cur32,
      st->ax0 && st->aty0, st->n, st->m, part);
  return (int)hipGetLastError();
}

int dpdlp_dev_anchor_products(const dpdlp_state_t *st, void *stream) {
  hipStream_t s = ST(stream);
  if (!st->ax0 || !st->aty0) return (int)hipErrorInvalidValue;
  EpiStore ea{st->ax0, 0};
  RET(launch_spmv(&st->A, st->x_init, ea, nullptr, 0, 0, s));
  EpiStore et{st->aty0, 0};
  RET(launch_spmv(&st->At, st->y_init, et, nullptr, 0, 0, s));
  if (st->y_cur32 && st->m > 0) {
    k_delta_f32<<<ew_grid(st->m), DPDLP_BS, 0, s>>>(st->y_cur, st->y_init, st->y_cur32, st->m);
    RET(hipGetLastError());
  }
  return 0;
}

int dpdlp_dev_power_init(const dpdlp_state_t *st, double *eig, double *next, void *stream) {
  hipStream_t s = ST(stream);
  int g = ew_grid(st->m);
  k_nrm_partials<<<g, DPDLP_BS, 0, s>>>(eig, st->m, st->part, st->part_stride,
                                        DPDLP_SL_PI + 1);
  RET(hipGetLastError());
  RET(finalize(st, g, DPDLP_SL_PI + 1, 3, 0u, DPDLP_SL_PI + 1, -1, -1, s));
  k_pow_normalize<<<g, DPDLP_BS, 0, s>>>(eig, next, st->m, st->scal, DPDLP_SL_PI + 1);
  return (int)hipGetLastError();
}

static hipError_t launch_pow_at(const dpdlp_state_t *st, const double *eig, double *next,
                                double *dual, int g, int pow_nparts, int pow_hist, int res_slot,
                                int prev_slot, int prev_nparts, int prev_hist, hipStream_t s) {
  const dpdlp_csr_t *A = &st->At;
  if (A->rows <= 0) return hipSuccess;
  switch (A->cap) {
  case 1024:
    k_pow_at<1024><<<A->nblk, DPDLP_BS, 0, s>>>(A->blk, A->rp, A->ci, A->v, eig, next, dual,
                                               st->m, g, st->scal, st->part, st->part_stride,
                                               pow_nparts, pow_hist, res_slot, prev_slot,
                                               prev_nparts, prev_hist);
    break;
  case 4096:
    k_pow_at<4096><<<A->nblk, DPDLP_BS, 0, s>>>(A->blk, A->rp, A->ci, A->v, eig, next, dual,
                                               st->m, g, st->scal, st->part, st->part_stride,
                                               pow_nparts, pow_hist, res_slot, prev_slot,
                                               prev_nparts, prev_hist);
    break;
  default:
    k_pow_at<2048><<<A->nblk, DPDLP_BS, 0, s>>>(A->blk, A->rp, A->ci, A->v, eig, next, dual,
                                               st->m, g, st->scal, st->part, st->part_stride,
                                               pow_nparts, pow_hist, res_slot, prev_slot,
                                               prev_nparts, prev_hist);
    break;
  }
  return hipGetLastError();
}

/* DPDLP_PI_FUSE=0 keeps the five-launch power step (for comparison) */
static bool pi_fuse_enabled() {
  static int enabled = -1;
  if (enabled < 0) {
    const char *e = getenv("DPDLP_PI_FUSE");
    enabled = (e && e[0] == '0') ? 0 : 1;
  }
  return enabled == 1;
}

int dpdlp_dev_power_steps(const dpdlp_state_t *st, double *eig, double *next, double *dual,
                          int iters, void *stream) {
  hipStream_t s = ST(stream);
  int g = ew_grid(st->m);
  /* Step t's update (sigma^2, residual, q <- z / |z|) runs inside step
   * t + 1's A' product (k_pow_at) when A' is untiled; the last step of the
   * batch gets a k_pow_update of its own.  Totals of partial sums are formed
   * inside the kernel that needs them when that costs each workgroup only a
   * few loads per thread (small LPs, where every launch's fixed cost
   * counts), else by one finalize launch after each A product.  Every
   * variant gives the same bits (wg_slot_sums keeps k_finalize's order, the
   * decomposition of the update is k_pow_update's).  Launches per step
   * besides the A product's: 1 (small LP, A' untiled), 2 (large LP, A'
   * untiled), 2 or 4 (A' tiled). */
  const int np = spmv_nparts(&st->A);
  const bool on = pi_fuse_enabled();
  const bool small = on && 4LL * np + 3LL * g <= 8LL * DPDLP_BS;
  const bool at = on && st->At.P <= 1;
  const int nit = iters < DPDLP_PI_BATCH ? iters : DPDLP_PI_BATCH;
  /* residual partials of step j go to set j & 1 */
  auto rset = [](int j) { return (j & 1) ? (int)DPDLP_SL_PI2 : (int)DPDLP_SL_PI + 4; };
  for (int it = 0; it < nit; ++it) {
    const int hist = DPDLP_SC_PI_HIST + 4 * it;
    if (it > 0 && at) {
      /* step it - 1's update folded into this A' product */
      const int ph = hist - 4;
      if (small) {
        RET(launch_pow_at(st, eig, next, dual, g, np, ph, rset(it - 1),
                          it >= 2 ? rset(it - 2) : -1, g, ph - 4 + 1, s));
      } else {
        /* sigma^2 and |z| of step it - 1 were totalled after its A product */
        RET(launch_pow_at(st, eig, next, dual, g, 0, ph, rset(it - 1), -1, 0, 0, s));
      }
    } else {
      EpiStore e1{dual, 0};
      RET(launch_spmv(&st->At, next, e1, nullptr, 0, 0, s));
    }
    EpiPow e2{eig, next, 1};
    RET(launch_spmv(&st->A, dual, e2, st->part, st->part_stride, DPDLP_SL_PI, s));
    if (at && !small) {
      /* this step's sigma^2 and |z| and the previous step's residual (from
       * this step's k_pow_at) in one launch */
      if (it >= 1)
        k_finalize_pair<<<7, DPDLP_BS, 0, s>>>(st->part, st->part_stride, st->scal, np,
                                               DPDLP_SL_PI, 4, hist, g, rset(it - 1),
                                               hist - 4 + 1);
      else
        k_finalize_pair<<<4, DPDLP_BS, 0, s>>>(st->part, st->part_stride, st->scal, np,
                                               DPDLP_SL_PI, 4, hist, 0, 0, 0);
      RET(hipGetLastError());
      if (it < nit - 1) continue; /* the next A' product updates */
      k_pow_update<<<g, DPDLP_BS, 0, s>>>(eig, next, st->m, st->scal, st->part,
                                          st->part_stride, 0, hist, rset(it), -1, -1);
      RET(hipGetLastError());
      RET(finalize(st, g, rset(it), 3, 0u, rset(it), hist + 1, -1, s));
      continue;
    }
    if (at && it < nit - 1) continue; /* the next A' product updates */
    if (small) {
      /* step it - 1's residual partials (from the last update, folded or
       * not) are still pending: this launch totals them */
      const int res = rset(it), prev = it >= 1 ? rset(it - 1) : -1;
      k_pow_update<<<g, DPDLP_BS, 0, s>>>(eig, next, st->m, st->scal, st->part,
                                          st->part_stride, np, hist, res, prev, hist - 4 + 1);
      RET(hipGetLastError());
      if (it == nit - 1) RET(finalize(st, g, res, 3, 0u, res, hist + 1, -1, s));
    } else {
      /* sigma^2 also recorded in the history (slot 0 of the 4) */
      RET(finalize(st, np, DPDLP_SL_PI, 4, 0u, DPDLP_SL_PI, hist, -1, s));
      k_pow_update<<<g, DPDLP_BS, 0, s>>>(eig, next, st->m, st->scal, st->part,
                                          st->part_stride, 0, hist, DPDLP_SL_PI + 4, -1, -1);
      RET(hipGetLastError());
      RET(finalize(st, g, DPDLP_SL_PI + 4, 3, 0u, DPDLP_SL_PI + 4, hist + 1, -1, s));
    }
  }
  return 0;
}

int dpdlp_dev_result(const dpdlp_state_t *st, double beta, double gamma, double *x_out,
                     double *y_out, double *rc_out, void *stream) {
  hipStream_t s = ST(stream);
  EpiRc e{rc_out, st->c, st->D_c, gamma, 0};
  RET(launch_spmv(&st->At, st->y_pdhg, e, nullptr, 0, 0, s));
  k_unscale<<<ew_grid((long long)st->n + st->m), DPDLP_BS, 0, s>>>(
      st->x_pdhg, st->y_pdhg, st->D_c, st->D_r, beta, gamma, x_out, y_out, st->n, st->m);
  return (int)hipGetLastError();
}

/* one Ruiz step: three launches (row maxima of A and A', factors and
 * vectors, matrix values) */
int dpdlp_dev_ruiz_step(dpdlp_state_t *st, double *rs, double *cs, void *stream) {
  hipStream_t s = ST(stream);
  row_reduce_both(st, 0, 1.0, 1.0, rs, cs, s);
  k_scale_factors_vectors<<<ew_grid((long long)st->n + st->m), DPDLP_BS, 0, s>>>(
      cs, st->D_c, st->c, st->lv, st->uv, st->n, rs, st->D_r, st->lc, st->uc, st->m);
  scale_values_both(st, rs, cs, s);
  return (int)hipGetLastError();
}

int dpdlp_dev_pock_chambolle(dpdlp_state_t *st, double alpha, double *rs, double *cs,
                             void *stream) {
  hipStream_t s = ST(stream);
  row_reduce_both(st, 1, alpha, 2.0 - alpha, rs, cs, s);
  k_scale_factors_vectors<<<ew_grid((long long)st->n + st->m), DPDLP_BS, 0, s>>>(
      cs, st->D_c, st->c, st->lv, st->uv, st->n, rs, st->D_r, st->lc, st->uc, st->m);
  scale_values_both(st, rs, cs, s);
  return (int)hipGetLastError();
}

int dpdlp_dev_bound_obj_sums(const dpdlp_state_t *st, void *stream) {
  hipStream_t s = ST(stream);
  int g = ew_grid((long long)st->m + st->n);
  k_bound_obj_sums<<<g, DPDLP_BS, 0, s>>>(st->lc, st->uc, st->m, st->c, st->n, st->part,
                                          st->part_stride, DPDLP_SL_SCAL);
  RET(hipGetLastError());
  return (int)finalize(st, g, DPDLP_SL_SCAL, 2, 0u, DPDLP_SL_SCAL, -1, -1, s);
}

int dpdlp_dev_bound_obj_apply_scal(dpdlp_state_t *st, void *stream) {
  k_bound_obj_apply_scal<<<ew_grid((long long)st->m + st->n), DPDLP_BS, 0, ST(stream)>>>(
      st->lc, st->uc, st->m, st->lv, st->uv, st->c, st->n, st->scal);
  return (int)hipGetLastError();
}

int dpdlp_dev_make_f32(const double *src, float *dst, long long count, void *stream) {
  if (count <= 0) return 0;
  k_to_f32<<<ew_grid(count), DPDLP_BS, 0, ST(stream)>>>(src, dst, count);
  return (int)hipGetLastError();
}

int dpdlp_dev_zero_finite(double *v, int count, void *stream) {
  if (count <= 0) return 0;
  k_zero_finite<<<ew_grid(count), DPDLP_BS, 0, ST(stream)>>>(v, count);
  return (int)hipGetLastError();
}

int dpdlp_dev_fill(double *v, double value, long long count, void *stream) {
  if (count <= 0) return 0;
  k_fill<<<ew_grid(count), DPDLP_BS, 0, ST(stream)>>>(v, value, count);
  return (int)hipGetLastError();
}

int dpdlp_dev_permute(const double *src, const int *perm, double *dst, long long count,
                      void *stream) {
  if (count <= 0) return 0;
  k_permute<double><<<ew_grid(count), DPDLP_BS, 0, ST(stream)>>>(src, perm, dst, count);
  return (int)hipGetLastError();
}

int dpdlp_dev_permute_f32(const float *src, const int *perm, float *dst, long long count,
                          void *stream) {
  if (count <= 0) return 0;
  k_permute<float><<<ew_grid(count), DPDLP_BS, 0, ST(stream)>>>(src, perm, dst, count);
  return (int)hipGetLastError();
}

int dpdlp_dev_reduce_blocks(long long n) {
  long long g = (n + DPDLP_BS - 1) / DPDLP_BS;
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return (int)g;
}

int dpdlp_dev_reduce(int op, const double *x, const double *y, long long n, double *part,
                     double *out, void *stream) {
  if (op < DPDLP_RED_NRM2 || op > DPDLP_RED_SUM || n < 0 || !part || !out ||
      (n > 0 && (!x || (op == DPDLP_RED_DOT && !y))))
    return (int)hipErrorInvalidValue;
  const int nb = dpdlp_dev_reduce_blocks(n);
  k_reduce_partials<<<nb, DPDLP_BS, 0, ST(stream)>>>(op, x, y, n, part, nb);
  k_reduce_final<<<1, DPDLP_BS, 0, ST(stream)>>>(op, part, nb, out);
  return (int)hipGetLastError();
}

int dpdlp_dev_gather_calib_values(const int *idx, const double *val, const double *x,
                                  double *y, long long count, int chunk, long long slice,
                                  int xcd_local, void *stream) {
  if (count <= 0 || chunk <= 0) return (int)hipErrorInvalidValue;
  long long g = (count + chunk - 1) / chunk;
  if (g > 0x7fffffffLL) return (int)hipErrorInvalidValue;
  if (val)
    k_gather_calib<true><<<(unsigned)g, DPDLP_BS, 0, ST(stream)>>>(idx, val, x, y, count, chunk,
                                                                   slice, xcd_local);
  else
    k_gather_calib<false><<<(unsigned)g, DPDLP_BS, 0, ST(stream)>>>(idx, val, x, y, count, chunk,
                                                                    slice, xcd_local);
  return (int)hipGetLastError();
}

int dpdlp_dev_gather_calib(const int *idx, const double *x, double *y, long long count,
                           int chunk, long long slice, int xcd_local, void *stream) {
  return dpdlp_dev_gather_calib_values(idx, NULL, x, y, count, chunk, slice, xcd_local, stream);
}

int dpdlp_dev_copy_calib(const void *src, void *dst, long long bytes, void *stream) {
  long long n4 = bytes / 32;
  if (n4 <= 0) return 0;
  long long g = (n4 + DPDLP_BS - 1) / DPDLP_BS;
  if (g > 65536) g = 65536;
  k_copy_calib<<<(int)g, DPDLP_BS, 0, ST(stream)>>>((const double4 *)src, (double4 *)dst, n4);
  return (int)hipGetLastError();
}

int dpdlp_dev_warm_start(double *a, double *b, double *c, const double *start, const double *D,
                         double s, int count, void *stream) {
  if (count <= 0) return 0;
  k_warm_start<<<ew_grid(count), DPDLP_BS, 0, ST(stream)>>>(a, b, c, start, D, s, count);
  return (int)hipGetLastError();
}

} /* extern "C" */
