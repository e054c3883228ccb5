/*
 * Stable CSR transpose on the device: A' as CSR with the entries of every A'
 * row in increasing row order of A, the layout csr2csc produces
 * (pdlp_core_op.cu:93-130).
 *
 * The nonzero positions k = 0..nnz-1 of A (row-major, hence already in the
 * order A' wants within each column) are sorted by their column index ci[k]
 * with a least-significant-digit radix sort, 8 bits per pass.  A pass:
 *   k_rs_hist     per tile of 4096 positions, a histogram of the digit;
 *                 stored digit-major (hist[d * ntiles + t]) so that one
 *                 exclusive scan turns it into every tile's output offsets;
 *   scan          exclusive scan of the histogram (three kernels);
 *   k_rs_scatter  each tile writes its positions to their offsets, ranking
 *                 equal digits in input order (wavefront ballots give the
 *                 rank inside a wavefront, per-digit wavefront counts in LDS
 *                 order the four wavefronts, a running offset per digit
 *                 orders the tile's 16 rounds) -- so every pass is stable
 *                 and the result is deterministic.
 * Keys are not stored: a pass reads ci[k] for the positions it moves.  The
 * positions ping-pong between the two int halves of A''s value array; the
 * last pass writes them to A''s index array, and k_tr_finish replaces each
 * by its row (binary search in A's row pointers) and gathers the value.
 * A''s row pointers (column counts of A) come from the host.
 */
#include "dev_common.h"
#include "dpdlp_dev.h"

namespace {

constexpr int RS_T = DPDLP_BS;  /* threads per tile */
constexpr int RS_ITEMS = 16;    /* rounds per tile */
constexpr long long RS_TILE = (long long)RS_T * RS_ITEMS;
constexpr int SCAN_ITEMS = 16;
constexpr long long SCAN_CHUNK = (long long)DPDLP_BS * SCAN_ITEMS;

__global__ __launch_bounds__(DPDLP_BS) void k_rs_hist(const int *__restrict__ ci,
                                                      const int *__restrict__ vin, long long N,
                                                      int shift, int ntiles,
                                                      int *__restrict__ hist) {
  __shared__ int cnt[256];
  const int t = blockIdx.x;
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)t * RS_TILE;
  for (int it = 0; it < RS_ITEMS; ++it) {
    const long long e = base + (long long)it * RS_T + threadIdx.x;
    if (e < N) {
      const int k = vin ? vin[e] : (int)e;
      atomicAdd(&cnt[(ci[k] >> shift) & 255], 1);
    }
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + t] = cnt[threadIdx.x];
}

/* exclusive scan of a 256-thread workgroup's values; returns the total */
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  int before = 0;
  for (int q = 0; q < w; ++q) before += lds[q];
  *total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return before + x - v;
}

/* in-place exclusive scan of each SCAN_CHUNK; chunk totals to bsum */
__global__ __launch_bounds__(DPDLP_BS) void k_scan_chunks(int *__restrict__ a, long long L,
                                                          int *__restrict__ bsum) {
  __shared__ int lds[4];
  const long long base = (long long)blockIdx.x * SCAN_CHUNK + (long long)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    v[i] = base + i < L ? a[base + i] : 0;
    s += v[i];
  }
  int total;
  int run = block_exclusive_scan(s, lds, &total);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < L) a[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

/* exclusive scan of the chunk totals by one workgroup */
__global__ __launch_bounds__(DPDLP_BS) void k_scan_totals(int *__restrict__ bsum, int nb) {
  __shared__ int lds[4];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += DPDLP_BS) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? bsum[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, lds, &total);
    if (i < nb) bsum[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(DPDLP_BS) void k_scan_add(int *__restrict__ a, long long L,
                                                       const int *__restrict__ bsum) {
  const int add = bsum[blockIdx.x];
  const long long base = (long long)blockIdx.x * SCAN_CHUNK;
  for (int i = threadIdx.x; i < SCAN_CHUNK; i += DPDLP_BS)
    if (base + i < L) a[base + i] += add;
}

__global__ __launch_bounds__(DPDLP_BS) void k_rs_scatter(const int *__restrict__ ci,
                                                         const int *__restrict__ vin,
                                                         int *__restrict__ vout, long long N,
                                                         int shift, int ntiles,
                                                         const int *__restrict__ hist) {
  __shared__ int run[256];
  __shared__ int wcnt[4][256];
  const int t = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  run[tid] = hist[(size_t)tid * ntiles + t];
  const unsigned long long below = (1ull << lane) - 1ull;
  const long long base = (long long)t * RS_TILE;
  for (int it = 0; it < RS_ITEMS; ++it) {
#pragma unroll
    for (int q = 0; q < 4; ++q) wcnt[q][tid] = 0;
    __syncthreads();
    const long long e = base + (long long)it * RS_T + tid;
    const bool valid = e < N;
    int k = 0, d = 0;
    if (valid) {
      k = vin ? vin[e] : (int)e;
      d = (ci[k] >> shift) & 255;
    }
    /* lanes of this wavefront holding the same digit */
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int bit = (d >> b) & 1;
      const unsigned long long on = __ballot(valid && bit);
      same &= bit ? on : ~on;
    }
    const int rank = __popcll(same & below);
    if (valid && rank == 0) wcnt[w][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      int pos = run[d] + rank;
      for (int q = 0; q < w; ++q) pos += wcnt[q][d];
      vout[pos] = k;
    }
    __syncthreads();
    run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
  }
}

/* positions -> (row of A, value): the row of position k is the last r with
 * rp[r] <= k (empty rows have rp[r] == rp[r + 1] and are skipped) */
__global__ void k_tr_finish(const int *__restrict__ rp, int m, const double *__restrict__ v,
                            int *__restrict__ t_ci, double *__restrict__ t_v, long long N) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const int k = t_ci[p];
    int lo = 0, hi = m; /* rp[lo] <= k < rp[hi] */
    while (hi - lo > 1) {
      const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
      if (rp[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    t_ci[p] = lo;
    t_v[p] = v[k];
  }
}

inline long long ntiles_of(long long nnz) { return (nnz + RS_TILE - 1) / RS_TILE; }

/* Column-panel tiles of a device CSR matrix under a given row-block plan,
 * the layout dpdlp_build_panels makes on the host: one workgroup per row
 * block.  Block b (rows r0..r1, R = r1 - r0) fills its own nonzero range
 * [rp[r0], rp[r1]) tile by tile (panel k = column / W); within a tile the
 * entries keep row order and, within a row, CSR order.  The threads count
 * every row's entries per panel into LDS, one thread per panel turns the
 * counts into exclusive offsets down the rows (the rofs entries: R + 1 per
 * tile), and each row's thread then writes its entries to their slots. */
constexpr int BP_RMAX = DPDLP_PANEL_RMAX;

__global__ __launch_bounds__(DPDLP_BS) void k_build_panels(
    const int *__restrict__ rp, const int *__restrict__ ci, const int *__restrict__ blk, int nblk,
    int P, long long W, int *__restrict__ toff, unsigned short *__restrict__ rofs,
    int *__restrict__ perm, int *__restrict__ pci) {
  __shared__ int cnt[BP_RMAX * 8];
  __shared__ int tot[8];
  const int b = blockIdx.x;
  const int r0 = blk[b], r1 = blk[b + 1], R = r1 - r0;
  if (R > BP_RMAX) return; /* the planner never makes such a block */
  for (int i = threadIdx.x; i < R; i += blockDim.x) {
    int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = rp[r0 + i]; e < rp[r0 + i + 1]; ++e) c[(int)(ci[e] / W)]++;
    for (int k = 0; k < P; ++k) cnt[i * P + k] = c[k];
  }
  __syncthreads();
  if ((int)threadIdx.x < P) {
    const int k = threadIdx.x;
    int run = 0;
    for (int i = 0; i < R; ++i) {
      const int v = cnt[i * P + k];
      cnt[i * P + k] = run;
      run += v;
    }
    tot[k] = run;
  }
  __syncthreads();
  int start[8];
  start[0] = rp[r0];
  for (int k = 1; k < P; ++k) start[k] = start[k - 1] + tot[k - 1];
  if ((int)threadIdx.x < P) toff[(size_t)b * P + threadIdx.x] = start[threadIdx.x];
  if (b == nblk - 1 && threadIdx.x == 0) toff[(size_t)nblk * P] = rp[r1];
  unsigned short *ro = rofs + (size_t)P * ((size_t)r0 + b);
  for (int q = threadIdx.x; q < P * (R + 1); q += blockDim.x) {
    const int k = q / (R + 1), i = q - k * (R + 1);
    ro[q] = (unsigned short)(i < R ? cnt[i * P + k] : tot[k]);
  }
  for (int i = threadIdx.x; i < R; i += blockDim.x) {
    int fill[8];
    for (int k = 0; k < P; ++k) fill[k] = start[k] + cnt[i * P + k];
    for (int e = rp[r0 + i]; e < rp[r0 + i + 1]; ++e) {
      const int c = ci[e];
      const int k = (int)(c / W);
      perm[fill[k]] = e;
      pci[fill[k]] = c;
      fill[k]++;
    }
  }
}

}  // namespace

extern "C" {

size_t dpdlp_dev_transpose_scratch_bytes(long long nnz) {
  const long long L = 256 * ntiles_of(nnz > 0 ? nnz : 1);
  const long long nb = (L + SCAN_CHUNK - 1) / SCAN_CHUNK;
  return (size_t)(L + nb + 64) * sizeof(int);
}

int dpdlp_dev_transpose(int m, int n, long long nnz, const int *rp, const int *ci,
                        const double *v, int *t_ci, double *t_v, void *scratch, void *stream) {
  if (nnz <= 0) return 0;
  if (nnz > 0x7fffffffLL || !scratch) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const long long ntiles = ntiles_of(nnz);
  const long long L = 256 * ntiles;
  const long long nb = (L + SCAN_CHUNK - 1) / SCAN_CHUNK;
  int *hist = (int *)scratch, *bsum = hist + L;
  int bits = 0;
  while (bits < 31 && (1LL << bits) < n) ++bits;
  const int passes = bits > 8 ? (bits + 7) / 8 : 1;
  /* the two position buffers are the halves of t_v's storage */
  int *half[2] = {(int *)t_v, (int *)t_v + nnz};
  const int *vin = nullptr;
  for (int p = 0; p < passes; ++p) {
    int *vout = p == passes - 1 ? t_ci : half[p & 1];
    const int shift = 8 * p;
    k_rs_hist<<<(unsigned)ntiles, DPDLP_BS, 0, s>>>(ci, vin, nnz, shift, (int)ntiles, hist);
    k_scan_chunks<<<(unsigned)nb, DPDLP_BS, 0, s>>>(hist, L, bsum);
    k_scan_totals<<<1, DPDLP_BS, 0, s>>>(bsum, (int)nb);
    k_scan_add<<<(unsigned)nb, DPDLP_BS, 0, s>>>(hist, L, bsum);
    k_rs_scatter<<<(unsigned)ntiles, DPDLP_BS, 0, s>>>(ci, vin, vout, nnz, shift, (int)ntiles,
                                                       hist);
    vin = vout;
  }
  k_tr_finish<<<dpdlp_ew_grid(nnz), DPDLP_BS, 0, s>>>(rp, m, v, t_ci, t_v, nnz);
  return (int)hipGetLastError();
}

int dpdlp_dev_build_panels(int rows, int cols, const int *rp, const int *ci, const int *blk,
                           int nblk, int P, int *toff, unsigned short *rofs, int *perm, int *pci,
                           void *stream) {
  (void)rows;
  if (P < 1 || P > 8 || nblk < 0) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (nblk == 0) return (int)hipMemsetAsync(toff, 0, sizeof(int), s);
  long long W = ((long long)cols + P - 1) / P;
  if (W < 1) W = 1;
  k_build_panels<<<(unsigned)nblk, DPDLP_BS, 0, s>>>(rp, ci, blk, nblk, P, W, toff, rofs, perm, pci);
  return (int)hipGetLastError();
}

}  /* extern "C" */
