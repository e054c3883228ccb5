/*
 * Device helpers shared by the gfx950 kernels: wave-64 and workgroup
 * reductions with a fixed combination order (so every reduction is
 * deterministic for a fixed launch shape), and the three-accumulator
 * safe-scaling sum of squares of Blue (Anderson, ACM TOMS 44(1), Alg. 978)
 * used for every Euclidean norm.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define DPDLP_WAVE 64
#define DPDLP_BS 256 /* threads per workgroup for every kernel here */

/* Blue's thresholds and scaling factors for IEEE binary64. */
#define BLUE_TSML 0x1p-511
#define BLUE_TBIG 0x1p+486
#define BLUE_SSML 0x1p+537
#define BLUE_SBIG 0x1p-538

__device__ __forceinline__ void blue_acc(double x, double &s, double &m, double &b) {
  /* branch-free (adds +0.0 to the two accumulators x does not belong to), so
   * the accumulators stay in registers */
  double ax = fabs(x);
  bool big = ax > BLUE_TBIG, sml = ax < BLUE_TSML;
  double tb = ax * BLUE_SBIG, ts = ax * BLUE_SSML;
  b += big ? tb * tb : 0.0;
  s += sml ? ts * ts : 0.0;
  m += (big || sml) ? 0.0 : ax * ax;
}

/* Combines the three Blue accumulators into the norm (host and device). */
__host__ __device__ inline double blue_combine(double asml, double amed, double abig) {
  double scl, sumsq;
  if (abig > 0.0) {
    if (amed > 0.0 || isnan(amed)) abig += (amed * BLUE_SBIG) * BLUE_SBIG;
    scl = 1.0 / BLUE_SBIG;
    sumsq = abig;
  } else if (asml > 0.0) {
    if (amed > 0.0 || isnan(amed)) {
      double ymed = sqrt(amed);
      double ysml = sqrt(asml) / BLUE_SSML;
      double ymin = ysml > ymed ? ymed : ysml;
      double ymax = ysml > ymed ? ysml : ymed;
      double r = ymin / ymax;
      scl = 1.0;
      sumsq = ymax * ymax * (1.0 + r * r);
    } else {
      scl = 1.0 / BLUE_SSML;
      sumsq = asml;
    }
  } else {
    scl = 1.0;
    sumsq = amed;
  }
  return scl * sqrt(sumsq);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, DPDLP_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, DPDLP_WAVE));
  return v;
}

/*
 * Reduces NS sums and NM maxima over the 256-thread workgroup and writes the
 * block's partials to part[slot * stride + index] (sums first, then maxima;
 * block_partials: index = blockIdx.x).  Every thread must call it.  `lds`
 * needs (NS+NM)*4 doubles.
 */
template <int NS, int NM>
__device__ __forceinline__ void block_partials_at(double (&s)[NS > 0 ? NS : 1],
                                                  double (&mx)[NM > 0 ? NM : 1],
                                                  double *lds, double *part, int stride,
                                                  int slot0, int index) {
  const int lane = threadIdx.x & (DPDLP_WAVE - 1);
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double v = wave_sum(s[k]);
    if (lane == 0) lds[k * 4 + wid] = v;
  }
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    double v = wave_max(mx[k]);
    if (lane == 0) lds[(NS + k) * 4 + wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS + NM) {
    int k = threadIdx.x;
    const double *q = lds + k * 4;
    double v = (k < NS) ? ((q[0] + q[1]) + (q[2] + q[3]))
                        : fmax(fmax(q[0], q[1]), fmax(q[2], q[3]));
    part[(size_t)(slot0 + k) * stride + index] = v;
  }
}

template <int NS, int NM>
__device__ __forceinline__ void block_partials(double (&s)[NS > 0 ? NS : 1],
                                               double (&mx)[NM > 0 ? NM : 1],
                                               double *lds, double *part,
                                               int stride, int slot0) {
  block_partials_at<NS, NM>(s, mx, lds, part, stride, slot0, blockIdx.x);
}

__device__ __forceinline__ double lfin(double v) { return isfinite(v) ? v : 0.0; }
