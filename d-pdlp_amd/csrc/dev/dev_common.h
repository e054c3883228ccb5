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

/* Wavefront reductions through DPP lane permutations (register-to-register
 * within 16-lane rows) and four lane reads across the rows, instead of a
 * shuffle tree whose every step of a double crosses the LDS permute twice.
 * Every lane of the wavefront must be active. */
enum {
  DPP_XOR1 = 0xB1,         /* quad_perm [1,0,3,2] */
  DPP_XOR2 = 0x4E,         /* quad_perm [2,3,0,1] */
  DPP_HALF_MIRROR = 0x141, /* lane i of an 8-lane half <- lane 7 - i */
  DPP_MIRROR = 0x140       /* lane i of a 16-lane row <- lane 15 - i */
};

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

/* Sum over aligned groups of g lanes (g = 1, 2, 4, ..., 64, the same in
 * every lane), in every lane of the group: pairs, quads, 8-lane halves
 * (lane i with 7 - i), 16-lane rows (lane i with 15 - i), then the rows'
 * sums read across.  Both lanes of a pair add the same two values, so every
 * lane of a group holds the same bits. */
__device__ __forceinline__ double group_sum(double v, int g) {
  if (g >= 2) v += dpp_f64<DPP_XOR1>(v);
  if (g >= 4) v += dpp_f64<DPP_XOR2>(v);
  if (g >= 8) v += dpp_f64<DPP_HALF_MIRROR>(v);
  if (g >= 16) v += dpp_f64<DPP_MIRROR>(v);
  if (g >= 32) {
    const double r0 = readlane_f64(v, 0), r1 = readlane_f64(v, 16);
    const double r2 = readlane_f64(v, 32), r3 = readlane_f64(v, 48);
    v = g == 32 ? ((threadIdx.x & 32) ? r2 + r3 : r0 + r1) : (r0 + r1) + (r2 + r3);
  }
  return v;
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum(v, DPDLP_WAVE); }

__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_f64<DPP_XOR1>(v));
  v = fmax(v, dpp_f64<DPP_XOR2>(v));
  v = fmax(v, dpp_f64<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp_f64<DPP_MIRROR>(v));
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
              fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
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
