/*
 * Shared device code of the single-LP and batched SpMV kernels: streaming
 * loads, the deterministic row reduction, the CSR-stream row-block body, and
 * the fused epilogues (one per reference kernel family).
 */
#pragma once

#include "dev_common.h"

/* ------------------------------------------------------------------------ */
/* CSR-stream SpMV with a fused epilogue                                     */
/* ------------------------------------------------------------------------ */

template <typename T>
__device__ __forceinline__ T ntload(const T *p) {
  return __builtin_nontemporal_load(p);
}

/* A gathered entry as a double; with SX, the entry times xs (the gathered
 * vector is then xs times the stored one, formed as that product). */
template <bool SX, typename XT>
__device__ __forceinline__ double gx(XT v, double xs) {
  return SX ? (double)v * xs : (double)v;
}

/* One thread's share of a long row [p0, p1): four independent gathers in
 * flight, four accumulators combined in a fixed order.  Positions are
 * unsigned so that stepping past a row that ends near 2^31 - 1 (the largest
 * nonzero count of the int32 CSR interface) cannot overflow. */
template <typename VT, typename XT, bool SX = false>
__device__ __forceinline__ double row_sum_wg(const int *__restrict__ ci,
                                             const VT *__restrict__ val,
                                             const XT *__restrict__ x, int p0, int p1,
                                             double xs = 1.0) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  const unsigned e = (unsigned)p1;
  unsigned k = (unsigned)p0 + threadIdx.x;
  for (; k + 3 * DPDLP_BS < e; k += 4 * DPDLP_BS) {
    int c0 = ntload(ci + k), c1 = ntload(ci + k + DPDLP_BS);
    int c2 = ntload(ci + k + 2 * DPDLP_BS), c3 = ntload(ci + k + 3 * DPDLP_BS);
    VT v0 = ntload(val + k), v1 = ntload(val + k + DPDLP_BS);
    VT v2 = ntload(val + k + 2 * DPDLP_BS), v3 = ntload(val + k + 3 * DPDLP_BS);
    XT x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
    a0 += (double)v0 * gx<SX>(x0, xs);
    a1 += (double)v1 * gx<SX>(x1, xs);
    a2 += (double)v2 * gx<SX>(x2, xs);
    a3 += (double)v3 * gx<SX>(x3, xs);
  }
  for (; k < e; k += DPDLP_BS) a0 += (double)ntload(val + k) * gx<SX>(x[ntload(ci + k)], xs);
  return (a0 + a1) + (a2 + a3);
}


/* Deterministic row sums of the staged products prod[srp[r] .. srp[r+1]) for
 * the R rows of a block.  Phase A: rows of at most RR_PER_LANE entries per
 * lane are summed by a group of g lanes (g = 256/R rounded down to a power of
 * two, at most 64; one lane per row when R > 256).  Phase B: every longer row
 * is summed by a whole wavefront (lanes strided, then wave_sum),
 * long rows dealt round-robin to the 4 wavefronts.  emit(r, sum) is called
 * once per row by exactly one lane.  The wavefronts take the long rows one
 * at a time (each with its own strided loads and reduction), so the
 * groups keep every row up to RR_PER_LANE entries per lane: a block of 200
 * rows of 20 entries is summed in about 1 us by single lanes, against about
 * 30 us by wavefronts. */
constexpr int RR_PER_LANE = 64;

template <class Emit>
__device__ __forceinline__ void rows_reduce(const double *__restrict__ prod,
                                            const int *__restrict__ srp, int R, Emit emit) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int g = 1;
  if (R <= DPDLP_BS) {
    g = 1 << (31 - __clz(DPDLP_BS / R));
    if (g > DPDLP_WAVE) g = DPDLP_WAVE;
  }
  const int lim = g * RR_PER_LANE;
  if (g == 1) {
    for (int r = tid; r < R; r += DPDLP_BS) {
      const int a = srp[r], e = srp[r + 1];
      if (e - a <= lim) {
        double acc = 0.0;
        for (int q = a; q < e; ++q) acc += prod[q];
        emit(r, acc);
      }
    }
  } else {
    const int grp = tid / g, ln = tid & (g - 1);
    double acc = 0.0;
    bool mine = false;
    if (grp < R) {
      const int a = srp[grp], e = srp[grp + 1];
      mine = (e - a) <= lim;
      if (mine)
        for (int q = a + ln; q < e; q += g) acc += prod[q];
    }
    acc = group_sum(acc, g);
    if (mine && ln == 0) emit(grp, acc);
  }
  int cnt = 0;
  for (int base = 0; base < R; base += DPDLP_WAVE) {
    const int r = base + lane;
    bool lng = false;
    if (r < R) lng = (srp[r + 1] - srp[r]) > lim;
    unsigned long long mask = __ballot(lng);
    while (mask) {
      const int j = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      if ((cnt & 3) == wid) {
        const int rr = base + j, a = srp[rr], e = srp[rr + 1];
        double acc = 0.0;
        for (int q = a + lane; q < e; q += DPDLP_WAVE) acc += prod[q];
        acc = wave_sum(acc);
        if (lane == 0) emit(rr, acc);
      }
      ++cnt;
    }
  }
}

/* Epilogue operands that do not depend on the row's sum are loaded at the
 * start of the row block, so their latency overlaps the matrix stream.
 * Epilogues with a Pre type and pre()/apply() opt in; the rest load inside
 * operator(). */
template <class Epi, class = void> struct EpiPre {
  struct T {};
  __device__ static T load(const Epi &, int) { return T{}; }
  __device__ static void apply(const Epi &e, int r, double ax, const T &, double *s,
                               double *mx) {
    e(r, ax, s, mx);
  }
};
template <class Epi> struct EpiPre<Epi, decltype((void)sizeof(typename Epi::Pre))> {
  using T = typename Epi::Pre;
  __device__ static T load(const Epi &e, int r) { return e.pre(r); }
  __device__ static void apply(const Epi &e, int r, double ax, const T &q, double *s,
                               double *mx) {
    e.apply(r, ax, q, s, mx);
  }
};

/* One CSR-stream row block [r0, r1): stream, gather, stage in LDS, reduce each
 * row (rows_reduce) and call epi(row, sum, s, mx) once per row.  The caller
 * owns the LDS arrays (prod[CAP], srp[257], ssum[256]) and the per-thread
 * partial accumulators s[], mx[].  A block of more than `cap` nonzeros (the
 * capacity its matrix was planned with, at most CAP: then it is one long row)
 * is summed by the whole workgroup without staging.  Every thread of the
 * workgroup must call it. */
template <typename VT, typename XT, int CAP, class Epi, bool SX = false>
__device__ __forceinline__ void spmv_block(const int *__restrict__ rp, const int *__restrict__ ci,
                                           const VT *__restrict__ val, const XT *__restrict__ x,
                                           int r0, int r1, const Epi &epi, double *s, double *mx,
                                           double *prod, int *srp, double *ssum, int cap = CAP,
                                           double xs = 1.0) {
  constexpr int UNR = 8;
  using PF = EpiPre<Epi>;
  const int tid = threadIdx.x;
  const int R = r1 - r0;
  typename PF::T pre;
  if (tid < R) pre = PF::load(epi, r0 + tid);
  const int p0 = rp[r0], p1 = rp[r1];
  const int nz = p1 - p0;
  if (nz > cap || nz > CAP) {
    /* one long row, whole workgroup */
    double acc = row_sum_wg<VT, XT, SX>(ci, val, x, p0, p1, xs);
    acc = wave_sum(acc);
    if ((tid & 63) == 0) ssum[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
      double tot = (ssum[0] + ssum[1]) + (ssum[2] + ssum[3]);
      PF::apply(epi, r0, tot, pre, s, mx);
    }
    return;
  }
  for (int t = tid; t <= R; t += DPDLP_BS) srp[t] = rp[r0 + t] - p0;
  for (int base = tid; base < nz; base += DPDLP_BS * UNR) {
    int cidx[UNR];
    VT vv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      int k = base + u * DPDLP_BS;
      if (k < nz) {
        cidx[u] = ntload(ci + p0 + k);
        vv[u] = ntload(val + p0 + k);
      }
    }
    XT xx[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      int k = base + u * DPDLP_BS;
      if (k < nz) xx[u] = x[cidx[u]];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      int k = base + u * DPDLP_BS;
      if (k < nz) prod[k] = (double)vv[u] * gx<SX>(xx[u], xs);
    }
  }
  __syncthreads();
  rows_reduce(prod, srp, R, [&](int r, double v) { ssum[r] = v; });
  __syncthreads();
  if (tid < R) PF::apply(epi, r0 + tid, ssum[tid], pre, s, mx);
}

/* ---- epilogues ---- */

struct EpiStore {
  static constexpr int NS = 0, NM = 0;
  double *y;
  int partials;
  __device__ void operator()(int r, double ax, double *, double *) const { y[r] = ax; }
};

/* Phase 1: pdlp_core_op.cu:393-418 then the x half of :445-462. */
struct EpiPrimal {
  static constexpr int NS = 3, NM = 0;
  double *x_cur;
  const double *x_init, *c, *lv, *uv;
  double *x_ref, *x_pdhg, *z, *dx;
  float *x_ref32;
  /* anchored FP32 shadows: aty0 = A' y_init is added to the FP32 product of
   * the gathered y_cur - y_init; delta32: x_ref32 holds x_ref - x_init */
  const double *aty0;
  int delta32;
  double tau, w, rho;
  int major, fpe, partials;
  /* the column's operands, loadable before its A'y is known */
  struct Pre {
    double xc, cj, l, u, xi, a0;
  };
  __device__ Pre pre(int j) const {
    return Pre{x_cur[j], c[j], lv[j], uv[j], x_init[j], aty0 ? aty0[j] : 0.0};
  }
  __device__ void apply(int j, double aty, const Pre &q, double *s, double *) const {
    double xc = q.xc;
    if (aty0) aty += q.a0;
    double t = xc - tau * (q.cj - aty);
    double p = fmax(q.l, fmin(t, q.u));
    double xr;
    if (major) {
      x_pdhg[j] = p;
      z[j] = (p - t) / tau;
    }
    xr = 2.0 * p - xc;
    /* the dual phase gathers exactly one of the two copies */
    if (x_ref) x_ref[j] = xr;
    if (x_ref32) x_ref32[j] = (float)(delta32 ? xr - q.xi : xr);
    if (fpe) {
      double d = xr - xc;
      dx[j] = d;
      blue_acc(d, s[0], s[1], s[2]);
    }
    double rr = rho * xr + (1.0 - rho) * xc;
    x_cur[j] = w * rr + (1.0 - w) * q.xi;
  }
  __device__ void operator()(int j, double aty, double *s, double *mx) const {
    apply(j, aty, pre(j), s, mx);
  }
};

/* Phase 2: pdlp_core_op.cu:420-443 then the y half of :445-462.  The
 * reflected dual itself is not stored: nothing reads it after this step. */
struct EpiDual {
  static constexpr int NS = 3, NM = 0;
  double *y_cur;
  const double *y_init, *lc, *uc;
  double *y_pdhg, *dy;
  float *y_cur32;
  /* anchored FP32 shadows: ax0 = A x_init is added to the FP32 product of
   * the gathered x_ref - x_init; delta32: y_cur32 holds y_cur - y_init */
  const double *ax0;
  int delta32;
  double sigma, w, rho;
  int major, fpe, partials;
  struct Pre {
    double yc, l, u, yi, a0;
  };
  __device__ Pre pre(int i) const {
    return Pre{y_cur[i], lc[i], uc[i], y_init[i], ax0 ? ax0[i] : 0.0};
  }
  __device__ void apply(int i, double ax, const Pre &q, double *s, double *) const {
    double yc = q.yc;
    if (ax0) ax += q.a0;
    double t = yc / sigma - ax;
    double p = fmax(-q.u, fmin(t, -q.l));
    double yr;
    if (major) {
      double yp = (t - p) * sigma;
      y_pdhg[i] = yp;
      yr = 2.0 * yp - yc;
    } else {
      yr = 2.0 * (t - p) * sigma - yc;
    }
    if (fpe) {
      double d = yr - yc;
      dy[i] = d;
      blue_acc(d, s[0], s[1], s[2]);
    }
    double rr = rho * yr + (1.0 - rho) * yc;
    double yn = w * rr + (1.0 - w) * q.yi;
    y_cur[i] = yn;
    if (y_cur32) y_cur32[i] = (float)(delta32 ? yn - q.yi : yn);
  }
  __device__ void operator()(int i, double ax, double *s, double *mx) const {
    apply(i, ax, pre(i), s, mx);
  }
};

/* A' dy for the fixed-point error, pdlp_core_op.cu:666-692. */
struct EpiFpe {
  static constexpr int NS = 1, NM = 0;
  double *dprod;
  const double *dx;
  int partials;
  __device__ void operator()(int j, double ax, double *s, double *) const {
    dprod[j] = ax;
    s[0] += ax * dx[j];
  }
};

/* Rows of compute_residual_kernel (utils.cu:503-524) + norms + restart dist. */
struct EpiResP {
  static constexpr int NS = 7, NM = 1;
  double *pprod;
  const double *lc, *uc, *D_r, *y, *y_init;
  int dist, partials;
  __device__ void operator()(int i, double ax, double *s, double *mx) const {
    pprod[i] = ax;
    double l = lc[i], u = uc[i];
    double cl = fmax(l, fmin(ax, u));
    double r = (ax - cl) * D_r[i];
    blue_acc(r, s[0], s[1], s[2]);
    mx[0] = fmax(mx[0], fabs(r));
    double yv = y[i];
    s[3] += fmax(yv, 0.0) * lfin(l) + fmin(yv, 0.0) * lfin(u);
    if (dist) blue_acc(yv - y_init[i], s[4], s[5], s[6]);
  }
};

/* Columns of compute_residual_kernel (utils.cu:525-530) + objectives. */
struct EpiResD {
  static constexpr int NS = 8, NM = 1;
  double *dprod;
  const double *c, *z, *D_c, *x_obj, *x_init;
  int dist, partials;
  __device__ void operator()(int j, double ax, double *s, double *mx) const {
    dprod[j] = ax;
    double cj = c[j], zj = z[j];
    double r = (cj - ax - zj) * D_c[j];
    blue_acc(r, s[0], s[1], s[2]);
    mx[0] = fmax(mx[0], fabs(r));
    double xv = x_obj[j];
    s[3] += cj * xv;
    s[4] += zj * xv;
    if (dist) blue_acc(xv - x_init[j], s[5], s[6], s[7]);
  }
};

/* A dx for the primal ray (utils.cu:750-793). */
struct EpiInfP {
  static constexpr int NS = 1, NM = 1;
  double *pprod, *pslack;
  const double *lc, *uc, *D_r, *dy;
  int partials;
  __device__ void operator()(int i, double ax, double *s, double *mx) const {
    pprod[i] = ax;
    double l = lc[i], u = uc[i], d = dy[i];
    s[0] += fmax(d, 0.0) * lfin(l) + fmin(d, 0.0) * lfin(u);
    double v = (fmax(0.0, -ax) * (double)(isfinite(l) ? 1 : 0) +
                fmax(0.0, ax) * (double)(isfinite(u) ? 1 : 0)) * D_r[i];
    pslack[i] = v;
    mx[0] = fmax(mx[0], fabs(v));
  }
};

/* A' dy for the dual ray (utils.cu:755-797). */
struct EpiInfD {
  static constexpr int NS = 2, NM = 1;
  double *dprod, *dslack;
  const double *lv, *uv, *D_c, *c, *dx;
  int partials;
  __device__ void operator()(int j, double ax, double *s, double *mx) const {
    dprod[j] = ax;
    double l = lv[j], u = uv[j], nd = -ax;
    s[0] += fmax(nd, 0.0) * lfin(l) + fmin(nd, 0.0) * lfin(u);
    s[1] += c[j] * dx[j];
    double v = (fmax(0.0, nd) * (double)(isfinite(l) ? 0 : 1) -
                fmin(0.0, nd) * (double)(isfinite(u) ? 0 : 1)) * D_c[j];
    dslack[j] = v;
    mx[0] = fmax(mx[0], fabs(v));
  }
};

/* A (A' q) in the power iteration: store, <q, AA'q> and |AA'q| partials. */
struct EpiPow {
  static constexpr int NS = 4, NM = 0;
  double *eig;
  const double *next;
  int partials;
  __device__ void operator()(int i, double ax, double *s, double *) const {
    eig[i] = ax;
    s[0] += next[i] * ax;
    blue_acc(ax, s[1], s[2], s[3]);
  }
};

/* Reduced cost (pdlp_core_op.cu:31-40). */
struct EpiRc {
  static constexpr int NS = 0, NM = 0;
  double *rc;
  const double *c, *D_c;
  double gamma;
  int partials;
  __device__ void operator()(int j, double ax, double *, double *) const {
    rc[j] = (c[j] - ax) * D_c[j] / gamma;
  }
};

/* Row max |a| (Ruiz) or sequential sum of |a|^p (Pock-Chambolle). */
struct EpiStoreRow {
  static constexpr int NS = 0, NM = 0;
  double *y;
  int partials;
  __device__ void operator()(int r, double ax, double *, double *) const { y[r] = ax; }
};

