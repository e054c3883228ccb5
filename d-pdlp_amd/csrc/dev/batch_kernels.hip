/*
 * Batched gfx950 kernels: one launch covers the row blocks (or element
 * blocks) of every LP of a batch.  A work item names its LP; the LP's device
 * pointers come from dpdlp_blp_t and its per-step parameters (step sizes,
 * Halpern weight, which epilogue outputs to write) from dpdlp_bctl_t, so LPs
 * of different sizes, restart states and convergence states share launches.
 *
 * Each LP is partitioned exactly as in a solo solve (its own row-block plan,
 * its own grid-stride element blocks), and its partial sums are reduced over
 * its own contiguous range in the same fixed order, so an LP's iterates are
 * bit-identical whether it runs alone or in a mix.  Work items of LPs whose
 * control flag is off return at once.
 */
#include "dpdlp_dev.h"
#include "spmv_core.h"

enum { K_PRIMAL, K_DUAL, K_FPE, K_RESP, K_RESD, K_INFP, K_INFD, K_PI_AT, K_PI_A };

/* whether an LP's work item runs for this kernel kind */
__device__ __forceinline__ bool kind_on(int kind, const dpdlp_bctl_t &c) {
  switch (kind) {
  case K_PRIMAL:
  case K_DUAL: return c.phase != 0;
  case K_FPE: return c.fpe != 0;
  case K_RESP:
  case K_RESD: return c.eval != 0;
  case K_INFP:
  case K_INFD: return c.inf != 0;
  default: return c.pi != 0;
  }
}

template <int KIND> struct BEpi;
template <> struct BEpi<K_PRIMAL> {
  using T = EpiPrimal;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &c) {
    T e;
    e.x_cur = L.x_cur; e.x_init = L.x_init; e.c = L.c; e.lv = L.lv; e.uv = L.uv;
    e.x_pdhg = L.x_pdhg; e.z = L.dual_slack; e.dx = L.dx;
    /* FP32 rounds write only the FP32 copies the next SpMV gathers */
    e.x_ref = c.use32 ? nullptr : L.x_ref;
    e.x_ref32 = c.use32 ? L.x_ref32 : nullptr;
    e.aty0 = nullptr; e.delta32 = 0; /* a batch keeps plain shadows */
    e.tau = c.tau; e.w = c.w; e.rho = c.rho; e.major = c.major; e.fpe = c.fpe; e.partials = c.fpe;
    return e;
  }
};
template <> struct BEpi<K_DUAL> {
  using T = EpiDual;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &c) {
    T e;
    e.y_cur = L.y_cur; e.y_init = L.y_init; e.lc = L.lc; e.uc = L.uc;
    e.y_pdhg = L.y_pdhg; e.dy = L.dy; e.y_cur32 = c.use32 ? L.y_cur32 : nullptr;
    e.ax0 = nullptr; e.delta32 = 0;
    e.sigma = c.sigma; e.w = c.w; e.rho = c.rho; e.major = c.major; e.fpe = c.fpe;
    e.partials = c.fpe;
    return e;
  }
};
template <> struct BEpi<K_FPE> {
  using T = EpiFpe;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &) {
    return T{L.dual_product, L.dx, 1};
  }
};
template <> struct BEpi<K_RESP> {
  using T = EpiResP;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &c) {
    T e;
    e.pprod = L.primal_product; e.lc = L.lc; e.uc = L.uc; e.D_r = L.D_r; e.y = L.y_pdhg;
    e.y_init = L.y_init; e.dist = c.restart_dist; e.partials = 1;
    return e;
  }
};
template <> struct BEpi<K_RESD> {
  using T = EpiResD;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &c) {
    T e;
    e.dprod = L.dual_product; e.c = L.c; e.z = L.dual_slack; e.D_c = L.D_c; e.x_obj = L.x_pdhg;
    e.x_init = L.x_init; e.dist = c.restart_dist; e.partials = 1;
    return e;
  }
};
template <> struct BEpi<K_INFP> {
  using T = EpiInfP;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &) {
    T e;
    e.pprod = L.primal_product; e.pslack = L.primal_slack; e.lc = L.lc; e.uc = L.uc;
    e.D_r = L.D_r; e.dy = L.dy; e.partials = 1;
    return e;
  }
};
template <> struct BEpi<K_INFD> {
  using T = EpiInfD;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &) {
    T e;
    e.dprod = L.dual_product; e.dslack = L.dual_slack; e.lv = L.lv; e.uv = L.uv; e.D_c = L.D_c;
    e.c = L.c; e.dx = L.dx; e.partials = 1;
    return e;
  }
};
template <> struct BEpi<K_PI_AT> {
  using T = EpiStore;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &) { return T{L.pi_dual, 0}; }
};
template <> struct BEpi<K_PI_A> {
  using T = EpiPow;
  __device__ static T make(const dpdlp_blp_t &L, const dpdlp_bctl_t &) {
    return T{L.pi_eig, L.pi_next, 1};
  }
};

/* the matrix and gathered vector of each kind */
template <int KIND> __device__ __forceinline__ bool on_At() {
  return KIND == K_PRIMAL || KIND == K_FPE || KIND == K_RESD || KIND == K_INFD || KIND == K_PI_AT;
}
template <int KIND> __device__ __forceinline__ const double *gathered(const dpdlp_blp_t &L) {
  switch (KIND) {
  case K_PRIMAL: return L.y_cur;
  case K_DUAL: return L.x_ref;
  case K_FPE: return L.dy;
  case K_RESP: return L.x_pdhg;
  case K_RESD: return L.y_pdhg;
  case K_INFP: return L.dx;
  case K_INFD: return L.dy;
  case K_PI_AT: return L.pi_next;
  default: return L.pi_dual;
  }
}

template <int KIND, int CAP>
__global__ __launch_bounds__(DPDLP_BS) void k_bspmv(const dpdlp_blp_t *__restrict__ lps,
                                                   const dpdlp_bctl_t *__restrict__ ctl,
                                                   const dpdlp_bblk_t *__restrict__ blist,
                                                   double *__restrict__ part, int stride,
                                                   int slot0) {
  using Epi = typename BEpi<KIND>::T;
  constexpr int NS = Epi::NS, NM = Epi::NM;
  __shared__ double prod[CAP];
  __shared__ int srp[DPDLP_BS + 1];
  __shared__ double ssum[DPDLP_BS];
  __shared__ double red[(NS + NM > 0 ? NS + NM : 1) * 4];
  const dpdlp_bblk_t bk = blist[blockIdx.x];
  const dpdlp_bctl_t c = ctl[bk.lp];
  if (!kind_on(KIND, c)) return;
  const dpdlp_blp_t &L = lps[bk.lp];
  const Epi epi = BEpi<KIND>::make(L, c);
  double s[NS > 0 ? NS : 1];
  double mx[NM > 0 ? NM : 1];
#pragma unroll
  for (int k = 0; k < (NS > 0 ? NS : 1); ++k) s[k] = 0.0;
#pragma unroll
  for (int k = 0; k < (NM > 0 ? NM : 1); ++k) mx[k] = 0.0;
  const bool T = on_At<KIND>();
  const int *rp = T ? L.At_rp : L.A_rp;
  const int *ci = T ? L.At_ci : L.A_ci;
  const bool f32 = (KIND == K_PRIMAL || KIND == K_DUAL) && c.use32;
  const int cap = T ? L.cap_At : L.cap_A;
  if (f32) {
    const float *v = T ? L.At_v32 : L.A_v32;
    const float *x = KIND == K_PRIMAL ? L.y_cur32 : L.x_ref32;
    spmv_block<float, float, CAP>(rp, ci, v, x, bk.a, bk.b, epi, s, mx, prod, srp, ssum, cap);
  } else {
    const double *v = T ? L.At_v : L.A_v;
    spmv_block<double, double, CAP>(rp, ci, v, gathered<KIND>(L), bk.a, bk.b, epi, s, mx, prod,
                                    srp, ssum, cap);
  }
  if (NS + NM > 0) {
    if (epi.partials) block_partials<NS, NM>(s, mx, red, part, stride, slot0);
  }
}

/* Per-LP fixed-order finalize over the LP's own [lo, hi) partials (same
 * summation pattern as the solo k_finalize).  sel: which control flag. */
__global__ __launch_bounds__(DPDLP_BS) void k_bfinalize(const double *__restrict__ part,
                                                       int stride, const int *__restrict__ ranges,
                                                       const dpdlp_blp_t *__restrict__ lps,
                                                       const dpdlp_bctl_t *__restrict__ ctl,
                                                       int sel, int slot0, unsigned maxmask,
                                                       int dst0, int hist) {
  __shared__ double red[4];
  const int lp = blockIdx.x, k = blockIdx.y;
  const dpdlp_bctl_t c = ctl[lp];
  if (!kind_on(sel, c)) return;
  const int lo = ranges[2 * lp], hi = ranges[2 * lp + 1];
  const int nparts = hi - lo;
  const bool ismax = (maxmask >> k) & 1u;
  const double *p = part + (size_t)(slot0 + k) * stride + lo;
  double acc = 0.0;
  if (ismax) {
    for (int i = threadIdx.x; i < nparts; i += DPDLP_BS) acc = fmax(acc, p[i]);
    acc = wave_max(acc);
  } else {
    for (int i = threadIdx.x; i < nparts; i += DPDLP_BS) acc += p[i];
    acc = wave_sum(acc);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = ismax ? fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))
                     : (red[0] + red[1]) + (red[2] + red[3]);
    double *S = lps[lp].scal;
    S[dst0 + k] = v;
    if (hist == 1) S[DPDLP_SC_PI_HIST + 4 * c.pi_slot + k] = v;       /* sigma^2 + 3 (overwritten) */
    if (hist == 2) S[DPDLP_SC_PI_HIST + 4 * c.pi_slot + 1 + k] = v;   /* residual sums */
  }
}

/* copies the 7 FPE slots to the restart / major groups per LP */
__global__ void k_bfpe_copy(const dpdlp_blp_t *__restrict__ lps, const dpdlp_bctl_t *ctl,
                            int nlp) {
  int lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp >= nlp) return;
  const dpdlp_bctl_t c = ctl[lp];
  if (!c.fpe) return;
  double *S = lps[lp].scal;
  for (int k = 0; k < 7; ++k) {
    double v = S[DPDLP_SL_FPE + k];
    if (c.fpe_to & 1) S[DPDLP_SC_FPE_RESTART + k] = v;
    if (c.fpe_to & 2) S[DPDLP_SC_FPE_MAJOR + k] = v;
  }
}

/* element blocks: (lp, j, G) = block j of the LP's G grid-stride blocks */
#define EW_LOOP(count)                                                          \
  for (long long i = (long long)bk.a * DPDLP_BS + threadIdx.x; i < (count);     \
       i += (long long)bk.b * DPDLP_BS)

__global__ void k_brestart_copy(const dpdlp_blp_t *__restrict__ lps, const dpdlp_bctl_t *ctl,
                                const dpdlp_bblk_t *__restrict__ el) {
  const dpdlp_bblk_t bk = el[blockIdx.x];
  const dpdlp_bctl_t c = ctl[bk.lp];
  if (!c.restart) return;
  const dpdlp_blp_t &L = lps[bk.lp];
  EW_LOOP((long long)L.n + L.m) {
    if (i < L.n) {
      if (c.restart & 1) {
        double v = L.x_pdhg[i];
        L.x_init[i] = v;
        L.x_cur[i] = v;
      }
    } else {
      int k = (int)(i - L.n);
      if (c.restart & 2) {
        double v = L.y_pdhg[k];
        L.y_init[k] = v;
        L.y_cur[k] = v;
        if (L.y_cur32) L.y_cur32[k] = (float)v;
      }
    }
  }
}

__global__ __launch_bounds__(DPDLP_BS) void k_binf_project(const dpdlp_blp_t *__restrict__ lps,
                                                          const dpdlp_bctl_t *ctl,
                                                          const dpdlp_bblk_t *__restrict__ el,
                                                          double *part, int stride, int slot0) {
  __shared__ double red[2 * 4];
  const dpdlp_bblk_t bk = el[blockIdx.x];
  const dpdlp_bctl_t c = ctl[bk.lp];
  if (!c.inf) return;
  const dpdlp_blp_t &L = lps[bk.lp];
  double s[1] = {0.0};
  double mx[2] = {0.0, 0.0};
  EW_LOOP((long long)L.n + L.m) {
    if (i < L.n) {
      double d = L.dx[i];
      if (isfinite(L.lv[i])) d = fmax(d, 0.0);
      if (isfinite(L.uv[i])) d = fmin(d, 0.0);
      L.dx[i] = d;
      mx[0] = fmax(mx[0], fabs(d));
    } else {
      int k = (int)(i - L.n);
      double d = L.dy[k];
      if (!isfinite(L.lc[k])) d = fmin(d, 0.0);
      if (!isfinite(L.uc[k])) d = fmax(d, 0.0);
      L.dy[k] = d;
      mx[1] = fmax(mx[1], fabs(d));
    }
  }
  block_partials<0, 2>(s, mx, red, part, stride, slot0);
}

__global__ void k_bscale_by_inv(const dpdlp_blp_t *__restrict__ lps, const dpdlp_bctl_t *ctl,
                                const dpdlp_bblk_t *__restrict__ el) {
  const dpdlp_bblk_t bk = el[blockIdx.x];
  const dpdlp_bctl_t c = ctl[bk.lp];
  if (!c.inf) return;
  const dpdlp_blp_t &L = lps[bk.lp];
  double a = L.scal[DPDLP_SL_INF];
  if (!(a > 0.0)) return;
  double sc = 1.0 / a;
  EW_LOOP(L.n) L.dx[i] *= sc;
}

__global__ __launch_bounds__(DPDLP_BS) void k_bpow_update(const dpdlp_blp_t *__restrict__ lps,
                                                         const dpdlp_bctl_t *ctl,
                                                         const dpdlp_bblk_t *__restrict__ el,
                                                         double *part, int stride, int slot0) {
  __shared__ double red[3 * 4];
  const dpdlp_bblk_t bk = el[blockIdx.x];
  const dpdlp_bctl_t c = ctl[bk.lp];
  if (!c.pi) return;
  const dpdlp_blp_t &L = lps[bk.lp];
  double s[3] = {0.0, 0.0, 0.0};
  double mx[1] = {0.0};
  const double neg = -L.scal[DPDLP_SL_PI];
  const double nrm = blue_combine(L.scal[DPDLP_SL_PI + 1], L.scal[DPDLP_SL_PI + 2],
                                  L.scal[DPDLP_SL_PI + 3]);
  const double inv = 1.0 / nrm;
  EW_LOOP(L.m) {
    double e = L.pi_eig[i];
    double r = L.pi_next[i] * neg + e;
    blue_acc(r, s[0], s[1], s[2]);
    L.pi_next[i] = e * inv;
  }
  block_partials<3, 0>(s, mx, red, part, stride, slot0);
}

/* ------------------------------------------------------------------------ */

#define ST(s) ((hipStream_t)(s))
#define RET(e)                                                                 \
  do {                                                                         \
    hipError_t _e = (e);                                                       \
    if (_e != hipSuccess) return (int)_e;                                      \
  } while (0)

template <int KIND>
static hipError_t launch_b(const dpdlp_batch_dev_t *b, int slot0, hipStream_t s) {
  const bool T = (KIND == K_PRIMAL || KIND == K_FPE || KIND == K_RESD || KIND == K_INFD ||
                  KIND == K_PI_AT);
  const dpdlp_bblk_t *list = T ? b->bAt : b->bA;
  int n = T ? b->nbAt : b->nbA;
  if (n <= 0) return hipSuccess;
  switch (T ? b->cap_At : b->cap) {
  case 1024:
    k_bspmv<KIND, 1024><<<n, DPDLP_BS, 0, s>>>(b->lps, b->ctl, list, b->part, b->stride, slot0);
    break;
  case 4096:
    k_bspmv<KIND, 4096><<<n, DPDLP_BS, 0, s>>>(b->lps, b->ctl, list, b->part, b->stride, slot0);
    break;
  default:
    k_bspmv<KIND, 2048><<<n, DPDLP_BS, 0, s>>>(b->lps, b->ctl, list, b->part, b->stride, slot0);
    break;
  }
  return hipGetLastError();
}

static hipError_t bfin(const dpdlp_batch_dev_t *b, const int *ranges, int sel, int slot0, int count,
                       unsigned maxmask, int dst0, int hist, hipStream_t s) {
  if (b->nlp <= 0) return hipSuccess;
  k_bfinalize<<<dim3(b->nlp, count), DPDLP_BS, 0, s>>>(b->part, b->stride, ranges, b->lps, b->ctl,
                                                       sel, slot0, maxmask, dst0, hist);
  return hipGetLastError();
}

extern "C" {

int dpdlp_bdev_restart_copy(const dpdlp_batch_dev_t *b, void *stream) {
  if (b->neNM <= 0) return 0;
  k_brestart_copy<<<b->neNM, DPDLP_BS, 0, ST(stream)>>>(b->lps, b->ctl, b->eNM);
  return (int)hipGetLastError();
}

int dpdlp_bdev_primal(const dpdlp_batch_dev_t *b, void *stream) {
  return (int)launch_b<K_PRIMAL>(b, DPDLP_SL_FPE, ST(stream));
}

int dpdlp_bdev_dual(const dpdlp_batch_dev_t *b, void *stream) {
  return (int)launch_b<K_DUAL>(b, DPDLP_SL_FPE + 3, ST(stream));
}

int dpdlp_bdev_phases(const dpdlp_batch_dev_t *b, void *stream) {
  RET(launch_b<K_PRIMAL>(b, DPDLP_SL_FPE, ST(stream)));
  RET(launch_b<K_DUAL>(b, DPDLP_SL_FPE + 3, ST(stream)));
  return 0;
}

int dpdlp_bdev_fpe(const dpdlp_batch_dev_t *b, void *stream) {
  hipStream_t s = ST(stream);
  RET(launch_b<K_FPE>(b, DPDLP_SL_FPE + 6, s));
  RET(bfin(b, b->rAt, K_FPE, DPDLP_SL_FPE, 3, 0u, DPDLP_SL_FPE, 0, s));
  RET(bfin(b, b->rA, K_FPE, DPDLP_SL_FPE + 3, 3, 0u, DPDLP_SL_FPE + 3, 0, s));
  RET(bfin(b, b->rAt, K_FPE, DPDLP_SL_FPE + 6, 1, 0u, DPDLP_SL_FPE + 6, 0, s));
  k_bfpe_copy<<<(b->nlp + 63) / 64, 64, 0, s>>>(b->lps, b->ctl, b->nlp);
  return (int)hipGetLastError();
}

int dpdlp_bdev_residual(const dpdlp_batch_dev_t *b, void *stream) {
  hipStream_t s = ST(stream);
  RET(launch_b<K_RESP>(b, DPDLP_SL_RESP, s));
  RET(bfin(b, b->rA, K_RESP, DPDLP_SL_RESP, 8, 1u << 7, DPDLP_SL_RESP, 0, s));
  RET(launch_b<K_RESD>(b, DPDLP_SL_RESD, s));
  RET(bfin(b, b->rAt, K_RESD, DPDLP_SL_RESD, 9, 1u << 8, DPDLP_SL_RESD, 0, s));
  return 0;
}

int dpdlp_bdev_infeasibility(const dpdlp_batch_dev_t *b, void *stream) {
  hipStream_t s = ST(stream);
  if (b->neNM > 0) {
    k_binf_project<<<b->neNM, DPDLP_BS, 0, s>>>(b->lps, b->ctl, b->eNM, b->part, b->stride,
                                                DPDLP_SL_INF);
    RET(hipGetLastError());
  }
  RET(bfin(b, b->rNM, K_INFP, DPDLP_SL_INF, 2, 3u, DPDLP_SL_INF, 0, s));
  if (b->neN > 0) {
    k_bscale_by_inv<<<b->neN, DPDLP_BS, 0, s>>>(b->lps, b->ctl, b->eN);
    RET(hipGetLastError());
  }
  RET(launch_b<K_INFP>(b, DPDLP_SL_INF + 2, s));
  RET(launch_b<K_INFD>(b, DPDLP_SL_INF + 4, s));
  RET(bfin(b, b->rA, K_INFP, DPDLP_SL_INF + 2, 2, 2u, DPDLP_SL_INF + 2, 0, s));
  RET(bfin(b, b->rAt, K_INFD, DPDLP_SL_INF + 4, 3, 4u, DPDLP_SL_INF + 4, 0, s));
  return 0;
}

int dpdlp_bdev_power_step(const dpdlp_batch_dev_t *b, void *stream) {
  hipStream_t s = ST(stream);
  RET(launch_b<K_PI_AT>(b, 0, s));
  RET(launch_b<K_PI_A>(b, DPDLP_SL_PI, s));
  RET(bfin(b, b->rA, K_PI_A, DPDLP_SL_PI, 4, 0u, DPDLP_SL_PI, 1, s));
  if (b->neM > 0) {
    k_bpow_update<<<b->neM, DPDLP_BS, 0, s>>>(b->lps, b->ctl, b->eM, b->part, b->stride,
                                              DPDLP_SL_PI + 4);
    RET(hipGetLastError());
  }
  RET(bfin(b, b->rM, K_PI_A, DPDLP_SL_PI + 4, 3, 0u, DPDLP_SL_PI + 4, 2, s));
  return 0;
}

} /* extern "C" */
