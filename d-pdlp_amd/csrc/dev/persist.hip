/*
 * Persistent solving: one workgroup runs one LP's whole solve loop on the
 * device, with no launch or host round trip per iteration.
 *
 * Workgroup b owns LP b of the launch: it runs the power iteration of
 * utils.cu:66-171 (from the normalised start vector the host left in
 * pi_next), then the PDHG loop of solver.cu:56-102 with the reference's
 * evaluation cadence, residuals (utils.cu:639-719), infeasibility rays
 * (utils.cu:721-813), termination test, restart test and PID primal weight
 * (pdlp_core_op.cu:577-636).  Every SpMV is a pass of the LP's CSR matrix in
 * which each wave streams its quarter of the entries and each lane sums a
 * row in storage order, with the same fused epilogues as the batched and
 * single-LP kernels (spmv_core.h); the
 * scalar decisions are the host controller's code compiled for the device
 * (control_impl.h).  Reductions combine per-thread partial sums over the
 * whole pass in a fixed order, so an LP's iterates do not depend on which
 * other LPs share the launch.
 *
 * Workgroups never wait on each other, so any number of LPs may share a
 * launch.  Each launch runs at most `budget` iterations per LP and saves the
 * LP's state in its descriptor; the host relaunches the LPs not yet done.
 */
#include "dpdlp_dev.h"
#include "spmv_core.h"

#include <stddef.h>

/* The descriptor and controller state are written by host C (batch.c) and
 * read here: both sides assert the same layout. */
static_assert(sizeof(hstate_t) == 224, "hstate_t layout");
static_assert(sizeof(dpdlp_blp_t) == 296, "dpdlp_blp_t layout");
static_assert(sizeof(dpdlp_plp_t) == 616, "dpdlp_plp_t layout");
static_assert(offsetof(dpdlp_plp_t, k) == 320 && offsetof(dpdlp_plp_t, h) == 360 &&
                  offsetof(dpdlp_plp_t, clk0) == 592 && offsetof(dpdlp_plp_t, stage) == 600 &&
                  offsetof(dpdlp_plp_t, iters_run) == 612,
              "dpdlp_plp_t layout");

#define CTL_FN static __device__ __forceinline__
#define CTL(x) dctl_##x
#include "control_impl.h"

namespace {

struct PShared {
  hstate_t h; /* the LP's controller state during the launch (thread 0) */
  double sigma_sq;
  long long clk0;
  int pi_iters, do_restart;
  double red[9 * 4];
  /* per-iteration control, written by thread 0 */
  double tau, sigma, w, rho;
  int major, fpe, fpe_to, restart, stop, ev_on, ev_rc, ev_inf, ev_final;
  int stage; /* the LP's stage; thread 0 writes it back to the descriptor */
  int cutA[5], cutAt[5]; /* each wave's rows of A and of A' */
};

/* LDS copy of the gathered vector when it has at most this many entries */
constexpr int XS_CAP = 6144;
/* a wave's block: at most this many entries (staged in LDS) and rows */
constexpr int WB_NZ = 512, WB_ROWS = 128;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Rows [0, rows) of a CSR matrix cut into 4 ranges of about equal nonzeros,
 * one per wave: cut[w] = first row whose start offset reaches w * nnz / 4. */
__device__ __forceinline__ void wave_ranges(const int *__restrict__ rp, int rows, int *cut) {
  const int t = threadIdx.x;
  if (t == 0) {
    cut[0] = 0;
    cut[4] = rows;
  } else if (t < 4) {
    const long long nnz = rp[rows];
    const long long target = nnz * t / 4;
    int lo = 0, hi = rows;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (rp[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    cut[t] = lo;
  }
}

/* One SpMV pass y = M x over all rows of a CSR matrix (cols columns) with a
 * fused epilogue epi(row, sum, s, mx).  Wave w walks rows [cut[w], cut[w+1])
 * in blocks of at most WB_ROWS rows and WB_NZ entries: it reads the block's
 * entries coalesced, gathers (from an LDS copy of x when x has at most
 * XS_CAP entries), stages the products in its own LDS slice, and each lane
 * then sums one row sequentially in storage order (the oracle's order) and
 * applies the epilogue, whose operands it loaded with the row pointers.  A
 * row longer than WB_NZ is summed by the whole wave (lane-strided, then
 * wave_sum).  Waves never wait on each other inside the pass; the
 * epilogue's per-thread partial sums accumulate over the pass. */
template <class Epi>
__device__ __forceinline__ void pass(const int *__restrict__ rp, const int *__restrict__ ci,
                                     const double *__restrict__ v, const double *x, int rows,
                                     int cols, const Epi &epi, double *s, double *mx,
                                     double *xs, double *wprod, const int *cut) {
  using PF = EpiPre<Epi>;
  const bool lds = cols <= XS_CAP;
  if (lds) {
    for (int i = threadIdx.x; i < cols; i += DPDLP_BS) xs[i] = x[i];
    __syncthreads();
  }
  const double *xg = lds ? xs : x;
  const int lane = threadIdx.x & (DPDLP_WAVE - 1), wid = threadIdx.x >> 6;
  double *prod = wprod + wid * WB_NZ;
  int r = cut[wid];
  const int rend = cut[wid + 1];
  while (r < rend) {
    /* row pointers of up to WB_ROWS rows: lane l holds rp[r + l] and
     * rp[r + 64 + l], every lane rp[r + 128] */
    const int ra = r + lane, rb = r + DPDLP_WAVE + lane, rc = r + WB_ROWS;
    const int pa = ra <= rend ? rp[ra] : 0;
    const int pb = rb <= rend ? rp[rb] : 0;
    const int pc = rc <= rend ? rp[rc] : 0;
    const int p0 = __shfl(pa, 0, DPDLP_WAVE);
    /* rows that end within WB_NZ entries of p0 form a prefix of the block */
    const bool fa = lane >= 1 && ra <= rend && pa - p0 <= WB_NZ;
    const bool fb = rb <= rend && pb - p0 <= WB_NZ;
    const int k = __popcll(__ballot(fa)) + __popcll(__ballot(fb)) +
                  ((rc <= rend && pc - p0 <= WB_NZ) ? 1 : 0);
    if (k == 0) {
      /* one row longer than WB_NZ: the whole wave */
      const int e1 = __shfl(pa, 1, DPDLP_WAVE);
      double acc = 0.0;
      for (unsigned q = (unsigned)p0 + lane; q < (unsigned)e1; q += DPDLP_WAVE)
        acc += v[q] * xg[ci[q]];
      acc = wave_sum(acc);
      if (lane == 0) epi(r, acc, s, mx);
      r += 1;
      continue;
    }
    /* every lane takes part in every shuffle (no shuffle under a
     * lane-dependent branch); the selections follow */
    const int nxa = __shfl(pa, (lane + 1) & (DPDLP_WAVE - 1), DPDLP_WAVE);
    const int nxb = __shfl(pb, (lane + 1) & (DPDLP_WAVE - 1), DPDLP_WAVE);
    const int pka = __shfl(pa, k & (DPDLP_WAVE - 1), DPDLP_WAVE);
    const int pkb = __shfl(pb, k & (DPDLP_WAVE - 1), DPDLP_WAVE);
    const int pb0 = __shfl(pb, 0, DPDLP_WAVE);
    const int p1 = k < DPDLP_WAVE ? pka : k < WB_ROWS ? pkb : pc;
    /* this lane's rows j = lane and j = 64 + lane: bounds and epilogue operands */
    const int sa = pa, ea = lane + 1 < DPDLP_WAVE ? nxa : pb0;
    const int sb = pb, eb = lane + 1 < DPDLP_WAVE ? nxb : pc;
    const bool ja = lane < k, jb = DPDLP_WAVE + lane < k;
    typename PF::T qa, qb;
    if (ja) qa = PF::load(epi, r + lane);
    if (jb) qb = PF::load(epi, r + DPDLP_WAVE + lane);
    /* stage the block's products, entries read coalesced */
    const int nz = p1 - p0;
    int cc[WB_NZ / DPDLP_WAVE];
    double vv[WB_NZ / DPDLP_WAVE];
#pragma unroll
    for (int u = 0; u < WB_NZ / DPDLP_WAVE; ++u) {
      const int q = lane + u * DPDLP_WAVE;
      cc[u] = q < nz ? ci[p0 + q] : 0;
      vv[u] = q < nz ? v[p0 + q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < WB_NZ / DPDLP_WAVE; ++u) {
      const int q = lane + u * DPDLP_WAVE;
      if (q < nz) prod[q] = vv[u] * xg[cc[u]];
    }
    wave_sync();
    if (ja) {
      double acc = 0.0;
      for (int q = sa - p0; q < ea - p0; ++q) acc += prod[q];
      PF::apply(epi, r + lane, acc, qa, s, mx);
    }
    if (jb) {
      double acc = 0.0;
      for (int q = sb - p0; q < eb - p0; ++q) acc += prod[q];
      PF::apply(epi, r + DPDLP_WAVE + lane, acc, qb, s, mx);
    }
    wave_sync(); /* the slice is refilled by the next block */
    r += k;
  }
  __syncthreads(); /* the pass's outputs are read by the next pass */
}

/* workgroup totals of NS sums and NM maxima into S[slot0 ...] (sums, then
 * maxima; thread k writes slot k), visible to every thread on return; every
 * thread must call it */
template <int NS, int NM>
__device__ __forceinline__ void wg_total(double (&s)[NS > 0 ? NS : 1],
                                         double (&mx)[NM > 0 ? NM : 1], PShared &sh, double *S,
                                         int slot0) {
  block_partials_at<NS, NM>(s, mx, sh.red, S, 1, slot0, 0);
  __syncthreads();
}

__device__ __forceinline__ void zero(double *a, int k) {
  for (int i = 0; i < k; ++i) a[i] = 0.0;
}

}  // namespace

__global__ __launch_bounds__(DPDLP_BS) void k_persist(dpdlp_plp_t *__restrict__ lps,
                                                     const pdhg_parameters_t *__restrict__ prmp,
                                                     double clk_hz, int budget) {
  __shared__ PShared sh;
  __shared__ double xs[XS_CAP];
  __shared__ double wprod[4 * WB_NZ];
  dpdlp_plp_t *P = lps + blockIdx.x;
  if (P->stage == DPDLP_PS_DONE) return;
  const pdhg_parameters_t &prm = *prmp;
  if (threadIdx.x == 0) {
    sh.stage = P->stage;
    sh.h = P->h;
    sh.sigma_sq = P->sigma_sq;
    sh.clk0 = P->clk0;
    sh.pi_iters = P->pi_iters;
    sh.do_restart = P->do_restart;
  }
  wave_ranges(P->L.A_rp, P->L.m, sh.cutA);
  wave_ranges(P->L.At_rp, P->L.n, sh.cutAt);
  __syncthreads();
  const dpdlp_blp_t &L = P->L;
  const int tid = threadIdx.x;
  const int m = L.m, n = L.n;
  double *S = L.scal;
  const termination_criteria_t *tc = &prm.termination_criteria;
  const int freq = prm.termination_evaluation_frequency;
  int used = 0;

  /* ---- power iteration (utils.cu:116-163): one step per pass pair ---- */
  while (sh.stage == DPDLP_PS_PI && used < budget) {
    double s[4], mx[1];
    {
      EpiStore e{L.pi_dual, 0};
      pass(L.At_rp, L.At_ci, L.At_v, L.pi_next, n, m, e, s, mx, xs, wprod, sh.cutAt);
    }
    zero(s, 4);
    mx[0] = 0.0;
    {
      EpiPow e{L.pi_eig, L.pi_next, 1};
      pass(L.A_rp, L.A_ci, L.A_v, L.pi_dual, m, n, e, s, mx, xs, wprod, sh.cutA);
    }
    wg_total<4, 0>(s, mx, sh, S, DPDLP_SL_PI);
    /* residual AA'q - sigma^2 q, then q <- AA'q / |AA'q| for the next step */
    const double neg = -S[DPDLP_SL_PI];
    const double inv = 1.0 / blue_combine(S[DPDLP_SL_PI + 1], S[DPDLP_SL_PI + 2],
                                          S[DPDLP_SL_PI + 3]);
    double r3[3] = {0.0, 0.0, 0.0}, rm[1] = {0.0};
    for (int i = tid; i < m; i += DPDLP_BS) {
      double e = L.pi_eig[i];
      double r = L.pi_next[i] * neg + e;
      blue_acc(r, r3[0], r3[1], r3[2]);
      L.pi_next[i] = e * inv;
    }
    wg_total<3, 0>(r3, rm, sh, S, DPDLP_SL_PI + 4);
    if (tid == 0) {
      sh.sigma_sq = S[DPDLP_SL_PI];
      sh.pi_iters++;
      if (dctl_blue(S + DPDLP_SL_PI + 4) < prm.sv_tol || sh.pi_iters >= prm.sv_max_iter) {
        double step = 0.998 / sqrt(sh.sigma_sq);
        dctl_init_state(&sh.h, &P->k, &prm, step);
        sh.do_restart = 0;
        sh.clk0 = wall_clock64();
        sh.stage = DPDLP_PS_PDHG;
      }
    }
    ++used;
    __syncthreads();
  }

  /* ---- PDHG loop (solver.cu:56-102) ---- */
  while (sh.stage == DPDLP_PS_PDHG && used < budget) {
    hstate_t *h = &sh.h;
    /* evaluation point at the top of iteration T = h->total */
    if (tid == 0) {
      int on = 0, rc = 0, inf = 0, fin = 0;
      if (h->total >= tc->iteration_limit) {
        on = 1;
        fin = 1;
      } else {
        rc = h->is_major || h->total == 0;
        if (rc || h->total % dctl_print_frequency(h->total) == 0) {
          on = 1;
          inf = h->is_major && h->total < 3 * freq;
        }
      }
      sh.ev_on = on;
      sh.ev_rc = rc;
      sh.ev_inf = inf;
      sh.ev_final = fin;
    }
    __syncthreads();
    if (sh.ev_on) {
      const int dist = sh.ev_rc && !sh.ev_final;
      {
        double s[7], mx[1];
        zero(s, 7);
        mx[0] = 0.0;
        EpiResP e;
        e.pprod = L.primal_product; e.lc = L.lc; e.uc = L.uc; e.D_r = L.D_r; e.y = L.y_pdhg;
        e.y_init = L.y_init; e.dist = dist; e.partials = 1;
        pass(L.A_rp, L.A_ci, L.A_v, L.x_pdhg, m, n, e, s, mx, xs, wprod, sh.cutA);
        wg_total<7, 1>(s, mx, sh, S, DPDLP_SL_RESP);
      }
      {
        double s[8], mx[1];
        zero(s, 8);
        mx[0] = 0.0;
        EpiResD e;
        e.dprod = L.dual_product; e.c = L.c; e.z = L.dual_slack; e.D_c = L.D_c;
        e.x_obj = L.x_pdhg; e.x_init = L.x_init; e.dist = dist; e.partials = 1;
        pass(L.At_rp, L.At_ci, L.At_v, L.y_pdhg, n, m, e, s, mx, xs, wprod, sh.cutAt);
        wg_total<8, 1>(s, mx, sh, S, DPDLP_SL_RESD);
      }
      if (sh.ev_inf) {
        /* ray projections (utils.cu:533-560) */
        double s0[1] = {0.0}, mx2[2] = {0.0, 0.0};
        for (int i = tid; i < n + m; i += DPDLP_BS) {
          if (i < n) {
            double d = L.dx[i];
            if (isfinite(L.lv[i])) d = fmax(d, 0.0);
            if (isfinite(L.uv[i])) d = fmin(d, 0.0);
            L.dx[i] = d;
            mx2[0] = fmax(mx2[0], fabs(d));
          } else {
            int k = i - n;
            double d = L.dy[k];
            if (!isfinite(L.lc[k])) d = fmin(d, 0.0);
            if (!isfinite(L.uc[k])) d = fmax(d, 0.0);
            L.dy[k] = d;
            mx2[1] = fmax(mx2[1], fabs(d));
          }
        }
        wg_total<0, 2>(s0, mx2, sh, S, DPDLP_SL_INF);
        const double a = S[DPDLP_SL_INF];
        if (a > 0.0) {
          const double sc = 1.0 / a;
          for (int i = tid; i < n; i += DPDLP_BS) L.dx[i] *= sc;
        }
        __syncthreads();
        {
          double s[1] = {0.0}, mx[1] = {0.0};
          EpiInfP e;
          e.pprod = L.primal_product; e.pslack = L.primal_slack; e.lc = L.lc; e.uc = L.uc;
          e.D_r = L.D_r; e.dy = L.dy; e.partials = 1;
          pass(L.A_rp, L.A_ci, L.A_v, L.dx, m, n, e, s, mx, xs, wprod, sh.cutA);
          wg_total<1, 1>(s, mx, sh, S, DPDLP_SL_INF + 2);
        }
        {
          double s[2] = {0.0, 0.0}, mx[1] = {0.0};
          EpiInfD e;
          e.dprod = L.dual_product; e.dslack = L.dual_slack; e.lv = L.lv; e.uv = L.uv;
          e.D_c = L.D_c; e.c = L.c; e.dx = L.dx; e.partials = 1;
          pass(L.At_rp, L.At_ci, L.At_v, L.dy, n, m, e, s, mx, xs, wprod, sh.cutAt);
          wg_total<2, 1>(s, mx, sh, S, DPDLP_SL_INF + 4);
        }
      }
      /* the reference loop's decisions at an evaluation point (solver.cu:60-84) */
      if (tid == 0) {
        int stop = 0;
        dctl_pending_fpe(h, S, 0);
        dctl_residual(h, &P->k, S, prm.optimality_norm);
        if (sh.ev_final) {
          h->reason = TERMINATION_REASON_ITERATION_LIMIT;
          stop = 1;
        } else {
          if (sh.ev_inf) dctl_infeasibility(h, &P->k, S);
          h->t_cum = (double)(wall_clock64() - sh.clk0) / clk_hz;
          dctl_check_termination(h, tc);
          if (h->reason != TERMINATION_REASON_UNSPECIFIED) {
            stop = 1;
          } else if (sh.ev_rc) {
            sh.do_restart = dctl_should_restart(h, &prm.restart_params, freq);
            if (sh.do_restart) {
              dctl_pid_update(h, &prm.restart_params, dctl_blue(S + DPDLP_SL_RESD + 5),
                              dctl_blue(S + DPDLP_SL_RESP + 4));
              h->inner = 0;
              h->fpe_last = INFINITY;
            }
          }
        }
        if (stop) sh.stage = DPDLP_PS_DONE;
        sh.stop = stop;
      }
      __syncthreads();
      if (sh.stop) break;
    }

    /* one iteration (solver.cu:85-102): this iteration's parameters */
    if (tid == 0) {
      const int is_major = ((h->total + 1) % freq) == 0;
      const int dr = sh.do_restart;
      const int major_variant =
          is_major || ((h->total + 2) % dctl_print_frequency(h->total + 2)) == 0;
      const int fpe = is_major || dr;
      sh.tau = h->step / h->omega;
      sh.sigma = h->step * h->omega;
      sh.w = (double)(h->inner + 1) / (h->inner + 2);
      sh.rho = prm.reflection_coefficient;
      sh.major = major_variant;
      sh.fpe = fpe;
      sh.fpe_to = (dr ? 1 : 0) | (is_major ? 2 : 0);
      sh.restart = dr;
      if (fpe) {
        if (dr) h->pend_rst = 1;
        if (is_major) h->pend_maj = 1;
      }
      h->is_major = is_major;
      h->inner++;
      h->total++;
      sh.do_restart = 0;
    }
    __syncthreads();
    if (sh.restart) {
      /* restart copies (pdlp_core_op.cu:621-632) */
      for (int i = tid; i < n + m; i += DPDLP_BS) {
        if (i < n) {
          double v = L.x_pdhg[i];
          L.x_init[i] = v;
          L.x_cur[i] = v;
        } else {
          int k = i - n;
          double v = L.y_pdhg[k];
          L.y_init[k] = v;
          L.y_cur[k] = v;
        }
      }
      __syncthreads();
    }
    {
      double s[3] = {0.0, 0.0, 0.0}, mx[1] = {0.0};
      EpiPrimal e;
      e.x_cur = L.x_cur; e.x_init = L.x_init; e.c = L.c; e.lv = L.lv; e.uv = L.uv;
      e.x_pdhg = L.x_pdhg; e.z = L.dual_slack; e.dx = L.dx; e.x_ref = L.x_ref;
      e.x_ref32 = nullptr; e.aty0 = nullptr; e.delta32 = 0;
      e.tau = sh.tau; e.w = sh.w; e.rho = sh.rho; e.major = sh.major; e.fpe = sh.fpe;
      e.partials = sh.fpe;
      pass(L.At_rp, L.At_ci, L.At_v, L.y_cur, n, m, e, s, mx, xs, wprod, sh.cutAt);
      if (sh.fpe) wg_total<3, 0>(s, mx, sh, S, DPDLP_SL_FPE);
    }
    {
      double s[3] = {0.0, 0.0, 0.0}, mx[1] = {0.0};
      EpiDual e;
      e.y_cur = L.y_cur; e.y_init = L.y_init; e.lc = L.lc; e.uc = L.uc; e.y_pdhg = L.y_pdhg;
      e.dy = L.dy; e.y_cur32 = nullptr; e.ax0 = nullptr; e.delta32 = 0;
      e.sigma = sh.sigma; e.w = sh.w; e.rho = sh.rho; e.major = sh.major; e.fpe = sh.fpe;
      e.partials = sh.fpe;
      pass(L.A_rp, L.A_ci, L.A_v, L.x_ref, m, n, e, s, mx, xs, wprod, sh.cutA);
      if (sh.fpe) wg_total<3, 0>(s, mx, sh, S, DPDLP_SL_FPE + 3);
    }
    if (sh.fpe) {
      /* fixed-point error finish (pdlp_core_op.cu:666-692): A'dy and <A'dy, dx> */
      double s[1] = {0.0}, mx[1] = {0.0};
      EpiFpe e{L.dual_product, L.dx, 1};
      pass(L.At_rp, L.At_ci, L.At_v, L.dy, n, m, e, s, mx, xs, wprod, sh.cutAt);
      wg_total<1, 0>(s, mx, sh, S, DPDLP_SL_FPE + 6);
      if (tid < 7) {
        const double v = S[DPDLP_SL_FPE + tid];
        if (sh.fpe_to & 1) S[DPDLP_SC_FPE_RESTART + tid] = v;
        if (sh.fpe_to & 2) S[DPDLP_SC_FPE_MAJOR + tid] = v;
      }
      __syncthreads();
    }
    ++used;
  }
  if (tid == 0) {
    P->stage = sh.stage;
    P->h = sh.h;
    P->sigma_sq = sh.sigma_sq;
    P->clk0 = sh.clk0;
    P->pi_iters = sh.pi_iters;
    P->do_restart = sh.do_restart;
    P->iters_run = used;
  }
}

extern "C" int dpdlp_pdev_run(dpdlp_plp_t *lps, int count, const pdhg_parameters_t *params,
                              double clk_hz, int budget, void *stream) {
  if (count <= 0) return 0;
  if (!lps || !params || budget < 1 || !(clk_hz > 0.0)) return (int)hipErrorInvalidValue;
  k_persist<<<count, DPDLP_BS, 0, (hipStream_t)stream>>>(lps, params, clk_hz, budget);
  return (int)hipGetLastError();
}
