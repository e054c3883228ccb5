/*
 * Batched solving: many LPs advance together through the single-LP algorithm
 * of solver.c (the reference's optimize(), src/solver.cu:32-127), each with
 * its own step size, primal weight, restart state and termination state.
 *
 * Time is cut into windows of W = gcd(10, termination_evaluation_frequency)
 * iterations.  Every evaluation point of the reference loop (T % f(T) == 0
 * with f(T) = 10, 100, ..., and the major points T % eval_freq == 0) is a
 * multiple of W, and an LP always starts its loop at a window boundary, so
 * all evaluations fall on window starts:
 *
 *   window start  one residual (+ infeasibility) launch for every LP that
 *                 evaluates, one 2-D device-to-host copy of every LP's scalar
 *                 slots (and power-iteration history), one synchronisation;
 *                 the host then runs the reference's per-LP decisions
 *                 (termination, restart, PID weight) from those values;
 *   W rounds      per round at most: restart copy, primal phase, dual phase,
 *                 fixed-point-error finish, power-iteration step, each ONE
 *                 launch over the row blocks of every LP in flight, driven by
 *                 a per-LP control record written for the whole window.
 *
 * Finished LPs leave at the window start where they terminate (result copied
 * out, state freed); new LPs join there too, so slots are backfilled.
 */
#define _POSIX_C_SOURCE 200809L
#include "control.h"
#include "cupdlpx.h"
#include "dpdlp_dev.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"
#include "solver_internal.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define BCHK(call)                                                             \
  do {                                                                         \
    hipError_t _e = (hipError_t)(call);                                        \
    if (_e != hipSuccess) {                                                    \
      dpdlp_set_error("HIP error %d (%s) at %s:%d", (int)_e,                   \
                      hipGetErrorString(_e), __FILE__, __LINE__);              \
      return -1;                                                               \
    }                                                                          \
  } while (0)

/* slot stages: FREE -> RESERVED (being set up, possibly on another thread)
 * -> READY (set up, waiting for a window start) -> PI -> PDHG -> FREE */
enum { S_FREE = 0, S_RESERVED, S_READY, S_PI, S_PDHG, S_ZOMBIE, S_PERSIST };
/* S_ZOMBIE: retired, but the work lists still name the slot (its items
 * return at once: the slot's control records are zero); it becomes FREE at
 * the next list rebuild.  S_PERSIST: the LP runs whole in the persistent
 * kernel (one workgroup, its own loop on the device), outside the work lists */


enum { NLIST = 5 };

/* the persistent kernel reads dpdlp_plp_t as written here (persist.hip
 * asserts the same layout on the device side) */
_Static_assert(sizeof(hstate_t) == 224, "hstate_t layout");
_Static_assert(sizeof(dpdlp_blp_t) == 296, "dpdlp_blp_t layout");
_Static_assert(sizeof(dpdlp_plp_t) == 616, "dpdlp_plp_t layout");
_Static_assert(offsetof(dpdlp_plp_t, k) == 320 && offsetof(dpdlp_plp_t, h) == 360 &&
                   offsetof(dpdlp_plp_t, clk0) == 592 && offsetof(dpdlp_plp_t, stage) == 600 &&
                   offsetof(dpdlp_plp_t, iters_run) == 612,
               "dpdlp_plp_t layout");

typedef struct {
  int stage; /* read and written with atomics: set-up threads move it */
  hipEvent_t ev_ready; /* end of the set-up work on the set-up stream */
  long long tag;
  dpdlp_resident_t *R;
  int own_R;
  ctx_t c;
  hstate_t h;
  double t_add, t_pi0, t_loop0, rescale_time, power_time;
  int pi_iters, pi_issued;
  double sigma_sq;
  int do_restart;
  int ev_on, ev_rc, ev_inf, ev_final, fresh;
  int persist; /* 1: in the persistent group (stage S_PERSIST until it retires) */
} bslot_t;

typedef struct {
  long long tag;
  cupdlpx_result_t *res;
  double latency;
} bdone_t;

/* a retired LP whose result copies are in flight (complete at the next
 * synchronisation); its arena and resident problem stay alive until then */
typedef struct {
  long long tag;
  cupdlpx_result_t *res;
  size_t off;
  int n, m;
  double t_add;
  char *arena;
  size_t arena_cap;
  dpdlp_resident_t *R;
  int own_R;
} bpend_t;

typedef struct {
  char *p;
  size_t cap;
} barena_t;

struct dpdlp_batch {
  pthread_mutex_t mu; /* slot reservation and READY hand-over */
  pdhg_parameters_t prm;
  dpdlp_batch_opts_t o;
  int W, freq, rbw;
  hipStream_t s;
  /* power-iteration steps run on their own stream, beside the PDHG phases
   * (disjoint LPs, disjoint partial-sum rows); ev_ctl orders them after the
   * window's control upload, ev_pi joins them before the next window start */
  hipStream_t s_pi;
  hipEvent_t ev_ctl, ev_pi;
  int pi_pending;
  int cap;
  bslot_t *sl;
  int nhw, in_flight, dirty;
  long long list_items, zombie_items;
  long long last_activation;
  int reserve_failed;
  dpdlp_blp_t *d_lps, *h_lps;
  dpdlp_bctl_t *d_ctl_eval, *h_ctl_eval;
  dpdlp_bctl_t *d_ctl_win, *h_ctl_win[2];
  hipEvent_t ev_win[2];
  int win_buf;
  double *d_scal, *h_scal, *h_rb;
  dpdlp_bblk_t *d_list, *h_list;
  long long list_cap;
  int *d_rng, *h_rng;
  double *d_part;
  long long part_len;
  double *d_normals;
  int normals_len;
  dpdlp_batch_dev_t bd, bd_pi; /* lists of the iterating / power-iteration LPs */
  bdone_t *done;
  int ndone, done_head, done_cap;
  /* two pinned staging sets: retirements of this window start go to set
   * scur; the other set (copies complete at the last synchronisation) is
   * completed on the host while the next window runs */
  bpend_t *pend[2];
  int npend[2];
  double *h_stage[2];
  size_t stage_cap[2], stage_off[2];
  int scur;
  barena_t *pool; /* recycled LP-state arenas (guarded by mu) */
  int npool;
  size_t pool_bytes;
  dpdlp_resident_t **dead; /* retired owned problems awaiting free (guarded by mu) */
  int ndead, dead_cap;
  long long windows, syncs, launches;
  size_t state_bytes;
  /* window graphs: captured launch sequences, keyed by the list epoch, the
   * slot count and which kernel kinds each round launches */
  struct {
    long long epoch;
    int nl;
    unsigned long long sig;
    hipGraphExec_t exec;
    long long used;
  } gcache[8];
  long long epoch, graph_replays, graph_captures;
  /* live phase timing */
  hipEvent_t pev[32][3];
  double pev_bytes[32][2];
  int npev_created, npev_used, psamples;
  double p_ms[2], p_bytes[2];
  /* persistent group: descriptors of its LPs (device and pinned host copy),
   * the slot of each, the launch in flight on s_pp (ev_pp: kernel and the
   * descriptors' copy back complete), ev_pp_go: set-up work on s done */
  hipStream_t s_pp;
  hipEvent_t ev_pp, ev_pp_go;
  dpdlp_plp_t *d_plp, *h_plp;
  int *pp_slot;
  int pp_n, pp_inflight;
  pdhg_parameters_t *d_prm;
  double clk_hz;
  long long pp_launches;
};

static int stage_of(const bslot_t *q) { return __atomic_load_n(&q->stage, __ATOMIC_ACQUIRE); }
static void set_stage(bslot_t *q, int st) { __atomic_store_n(&q->stage, st, __ATOMIC_RELEASE); }
static int active(const bslot_t *q) {
  int st = stage_of(q);
  return st == S_PI || st == S_PDHG;
}

void dpdlp_batch_default_opts(dpdlp_batch_opts_t *o) {
  memset(o, 0, sizeof(*o));
  o->max_lps = 1024;
  o->precision = DPDLP_PREC_FP64;
  o->mixed_switch_tol = 1e-6;
  o->spmv_cap = 4096;
  o->solo_lps = 4;
  o->persist_budget = 250;
}

static int gcd(int a, int b) {
  while (b) {
    int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

void dpdlp_batch_free(dpdlp_batch_t *b) {
  if (!b) return;
  if (b->s) hipStreamSynchronize(b->s);
  if (b->s_pi) hipStreamSynchronize(b->s_pi);
  if (b->s_pp) hipStreamSynchronize(b->s_pp);
  for (int i = 0; i < b->cap && b->sl; ++i) {
    bslot_t *q = &b->sl[i];
    if (stage_of(q) == S_FREE) continue;
    dpdlp_i_ctx_free(&q->c);
    if (q->own_R) dpdlp_resident_free(q->R);
  }
  for (int i = b->done_head; i < b->ndone; ++i) cupdlpx_result_free(b->done[i].res);
  free(b->done);
  for (int g = 0; g < 2; ++g) {
    for (int i = 0; i < b->npend[g]; ++i) {
      cupdlpx_result_free(b->pend[g][i].res);
      if (b->pend[g][i].arena) hipFree(b->pend[g][i].arena);
      if (b->pend[g][i].own_R) dpdlp_resident_free(b->pend[g][i].R);
    }
    free(b->pend[g]);
    if (b->h_stage[g]) hipHostFree(b->h_stage[g]);
  }
  for (int i = 0; i < b->npool; ++i) hipFree(b->pool[i].p);
  free(b->pool);
  for (int i = 0; i < b->ndead; ++i) dpdlp_resident_free(b->dead[i]);
  free(b->dead);
  for (int i = 0; b->sl && i < b->cap; ++i)
    if (b->sl[i].ev_ready) hipEventDestroy(b->sl[i].ev_ready);
  free(b->sl);
  void *dev[] = {b->d_lps,  b->d_ctl_eval, b->d_ctl_win, b->d_scal, b->d_list,
                 b->d_rng,  b->d_part,     b->d_normals, b->d_plp,  b->d_prm};
  for (size_t k = 0; k < sizeof(dev) / sizeof(dev[0]); ++k)
    if (dev[k]) hipFree(dev[k]);
  void *host[] = {b->h_lps,  b->h_ctl_eval, b->h_ctl_win[0], b->h_ctl_win[1],
                  b->h_scal, b->h_rb,       b->h_list,       b->h_rng,
                  b->h_plp};
  for (size_t k = 0; k < sizeof(host) / sizeof(host[0]); ++k)
    if (host[k]) hipHostFree(host[k]);
  for (int k = 0; k < 2; ++k)
    if (b->ev_win[k]) hipEventDestroy(b->ev_win[k]);
  for (int k = 0; k < b->npev_created; ++k)
    for (int q = 0; q < 3; ++q) hipEventDestroy(b->pev[k][q]);
  for (int k = 0; k < 8; ++k)
    if (b->gcache[k].exec) hipGraphExecDestroy(b->gcache[k].exec);
  if (b->s) hipStreamDestroy(b->s);
  if (b->s_pi) hipStreamDestroy(b->s_pi);
  if (b->ev_ctl) hipEventDestroy(b->ev_ctl);
  if (b->ev_pi) hipEventDestroy(b->ev_pi);
  if (b->s_pp) hipStreamDestroy(b->s_pp);
  if (b->ev_pp) hipEventDestroy(b->ev_pp);
  if (b->ev_pp_go) hipEventDestroy(b->ev_pp_go);
  free(b->pp_slot);
  pthread_mutex_destroy(&b->mu);
  free(b);
}

#define CCHK(call)                                                             \
  do {                                                                         \
    hipError_t _e = (hipError_t)(call);                                        \
    if (_e != hipSuccess) {                                                    \
      dpdlp_set_error("HIP error %d (%s) creating a batch", (int)_e,           \
                      hipGetErrorString(_e));                                  \
      dpdlp_batch_free(b);                                                     \
      return NULL;                                                             \
    }                                                                          \
  } while (0)

dpdlp_batch_t *dpdlp_batch_create(const pdhg_parameters_t *params,
                                  const dpdlp_batch_opts_t *o) {
  dpdlp_batch_t *b = (dpdlp_batch_t *)calloc(1, sizeof(*b));
  if (!b) return NULL;
  pthread_mutex_init(&b->mu, NULL);
  if (params)
    b->prm = *params;
  else
    set_default_parameters(&b->prm);
  if (o)
    b->o = *o;
  else
    dpdlp_batch_default_opts(&b->o);
  if (b->o.max_lps <= 0) b->o.max_lps = 1024;
  if (b->o.spmv_cap != 1024 && b->o.spmv_cap != 2048 && b->o.spmv_cap != 4096)
    b->o.spmv_cap = 4096;
  if (b->o.spmv_cap_At != 1024 && b->o.spmv_cap_At != 2048 && b->o.spmv_cap_At != 4096)
    b->o.spmv_cap_At = b->o.spmv_cap;
  if (b->o.persist_budget <= 0) b->o.persist_budget = 250;
  if (b->prm.termination_evaluation_frequency <= 0) {
    dpdlp_set_error("termination_evaluation_frequency must be positive");
    pthread_mutex_destroy(&b->mu);
    free(b);
    return NULL;
  }
  b->freq = b->prm.termination_evaluation_frequency;
  b->W = gcd(10, b->freq);
  b->rbw = DPDLP_SC_PI_HIST + 4 * b->W;
  b->cap = b->o.max_lps;
  b->sl = (bslot_t *)calloc((size_t)b->cap, sizeof(bslot_t));
  b->done_cap = 64;
  b->done = (bdone_t *)malloc((size_t)b->done_cap * sizeof(bdone_t));
  b->pend[0] = (bpend_t *)calloc((size_t)b->cap, sizeof(bpend_t));
  b->pend[1] = (bpend_t *)calloc((size_t)b->cap, sizeof(bpend_t));
  b->pool = (barena_t *)calloc((size_t)b->cap, sizeof(barena_t));
  b->dead_cap = 64;
  b->dead = (dpdlp_resident_t **)calloc((size_t)b->dead_cap, sizeof(*b->dead));
  if (!b->sl || !b->done || !b->pend[0] || !b->pend[1] || !b->pool || !b->dead) {
    dpdlp_batch_free(b);
    return NULL;
  }
  size_t C = (size_t)b->cap;
  CCHK(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
  CCHK(hipStreamCreateWithFlags(&b->s_pi, hipStreamNonBlocking));
  CCHK(hipEventCreateWithFlags(&b->ev_ctl, hipEventDisableTiming));
  CCHK(hipEventCreateWithFlags(&b->ev_pi, hipEventDisableTiming));
  CCHK(hipMalloc((void **)&b->d_lps, C * sizeof(dpdlp_blp_t)));
  CCHK(hipHostMalloc((void **)&b->h_lps, C * sizeof(dpdlp_blp_t), 0));
  CCHK(hipMalloc((void **)&b->d_ctl_eval, C * sizeof(dpdlp_bctl_t)));
  CCHK(hipHostMalloc((void **)&b->h_ctl_eval, C * sizeof(dpdlp_bctl_t), 0));
  CCHK(hipMalloc((void **)&b->d_ctl_win, C * b->W * sizeof(dpdlp_bctl_t)));
  for (int k = 0; k < 2; ++k) {
    CCHK(hipHostMalloc((void **)&b->h_ctl_win[k], C * b->W * sizeof(dpdlp_bctl_t), 0));
    CCHK(hipEventCreateWithFlags(&b->ev_win[k], hipEventDisableTiming));
  }
  CCHK(hipMalloc((void **)&b->d_scal, C * DPDLP_SCAL_LEN * sizeof(double)));
  CCHK(hipMemset(b->d_scal, 0, C * DPDLP_SCAL_LEN * sizeof(double)));
  CCHK(hipHostMalloc((void **)&b->h_scal, C * DPDLP_SCAL_LEN * sizeof(double), 0));
  CCHK(hipHostMalloc((void **)&b->h_rb, C * b->rbw * sizeof(double), 0));
  CCHK(hipMalloc((void **)&b->d_rng, C * 2 * NLIST * sizeof(int)));
  CCHK(hipHostMalloc((void **)&b->h_rng, C * 2 * NLIST * sizeof(int), 0));
  memset(b->h_lps, 0, C * sizeof(dpdlp_blp_t));
  if (b->o.persist_max_nnz > 0) {
    CCHK(hipStreamCreateWithFlags(&b->s_pp, hipStreamNonBlocking));
    CCHK(hipEventCreateWithFlags(&b->ev_pp, hipEventDisableTiming));
    CCHK(hipEventCreateWithFlags(&b->ev_pp_go, hipEventDisableTiming));
    CCHK(hipMalloc((void **)&b->d_plp, C * sizeof(dpdlp_plp_t)));
    CCHK(hipHostMalloc((void **)&b->h_plp, C * sizeof(dpdlp_plp_t), 0));
    CCHK(hipMalloc((void **)&b->d_prm, sizeof(pdhg_parameters_t)));
    CCHK(hipMemcpy(b->d_prm, &b->prm, sizeof(pdhg_parameters_t), hipMemcpyHostToDevice));
    int dev = 0, khz = 0;
    CCHK(hipGetDevice(&dev));
    CCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    b->clk_hz = khz > 0 ? 1e3 * khz : 1e8;
    b->pp_slot = (int *)calloc(C, sizeof(int));
    if (!b->pp_slot) {
      dpdlp_batch_free(b);
      return NULL;
    }
  }
  if (b->o.profile_phases) {
    for (int k = 0; k < 32; ++k) {
      for (int q = 0; q < 3; ++q) CCHK(hipEventCreate(&b->pev[k][q]));
      b->npev_created = k + 1;
    }
  }
  b->state_bytes = C * (sizeof(dpdlp_blp_t) + sizeof(dpdlp_bctl_t) * (1 + b->W) +
                        DPDLP_SCAL_LEN * sizeof(double) + 2 * NLIST * sizeof(int));
  return b;
}

int dpdlp_batch_in_flight(const dpdlp_batch_t *b) {
  if (!b) return 0;
  pthread_mutex_lock((pthread_mutex_t *)&b->mu);
  int k = b->in_flight;
  pthread_mutex_unlock((pthread_mutex_t *)&b->mu);
  return k;
}

size_t dpdlp_batch_bytes(const dpdlp_batch_t *b) {
  if (!b) return 0;
  size_t s = b->state_bytes + (size_t)b->list_cap * sizeof(dpdlp_bblk_t) +
             (size_t)b->part_len * sizeof(double) + (size_t)b->normals_len * sizeof(double);
  for (int i = 0; i < b->nhw; ++i)
    if (active(&b->sl[i])) s += b->sl[i].c.arena_len;
  return s;
}

long long dpdlp_batch_graph_replays(const dpdlp_batch_t *b) { return b ? b->graph_replays : 0; }

void dpdlp_batch_counters(const dpdlp_batch_t *b, long long *windows, long long *syncs,
                          long long *launches) {
  if (windows) *windows = b ? b->windows : 0;
  if (syncs) *syncs = b ? b->syncs : 0;
  if (launches) *launches = b ? b->launches : 0;
}

/* the power-iteration start vector: draws of std::mt19937(1) + normal
 * distribution; a solve of m rows uses the first m (utils.cu:80-92) */
static int ensure_normals(dpdlp_batch_t *b, int m) {
  if (m <= b->normals_len) return 0;
  int len = m > 2 * b->normals_len ? m : 2 * b->normals_len;
  double *h = (double *)malloc((size_t)len * sizeof(double));
  if (!h) return -1;
  dpdlp_normal_fill(1u, len, h);
  BCHK(hipStreamSynchronize(b->s));
  if (b->d_normals) hipFree(b->d_normals);
  b->d_normals = NULL;
  b->normals_len = 0;
  if (hipMalloc((void **)&b->d_normals, (size_t)len * sizeof(double)) != hipSuccess ||
      hipMemcpy(b->d_normals, h, (size_t)len * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess) {
    free(h);
    dpdlp_set_error("device allocation failed (power-iteration start vector)");
    return -1;
  }
  free(h);
  b->normals_len = len;
  return 0;
}

static void fill_blp(dpdlp_blp_t *L, const ctx_t *c) {
  const dpdlp_state_t *st = &c->st;
  memset(L, 0, sizeof(*L));
  L->m = st->m;
  L->n = st->n;
  L->A_rp = st->A.rp;
  L->A_ci = st->A.ci;
  L->At_rp = st->At.rp;
  L->At_ci = st->At.ci;
  L->A_v = st->A.v;
  L->At_v = st->At.v;
  L->A_v32 = st->A.v32;
  L->At_v32 = st->At.v32;
  L->cap_A = st->A.cap;
  L->cap_At = st->At.cap;
  L->lv = st->lv;
  L->uv = st->uv;
  L->c = st->c;
  L->lc = st->lc;
  L->uc = st->uc;
  L->D_r = st->D_r;
  L->D_c = st->D_c;
  L->x_init = st->x_init;
  L->x_cur = st->x_cur;
  L->x_pdhg = st->x_pdhg;
  L->x_ref = st->x_ref;
  L->dual_product = st->dual_product;
  L->dual_slack = st->dual_slack;
  L->dx = st->dx;
  L->y_init = st->y_init;
  L->y_cur = st->y_cur;
  L->y_pdhg = st->y_pdhg;
  L->y_ref = st->y_ref;
  L->primal_product = st->primal_product;
  L->primal_slack = st->primal_slack;
  L->dy = st->dy;
  L->x_ref32 = st->x_ref32;
  L->y_cur32 = st->y_cur32;
  L->pi_eig = c->pi_eig;
  L->pi_next = c->pi_next;
  L->pi_dual = c->pi_dual;
  L->scal = st->scal;
}

static void start_loop(dpdlp_batch_t *b, bslot_t *q) {
  double step = q->R->nnz == 0 ? 1.0 : 0.998 / sqrt(q->sigma_sq);
  dpdlp_i_init_hstate(&q->h, &q->c, &b->prm, step);
  q->power_time = dpdlp_wall() - q->t_pi0;
  q->t_loop0 = dpdlp_wall();
  set_stage(q, S_PDHG);
  b->dirty = 1; /* the LP moves to the iterating lists */
  q->do_restart = 0;
  q->fresh = 1;
}

/* smallest pooled arena of at least `need` bytes (and at most 4x), or none */
static void take_arena(dpdlp_batch_t *b, size_t need, char **p, size_t *cap) {
  *p = NULL;
  *cap = 0;
  pthread_mutex_lock(&b->mu);
  int best = -1;
  for (int i = 0; i < b->npool; ++i)
    if (b->pool[i].cap >= need && b->pool[i].cap <= 4 * need &&
        (best < 0 || b->pool[i].cap < b->pool[best].cap))
      best = i;
  if (best >= 0) {
    *p = b->pool[best].p;
    *cap = b->pool[best].cap;
    b->pool_bytes -= *cap;
    b->pool[best] = b->pool[--b->npool];
  }
  pthread_mutex_unlock(&b->mu);
}

/* pools an arena; returns one to free instead when the pool is full */
static char *put_arena(dpdlp_batch_t *b, char *p, size_t cap) {
  char *out = NULL;
  pthread_mutex_lock(&b->mu);
  if (b->npool < b->cap) {
    b->pool[b->npool].p = p;
    b->pool[b->npool].cap = cap;
    b->npool++;
    b->pool_bytes += cap;
  } else {
    out = p;
  }
  pthread_mutex_unlock(&b->mu);
  return out;
}

/* Frees retired problems and trims the arena pool to `keep_bytes`; device
 * frees may synchronise the device, so a set-up thread calls this, not the
 * stepping thread. */
void dpdlp_i_batch_reap(dpdlp_batch_t *b, size_t keep_bytes) {
  for (;;) {
    dpdlp_resident_t *R = NULL;
    char *a = NULL;
    pthread_mutex_lock(&b->mu);
    if (b->ndead > 0) {
      R = b->dead[--b->ndead];
    } else if (b->pool_bytes > keep_bytes && b->npool > 0) {
      int big = 0;
      for (int i = 1; i < b->npool; ++i)
        if (b->pool[i].cap > b->pool[big].cap) big = i;
      a = b->pool[big].p;
      b->pool_bytes -= b->pool[big].cap;
      b->pool[big] = b->pool[--b->npool];
    }
    pthread_mutex_unlock(&b->mu);
    if (R && R->pooled) {
      /* the problem's device blob goes back to the pool */
      size_t cap = 0;
      char *blob = dpdlp_i_resident_release(R, &cap);
      char *f = blob ? put_arena(b, blob, cap) : NULL;
      if (f) hipFree(f);
    } else if (R) {
      dpdlp_resident_free(R);
    } else if (a) {
      hipFree(a);
    } else {
      break;
    }
  }
}

int dpdlp_i_batch_cap_At(const dpdlp_batch_t *b) { return b->o.spmv_cap_At; }

char *dpdlp_i_batch_get_mem(dpdlp_batch_t *b, size_t need, size_t *cap) {
  char *p = NULL;
  take_arena(b, need, &p, cap);
  if (p) return p;
  if (hipMalloc((void **)&p, need) != hipSuccess) return NULL;
  *cap = need;
  return p;
}

void dpdlp_i_batch_put_mem(dpdlp_batch_t *b, char *p, size_t cap) {
  char *f = p ? put_arena(b, p, cap) : NULL;
  if (f) hipFree(f);
}

size_t dpdlp_i_batch_pool_bytes(dpdlp_batch_t *b) {
  pthread_mutex_lock(&b->mu);
  size_t v = b->pool_bytes;
  pthread_mutex_unlock(&b->mu);
  return v;
}

int dpdlp_i_batch_reserve(dpdlp_batch_t *b) {
  pthread_mutex_lock(&b->mu);
  int i = 0;
  while (i < b->cap && stage_of(&b->sl[i]) != S_FREE) ++i;
  if (i < b->cap)
    set_stage(&b->sl[i], S_RESERVED);
  else
    b->reserve_failed = 1; /* retired slots are reclaimed at the next rebuild */
  pthread_mutex_unlock(&b->mu);
  return i < b->cap ? i : -1;
}

void dpdlp_i_batch_unreserve(dpdlp_batch_t *b, int i) {
  pthread_mutex_lock(&b->mu);
  set_stage(&b->sl[i], S_FREE);
  pthread_mutex_unlock(&b->mu);
}

/* Sets up a reserved slot (any thread, on `stream`; NULL = the batch's
 * stream): the same scaled problem a single-LP solve would build.  The slot
 * joins at the next window start. */
int dpdlp_i_batch_prepare_issue(dpdlp_batch_t *b, int i, dpdlp_resident_t *R,
                                int take_ownership, long long tag, void *stream) {
  /* any untiled layout whose row blocks fit the batch's kernels: capacities
   * at most the batch's (its kernels stage that many products in LDS) */
  if (R->cap > b->o.spmv_cap || R->At_cap > b->o.spmv_cap_At || R->A_P != 1 || R->At_P != 1 ||
      R->h_row_perm) {
    dpdlp_set_error("batch LPs need untiled A and A' in the caller's row order with row-block "
                    "capacities at most %d and %d (dpdlp_upload_caps(prob, %d, %d, 1, 1) or "
                    "dpdlp_upload_batch_layout)",
                    b->o.spmv_cap, b->o.spmv_cap_At, b->o.spmv_cap, b->o.spmv_cap_At);
    dpdlp_i_batch_unreserve(b, i);
    return -1;
  }
  bslot_t *q = &b->sl[i];
  memset(&q->tag, 0, sizeof(*q) - offsetof(bslot_t, tag));
  q->t_add = dpdlp_wall();
  q->tag = tag;
  q->R = R;
  q->own_R = take_ownership;
  hipStream_t s = stream ? (hipStream_t)stream : b->s;
  take_arena(b, dpdlp_i_ctx_footprint(R, b->o.precision == DPDLP_PREC_MIXED, 0), &q->c.arena,
             &q->c.arena_cap);
  if (dpdlp_i_ctx_setup_issue(&q->c, R, &b->prm, b->o.precision == DPDLP_PREC_MIXED, s,
                              b->d_scal + (size_t)i * DPDLP_SCAL_LEN,
                              b->h_scal + (size_t)i * DPDLP_SCAL_LEN)) {
    dpdlp_i_ctx_free(&q->c);
    memset(&q->tag, 0, sizeof(*q) - offsetof(bslot_t, tag));
    dpdlp_i_batch_unreserve(b, i);
    return -1;
  }
  return 0;
}

int dpdlp_i_batch_prepare_complete(dpdlp_batch_t *b, int i) {
  bslot_t *q = &b->sl[i];
  hipStream_t s = q->c.s;
  /* the batch's stream waits for the rest of the set-up (warm start, FP32
   * copies) before the LP's first kernels there */
  if (dpdlp_i_ctx_setup_complete(&q->c, q->R, &b->prm, &q->rescale_time) ||
      (!q->ev_ready &&
       hipEventCreateWithFlags(&q->ev_ready, hipEventDisableTiming) != hipSuccess) ||
      hipEventRecord(q->ev_ready, s) != hipSuccess) {
    if (!dpdlp_last_error()[0]) dpdlp_set_error("set-up event failed");
    dpdlp_i_ctx_free(&q->c);
    memset(&q->tag, 0, sizeof(*q) - offsetof(bslot_t, tag));
    dpdlp_i_batch_unreserve(b, i);
    return -1;
  }
  pthread_mutex_lock(&b->mu);
  set_stage(q, S_READY);
  b->in_flight++;
  pthread_mutex_unlock(&b->mu);
  return 0;
}

/* drops an issued set-up whose stream has been synchronised */
void dpdlp_i_batch_abandon(dpdlp_batch_t *b, int i) {
  dpdlp_i_ctx_free(&b->sl[i].c);
  memset(&b->sl[i].tag, 0, sizeof(bslot_t) - offsetof(bslot_t, tag));
  dpdlp_i_batch_unreserve(b, i);
}

int dpdlp_i_batch_prepare(dpdlp_batch_t *b, int i, dpdlp_resident_t *R, int take_ownership,
                          long long tag, void *stream) {
  if (dpdlp_i_batch_prepare_issue(b, i, R, take_ownership, tag, stream)) return -1;
  if (hipStreamSynchronize(b->sl[i].c.s) != hipSuccess) {
    dpdlp_set_error("device sync failed during set-up");
    dpdlp_i_ctx_free(&b->sl[i].c);
    memset(&b->sl[i].tag, 0, sizeof(bslot_t) - offsetof(bslot_t, tag));
    dpdlp_i_batch_unreserve(b, i);
    return -1;
  }
  return dpdlp_i_batch_prepare_complete(b, i);
}

/* READY slots join (main thread, at a window start): power iteration start
 * (utils.cu:80-135) or, without nonzeros, the loop directly */
static int activate_ready(dpdlp_batch_t *b) {
  /* joins are gathered (each one rebuilds the lists): at least 16 waiting,
   * 4 windows since the last join, or nothing else to run */
  int ready = 0, busy = 0;
  for (int i = 0; i < b->cap; ++i) {
    int st = stage_of(&b->sl[i]);
    ready += st == S_READY;
    busy += st == S_PI || st == S_PDHG;
  }
  if (ready == 0) return 0;
  if (busy > 0 && ready < 16 && b->windows - b->last_activation < 4) return 0;
  b->last_activation = b->windows;
  for (int i = 0; i < b->cap; ++i) {
    bslot_t *q = &b->sl[i];
    if (stage_of(q) != S_READY) continue;
    dpdlp_resident_t *R = q->R;
    q->c.s = b->s; /* it may have been set up on another stream */
    if (hipStreamWaitEvent(b->s, q->ev_ready, 0) != hipSuccess) {
      dpdlp_set_error("waiting for an LP's set-up failed");
      return -1;
    }
    fill_blp(&b->h_lps[i], &q->c);
    q->t_pi0 = dpdlp_wall();
    q->sigma_sq = 1.0;
    if (i + 1 > b->nhw) b->nhw = i + 1;
    b->dirty = 1;
    if (R->nnz > 0 && b->prm.sv_max_iter > 0) {
      if (ensure_normals(b, R->m)) return -1;
      if (hipMemcpyAsync(q->c.pi_eig, b->d_normals, (size_t)R->m * sizeof(double),
                         hipMemcpyDeviceToDevice, b->s) != hipSuccess ||
          dpdlp_dev_power_init(&q->c.st, q->c.pi_eig, q->c.pi_next, b->s)) {
        dpdlp_set_error("power-iteration start failed on the device");
        return -1;
      }
      if (b->o.persist_max_nnz > 0 && R->nnz <= b->o.persist_max_nnz &&
          R->cap <= DPDLP_PERSIST_CAP && R->At_cap <= DPDLP_PERSIST_CAP) {
        /* the whole solve runs in the persistent kernel (FP64) */
        q->persist = 1;
        q->c.use32 = 0;
        set_stage(q, S_PERSIST);
      } else {
        set_stage(q, S_PI);
      }
    } else {
      start_loop(b, q);
      q->fresh = 0; /* iteration 0 is evaluated with the others below */
    }
  }
  return 0;
}

static int rebuild_lists(dpdlp_batch_t *b);

int dpdlp_batch_add(dpdlp_batch_t *b, dpdlp_resident_t *R, int take_ownership, long long tag) {
  if (!b || !R) {
    dpdlp_set_error("NULL batch or problem");
    return -1;
  }
  int i = dpdlp_i_batch_reserve(b);
  if (i < 0 && b->zombie_items > 0) {
    /* reclaim retired slots (this is the stepping thread) */
    if (rebuild_lists(b)) return -1;
    i = dpdlp_i_batch_reserve(b);
  }
  if (i < 0) {
    dpdlp_set_error("batch full (%d LPs)", b->cap);
    return -1;
  }
  return dpdlp_i_batch_prepare(b, i, R, take_ownership, tag, NULL);
}

/* ---- block lists ---- */

static int rebuild_lists(dpdlp_batch_t *b) {
  pthread_mutex_lock(&b->mu);
  for (int i = 0; i < b->nhw; ++i)
    if (stage_of(&b->sl[i]) == S_ZOMBIE) set_stage(&b->sl[i], S_FREE);
  b->reserve_failed = 0;
  pthread_mutex_unlock(&b->mu);
  while (b->nhw > 0 && stage_of(&b->sl[b->nhw - 1]) == S_FREE) b->nhw--;
  b->zombie_items = 0;
  /* every list is split into the LPs iterating (group 0) and the LPs still in
   * power iteration (group 1), so each launch only covers the LPs it serves */
  long long cnt[2][NLIST];
  memset(cnt, 0, sizeof(cnt));
  for (int i = 0; i < b->nhw; ++i) {
    bslot_t *q = &b->sl[i];
    if (!active(q)) continue;
    const int g = stage_of(q) == S_PI;
    const dpdlp_resident_t *R = q->R;
    cnt[g][0] += R->A_nblk;
    cnt[g][1] += R->At_nblk;
    cnt[g][2] += dpdlp_ew_grid(R->n);
    cnt[g][3] += dpdlp_ew_grid(R->m);
    cnt[g][4] += dpdlp_ew_grid((long long)R->n + R->m);
  }
  long long total = 0, need_stride = 1;
  for (int g = 0; g < 2; ++g)
    for (int k = 0; k < NLIST; ++k) {
      total += cnt[g][k];
      if (cnt[g][k] > need_stride) need_stride = cnt[g][k];
    }
  b->list_items = total;
  if (need_stride > 0x7fffffffLL) {
    dpdlp_set_error("batch too large (%lld work items in one list)", need_stride);
    return -1;
  }
  /* pinned buffers and the part array are rewritten: drain the streams */
  BCHK(hipStreamSynchronize(b->s_pi));
  BCHK(hipStreamSynchronize(b->s));
  b->syncs++;
  if (total > b->list_cap) {
    long long nc = total + total / 4 + 1024;
    if (b->d_list) hipFree(b->d_list);
    if (b->h_list) hipHostFree(b->h_list);
    b->d_list = NULL;
    b->h_list = NULL;
    b->list_cap = 0;
    BCHK(hipMalloc((void **)&b->d_list, (size_t)nc * sizeof(dpdlp_bblk_t)));
    BCHK(hipHostMalloc((void **)&b->h_list, (size_t)nc * sizeof(dpdlp_bblk_t), 0));
    b->list_cap = nc;
  }
  long long plen = (long long)DPDLP_SLOTS * need_stride;
  if (plen > b->part_len) {
    long long nl = plen + plen / 4;
    if (b->d_part) hipFree(b->d_part);
    b->d_part = NULL;
    b->part_len = 0;
    BCHK(hipMalloc((void **)&b->d_part, (size_t)nl * sizeof(double)));
    b->part_len = nl;
  }
  /* kind k occupies [off[0][k], off[0][k] + cnt[0][k] + cnt[1][k]): group 0 then group 1 */
  long long off[2][NLIST];
  long long run = 0;
  for (int k = 0; k < NLIST; ++k) {
    off[0][k] = run;
    off[1][k] = run + cnt[0][k];
    run += cnt[0][k] + cnt[1][k];
  }
  long long pos[2][NLIST];
  memset(pos, 0, sizeof(pos));
  int *rng = b->h_rng;
  memset(rng, 0, (size_t)b->cap * 2 * NLIST * sizeof(int));
  dpdlp_bblk_t *L = b->h_list;
  for (int i = 0; i < b->nhw; ++i) {
    bslot_t *q = &b->sl[i];
    if (!active(q)) continue;
    const int g = stage_of(q) == S_PI;
    const dpdlp_resident_t *R = q->R;
    for (int k = 0; k < NLIST; ++k) rng[(size_t)k * 2 * b->cap + 2 * i] = (int)pos[g][k];
    for (int j = 0; j < R->A_nblk; ++j) {
      dpdlp_bblk_t *e = &L[off[g][0] + pos[g][0]++];
      e->lp = i;
      e->a = R->h_A_blk[j];
      e->b = R->h_A_blk[j + 1];
      e->pad = 0;
    }
    for (int j = 0; j < R->At_nblk; ++j) {
      dpdlp_bblk_t *e = &L[off[g][1] + pos[g][1]++];
      e->lp = i;
      e->a = R->h_At_blk[j];
      e->b = R->h_At_blk[j + 1];
      e->pad = 0;
    }
    long long cnts[3] = {R->n, R->m, (long long)R->n + R->m};
    for (int k = 0; k < 3; ++k) {
      int G = dpdlp_ew_grid(cnts[k]);
      for (int j = 0; j < G; ++j) {
        dpdlp_bblk_t *e = &L[off[g][2 + k] + pos[g][2 + k]++];
        e->lp = i;
        e->a = j;
        e->b = G;
        e->pad = 0;
      }
    }
    for (int k = 0; k < NLIST; ++k) rng[(size_t)k * 2 * b->cap + 2 * i + 1] = (int)pos[g][k];
  }
  if (total > 0)
    BCHK(hipMemcpyAsync(b->d_list, b->h_list, (size_t)total * sizeof(dpdlp_bblk_t),
                        hipMemcpyHostToDevice, b->s));
  BCHK(hipMemcpyAsync(b->d_rng, b->h_rng, (size_t)b->cap * 2 * NLIST * sizeof(int),
                      hipMemcpyHostToDevice, b->s));
  BCHK(hipMemcpyAsync(b->d_lps, b->h_lps, (size_t)b->nhw * sizeof(dpdlp_blp_t),
                      hipMemcpyHostToDevice, b->s));
  for (int g = 0; g < 2; ++g) {
    dpdlp_batch_dev_t *d = g ? &b->bd_pi : &b->bd;
    d->lps = b->d_lps;
    d->bA = b->d_list + off[g][0];
    d->bAt = b->d_list + off[g][1];
    d->eN = b->d_list + off[g][2];
    d->eM = b->d_list + off[g][3];
    d->eNM = b->d_list + off[g][4];
    d->nbA = (int)cnt[g][0];
    d->nbAt = (int)cnt[g][1];
    d->neN = (int)cnt[g][2];
    d->neM = (int)cnt[g][3];
    d->neNM = (int)cnt[g][4];
    d->rA = b->d_rng;
    d->rAt = b->d_rng + 2 * b->cap;
    d->rN = b->d_rng + 4 * b->cap;
    d->rM = b->d_rng + 6 * b->cap;
    d->rNM = b->d_rng + 8 * b->cap;
    d->part = b->d_part;
    d->stride = (int)need_stride;
    d->cap = b->o.spmv_cap;
    d->cap_At = b->o.spmv_cap_At;
  }
  b->dirty = 0;
  b->epoch++;
  for (int k = 0; k < 8; ++k)
    if (b->gcache[k].exec) {
      hipGraphExecDestroy(b->gcache[k].exec);
      b->gcache[k].exec = NULL;
    }
  return 0;
}

/* ---- evaluation points ---- */

/* marks what slot q evaluates at the top of its iteration T = h.total */
static void mark_eval(dpdlp_batch_t *b, bslot_t *q) {
  hstate_t *h = &q->h;
  const termination_criteria_t *tc = &b->prm.termination_criteria;
  q->ev_on = q->ev_rc = q->ev_inf = q->ev_final = 0;
  if (h->total >= tc->iteration_limit) {
    q->ev_on = 1;
    q->ev_final = 1;
    return;
  }
  int rc = h->is_major || h->total == 0;
  if (rc || h->total % dpdlp_print_frequency(h->total) == 0) {
    q->ev_on = 1;
    q->ev_rc = rc;
    q->ev_inf = h->is_major && h->total < 3 * b->freq;
  }
}

/* sampled phase events are complete after a synchronisation */
static void harvest_phase_events(dpdlp_batch_t *b) {
  for (int k = 0; k < b->npev_used; ++k) {
    float a = 0.f, c = 0.f;
    if (hipEventElapsedTime(&a, b->pev[k][0], b->pev[k][1]) == hipSuccess &&
        hipEventElapsedTime(&c, b->pev[k][1], b->pev[k][2]) == hipSuccess) {
      b->p_ms[0] += a;
      b->p_ms[1] += c;
      b->p_bytes[0] += b->pev_bytes[k][0];
      b->p_bytes[1] += b->pev_bytes[k][1];
      b->psamples++;
    }
  }
  b->npev_used = 0;
}

void dpdlp_batch_phase_stats(const dpdlp_batch_t *b, int *samples, double *primal_ms,
                             double *primal_bytes, double *dual_ms, double *dual_bytes) {
  if (samples) *samples = b ? b->psamples : 0;
  if (primal_ms) *primal_ms = b ? b->p_ms[0] : 0.0;
  if (primal_bytes) *primal_bytes = b ? b->p_bytes[0] : 0.0;
  if (dual_ms) *dual_ms = b ? b->p_ms[1] : 0.0;
  if (dual_bytes) *dual_bytes = b ? b->p_bytes[1] : 0.0;
}

/* Algorithmic bytes of one LP's two phase kernels in a round: matrix stream
 * (values + indices), row pointers, the gathered vector read once, and the
 * epilogue's per-row vector reads and writes (as the epilogues do them). */
static void phase_bytes(const dpdlp_resident_t *R, const dpdlp_bctl_t *c, double *p1,
                        double *p2) {
  double m = R->m, n = R->n, nnz = R->nnz;
  double sv = c->use32 ? 4 : 8, sx = c->use32 ? 4 : 8;
  double x_rw = 40 + 16 + (c->use32 ? 4 : 0) + (c->major ? 16 : 0) + (c->fpe ? 8 : 0);
  double y_rw = 32 + 8 + (c->use32 ? 4 : 0) + (c->major ? 8 : 0) + (c->fpe ? 8 : 0);
  *p1 = nnz * (sv + 4) + 4 * (n + 1) + sx * m + x_rw * n;
  *p2 = nnz * (sv + 4) + 4 * (m + 1) + sx * n + y_rw * m;
}

/* launches the residual (and infeasibility) kernels of the marked slots and
 * reads every slot's scalar slots and power-iteration history back: one
 * synchronisation */
static int flush_pending(dpdlp_batch_t *b);

/* the batch stream waits for the last window's power-iteration steps */
static int join_pi(dpdlp_batch_t *b) {
  if (!b->pi_pending) return 0;
  BCHK(hipStreamWaitEvent(b->s, b->ev_pi, 0));
  b->pi_pending = 0;
  return 0;
}

static int eval_and_read(dpdlp_batch_t *b, int any_eval, int any_inf) {
  if (join_pi(b)) return -1;
  if (b->dirty && rebuild_lists(b)) return -1;
  if (any_eval) {
    memset(b->h_ctl_eval, 0, (size_t)b->nhw * sizeof(dpdlp_bctl_t));
    for (int i = 0; i < b->nhw; ++i) {
      bslot_t *q = &b->sl[i];
      if (stage_of(q) != S_PDHG || !q->ev_on) continue;
      dpdlp_bctl_t *c = &b->h_ctl_eval[i];
      c->eval = 1;
      c->restart_dist = q->ev_rc && !q->ev_final;
      c->inf = q->ev_inf;
    }
    BCHK(hipMemcpyAsync(b->d_ctl_eval, b->h_ctl_eval, (size_t)b->nhw * sizeof(dpdlp_bctl_t),
                        hipMemcpyHostToDevice, b->s));
    b->bd.ctl = b->d_ctl_eval;
    b->bd.nlp = b->nhw;
    if (dpdlp_bdev_residual(&b->bd, b->s)) return -1;
    b->launches += 4;
    if (any_inf) {
      if (dpdlp_bdev_infeasibility(&b->bd, b->s)) return -1;
      b->launches += 7;
    }
  }
  BCHK(hipMemcpy2DAsync(b->h_rb, (size_t)b->rbw * sizeof(double), b->d_scal,
                        (size_t)DPDLP_SCAL_LEN * sizeof(double), (size_t)b->rbw * sizeof(double),
                        (size_t)b->nhw, hipMemcpyDeviceToHost, b->s));
  BCHK(hipStreamSynchronize(b->s));
  b->syncs++;
  if (b->npev_used) harvest_phase_events(b);
  return 0;
}

static int push_done(dpdlp_batch_t *b, long long tag, cupdlpx_result_t *res, double lat) {
  if (b->done_head > 0 && b->done_head == b->ndone) b->done_head = b->ndone = 0;
  if (b->ndone == b->done_cap) {
    int nc = b->done_cap * 2;
    bdone_t *d = (bdone_t *)realloc(b->done, (size_t)nc * sizeof(bdone_t));
    if (!d) return -1;
    b->done = d;
    b->done_cap = nc;
  }
  b->done[b->ndone].tag = tag;
  b->done[b->ndone].res = res;
  b->done[b->ndone].latency = lat;
  b->ndone++;
  return 0;
}

/* completes the retired LPs whose result copies were issued before the last
 * synchronisation: results become available, arenas go back to the pool */
static int flush_set(dpdlp_batch_t *b, int g) {
  double now = dpdlp_wall();
  for (int k = 0; k < b->npend[g]; ++k) {
    bpend_t *e = &b->pend[g][k];
    cupdlpx_result_t *r = e->res;
    const double *st = b->h_stage[g] + e->off;
    if (e->n > 0) {
      memcpy(r->primal_solution, st, (size_t)e->n * sizeof(double));
      memcpy(r->reduced_cost, st + e->n, (size_t)e->n * sizeof(double));
    }
    if (e->m > 0) memcpy(r->dual_solution, st + 2 * (size_t)e->n, (size_t)e->m * sizeof(double));
    if (push_done(b, e->tag, r, now - e->t_add)) return -1;
    e->res = NULL;
    char *drop = put_arena(b, e->arena, e->arena_cap);
    if (drop) hipFree(drop);
    dpdlp_resident_t *free_now = NULL;
    pthread_mutex_lock(&b->mu);
    if (e->own_R) {
      /* the list grows as needed: retirements can outpace the set-up
       * thread that frees them */
      if (b->ndead == b->dead_cap) {
        int nc = 2 * b->dead_cap;
        dpdlp_resident_t **d =
            (dpdlp_resident_t **)realloc(b->dead, (size_t)nc * sizeof(*b->dead));
        if (d) {
          b->dead = d;
          b->dead_cap = nc;
        }
      }
      if (b->ndead < b->dead_cap)
        b->dead[b->ndead++] = e->R;
      else
        free_now = e->R;
    }
    b->in_flight--;
    pthread_mutex_unlock(&b->mu);
    if (free_now) dpdlp_resident_free(free_now);
  }
  b->npend[g] = 0;
  b->stage_off[g] = 0;
  return 0;
}

/* completes every retired LP (only after a synchronisation) */
static int flush_pending(dpdlp_batch_t *b) {
  if (flush_set(b, 0) || flush_set(b, 1)) return -1;
  return 0;
}

/* an LP terminated: polish (as the single-LP path does), issue its result
 * copies into pinned staging, free its slot; the result completes at the
 * next synchronisation (flush_pending) */
static int finish(dpdlp_batch_t *b, int i) {
  bslot_t *q = &b->sl[i];
  hstate_t *h = &q->h;
  double polish_time = 0.0;
  int polish_iters = 0;
  if (b->prm.feasibility_polishing && h->reason != TERMINATION_REASON_DUAL_INFEASIBLE &&
      h->reason != TERMINATION_REASON_PRIMAL_INFEASIBLE) {
    if (dpdlp_i_feasibility_polish(&q->c, h, &b->prm, &polish_time, &polish_iters)) {
      dpdlp_set_error("feasibility polishing failed on the device");
      return -1;
    }
  }
  int n = q->c.st.n, m = q->c.st.m;
  size_t need = 2 * (size_t)n + (size_t)m;
  const int g = b->scur;
  if (b->stage_off[g] + need > b->stage_cap[g]) {
    /* staging full: complete what is in flight first */
    BCHK(hipStreamSynchronize(b->s));
    b->syncs++;
    if (flush_pending(b)) return -1;
    if (need > b->stage_cap[g]) {
      size_t nc = need > 2 * b->stage_cap[g] ? need : 2 * b->stage_cap[g];
      if (nc < (1u << 20)) nc = 1u << 20;
      if (b->h_stage[g]) hipHostFree(b->h_stage[g]);
      b->h_stage[g] = NULL;
      b->stage_cap[g] = 0;
      BCHK(hipHostMalloc((void **)&b->h_stage[g], nc * sizeof(double), 0));
      b->stage_cap[g] = nc;
    }
  }
  cupdlpx_result_t *res = (cupdlpx_result_t *)calloc(1, sizeof(*res));
  if (!res) return -1;
  res->primal_solution = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
  res->dual_solution = (double *)malloc((size_t)(m ? m : 1) * sizeof(double));
  res->reduced_cost = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
  if (!res->primal_solution || !res->dual_solution || !res->reduced_cost) {
    cupdlpx_result_free(res);
    return -1;
  }
  dpdlp_i_result_scalars(res, &q->c, h, q->rescale_time, polish_time, polish_iters);
  double *stg = b->h_stage[g] + b->stage_off[g];
  if (dpdlp_i_result_launch(&q->c, stg, stg + 2 * (size_t)n, stg + n)) {
    cupdlpx_result_free(res);
    return -1;
  }
  bpend_t *e = &b->pend[g][b->npend[g]++];
  e->tag = q->tag;
  e->res = res;
  e->off = b->stage_off[g];
  e->n = n;
  e->m = m;
  e->t_add = q->t_add;
  e->arena = q->c.arena;
  e->arena_cap = q->c.arena_cap;
  e->R = q->R;
  e->own_R = q->own_R;
  b->stage_off[g] += need;
  /* the arena moves to the pending entry; the slot's other allocations (if
   * polishing made any) are freed now; no stream synchronisation here */
  if (!q->persist) {
    const dpdlp_resident_t *R = q->R;
    b->zombie_items += R->A_nblk + R->At_nblk + dpdlp_ew_grid(R->n) + dpdlp_ew_grid(R->m) +
                       dpdlp_ew_grid((long long)R->n + R->m);
  } else {
    /* no list items, but the slot is reclaimed by the same rebuild */
    b->zombie_items += 1;
  }
  q->c.arena = NULL;
  q->c.s = NULL;
  dpdlp_i_ctx_free(&q->c);
  memset(&q->tag, 0, sizeof(*q) - offsetof(bslot_t, tag));
  memset(&b->h_lps[i], 0, sizeof(dpdlp_blp_t));
  pthread_mutex_lock(&b->mu);
  set_stage(q, S_ZOMBIE);
  pthread_mutex_unlock(&b->mu);
  /* the lists drop the slot at the next rebuild: when a quarter of their
   * items are retired ones, when slots run out, or when LPs join */
  if (b->zombie_items * 4 > b->list_items) b->dirty = 1;
  return 0;
}

/* the reference loop's decisions at an evaluation point (solver.cu:60-84) */
static int process_eval(dpdlp_batch_t *b, int i) {
  bslot_t *q = &b->sl[i];
  hstate_t *h = &q->h;
  const double *S = b->h_rb + (size_t)i * b->rbw;
  const pdhg_parameters_t *p = &b->prm;
  q->ev_on = 0;
  ctl_pending_fpe(h, S, 0);
  ctl_residual(h, &q->c.k, S, p->optimality_norm);
  if (q->ev_final) {
    h->reason = TERMINATION_REASON_ITERATION_LIMIT;
    return finish(b, i);
  }
  if (q->ev_inf) ctl_infeasibility(h, &q->c.k, S);
  h->t_cum = dpdlp_wall() - q->t_loop0;
  double kkt = fmax(h->rel_p, fmax(h->rel_d, h->rel_gap));
  if (q->c.use32 &&
      !ctl_mixed_keep_fp32(h, kkt, b->o.mixed_switch_tol,
                           fmax(p->termination_criteria.eps_optimal_relative,
                                p->termination_criteria.eps_feasible_relative),
                           0))
    q->c.use32 = 0;
  ctl_check_termination(h, &p->termination_criteria);
  if (h->reason != TERMINATION_REASON_UNSPECIFIED) return finish(b, i);
  if (q->ev_rc) {
    q->do_restart = ctl_should_restart(h, &p->restart_params, b->freq);
    if (q->do_restart) {
      ctl_pid_update(h, &p->restart_params, dpdlp_blue(S + DPDLP_SL_RESD + 5),
                     dpdlp_blue(S + DPDLP_SL_RESP + 4));
      h->inner = 0;
      h->fpe_last = INFINITY;
    }
  }
  return 0;
}

/* power-iteration history of the last window (utils.cu:66-171) */
static void process_pi(dpdlp_batch_t *b, int i) {
  bslot_t *q = &b->sl[i];
  const double *S = b->h_rb + (size_t)i * b->rbw;
  int done = 0;
  for (int k = 0; k < q->pi_issued; ++k) {
    const double *hh = S + DPDLP_SC_PI_HIST + 4 * k;
    q->sigma_sq = hh[0];
    q->pi_iters++;
    if (dpdlp_blue(hh + 1) < b->prm.sv_tol) {
      done = 1;
      break;
    }
  }
  q->pi_issued = 0;
  if (done || q->pi_iters >= b->prm.sv_max_iter) start_loop(b, q);
}

/* ---- one window ---- */

/* the window's launches, round by round; ctl (host copy) non-NULL: time the
 * two phase kernels of one round with events */
static int issue_rounds(dpdlp_batch_t *b, int nl, int any_restart, const int *any_phase,
                        const int *any_fpe, const int *any_pi, const dpdlp_bctl_t *ctl,
                        const dpdlp_bctl_t *solo, hipStream_t spi) {
  const int W = b->W;
  int pi_started = 0;
  for (int r = 0; r < W; ++r) {
    b->bd.ctl = b->d_ctl_win + (size_t)r * nl;
    b->bd.nlp = nl;
    if (solo) {
      /* few LPs iterate: their own single-LP launches (kernel arguments, no
       * work lists), the same arithmetic as the batched kernels */
      for (int i = 0; i < nl; ++i) {
        const dpdlp_bctl_t *c = &solo[(size_t)r * nl + i];
        if (!c->phase) continue;
        const dpdlp_state_t *st = &b->sl[i].c.st;
        if (c->restart && dpdlp_dev_restart_copy(st, c->restart, b->s)) return -1;
        if (dpdlp_dev_primal_phase(st, c->tau, c->w, c->rho, c->major, c->fpe, c->use32, b->s) ||
            dpdlp_dev_dual_phase(st, c->sigma, c->w, c->rho, c->major, c->fpe, c->use32, b->s))
          return -1;
        if (c->fpe && dpdlp_dev_fpe_finish(st, c->fpe_to & 1, (c->fpe_to >> 1) & 1, b->s))
          return -1;
        b->launches += 2 + (c->restart ? 1 : 0) + (c->fpe ? 4 : 0);
      }
      if (any_pi[r]) {
        if (spi != b->s && !pi_started) BCHK(hipStreamWaitEvent(spi, b->ev_ctl, 0));
        pi_started = 1;
        b->bd_pi.ctl = b->bd.ctl;
        b->bd_pi.nlp = nl;
        if (dpdlp_bdev_power_step(&b->bd_pi, spi)) return -1;
        b->launches += 5;
      }
      continue;
    }
    if (r == 0 && any_restart) {
      if (dpdlp_bdev_restart_copy(&b->bd, b->s)) return -1;
      b->launches += 1;
    }
    if (any_phase[r]) {
      int timed = ctl && b->npev_used < b->npev_created && r == (W > 1 ? 1 : 0);
      if (timed) {
        hipEvent_t *ev = b->pev[b->npev_used];
        double s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < nl; ++i) {
          const dpdlp_bctl_t *c = &ctl[(size_t)r * nl + i];
          if (!c->phase) continue;
          double a1, a2;
          phase_bytes(b->sl[i].R, c, &a1, &a2);
          s1 += a1;
          s2 += a2;
        }
        b->pev_bytes[b->npev_used][0] = s1;
        b->pev_bytes[b->npev_used][1] = s2;
        b->npev_used++;
        BCHK(hipEventRecord(ev[0], b->s));
        if (dpdlp_bdev_primal(&b->bd, b->s)) return -1;
        BCHK(hipEventRecord(ev[1], b->s));
        if (dpdlp_bdev_dual(&b->bd, b->s)) return -1;
        BCHK(hipEventRecord(ev[2], b->s));
      } else {
        if (dpdlp_bdev_phases(&b->bd, b->s)) return -1;
      }
      b->launches += 2;
    }
    if (any_fpe[r]) {
      if (dpdlp_bdev_fpe(&b->bd, b->s)) return -1;
      b->launches += 5;
    }
    if (any_pi[r]) {
      if (spi != b->s && !pi_started) BCHK(hipStreamWaitEvent(spi, b->ev_ctl, 0));
      pi_started = 1;
      b->bd_pi.ctl = b->bd.ctl;
      b->bd_pi.nlp = nl;
      if (dpdlp_bdev_power_step(&b->bd_pi, spi)) return -1;
      b->launches += 5;
    }
  }
  if (pi_started && spi != b->s) {
    BCHK(hipEventRecord(b->ev_pi, spi));
    b->pi_pending = 1;
  }
  return 0;
}

static int launch_window(dpdlp_batch_t *b) {
  const int W = b->W, nl = b->nhw;
  const pdhg_parameters_t *p = &b->prm;
  const int limit = p->termination_criteria.iteration_limit;
  int k = b->win_buf;
  BCHK(hipEventSynchronize(b->ev_win[k]));
  dpdlp_bctl_t *ctl = b->h_ctl_win[k];
  memset(ctl, 0, (size_t)W * nl * sizeof(dpdlp_bctl_t));
  int any_phase[10] = {0}, any_fpe[10] = {0}, any_pi[10] = {0}, any_restart = 0;
  for (int i = 0; i < nl; ++i) {
    bslot_t *q = &b->sl[i];
    int st = stage_of(q);
    if (st == S_PDHG) {
      hstate_t *h = &q->h;
      for (int r = 0; r < W && h->total < limit; ++r) {
        dpdlp_bctl_t *c = &ctl[(size_t)r * nl + i];
        int is_major = ((h->total + 1) % b->freq) == 0;
        int dr = (r == 0) && q->do_restart;
        int major_variant =
            is_major || ((h->total + 2) % dpdlp_print_frequency(h->total + 2)) == 0;
        int fpe = is_major || dr;
        c->tau = h->step / h->omega;
        c->sigma = h->step * h->omega;
        c->w = (double)(h->inner + 1) / (h->inner + 2);
        c->rho = p->reflection_coefficient;
        c->phase = 1;
        c->major = major_variant;
        c->fpe = fpe;
        c->fpe_to = (dr ? 1 : 0) | (is_major ? 2 : 0);
        c->restart = dr ? 3 : 0;
        c->use32 = q->c.use32;
        if (fpe) {
          if (dr) h->pend_rst = 1;
          if (is_major) h->pend_maj = 1;
        }
        h->is_major = is_major;
        h->inner++;
        h->total++;
        any_phase[r] = 1;
        any_fpe[r] |= fpe;
        any_restart |= dr;
      }
      q->do_restart = 0;
    } else if (st == S_PI) {
      int left = p->sv_max_iter - q->pi_iters;
      int kk = left < W ? left : W;
      for (int r = 0; r < kk; ++r) {
        dpdlp_bctl_t *c = &ctl[(size_t)r * nl + i];
        c->pi = 1;
        c->pi_slot = r;
        any_pi[r] = 1;
      }
      q->pi_issued = kk;
    }
  }
  if (join_pi(b)) return -1;
  BCHK(hipMemcpyAsync(b->d_ctl_win, ctl, (size_t)W * nl * sizeof(dpdlp_bctl_t),
                      hipMemcpyHostToDevice, b->s));
  BCHK(hipEventRecord(b->ev_win[k], b->s));
  BCHK(hipEventRecord(b->ev_ctl, b->s));
  b->win_buf ^= 1;
  int n_iter = 0;
  for (int i = 0; i < nl; ++i) n_iter += stage_of(&b->sl[i]) == S_PDHG;
  /* few iterating LPs: their single-LP launches (shorter dependency chain per
   * workgroup, no work lists) */
  const dpdlp_bctl_t *solo = n_iter <= b->o.solo_lps ? ctl : NULL;
  /* phase timing samples every 4th batched window, power-iteration steps
   * running beside it included (the launch population a counter profile of
   * the same kernels averages over) */
  int timed_window = !solo && b->npev_used < b->npev_created && (b->windows & 3) == 0;
  unsigned long long sig = (unsigned long long)any_restart;
  for (int r = 0; r < W; ++r)
    sig |= (unsigned long long)(any_phase[r] | (any_fpe[r] << 1) | (any_pi[r] << 2)) << (1 + 3 * r);
  if (b->o.use_graph && !timed_window && !solo) {
    int slot = -1, lru = 0;
    for (int g = 0; g < 8; ++g) {
      if (b->gcache[g].exec && b->gcache[g].epoch == b->epoch && b->gcache[g].nl == nl &&
          b->gcache[g].sig == sig) {
        slot = g;
        break;
      }
      if (!b->gcache[g].exec || b->gcache[g].used < b->gcache[lru].used) lru = g;
    }
    if (slot < 0) {
      /* capture this window's launches once; later windows with the same
       * shape replay it (their per-LP parameters live in d_ctl_win) */
      hipGraph_t graph = NULL;
      BCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
      int e = issue_rounds(b, nl, any_restart, any_phase, any_fpe, any_pi, NULL, NULL, b->s);
      hipError_t ce = hipStreamEndCapture(b->s, &graph);
      if (e || ce != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        dpdlp_set_error("window graph capture failed");
        return -1;
      }
      if (b->gcache[lru].exec) hipGraphExecDestroy(b->gcache[lru].exec);
      b->gcache[lru].exec = NULL;
      hipError_t ie = hipGraphInstantiate(&b->gcache[lru].exec, graph, NULL, NULL, 0);
      hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        b->gcache[lru].exec = NULL;
        dpdlp_set_error("window graph instantiation failed: %s", hipGetErrorString(ie));
        return -1;
      }
      b->gcache[lru].epoch = b->epoch;
      b->gcache[lru].nl = nl;
      b->gcache[lru].sig = sig;
      slot = lru;
      b->graph_captures++;
    }
    b->gcache[slot].used = b->windows;
    BCHK(hipGraphLaunch(b->gcache[slot].exec, b->s));
    b->graph_replays++;
  } else {
    if (issue_rounds(b, nl, any_restart, any_phase, any_fpe, any_pi,
                     timed_window ? ctl : NULL, solo, b->s_pi))
      return -1;
  }
  b->windows++;
  return 0;
}

/* ---- persistent group ---- */

/* Completes the persistent launch in flight (block: wait for it).  LPs that
 * terminated retire (result copies queued on the batch stream, after the
 * launch); the others keep their descriptors for the next launch. */
static int pp_complete(dpdlp_batch_t *b, int block) {
  if (!b->pp_inflight) return 0;
  if (block) {
    BCHK(hipEventSynchronize(b->ev_pp));
  } else {
    hipError_t e = hipEventQuery(b->ev_pp);
    if (e == hipErrorNotReady) return 0;
    BCHK(e);
  }
  b->pp_inflight = 0;
  BCHK(hipStreamWaitEvent(b->s, b->ev_pp, 0));
  int keep = 0;
  for (int k = 0; k < b->pp_n; ++k) {
    const int i = b->pp_slot[k];
    bslot_t *q = &b->sl[i];
    if (b->h_plp[k].stage == DPDLP_PS_DONE) {
      q->h = b->h_plp[k].h;
      q->sigma_sq = b->h_plp[k].sigma_sq;
      q->pi_iters = b->h_plp[k].pi_iters;
      if (finish(b, i)) return -1;
    } else {
      if (keep != k) {
        b->h_plp[keep] = b->h_plp[k];
        b->pp_slot[keep] = i;
      }
      keep++;
    }
  }
  b->pp_n = keep;
  return 0;
}

/* Adds the group's new LPs and launches the next round of the group on its
 * own stream (after the set-up and power-iteration starts queued on the
 * batch stream). */
static int pp_launch(dpdlp_batch_t *b) {
  if (b->pp_inflight) return 0;
  for (int i = 0; i < b->nhw; ++i) {
    bslot_t *q = &b->sl[i];
    if (stage_of(q) != S_PERSIST || q->persist != 1) continue;
    dpdlp_plp_t *D = &b->h_plp[b->pp_n];
    memset(D, 0, sizeof(*D));
    D->L = b->h_lps[i];
    D->A_blk = q->c.st.A.blk;
    D->At_blk = q->c.st.At.blk;
    D->A_nblk = q->c.st.A.nblk;
    D->At_nblk = q->c.st.At.nblk;
    D->k = q->c.k;
    D->sigma_sq = 1.0;
    D->stage = DPDLP_PS_PI;
    b->pp_slot[b->pp_n++] = i;
    q->persist = 2;
  }
  if (b->pp_n == 0) return 0;
  const size_t bytes = (size_t)b->pp_n * sizeof(dpdlp_plp_t);
  BCHK(hipEventRecord(b->ev_pp_go, b->s));
  BCHK(hipStreamWaitEvent(b->s_pp, b->ev_pp_go, 0));
  BCHK(hipMemcpyAsync(b->d_plp, b->h_plp, bytes, hipMemcpyHostToDevice, b->s_pp));
  /* at most persist_budget iterations (power-iteration steps and PDHG
   * iterations) per LP: each launch ends within a bounded time, and the host
   * retires finished LPs and admits new ones between launches */
  if (dpdlp_pdev_run(b->d_plp, b->pp_n, b->d_prm, b->clk_hz, b->o.persist_budget, b->s_pp)) {
    dpdlp_set_error("persistent kernel launch failed");
    return -1;
  }
  BCHK(hipMemcpyAsync(b->h_plp, b->d_plp, bytes, hipMemcpyDeviceToHost, b->s_pp));
  BCHK(hipEventRecord(b->ev_pp, b->s_pp));
  b->pp_inflight = 1;
  b->pp_launches++;
  b->launches++;
  return 0;
}

static int n_active(const dpdlp_batch_t *b) {
  int k = 0;
  for (int i = 0; i < b->nhw; ++i) k += active(&b->sl[i]);
  return k;
}

int dpdlp_batch_step(dpdlp_batch_t *b) {
  if (!b) return -1;
  if (activate_ready(b)) return -1;
  pthread_mutex_lock(&b->mu);
  if (b->reserve_failed && b->zombie_items > 0) b->dirty = 1;
  pthread_mutex_unlock(&b->mu);
  if (b->o.persist_max_nnz > 0) {
    /* the persistent group runs beside the windows; with no window to run,
     * wait for its launch instead of polling */
    if (pp_complete(b, n_active(b) == 0) || pp_launch(b)) return -1;
  }
  if (n_active(b) == 0) {
    if (b->npend[0] || b->npend[1]) {
      BCHK(hipStreamSynchronize(b->s));
      b->syncs++;
      if (flush_pending(b)) return -1;
    }
    if (b->dirty && b->zombie_items > 0 && rebuild_lists(b)) return -1;
    return dpdlp_batch_in_flight(b);
  }
  /* evaluation points and power-iteration reads at this window start */
  int any_eval = 0, any_inf = 0, any_read = 0;
  for (int i = 0; i < b->nhw; ++i) {
    bslot_t *q = &b->sl[i];
    if (stage_of(q) == S_PDHG) {
      mark_eval(b, q);
      any_eval |= q->ev_on;
      any_inf |= q->ev_inf;
    } else if (stage_of(q) == S_PI && q->pi_issued > 0) {
      any_read = 1;
    }
  }
  if (any_eval || any_read) {
    if (eval_and_read(b, any_eval, any_inf)) return -1;
    /* every result copy issued so far is complete: the set that collected the
     * last window start's retirements is completed after this window is
     * launched; this window start's retirements go to the other set */
    b->scur ^= 1;
    if (b->npend[b->scur] && flush_set(b, b->scur)) return -1;
    int any_fresh = 0;
    for (int i = 0; i < b->nhw; ++i) {
      bslot_t *q = &b->sl[i];
      if (stage_of(q) == S_PI && q->pi_issued > 0) {
        process_pi(b, i);
        any_fresh |= q->fresh;
      }
    }
    for (int i = 0; i < b->nhw; ++i) {
      bslot_t *q = &b->sl[i];
      if (stage_of(q) == S_PDHG && q->ev_on && !q->fresh)
        if (process_eval(b, i)) return -1;
    }
    if (any_fresh) {
      /* LPs whose step size was just fixed evaluate their iteration 0 now */
      any_inf = 0;
      for (int i = 0; i < b->nhw; ++i) {
        bslot_t *q = &b->sl[i];
        if (stage_of(q) == S_PDHG) {
          if (q->fresh)
            mark_eval(b, q);
          else
            q->ev_on = 0;
          any_inf |= q->ev_on && q->ev_inf;
        }
      }
      if (eval_and_read(b, 1, any_inf)) return -1;
      for (int i = 0; i < b->nhw; ++i) {
        bslot_t *q = &b->sl[i];
        if (stage_of(q) == S_PDHG && q->fresh) {
          q->fresh = 0;
          if (q->ev_on && process_eval(b, i)) return -1;
        }
      }
    }
  }
  if (n_active(b) > 0) {
    if (b->dirty && rebuild_lists(b)) return -1;
    if (launch_window(b)) return -1;
    /* host copies of completed results overlap the window on the device */
    if (b->npend[b->scur ^ 1] && flush_set(b, b->scur ^ 1)) return -1;
  } else if (b->npend[0] || b->npend[1]) {
    BCHK(hipStreamSynchronize(b->s));
    b->syncs++;
    if (flush_pending(b)) return -1;
  }
  return dpdlp_batch_in_flight(b);
}

int dpdlp_batch_take(dpdlp_batch_t *b, long long *tag, cupdlpx_result_t **res,
                     double *latency_sec) {
  if (!b || b->done_head >= b->ndone) return 0;
  bdone_t *d = &b->done[b->done_head++];
  if (tag) *tag = d->tag;
  if (res)
    *res = d->res;
  else
    cupdlpx_result_free(d->res);
  if (latency_sec) *latency_sec = d->latency;
  return 1;
}

typedef struct {
  dpdlp_batch_t *b;
  dpdlp_resident_t *const *res;
  int count;
  hipStream_t s;
  int done, err;
} bloader_t;

/* sets the LPs up on its own stream while the batch already steps: set-ups
 * are issued in groups and completed after one synchronisation per group */
enum { LOAD_GROUP = 32 };

static void *solve_loader(void *arg) {
  bloader_t *L = (bloader_t *)arg;
  int err = 0;
  for (int i0 = 0; i0 < L->count && !err; i0 += LOAD_GROUP) {
    int slots[LOAD_GROUP], ns = 0;
    for (int i = i0; i < L->count && i < i0 + LOAD_GROUP; ++i) {
      int slot = dpdlp_i_batch_reserve(L->b);
      if (slot < 0 || dpdlp_i_batch_prepare_issue(L->b, slot, L->res[i], 0, i, L->s)) {
        err = 1;
        break;
      }
      slots[ns++] = slot;
    }
    if (hipStreamSynchronize((hipStream_t)L->s) != hipSuccess) err = 1;
    for (int k = 0; k < ns; ++k) {
      if (err)
        dpdlp_i_batch_abandon(L->b, slots[k]);
      else if (dpdlp_i_batch_prepare_complete(L->b, slots[k]))
        err = 1;
    }
  }
  __atomic_store_n(&L->err, err, __ATOMIC_RELEASE);
  __atomic_store_n(&L->done, 1, __ATOMIC_RELEASE);
  return NULL;
}

int dpdlp_solve_batch(dpdlp_resident_t *const *res, int count, const pdhg_parameters_t *params,
                      dpdlp_batch_opts_t *o, cupdlpx_result_t **results) {
  if (count < 0 || (count > 0 && (!res || !results))) {
    dpdlp_set_error("invalid batch arguments");
    return -1;
  }
  dpdlp_batch_opts_t local;
  if (o)
    local = *o;
  else
    dpdlp_batch_default_opts(&local);
  if (local.max_lps < count) local.max_lps = count;
  const int cap_At = (local.spmv_cap_At == 1024 || local.spmv_cap_At == 2048 ||
                      local.spmv_cap_At == 4096)
                         ? local.spmv_cap_At
                         : local.spmv_cap;
  for (int i = 0; i < count; ++i) results[i] = NULL;
  for (int i = 0; i < count; ++i)
    if (!res[i] || res[i]->cap > local.spmv_cap || res[i]->At_cap > cap_At || res[i]->A_P != 1 ||
        res[i]->At_P != 1 || res[i]->h_row_perm) {
      dpdlp_set_error("batch LPs need untiled A and A' in the caller's row order with row-block "
                      "capacities at most %d and %d (dpdlp_upload_caps(prob, %d, %d, 1, 1) or "
                      "dpdlp_upload_batch_layout)",
                      local.spmv_cap, cap_At, local.spmv_cap, cap_At);
      return -1;
    }
  dpdlp_batch_t *b = dpdlp_batch_create(params, &local);
  if (!b) return -1;
  bloader_t L;
  memset(&L, 0, sizeof(L));
  L.b = b;
  L.res = res;
  L.count = count;
  pthread_t th;
  int started = 0, got = 0, failed = 0;
  if (hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking) != hipSuccess) {
    dpdlp_set_error("stream creation failed");
    failed = 1;
  } else if (pthread_create(&th, NULL, solve_loader, &L) != 0) {
    dpdlp_set_error("thread creation failed");
    failed = 1;
  } else {
    started = 1;
  }
  while (!failed) {
    int ld = __atomic_load_n(&L.done, __ATOMIC_ACQUIRE);
    int k = dpdlp_batch_in_flight(b);
    if (k > 0) {
      if (dpdlp_batch_step(b) < 0) failed = 1;
    } else if (ld) {
      break;
    } else {
      struct timespec ts = {0, 50000};
      nanosleep(&ts, NULL);
    }
  }
  if (started) pthread_join(th, NULL);
  if (L.err) failed = 1;
  long long tag;
  cupdlpx_result_t *r;
  while (dpdlp_batch_take(b, &tag, &r, NULL)) {
    if (!failed && tag >= 0 && tag < count && !results[tag]) {
      results[tag] = r;
      got++;
    } else {
      cupdlpx_result_free(r);
    }
  }
  if (o) {
    dpdlp_batch_counters(b, &o->windows, &o->syncs, &o->launches);
    o->graph_replays = b->graph_replays;
    dpdlp_batch_phase_stats(b, &o->phase_samples, &o->primal_ms, &o->primal_bytes, &o->dual_ms,
                            &o->dual_bytes);
  }
  dpdlp_batch_free(b);
  if (L.s) hipStreamDestroy(L.s);
  if (failed) {
    for (int i = 0; i < count; ++i) {
      cupdlpx_result_free(results[i]);
      results[i] = NULL;
    }
    return -1;
  }
  return got;
}
