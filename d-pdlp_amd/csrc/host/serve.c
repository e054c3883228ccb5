/*
 * Serving: a request stream of LPs solved with continuous batching.
 *
 * Requests arrive as a Poisson process.  A loader thread admits each request
 * once the device bytes of the LPs in flight plus its own stay within the
 * HBM budget and a batch slot is free, uploads it, builds its scaled problem
 * on its own HIP stream and hands the slot to the batch; the main thread
 * advances the batch one window at a time, retires finished LPs (result
 * copied out, slot and memory freed) and so makes room for the next
 * admissions.  Requests of at least large_lane_nnz nonzeros queue for the
 * large-LP lane instead, which takes their memory only when it starts them,
 * so a backlog of large requests never holds back the small ones.  Every
 * request's result is the one a solve of that LP alone would give (the batch
 * guarantees it bit for bit).
 *
 * The comparison mode runs the same stream through the plain loop: one
 * request at a time, in arrival order, each uploaded and solved by the
 * single-LP path in FP64.
 */
#define _POSIX_C_SOURCE 200809L
#include "cupdlpx.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"
#include "solver_internal.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

void dpdlp_serve_default_spec(dpdlp_serve_spec_t *s) {
  memset(s, 0, sizeof(*s));
  s->rate = 50.0;
  s->duration_sec = 60.0;
  s->drain_sec = 60.0;
  s->seed = 4000;
  s->pool_small = 64;
  s->pool_large = 2;
  s->large_frac = 0.05;
  s->large_m = 1000000;
  s->large_nnz = 10000000;
  s->m_lo = 500;
  s->m_hi = 200000;
  s->member_seed = 3000;
  s->hbm_budget_gb = 200.0;
  s->max_lps = 512;
  s->precision = DPDLP_PREC_MIXED;
  s->mode = 0;
  s->spmv_cap = 4096;
  s->large_lane_nnz = 4000000;
  s->persist_max_nnz = 0;
}

/* SplitMix64 stream for arrival gaps and request draws */
static uint64_t sm_next(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double sm_u01(uint64_t *x) { return (double)(sm_next(x) >> 11) * 0x1p-53; }

/* device bytes of one request in flight: resident problem + solver state */
static double est_bytes(const lp_problem_t *p, int cap, int mixed) {
  double m = p->num_constraints, n = p->num_variables, nnz = p->constraint_matrix_num_nonzeros;
  double blocks = nnz / cap * 2.0 + (m + n) / 256.0 + 2.0;
  double stride = blocks > 2048 ? blocks : 2048;
  double res = nnz * 24 + (m + n + 2) * 4 + blocks * 4 + (3 * n + 2 * m) * 8;
  double st = nnz * 16 + (13 * n + 12 * m) * 8 + 44.0 * stride * 8 + 64 * 256;
  if (mixed) st += nnz * 8 + (n + m) * 4;
  return res + st;
}

static void sleep_until(double t) {
  for (;;) {
    double now = dpdlp_wall();
    if (now >= t) return;
    double d = t - now;
    if (d > 0.002) d = 0.002;
    struct timespec ts = {0, (long)(d * 1e9)};
    nanosleep(&ts, NULL);
  }
}

typedef struct {
  const dpdlp_serve_spec_t *spec;
  dpdlp_batch_t *b;
  lp_problem_t **pool;
  int npool;
  const int *req_pool;
  const double *arrival; /* absolute wall times */
  const double *bytes;   /* per request */
  int nreq;
  double t_cap;
  hipStream_t stream;
  pthread_mutex_t mu;
  pthread_cond_t cv;
  double in_use, budget;
  int admitted_live;
  int stop;
  int loader_done;
  int loader_error;
  double peak_bytes;
  int max_in_flight;
  double *t_admit;
  /* large-LP lane: requests of at least large_lane_nnz nonzeros are solved
   * one at a time by the single-LP path (XCD column panels, tuned block
   * size) on its own thread and stream, beside the batch */
  int *lq;
  int lq_head, lq_tail, lane_on, lane_busy;
  const pdhg_parameters_t *prm;
  double *t_done;
  cupdlpx_result_t **res;
  int completed;
} srv_t;

/* Uploads of one thread: device blobs from the batch's pool, one pinned
 * staging buffer (grown, never shrunk), one stream for the copy. */
typedef struct {
  dpdlp_batch_t *b;
  char *stage;
  size_t stage_cap;
  hipStream_t s;
} uctx_t;

static char *u_get_blob(void *ctx, size_t need, size_t *cap) {
  return dpdlp_i_batch_get_mem(((uctx_t *)ctx)->b, need, cap);
}

static void u_put_blob(void *ctx, char *p, size_t cap) {
  dpdlp_i_batch_put_mem(((uctx_t *)ctx)->b, p, cap);
}

static char *u_get_stage(void *ctx, size_t need) {
  uctx_t *u = (uctx_t *)ctx;
  if (need > u->stage_cap) {
    size_t nc = need > 2 * u->stage_cap ? need : 2 * u->stage_cap;
    if (nc < ((size_t)64 << 20)) nc = (size_t)64 << 20;
    char *p = NULL;
    if (hipHostMalloc((void **)&p, nc, 0) != hipSuccess) return NULL;
    if (u->stage) hipHostFree(u->stage);
    u->stage = p;
    u->stage_cap = nc;
  }
  return u->stage;
}

static int uctx_init(uctx_t *u, dpdlp_uploader_t *up, dpdlp_batch_t *b) {
  memset(u, 0, sizeof(*u));
  u->b = b;
  if (hipStreamCreateWithFlags(&u->s, hipStreamNonBlocking) != hipSuccess) return -1;
  up->get_blob = u_get_blob;
  up->put_blob = u_put_blob;
  up->get_stage = u_get_stage;
  up->ctx = u;
  up->stream = u->s;
  return 0;
}

static void uctx_free(uctx_t *u) {
  if (u->s) hipStreamSynchronize(u->s);
  if (u->stage) hipHostFree(u->stage);
  if (u->s) hipStreamDestroy(u->s);
  memset(u, 0, sizeof(*u));
}

static void *lane_main(void *arg) {
  srv_t *S = (srv_t *)arg;
  const dpdlp_serve_spec_t *sp = S->spec;
  uctx_t uc;
  dpdlp_uploader_t up;
  /* host loops stay on this thread: worker threads would take the cores the
   * batch's stepping thread needs (200 arrivals/s: 39.7 served LPs/s with
   * 16 workers here and in the loader, 44.7 with none) */
  dpdlp_host_threads_cap(1);
  if (uctx_init(&uc, &up, S->b)) {
    pthread_mutex_lock(&S->mu);
    S->loader_error = 1;
    S->stop = 1;
    pthread_cond_broadcast(&S->cv);
    pthread_mutex_unlock(&S->mu);
    return NULL;
  }
  for (;;) {
    pthread_mutex_lock(&S->mu);
    while (!S->stop && S->lq_head == S->lq_tail && !S->loader_done)
      pthread_cond_wait(&S->cv, &S->mu);
    if (S->stop || (S->lq_head == S->lq_tail && S->loader_done)) {
      pthread_mutex_unlock(&S->mu);
      break;
    }
    int k = S->lq[S->lq_head++];
    S->lane_busy = 1;
    /* a lane request takes its memory when the lane starts it, so queued
     * large requests never hold back the small ones */
    while (!S->stop && S->admitted_live > 0 && S->in_use + S->bytes[k] > S->budget)
      pthread_cond_wait(&S->cv, &S->mu);
    if (S->stop) {
      S->lane_busy = 0;
      pthread_mutex_unlock(&S->mu);
      break;
    }
    S->in_use += S->bytes[k];
    S->admitted_live++;
    if (S->in_use > S->peak_bytes) S->peak_bytes = S->in_use;
    if (S->admitted_live > S->max_in_flight) S->max_in_flight = S->admitted_live;
    S->t_admit[k] = dpdlp_wall();
    pthread_mutex_unlock(&S->mu);
    pdhg_parameters_t prm = *S->prm;
    double left = S->t_cap - dpdlp_wall();
    if (left < prm.termination_criteria.time_sec_limit)
      prm.termination_criteria.time_sec_limit = left > 0 ? left : 0.0;
    dpdlp_solve_opts_t o;
    dpdlp_default_opts(&o);
    o.quiet = 1;
    o.precision = sp->precision;
    dpdlp_resident_t *R = dpdlp_i_upload_into(S->pool[S->req_pool[k]], 0, 0, 0, 0, &up);
    cupdlpx_result_t *r = R ? dpdlp_solve_resident(R, &prm, &o) : NULL;
    if (R) {
      size_t cap = 0;
      char *blob = dpdlp_i_resident_release(R, &cap);
      dpdlp_i_batch_put_mem(S->b, blob, cap);
    }
    double now = dpdlp_wall();
    pthread_mutex_lock(&S->mu);
    if (!r) {
      S->loader_error = 1;
      S->stop = 1;
    } else if (r->termination_reason == TERMINATION_REASON_TIME_LIMIT && now >= S->t_cap) {
      cupdlpx_result_free(r); /* cut by the end of the run: unfinished */
    } else {
      S->t_done[k] = now;
      S->res[k] = r;
      S->completed++;
    }
    S->in_use -= S->bytes[k];
    S->admitted_live--;
    S->lane_busy = 0;
    pthread_cond_broadcast(&S->cv);
    pthread_mutex_unlock(&S->mu);
  }
  uctx_free(&uc);
  return NULL;
}

static void *loader_main(void *arg) {
  srv_t *S = (srv_t *)arg;
  const dpdlp_serve_spec_t *sp = S->spec;
  double budget = sp->hbm_budget_gb * 1e9;
  uctx_t uc;
  dpdlp_uploader_t up;
  dpdlp_host_threads_cap(1); /* as the large-LP lane */
  if (uctx_init(&uc, &up, S->b)) {
    pthread_mutex_lock(&S->mu);
    S->loader_error = 1;
    S->stop = 1;
    S->loader_done = 1;
    pthread_cond_broadcast(&S->cv);
    pthread_mutex_unlock(&S->mu);
    return NULL;
  }
  for (int k = 0; k < S->nreq; ++k) {
    /* retired problems and surplus pooled arenas are freed here, off the
     * stepping thread */
    dpdlp_i_batch_reap(S->b, (size_t)(0.1 * budget));
    sleep_until(S->arrival[k]);
    pthread_mutex_lock(&S->mu);
    const lp_problem_t *P = S->pool[S->req_pool[k]];
    if (S->lane_on && P->constraint_matrix_num_nonzeros >= sp->large_lane_nnz) {
      S->lq[S->lq_tail++] = k; /* admitted by the lane when it starts it */
      pthread_cond_broadcast(&S->cv);
      pthread_mutex_unlock(&S->mu);
      continue;
    }
    while (!S->stop && S->admitted_live > 0 &&
           (S->in_use + S->bytes[k] > budget || S->admitted_live >= sp->max_lps))
      pthread_cond_wait(&S->cv, &S->mu);
    if (S->stop) {
      pthread_mutex_unlock(&S->mu);
      break;
    }
    S->in_use += S->bytes[k];
    S->admitted_live++;
    if (S->in_use > S->peak_bytes) S->peak_bytes = S->in_use;
    if (S->admitted_live > S->max_in_flight) S->max_in_flight = S->admitted_live;
    pthread_mutex_unlock(&S->mu);
    /* a slot of a just-retired LP is reclaimed at the batch's next window */
    int slot;
    while ((slot = dpdlp_i_batch_reserve(S->b)) < 0) {
      struct timespec ts = {0, 100000};
      nanosleep(&ts, NULL);
      pthread_mutex_lock(&S->mu);
      int stop = S->stop;
      pthread_mutex_unlock(&S->mu);
      if (stop) break;
    }
    if (slot < 0) break;
    dpdlp_resident_t *R = dpdlp_i_upload_into(S->pool[S->req_pool[k]], sp->spmv_cap,
                                              dpdlp_i_batch_cap_At(S->b), 1, 1, &up);
    if (!R || dpdlp_i_batch_prepare(S->b, slot, R, 1, k, S->stream)) {
      if (!R) dpdlp_i_batch_unreserve(S->b, slot);
      if (R) {
        size_t cap = 0;
        char *blob = dpdlp_i_resident_release(R, &cap);
        dpdlp_i_batch_put_mem(S->b, blob, cap);
      }
      pthread_mutex_lock(&S->mu);
      S->loader_error = 1;
      S->stop = 1;
      pthread_mutex_unlock(&S->mu);
      break;
    }
    S->t_admit[k] = dpdlp_wall();
  }
  uctx_free(&uc);
  pthread_mutex_lock(&S->mu);
  S->loader_done = 1;
  pthread_cond_broadcast(&S->cv);
  pthread_mutex_unlock(&S->mu);
  return NULL;
}

static int cmp_dbl(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y ? 1 : 0;
}

static double pct(const double *sorted, int n, double q) {
  if (n <= 0) return NAN;
  int k = (int)ceil(q * n) - 1;
  if (k < 0) k = 0;
  if (k >= n) k = n - 1;
  return sorted[k];
}

int dpdlp_serve_run(const dpdlp_serve_spec_t *spec, const pdhg_parameters_t *params,
                    dpdlp_serve_report_t *rep, double *latency_out, int *pool_index_out,
                    cupdlpx_result_t **results_out, int max_requests) {
  dpdlp_serve_spec_t sp;
  if (spec)
    sp = *spec;
  else
    dpdlp_serve_default_spec(&sp);
  pdhg_parameters_t prm;
  if (params)
    prm = *params;
  else
    set_default_parameters(&prm);
  if (!rep || sp.rate <= 0 || sp.duration_sec < 0 || sp.pool_small + sp.pool_large <= 0 ||
      sp.max_lps <= 0) {
    dpdlp_set_error("invalid serving spec");
    return -1;
  }
  memset(rep, 0, sizeof(*rep));
  int rc = -1;
  double tg = dpdlp_wall();
  /* ---- request pool (host), generated before the clock starts ---- */
  int npool = sp.pool_small + (sp.large_frac > 0 ? sp.pool_large : 0);
  lp_problem_t **pool = (lp_problem_t **)calloc((size_t)npool, sizeof(*pool));
  double *pool_bytes = (double *)calloc((size_t)npool, sizeof(double));
  int *req_pool = NULL;
  double *arrival = NULL, *bytes = NULL, *t_done = NULL, *lat = NULL, *t_admit = NULL;
  cupdlpx_result_t **res = NULL;
  dpdlp_batch_t *b = NULL;
  srv_t S;
  memset(&S, 0, sizeof(S));
  pthread_t th, th_lane;
  int loader_started = 0, lane_started = 0;
  if (!pool || !pool_bytes) goto out;
  for (int k = 0; k < sp.pool_small; ++k) {
    dpdlp_gen_spec_t g;
    dpdlp_batch_member_spec(k, sp.member_seed, sp.m_lo, sp.m_hi, &g);
    pool[k] = dpdlp_generate(&g, NULL);
    if (!pool[k]) goto out;
  }
  for (int j = 0; j < npool - sp.pool_small; ++j) {
    dpdlp_gen_spec_t g;
    dpdlp_batch_member_spec(j, sp.member_seed + 100000, sp.large_m, sp.large_m, &g);
    double lam = (double)sp.large_nnz / sp.large_m - 1.0;
    g.poisson_lambda = lam > 0.5 ? lam : 0.5;
    pool[sp.pool_small + j] = dpdlp_generate(&g, NULL);
    if (!pool[sp.pool_small + j]) goto out;
  }
  for (int k = 0; k < npool; ++k)
    pool_bytes[k] = est_bytes(pool[k], sp.spmv_cap, sp.precision == DPDLP_PREC_MIXED);
  rep->pool_gen_sec = dpdlp_wall() - tg;

  /* ---- arrivals ---- */
  uint64_t rs = sp.seed;
  int cap_req = (int)(sp.rate * sp.duration_sec * 1.5 + 64);
  if (max_requests > 0 && cap_req > max_requests) cap_req = max_requests;
  arrival = (double *)malloc((size_t)cap_req * sizeof(double));
  req_pool = (int *)malloc((size_t)cap_req * sizeof(int));
  if (!arrival || !req_pool) goto out;
  int nreq = 0;
  double t = 0.0;
  for (;;) {
    t += -log(1.0 - sm_u01(&rs)) / sp.rate;
    if (t >= sp.duration_sec || nreq >= cap_req) break;
    arrival[nreq] = t;
    double u = sm_u01(&rs);
    int large = npool > sp.pool_small && u < sp.large_frac;
    double v = sm_u01(&rs);
    req_pool[nreq] = large ? sp.pool_small + (int)(v * (npool - sp.pool_small))
                           : (int)(v * sp.pool_small);
    if (req_pool[nreq] >= npool) req_pool[nreq] = npool - 1;
    nreq++;
  }
  rep->arrived = nreq;
  bytes = (double *)malloc((size_t)(nreq ? nreq : 1) * sizeof(double));
  t_done = (double *)malloc((size_t)(nreq ? nreq : 1) * sizeof(double));
  t_admit = (double *)malloc((size_t)(nreq ? nreq : 1) * sizeof(double));
  lat = (double *)malloc((size_t)(nreq ? nreq : 1) * sizeof(double));
  res = (cupdlpx_result_t **)calloc((size_t)(nreq ? nreq : 1), sizeof(*res));
  if (!bytes || !t_done || !lat || !res || !t_admit) goto out;
  for (int k = 0; k < nreq; ++k) {
    bytes[k] = pool_bytes[req_pool[k]];
    t_done[k] = INFINITY;
    t_admit[k] = INFINITY;
  }

  double t0 = dpdlp_wall();
  double t_cap = t0 + sp.duration_sec + sp.drain_sec;
  for (int k = 0; k < nreq; ++k) arrival[k] += t0;
  int completed = 0;

  if (sp.mode == 1) {
    /* plain loop: one request at a time, single-LP FP64 path */
    dpdlp_solve_opts_t o;
    dpdlp_default_opts(&o);
    o.quiet = 1;
    o.precision = DPDLP_PREC_FP64;
    int inflight = 0;
    for (int k = 0; k < nreq; ++k) {
      sleep_until(arrival[k]);
      if (dpdlp_wall() > t_cap) break;
      t_admit[k] = dpdlp_wall();
      dpdlp_resident_t *R = dpdlp_upload(pool[req_pool[k]], 0);
      if (!R) goto out;
      cupdlpx_result_t *r = dpdlp_solve_resident(R, &prm, &o);
      dpdlp_resident_free(R);
      if (!r) goto out;
      t_done[k] = dpdlp_wall();
      res[k] = r;
      completed++;
      inflight = 1;
      if (bytes[k] > rep->peak_gb * 1e9) rep->peak_gb = bytes[k] / 1e9;
    }
    rep->max_in_flight = inflight;
  } else {
    dpdlp_batch_opts_t bo;
    dpdlp_batch_default_opts(&bo);
    bo.max_lps = sp.max_lps;
    bo.precision = sp.precision;
    bo.spmv_cap = sp.spmv_cap;
    bo.persist_max_nnz = sp.persist_max_nnz;
    b = dpdlp_batch_create(&prm, &bo);
    if (!b) goto out;
    S.spec = &sp;
    S.b = b;
    S.pool = pool;
    S.npool = npool;
    S.req_pool = req_pool;
    S.arrival = arrival;
    S.bytes = bytes;
    S.nreq = nreq;
    S.t_cap = t_cap;
    S.budget = sp.hbm_budget_gb * 1e9;
    S.t_admit = t_admit;
    pthread_mutex_init(&S.mu, NULL);
    pthread_cond_init(&S.cv, NULL);
    if (hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking) != hipSuccess) goto out;
    S.prm = &prm;
    S.t_done = t_done;
    S.res = res;
    S.lane_on = sp.large_lane_nnz > 0;
    S.lq = (int *)malloc((size_t)(nreq ? nreq : 1) * sizeof(int));
    if (!S.lq) goto out;
    if (pthread_create(&th, NULL, loader_main, &S) != 0) goto out;
    loader_started = 1;
    if (S.lane_on) {
      if (pthread_create(&th_lane, NULL, lane_main, &S) != 0) {
        S.lane_on = 0;
        goto out;
      }
      lane_started = 1;
    }
    for (;;) {
      int inflight = dpdlp_batch_in_flight(b);
      if (inflight > 0) {
        if (dpdlp_batch_step(b) < 0) {
          pthread_mutex_lock(&S.mu);
          S.stop = 1;
          pthread_cond_broadcast(&S.cv);
          pthread_mutex_unlock(&S.mu);
          goto out;
        }
      } else {
        /* nothing iterating: still step, so retired slots are reclaimed for
         * the loader */
        if (dpdlp_batch_step(b) < 0) {
          pthread_mutex_lock(&S.mu);
          S.stop = 1;
          pthread_cond_broadcast(&S.cv);
          pthread_mutex_unlock(&S.mu);
          goto out;
        }
        struct timespec ts = {0, 200000};
        nanosleep(&ts, NULL);
      }
      long long tag;
      cupdlpx_result_t *r;
      int freed = 0;
      double freed_bytes = 0.0;
      int got = 0;
      while (dpdlp_batch_take(b, &tag, &r, NULL)) {
        double now = dpdlp_wall();
        if (tag >= 0 && tag < nreq) {
          t_done[tag] = now;
          res[tag] = r;
          freed_bytes += bytes[tag];
          got++;
        } else {
          cupdlpx_result_free(r);
        }
        freed++;
      }
      pthread_mutex_lock(&S.mu);
      S.completed += got;
      if (freed) {
        S.in_use -= freed_bytes;
        S.admitted_live -= freed;
        pthread_cond_broadcast(&S.cv);
      }
      int done = S.loader_done && dpdlp_batch_in_flight(b) == 0 && S.admitted_live == 0 &&
                 S.lq_head == S.lq_tail && !S.lane_busy;
      int err = S.loader_error;
      if (dpdlp_wall() > t_cap) S.stop = 1;
      int stop = S.stop;
      pthread_mutex_unlock(&S.mu);
      if (err) goto out;
      if (done || (stop && S.loader_done)) break;
      if (stop) {
        pthread_mutex_lock(&S.mu);
        pthread_cond_broadcast(&S.cv);
        pthread_mutex_unlock(&S.mu);
      }
    }
    pthread_join(th, NULL);
    loader_started = 0;
    if (lane_started) {
      pthread_join(th_lane, NULL);
      lane_started = 0;
    }
    completed = S.completed;
    rep->max_in_flight = S.max_in_flight;
    rep->peak_gb = S.peak_bytes / 1e9;
    dpdlp_batch_counters(b, &rep->windows, &rep->syncs, NULL);
  }

  /* ---- report ---- */
  {
    double t_last = t0;
    int optimal = 0;
    for (int k = 0; k < nreq; ++k) {
      lat[k] = t_done[k] - arrival[k];
      if (isfinite(t_done[k]) && t_done[k] > t_last) t_last = t_done[k];
      if (res[k] && res[k]->termination_reason == TERMINATION_REASON_OPTIMAL) optimal++;
    }
    rep->completed = completed;
    rep->unfinished = nreq - completed;
    rep->optimal = optimal;
    rep->makespan_sec = t_last - t0;
    rep->lps_per_sec = rep->makespan_sec > 0 ? completed / rep->makespan_sec : 0.0;
    double *sorted = (double *)malloc((size_t)(nreq ? nreq : 1) * sizeof(double));
    if (!sorted) goto out;
    memcpy(sorted, lat, (size_t)nreq * sizeof(double));
    qsort(sorted, (size_t)nreq, sizeof(double), cmp_dbl);
    rep->p50_sec = pct(sorted, nreq, 0.50);
    rep->p90_sec = pct(sorted, nreq, 0.90);
    rep->p99_sec = pct(sorted, nreq, 0.99);
    rep->max_sec = nreq ? sorted[nreq - 1] : NAN;
    double sum = 0.0;
    int cnt = 0;
    for (int k = 0; k < nreq; ++k)
      if (isfinite(lat[k])) {
        sum += lat[k];
        cnt++;
      }
    rep->mean_sec = cnt ? sum / cnt : NAN;
    free(sorted);
    if (latency_out)
      for (int k = 0; k < nreq && (max_requests <= 0 || k < max_requests); ++k)
        latency_out[k] = lat[k];
    if (pool_index_out)
      for (int k = 0; k < nreq && (max_requests <= 0 || k < max_requests); ++k)
        pool_index_out[k] = req_pool[k];
    if (results_out) {
      for (int k = 0; k < nreq && (max_requests <= 0 || k < max_requests); ++k) {
        results_out[k] = res[k];
        res[k] = NULL;
      }
    }
  }
  rc = 0;
out:
  if (loader_started || lane_started) {
    pthread_mutex_lock(&S.mu);
    S.stop = 1;
    pthread_cond_broadcast(&S.cv);
    pthread_mutex_unlock(&S.mu);
    if (loader_started) pthread_join(th, NULL);
    if (lane_started) pthread_join(th_lane, NULL);
  }
  free(S.lq);
  if (b) dpdlp_batch_free(b);
  if (S.stream) hipStreamDestroy(S.stream);
  if (sp.mode != 1 && S.b) {
    pthread_mutex_destroy(&S.mu);
    pthread_cond_destroy(&S.cv);
  }
  if (res)
    for (int k = 0; k < rep->arrived; ++k) cupdlpx_result_free(res[k]);
  free(res);
  free(lat);
  free(t_done);
  free(t_admit);
  free(bytes);
  free(arrival);
  free(req_pool);
  if (pool)
    for (int k = 0; k < npool; ++k)
      if (pool[k]) lp_problem_free(pool[k]);
  free(pool);
  free(pool_bytes);
  return rc;
}
