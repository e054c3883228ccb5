/*
 * Internal pieces of the solve driver shared by the single-LP path
 * (solver.c) and the batched / served path (batch.c).
 */
#pragma once

#include "control.h"
#include "dpdlp_dev.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"

#include <hip/hip_runtime_api.h>

#define MAX_ALLOCS 128
typedef struct {
  hipStream_t s;
  dpdlp_state_t st;
  double *h_scal;
  lpk_t k;
  void *allocs[MAX_ALLOCS];
  int nalloc;
  int mixed, use32;
  int owns_stream, owns_scal;
  char *arena; /* one device allocation for the LP state (zeroed) */
  size_t arena_off, arena_len, arena_cap;
  double *pi_eig, *pi_next, *pi_dual, *rs, *cs;
  const int *row_perm; /* the resident problem's row renumbering or NULL */
  double *w_A_v, *w_At_v;
  float *w_A_v32, *w_At_v32;
  /* live phase timing */
  hipEvent_t ev[64][3];
  int nev_created, nev_used;
  double p1_ms, p2_ms;
  double t_issue; /* wall time the set-up was issued */
  int nsamp;
} ctx_t;


/* Allocates one LP's device state from a resident problem and runs the
 * device scaling, warm start and FP32 copies.  stream: NULL = own stream;
 * scal_dev/scal_host: external slot arrays (NULL = own).  Returns 0 or -1. */
int dpdlp_i_ctx_setup(ctx_t *c, dpdlp_resident_t *R, const pdhg_parameters_t *params, int mixed,
                      hipStream_t stream, double *scal_dev, double *scal_host,
                      double *rescale_time);
/* The same in two halves: issue queues every set-up operation on the stream
 * (the scalars' copy back included) without waiting; complete, called once
 * the stream has been synchronised, reads beta/gamma and queues the rest
 * (warm start, FP32 copies). */
int dpdlp_i_ctx_setup_issue(ctx_t *c, dpdlp_resident_t *R, const pdhg_parameters_t *params,
                            int mixed, hipStream_t stream, double *scal_dev, double *scal_host);
int dpdlp_i_ctx_setup_complete(ctx_t *c, dpdlp_resident_t *R, const pdhg_parameters_t *params,
                               double *rescale_time);
/* Unscaled result of a finished LP (D2H, synchronous on c->s). */
cupdlpx_result_t *dpdlp_i_ctx_result(ctx_t *c, const hstate_t *h, double rescale_time,
                                     double polish_time, int polish_iters);
/* The same in two halves: launch the unscale kernels and the device-to-host
 * copies into caller buffers (asynchronous on c->s), and fill the scalars. */
int dpdlp_i_result_launch(ctx_t *c, double *x_host, double *y_host, double *rc_host);
void dpdlp_i_result_scalars(cupdlpx_result_t *res, const ctx_t *c, const hstate_t *h,
                            double rescale_time, double polish_time, int polish_iters);
void dpdlp_i_ctx_free(ctx_t *c);
/* Initial control state after the step size is known. */
void dpdlp_i_init_hstate(hstate_t *h, const ctx_t *c, const pdhg_parameters_t *params,
                         double step);
/* Device bytes of one LP's solver state (the arena dpdlp_i_ctx_setup makes). */
size_t dpdlp_i_ctx_footprint(const dpdlp_resident_t *R, int mixed, int own_scal);
/* Feasibility polishing after the main loop (pdlp_core_op.cu:862-1030). */
int dpdlp_i_feasibility_polish(ctx_t *c, hstate_t *h, const pdhg_parameters_t *params,
                               double *ptime, int *piters);

/* batch.c: slot hand-over for set-up on another thread (the serving loop) */
typedef struct dpdlp_batch dpdlp_batch_t;
int dpdlp_i_batch_reserve(dpdlp_batch_t *b);
void dpdlp_i_batch_unreserve(dpdlp_batch_t *b, int slot);
int dpdlp_i_batch_prepare(dpdlp_batch_t *b, int slot, dpdlp_resident_t *R, int take_ownership,
                          long long tag, void *stream);
/* dpdlp_i_batch_prepare in two halves around a synchronisation of the
 * set-up stream (many issues may share one synchronisation); abandon drops
 * an issued set-up after that synchronisation instead of completing it */
int dpdlp_i_batch_prepare_issue(dpdlp_batch_t *b, int slot, dpdlp_resident_t *R,
                                int take_ownership, long long tag, void *stream);
int dpdlp_i_batch_prepare_complete(dpdlp_batch_t *b, int slot);
void dpdlp_i_batch_abandon(dpdlp_batch_t *b, int slot);
/* frees retired problems and trims pooled arenas to keep_bytes (call from a
 * set-up thread: device frees may synchronise the device) */
void dpdlp_i_batch_reap(dpdlp_batch_t *b, size_t keep_bytes);
/* device memory from / back to the batch's pool of recycled allocations */
char *dpdlp_i_batch_get_mem(dpdlp_batch_t *b, size_t need, size_t *cap);
/* the row-block capacity the batch's LPs use for A' */
int dpdlp_i_batch_cap_At(const dpdlp_batch_t *b);
void dpdlp_i_batch_put_mem(dpdlp_batch_t *b, char *p, size_t cap);
size_t dpdlp_i_batch_pool_bytes(dpdlp_batch_t *b);
