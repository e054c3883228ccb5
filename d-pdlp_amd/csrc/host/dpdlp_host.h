/*
 * Internal declarations shared by the host C sources of the library.
 */
#pragma once

#include "cupdlpx_types.h"
#include <stdarg.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A problem resident in HBM (unscaled), see dpdlp_upload(). */
struct dpdlp_resident {
  int m, n, nnz, cap; /* cap: row-block capacity of A */
  int At_cap;         /* of A' */
  int rows_A, rows_At; /* most rows per untiled row block of A, A' */
  int *A_rp, *A_ci, *At_rp, *At_ci, *A_blk, *At_blk;
  int A_nblk, At_nblk;
  double *A_v, *At_v;
  /* column-panel tiles (P == 1: none) */
  int A_P, At_P;
  int *A_toff, *At_toff, *A_perm, *At_perm, *A_pci, *At_pci;
  unsigned short *A_rofs, *At_rofs;
  double *A_pv, *At_pv; /* unscaled tiled values */
  double *c, *lv, *uv, *lc, *uc;
  double c0, obj_l2, obj_linf, bnd_l2, bnd_linf;
  double *h_pstart, *h_dstart;
  int *h_A_blk, *h_At_blk; /* host copies of the row-block plans */
  int *h_row_perm;         /* rows renumbered: perm[new] = old (NULL: identity) */
  int *h_perm_rp, *h_perm_ci; /* upload-time copies (freed after upload) */
  double *h_perm_v, *h_perm_lc, *h_perm_uc;
  char *blob; /* the one device allocation every device array above lives in */
  size_t blob_cap;
  int pooled; /* blob came from an uploader's pool */
  double upload_time;
};

/* Source of a resident's device blob and pinned staging buffer (uploads
 * that must not allocate or free device memory while the device is busy). */
typedef struct {
  char *(*get_blob)(void *ctx, size_t need, size_t *cap);
  void (*put_blob)(void *ctx, char *p, size_t cap);
  char *(*get_stage)(void *ctx, size_t need);
  void *ctx;
  void *stream; /* hipStream_t of the one asynchronous copy */
} dpdlp_uploader_t;

/* dpdlp_upload_ex through an uploader */
struct dpdlp_resident *dpdlp_i_upload_into(const lp_problem_t *p, int spmv_cap, int spmv_cap_At,
                                           int panels_A, int panels_At,
                                           const dpdlp_uploader_t *u);
/* frees the host side of r and returns its device blob (and its size) */
char *dpdlp_i_resident_release(struct dpdlp_resident *r, size_t *cap);

/* Device descriptors of a resident problem's unscaled matrices (xpart must be
 * provided by the caller when P > 1, dpdlp_resident_xpart_len doubles,
 * zeroed before its first use: it ends with the tiled SpMVs' tickets). */
void dpdlp_resident_csr(const struct dpdlp_resident *r, int transpose, void *csr_out,
                        double *xpart);
/* Doubles of scratch the column-panel SpMVs of r need: the partial row sums,
 * then one 4-byte ticket per row block of each tiled matrix. */
long long dpdlp_resident_xpart_len(const struct dpdlp_resident *r);

/* logging.c: the reference's console formats (src/utils.cu:191-501) */
const char *dpdlp_termination_string(termination_reason_t r);
int dpdlp_print_frequency(int iter);
void dpdlp_print_initial_info(const pdhg_parameters_t *p, const lp_problem_t *prob);
void dpdlp_print_table_header(void);
void dpdlp_print_iteration(int total, double t, double pobj, double dobj, double abs_p,
                           double abs_d, double gap, double rel_p, double rel_d,
                           double rel_gap);
void dpdlp_print_final(const cupdlpx_result_t *r, const pdhg_parameters_t *p);
void dpdlp_print_polish_header(int primal, const pdhg_parameters_t *p);
void dpdlp_print_polish_iteration(int primal, int total, double t, double obj, double abs_r,
                                  double rel_r);
void dpdlp_print_polish_final(termination_reason_t pr, int piters, double ptime,
                              termination_reason_t dr, int diters, double dtime,
                              double prel, double drel, double gap);

/* util.c: error text, wall clock */
void dpdlp_set_error(const char *fmt, ...);
double dpdlp_wall(void);

/* util.c: host worker threads for the data-parallel loops of ingestion and
 * upload.  dpdlp_host_threads() is DPDLP_HOST_THREADS when set, else
 * OMP_NUM_THREADS when set, else the CPUs this process may run on, clamped
 * to [1, 32].  dpdlp_par_for runs fn(ctx, t) for every
 * t in [0, ntask) on up to that many threads (the caller's included; tasks
 * are taken in increasing order) and returns when all have finished. */
int dpdlp_host_threads(void);
void dpdlp_par_for(int ntask, void (*fn)(void *ctx, int task), void *ctx);
/* caps dpdlp_host_threads() for the calling thread only (0: no cap): the
 * serving loader and large-LP lane run their loops on their own thread, so
 * the thread that steps the batch keeps its core */
void dpdlp_host_threads_cap(int max_threads);

/* util.c: stable CSR transpose on the host */
int dpdlp_host_transpose(int m, int n, const int *rp, const int *ci, const double *v,
                         int *t_rp, int *t_ci, double *t_v);
/* A''s row pointers only (the device transposes the entries) */
void dpdlp_host_col_ptr(int m, int n, const int *rp, const int *ci, int *t_rp);

/* rng.cc: std::mt19937(seed) + std::normal_distribution<double>(0, 1) */
void dpdlp_normal_fill(uint32_t seed, int count, double *out);

#ifdef __cplusplus
}
#endif
