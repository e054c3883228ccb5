/*
 * Kernel timing helpers for the benchmark: a standalone SpMV on a resident
 * problem timed with HIP events on the stream it is launched on.
 */
#include "dpdlp_dev.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"
#include "solver_internal.h"

#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>


void dpdlp_resident_info(const dpdlp_resident_t *r, int *m, int *n, int *nnz, int *nblk_A,
                         int *nblk_At) {
  if (m) *m = r->m;
  if (n) *n = r->n;
  if (nnz) *nnz = r->nnz;
  if (nblk_A) *nblk_A = r->A_nblk;
  if (nblk_At) *nblk_At = r->At_nblk;
}

void dpdlp_resident_layout(const dpdlp_resident_t *r, int *panels_A, int *panels_At, int *cap_A,
                           int *cap_At, int *renumbered) {
  if (panels_A) *panels_A = r->A_P;
  if (panels_At) *panels_At = r->At_P;
  if (cap_A) *cap_A = r->cap;
  if (cap_At) *cap_At = r->At_cap;
  if (renumbered) *renumbered = r->h_row_perm != NULL;
}

double dpdlp_time_spmv(dpdlp_resident_t *r, int transpose, int reps, int use32) {
  if (!r || reps <= 0) return -1.0;
  dpdlp_csr_t A;
  double *xpart = NULL;
  long long xl = dpdlp_resident_xpart_len(r);
  if (xl > 0 && hipMalloc((void **)&xpart, (size_t)xl * sizeof(double)) != hipSuccess) return -1;
  if (xl > 0 && hipMemset(xpart, 0, (size_t)xl * sizeof(double)) != hipSuccess) {
    hipFree(xpart);
    return -1;
  }
  dpdlp_resident_csr(r, transpose, &A, xpart);
  hipStream_t s = NULL;
  hipEvent_t e0 = NULL, e1 = NULL;
  double *x = NULL, *y = NULL, *xp = NULL;
  float *x32 = NULL, *v32 = NULL, *pv32 = NULL;
  double ms = -1.0;
  if (hipStreamCreate(&s) != hipSuccess) return -1.0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto out;
  if (hipMalloc((void **)&x, (size_t)(A.cols ? A.cols : 1) * 8) != hipSuccess) goto out;
  if (hipMalloc((void **)&y, (size_t)(A.rows ? A.rows : 1) * 8) != hipSuccess) goto out;
  if (dpdlp_dev_fill(x, 1.0, A.cols, s)) goto out;
  if (use32) {
    if (hipMalloc((void **)&x32, (size_t)(A.cols ? A.cols : 1) * 4) != hipSuccess) goto out;
    if (hipMalloc((void **)&v32, (size_t)(A.nnz ? A.nnz : 1) * 4) != hipSuccess) goto out;
    if (dpdlp_dev_make_f32(x, x32, A.cols, s)) goto out;
    if (dpdlp_dev_make_f32(A.v, v32, A.nnz, s)) goto out;
    A.v32 = v32;
    if (A.P > 1) {
      if (hipMalloc((void **)&pv32, (size_t)(A.nnz ? A.nnz : 1) * 4) != hipSuccess) goto out;
      if (dpdlp_dev_make_f32(A.pv, pv32, A.nnz, s)) goto out;
      A.pv32 = pv32;
    }
  }
  /* warm up */
  for (int k = 0; k < 3; ++k)
    if ((use32 ? dpdlp_dev_spmv_f32(&A, x32, y, s) : dpdlp_dev_spmv(&A, x, y, s))) goto out;
  if (hipEventRecord(e0, s) != hipSuccess) goto out;
  for (int k = 0; k < reps; ++k)
    if ((use32 ? dpdlp_dev_spmv_f32(&A, x32, y, s) : dpdlp_dev_spmv(&A, x, y, s))) goto out;
  if (hipEventRecord(e1, s) != hipSuccess) goto out;
  if (hipEventSynchronize(e1) != hipSuccess) goto out;
  {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = (double)t / reps;
  }
out:
  free(xp);
  if (s) hipStreamSynchronize(s);
  if (x) hipFree(x);
  if (y) hipFree(y);
  if (x32) hipFree(x32);
  if (v32) hipFree(v32);
  if (pv32) hipFree(pv32);
  if (xpart) hipFree(xpart);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  return ms;
}

int dpdlp_resident_transpose(const dpdlp_resident_t *r, int *t_rp, int *t_ci, double *t_v) {
  if (!r || !t_rp || !t_ci || !t_v || r->At_P > 1) return -1;
  if (hipMemcpy(t_rp, r->At_rp, (size_t)(r->n + 1) * sizeof(int), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return -1;
  if (r->nnz > 0 &&
      (hipMemcpy(t_ci, r->At_ci, (size_t)r->nnz * sizeof(int), hipMemcpyDeviceToHost) !=
           hipSuccess ||
       hipMemcpy(t_v, r->At_v, (size_t)r->nnz * sizeof(double), hipMemcpyDeviceToHost) !=
           hipSuccess))
    return -1;
  return 0;
}

int dpdlp_resident_spmv(dpdlp_resident_t *r, int transpose, const double *xh, double *yh,
                        int use32) {
  return dpdlp_resident_spmv_ex(r, transpose, xh, yh, use32, 1, 0);
}

int dpdlp_resident_spmv_ex(dpdlp_resident_t *r, int transpose, const double *xh, double *yh,
                           int use32, int reps, int two_kernel) {
  if (!r || !xh || !yh || reps < 1) return -1;
  dpdlp_csr_t A;
  double *xpart = NULL;
  long long xl = dpdlp_resident_xpart_len(r);
  if (xl > 0 && hipMalloc((void **)&xpart, (size_t)xl * sizeof(double)) != hipSuccess) return -1;
  if (xl > 0 && hipMemset(xpart, 0, (size_t)xl * sizeof(double)) != hipSuccess) {
    hipFree(xpart);
    return -1;
  }
  dpdlp_resident_csr(r, transpose, &A, xpart);
  if (two_kernel) A.tick = NULL;
  double *x = NULL, *y = NULL, *xp = NULL;
  float *x32 = NULL, *v32 = NULL, *pv32 = NULL;
  int rc = -1;
  hipStream_t s = NULL;
  if (hipStreamCreate(&s) != hipSuccess) return -1;
  if (hipMalloc((void **)&x, (size_t)(A.cols ? A.cols : 1) * 8) != hipSuccess) goto out;
  if (hipMalloc((void **)&y, (size_t)(A.rows ? A.rows : 1) * 8) != hipSuccess) goto out;
  if (transpose && r->h_row_perm && A.cols) {
    /* A' x takes x in the caller's row order: renumber it */
    xp = (double *)malloc((size_t)A.cols * sizeof(double));
    if (!xp) goto out;
    for (int i = 0; i < A.cols; ++i) xp[i] = xh[r->h_row_perm[i]];
  }
  if (A.cols &&
      hipMemcpy(x, xp ? xp : xh, (size_t)A.cols * 8, hipMemcpyHostToDevice) != hipSuccess)
    goto out;
  if (use32) {
    if (hipMalloc((void **)&x32, (size_t)(A.cols ? A.cols : 1) * 4) != hipSuccess) goto out;
    if (hipMalloc((void **)&v32, (size_t)(A.nnz ? A.nnz : 1) * 4) != hipSuccess) goto out;
    if (dpdlp_dev_make_f32(x, x32, A.cols, s)) goto out;
    if (dpdlp_dev_make_f32(A.v, v32, A.nnz, s)) goto out;
    A.v32 = v32;
    if (A.P > 1) {
      if (hipMalloc((void **)&pv32, (size_t)(A.nnz ? A.nnz : 1) * 4) != hipSuccess) goto out;
      if (dpdlp_dev_make_f32(A.pv, pv32, A.nnz, s)) goto out;
      A.pv32 = pv32;
    }
    for (int k = 0; k < reps; ++k)
      if (dpdlp_dev_spmv_f32(&A, x32, y, s)) goto out;
  } else {
    for (int k = 0; k < reps; ++k)
      if (dpdlp_dev_spmv(&A, x, y, s)) goto out;
  }
  if (hipStreamSynchronize(s) != hipSuccess) goto out;
  if (A.rows && hipMemcpy(yh, y, (size_t)A.rows * 8, hipMemcpyDeviceToHost) != hipSuccess)
    goto out;
  if (!transpose && r->h_row_perm && A.rows) {
    /* A x comes out in the renumbered row order: back to the caller's */
    double *t = (double *)malloc((size_t)A.rows * sizeof(double));
    if (!t) goto out;
    for (int i = 0; i < A.rows; ++i) t[r->h_row_perm[i]] = yh[i];
    memcpy(yh, t, (size_t)A.rows * sizeof(double));
    free(t);
  }
  rc = 0;
out:
  if (s) hipStreamSynchronize(s);
  if (x) hipFree(x);
  if (y) hipFree(y);
  if (x32) hipFree(x32);
  if (v32) hipFree(v32);
  if (pv32) hipFree(pv32);
  if (xpart) hipFree(xpart);
  if (s) hipStreamDestroy(s);
  return rc;
}

int dpdlp_scaling_host(const lp_problem_t *p, const pdhg_parameters_t *params, double *val,
                       double *c, double *lv, double *uv, double *lc, double *uc, double *D_r,
                       double *D_c, double *beta_gamma) {
  dpdlp_resident_t *R = dpdlp_upload_ex(p, 0, 1, 1);
  if (!R) return -1;
  ctx_t cx;
  memset(&cx, 0, sizeof(cx));
  double rt = 0.0;
  int rc = -1;
  const size_t m = (size_t)R->m, n = (size_t)R->n, nnz = (size_t)R->nnz;
  if (dpdlp_i_ctx_setup(&cx, R, params, 0, NULL, NULL, NULL, &rt)) goto out;
  if ((nnz && hipMemcpy(val, cx.w_A_v, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess) ||
      (n && (hipMemcpy(c, cx.st.c, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(lv, cx.st.lv, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(uv, cx.st.uv, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(D_c, cx.st.D_c, n * 8, hipMemcpyDeviceToHost) != hipSuccess)) ||
      (m && (hipMemcpy(lc, cx.st.lc, m * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(uc, cx.st.uc, m * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(D_r, cx.st.D_r, m * 8, hipMemcpyDeviceToHost) != hipSuccess)))
    goto out;
  beta_gamma[0] = cx.k.beta;
  beta_gamma[1] = cx.k.gamma;
  rc = 0;
out:
  dpdlp_i_ctx_free(&cx);
  dpdlp_resident_free(R);
  return rc;
}

int dpdlp_reduce_host(int op, const double *xh, const double *yh, long long n, double *out) {
  double *x = NULL, *y = NULL, *part = NULL, *o = NULL;
  int rc = -1;
  size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
  if (!out || n < 0) return -1;
  if (hipMalloc((void **)&x, bytes) != hipSuccess || hipMalloc((void **)&y, bytes) != hipSuccess ||
      hipMalloc((void **)&part, 4 * (size_t)dpdlp_dev_reduce_blocks(n) * sizeof(double)) !=
          hipSuccess ||
      hipMalloc((void **)&o, sizeof(double)) != hipSuccess)
    goto out;
  if (n > 0 && (hipMemcpy(x, xh, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                (yh && hipMemcpy(y, yh, (size_t)n * sizeof(double), hipMemcpyHostToDevice) !=
                           hipSuccess)))
    goto out;
  if (dpdlp_dev_reduce(op, x, y, n, part, o, NULL) != 0) goto out;
  if (hipMemcpy(out, o, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) goto out;
  rc = 0;
out:
  if (x) hipFree(x);
  if (y) hipFree(y);
  if (part) hipFree(part);
  if (o) hipFree(o);
  return rc;
}

double dpdlp_gather_calibration_values(long long slice_bytes, long long gathers, int reps,
                                       int xcd_local, int chunk, int with_values) {
  long long slice = slice_bytes / 8;
  if (slice < 1 || slice > 0x7fffffffLL || gathers < 1 || reps < 1 || chunk < 1) return -1.0;
  long long xlen = xcd_local ? 8 * slice : slice;
  long long blocks = (gathers + chunk - 1) / chunk;
  int *h = (int *)malloc((size_t)gathers * sizeof(int));
  int *idx = NULL;
  double *x = NULL, *y = NULL, *val = NULL;
  hipEvent_t e0 = NULL, e1 = NULL;
  double rate = -1.0;
  if (!h) return -1.0;
  uint64_t st = 0x9e3779b97f4a7c15ull; /* xorshift64: fixed random columns */
  for (long long k = 0; k < gathers; ++k) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    h[k] = (int)(st % (uint64_t)slice);
  }
  if (hipMalloc((void **)&idx, (size_t)gathers * sizeof(int)) != hipSuccess) goto out;
  if (hipMalloc((void **)&x, (size_t)xlen * sizeof(double)) != hipSuccess) goto out;
  if (hipMalloc((void **)&y, (size_t)blocks * 256 * sizeof(double)) != hipSuccess) goto out;
  if (with_values) {
    if (hipMalloc((void **)&val, (size_t)gathers * sizeof(double)) != hipSuccess) goto out;
    if (hipMemset(val, 0, (size_t)gathers * sizeof(double)) != hipSuccess) goto out;
  }
  if (hipMemcpy(idx, h, (size_t)gathers * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(x, 0, (size_t)xlen * sizeof(double)) != hipSuccess)
    goto out;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto out;
  if (dpdlp_dev_gather_calib_values(idx, val, x, y, gathers, chunk, slice, xcd_local, NULL))
    goto out;
  if (hipEventRecord(e0, NULL) != hipSuccess) goto out;
  for (int k = 0; k < reps; ++k)
    if (dpdlp_dev_gather_calib_values(idx, val, x, y, gathers, chunk, slice, xcd_local, NULL))
      goto out;
  if (hipEventRecord(e1, NULL) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto out;
  {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) == hipSuccess && t > 0.f)
      rate = (double)gathers * reps / ((double)t * 1e-3) / 1e9;
  }
out:
  hipDeviceSynchronize();
  free(h);
  if (idx) hipFree(idx);
  if (x) hipFree(x);
  if (y) hipFree(y);
  if (val) hipFree(val);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  return rate;
}

double dpdlp_gather_calibration_ex(long long slice_bytes, long long gathers, int reps,
                                   int xcd_local, int chunk) {
  return dpdlp_gather_calibration_values(slice_bytes, gathers, reps, xcd_local, chunk, 0);
}

double dpdlp_gather_calibration(long long slice_bytes, long long gathers, int reps) {
  return dpdlp_gather_calibration_ex(slice_bytes, gathers, reps, 1, 2048);
}

double dpdlp_copy_calibration(long long bytes, int reps) {
  void *a = NULL, *b = NULL;
  hipEvent_t e0 = NULL, e1 = NULL;
  double ms = -1.0;
  if (hipMalloc(&a, (size_t)bytes) != hipSuccess) return -1.0;
  if (hipMalloc(&b, (size_t)bytes) != hipSuccess) goto out;
  if (hipMemset(a, 0, (size_t)bytes) != hipSuccess) goto out;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto out;
  if (dpdlp_dev_copy_calib(a, b, bytes, NULL)) goto out;
  if (hipEventRecord(e0, NULL) != hipSuccess) goto out;
  for (int k = 0; k < reps; ++k)
    if (dpdlp_dev_copy_calib(a, b, bytes, NULL)) goto out;
  if (hipEventRecord(e1, NULL) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto out;
  {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = (double)t / reps;
  }
out:
  hipDeviceSynchronize();
  if (a) hipFree(a);
  if (b) hipFree(b);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  return ms;
}

int dpdlp_resident_device_csr(const dpdlp_resident_t *r, int transpose, void *csr_out) {
  if (!r || !csr_out) return -1;
  dpdlp_resident_csr(r, transpose, csr_out, NULL);
  return 0;
}
