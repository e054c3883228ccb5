/*
 * Batch layout autotuner: per LP-size class, the row-block layout of A and A'
 * (row-block capacity in nonzeros, and most rows per block, which sets the
 * lane-group width of the row reduction) that runs the batch's own phase
 * kernels fastest, timed on a sample of the class and cached by signature.
 *
 * A size class bins an LP by its nonzeros (factors of 4) and by the mean row
 * lengths of A and A' (factors of 2).  For each class not yet in the cache, up
 * to AT_SAMPLE members of the class (the first in index order) are uploaded
 * with every candidate layout in turn and solved as one batch for a short
 * fixed number of iterations with live phase timing: the dual phase (A x)
 * times the A layout, the primal phase (A' y) the A' layout, so one batch
 * per candidate times both.  The cache is a text file of lines
 *   dpdlp-layout-v1 <device> <class> <cap_A> <rows_A> <cap_At> <rows_At> <us_A> <us_At>
 * keyed by the device name, its compute-unit count, the precision mode and
 * the class; a class found there is not timed again.
 */
#include "cupdlpx.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define AT_SAMPLE 16
#define AT_MAX_CLASSES 256

static const int at_caps[3] = {1024, 2048, 4096};
static const int at_rows[3] = {64, 128, 256};

static int ilog(double v, double base) {
  if (!(v > 0)) return 0;
  return (int)floor(log(v) / log(base) + 1e-9);
}

int dpdlp_batch_size_class(const lp_problem_t *p) {
  if (!p) return -1;
  const int m = p->num_constraints, n = p->num_variables;
  const long long nnz = p->constraint_matrix_num_nonzeros;
  const int bn = ilog((double)(nnz > 0 ? nnz : 1), 4.0);
  const int ba = ilog(m > 0 ? (double)nnz / m : 1.0, 2.0);
  const int bt = ilog(n > 0 ? (double)nnz / n : 1.0, 2.0);
  const int ca = ba < 0 ? 0 : (ba > 15 ? 15 : ba), ct = bt < 0 ? 0 : (bt > 15 ? 15 : bt);
  return bn * 256 + ca * 16 + ct;
}

static void device_signature(char *out, size_t len) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    snprintf(out, len, "unknown");
    return;
  }
  snprintf(out, len, "%.64s:%dcu", prop.gcnArchName, prop.multiProcessorCount);
  for (char *c = out; *c; ++c)
    if (*c == ' ') *c = '_';
}

typedef struct {
  int cls;
  dpdlp_batch_layout_t L;
  int have;
} at_entry_t;

static int cache_lookup(const char *path, const char *dev, at_entry_t *E, int ne) {
  FILE *f = path ? fopen(path, "r") : NULL;
  int found = 0;
  if (!f) return 0;
  char line[512], tag[64], d[200];
  while (fgets(line, sizeof line, f)) {
    int cls, ca, ra, ct, rt;
    double ua, ut;
    if (sscanf(line, "%63s %199s %d %d %d %d %d %lf %lf", tag, d, &cls, &ca, &ra, &ct, &rt, &ua,
               &ut) != 9)
      continue;
    if (strcmp(tag, "dpdlp-layout-v1") || strcmp(d, dev)) continue;
    for (int k = 0; k < ne; ++k)
      if (E[k].cls == cls && !E[k].have) {
        E[k].L = (dpdlp_batch_layout_t){ca, ra, ct, rt};
        E[k].have = 1;
        ++found;
      }
  }
  fclose(f);
  return found;
}

/* Times one candidate layout on a class sample: mean ms per sampled launch
 * of the dual (A) and primal (A') phases, or -1. */
static int time_candidate(const lp_problem_t *const *sample, int ns, const dpdlp_batch_layout_t *L,
                          const pdhg_parameters_t *prm, int precision, double *ms_A,
                          double *ms_At) {
  dpdlp_resident_t *R[AT_SAMPLE];
  cupdlpx_result_t *res[AT_SAMPLE];
  int nr = 0, rc = -1;
  for (; nr < ns; ++nr) {
    R[nr] = dpdlp_upload_batch_layout(sample[nr], L);
    if (!R[nr]) goto out;
  }
  dpdlp_batch_opts_t o;
  dpdlp_batch_default_opts(&o);
  o.max_lps = ns;
  o.precision = precision;
  o.spmv_cap = 4096;
  o.spmv_cap_At = 4096;
  o.profile_phases = 1;
  o.solo_lps = 0;
  memset(res, 0, sizeof res);
  if (dpdlp_solve_batch(R, ns, prm, &o, res) != ns) goto out;
  for (int i = 0; i < ns; ++i) cupdlpx_result_free(res[i]);
  if (o.phase_samples <= 0) goto out;
  *ms_A = o.dual_ms / o.phase_samples;
  *ms_At = o.primal_ms / o.phase_samples;
  rc = 0;
out:
  for (int i = 0; i < nr; ++i) dpdlp_resident_free(R[i]);
  return rc;
}

int dpdlp_batch_autotune(const lp_problem_t *const *probs, int count,
                         const pdhg_parameters_t *params, int precision, const char *cache_path,
                         dpdlp_batch_layout_t *layouts, dpdlp_autotune_report_t *rep) {
  if (!probs || count < 0 || !layouts || !params) {
    dpdlp_set_error("dpdlp_batch_autotune: bad arguments");
    return -1;
  }
  double t0 = dpdlp_wall();
  dpdlp_autotune_report_t r0;
  memset(&r0, 0, sizeof r0);
  int *cls = (int *)malloc((size_t)(count ? count : 1) * sizeof(int));
  at_entry_t *E = (at_entry_t *)calloc(AT_MAX_CLASSES, sizeof(at_entry_t));
  int ne = 0, rc = -1;
  if (!cls || !E) {
    dpdlp_set_error("host allocation failed");
    goto out;
  }
  for (int i = 0; i < count; ++i) {
    cls[i] = dpdlp_batch_size_class(probs[i]);
    int k = 0;
    while (k < ne && E[k].cls != cls[i]) ++k;
    if (k == ne) {
      if (ne == AT_MAX_CLASSES) {
        dpdlp_set_error("dpdlp_batch_autotune: more than %d size classes", AT_MAX_CLASSES);
        goto out;
      }
      E[ne++].cls = cls[i];
    }
  }
  r0.classes = ne;
  char dev[200];
  device_signature(dev, sizeof dev);
  {
    /* FP32 and FP64 matrix values time differently: separate cache keys */
    size_t l = strlen(dev);
    snprintf(dev + l, sizeof dev - l, ":%s", precision == DPDLP_PREC_MIXED ? "mixed" : "fp64");
  }
  r0.cached = cache_lookup(cache_path, dev, E, ne);

  /* a short run: the layout's effect shows in every phase launch */
  pdhg_parameters_t prm = *params;
  prm.termination_criteria.iteration_limit = 200;
  prm.sv_max_iter = 30;
  prm.feasibility_polishing = false;
  prm.verbose = false;
  for (int k = 0; k < ne; ++k) {
    if (E[k].have) continue;
    const lp_problem_t *sample[AT_SAMPLE];
    int ns = 0;
    for (int i = 0; i < count && ns < AT_SAMPLE; ++i)
      if (cls[i] == E[k].cls) sample[ns++] = probs[i];
    double bestA = 1e300, bestAt = 1e300;
    dpdlp_batch_layout_t best = {4096, 256, 4096, 256};
    for (int ci = 0; ci < 3; ++ci)
      for (int ri = 0; ri < 3; ++ri) {
        dpdlp_batch_layout_t L = {at_caps[ci], at_rows[ri], at_caps[ci], at_rows[ri]};
        double ta, tt;
        if (time_candidate(sample, ns, &L, &prm, precision, &ta, &tt)) goto out;
        if (ta < bestA) {
          bestA = ta;
          best.cap_A = L.cap_A;
          best.rows_A = L.rows_A;
        }
        if (tt < bestAt) {
          bestAt = tt;
          best.cap_At = L.cap_At;
          best.rows_At = L.rows_At;
        }
      }
    E[k].L = best;
    E[k].have = 1;
    r0.tuned++;
    if (cache_path) {
      FILE *f = fopen(cache_path, "a");
      if (f) {
        fprintf(f, "dpdlp-layout-v1 %s %d %d %d %d %d %.3f %.3f\n", dev, E[k].cls, best.cap_A,
                best.rows_A, best.cap_At, best.rows_At, bestA * 1e3, bestAt * 1e3);
        fclose(f);
      }
    }
  }
  for (int i = 0; i < count; ++i)
    for (int k = 0; k < ne; ++k)
      if (E[k].cls == cls[i]) layouts[i] = E[k].L;
  rc = 0;
out:
  r0.seconds = dpdlp_wall() - t0;
  if (rep) *rep = r0;
  free(cls);
  free(E);
  return rc;
}
