/*
 * Host-code harness built with AddressSanitizer + UndefinedBehaviorSanitizer
 * by tests/test_sanitizers.py (CPU only; no device code is linked).
 *
 * Exercises the host C of the library that runs without a device: the MPS
 * reader (every fixture given on the command line, plus missing and
 * malformed files), create_lp_problem's four input layouts and its error
 * paths, the LP generator, the stable transpose, the row-block and
 * column-panel planners, and the FP64 oracle solve.  Exit status 0 and a
 * clean sanitizer report are the pass condition.
 *
 * The two solver symbols api.c references (the device solve) abort if
 * reached: nothing here may call them.
 */
#include "cupdlpx.h"
#include "dpdlp_dev.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"
#include "mps_parser.h"
#include "pdlp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

cupdlpx_result_t *dpdlp_solve_ex(const lp_problem_t *p, const pdhg_parameters_t *q,
                                 dpdlp_solve_opts_t *o) {
  (void)p;
  (void)q;
  (void)o;
  fprintf(stderr, "harness: device solve reached\n");
  abort();
}

void dpdlp_default_opts(dpdlp_solve_opts_t *o) { memset(o, 0, sizeof(*o)); }

#define CHECK(c)                                                                   \
  do {                                                                             \
    if (!(c)) {                                                                    \
      fprintf(stderr, "harness: check failed at line %d: %s\n", __LINE__, #c);     \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

static int layouts(void) {
  /* the C API example: 3 x 2 */
  const double A[6] = {1, 1, 1, -1, 1, 0};
  const double c[2] = {1, 1}, cl[3] = {3, -INFINITY, 1}, cu[3] = {3, 0, INFINITY};
  const double vl[2] = {0, 0}, vu[2] = {INFINITY, INFINITY};
  matrix_desc_t d;
  memset(&d, 0, sizeof(d));
  d.m = 3;
  d.n = 2;
  d.fmt = matrix_dense;
  d.zero_tolerance = 1e-12;
  d.data.dense.A = A;
  lp_problem_t *p = create_lp_problem(c, &d, cl, cu, vl, vu, NULL);
  CHECK(p && p->constraint_matrix_num_nonzeros == 5);
  const double x0[2] = {1, 2}, y0[3] = {1, -1, 0};
  set_start_values(p, x0, y0);
  set_start_values(p, NULL, y0);
  lp_problem_free(p);

  const int rp[4] = {0, 2, 4, 5}, ci[5] = {0, 1, 0, 1, 0};
  const double v[5] = {1, 1, 1, -1, 1};
  d.fmt = matrix_csr;
  d.data.csr.nnz = 5;
  d.data.csr.row_ptr = rp;
  d.data.csr.col_ind = ci;
  d.data.csr.vals = v;
  p = create_lp_problem(c, &d, NULL, NULL, NULL, NULL, NULL); /* defaults */
  CHECK(p);
  lp_problem_free(p);

  const int cp[3] = {0, 3, 5}, ri[5] = {0, 1, 2, 0, 1};
  const double cv[5] = {1, 1, 1, 1, -1};
  d.fmt = matrix_csc;
  d.zero_tolerance = 0.0;
  d.data.csc.nnz = 5;
  d.data.csc.col_ptr = cp;
  d.data.csc.row_ind = ri;
  d.data.csc.vals = cv;
  p = create_lp_problem(c, &d, cl, cu, vl, vu, NULL);
  CHECK(p && p->constraint_matrix_num_nonzeros == 5);
  lp_problem_free(p);

  const int qr[5] = {2, 0, 1, 0, 1}, qc[5] = {0, 0, 0, 1, 1};
  d.fmt = matrix_coo;
  d.data.coo.nnz = 5;
  d.data.coo.row_ind = qr;
  d.data.coo.col_ind = qc;
  d.data.coo.vals = cv;
  p = create_lp_problem(c, &d, cl, cu, vl, vu, NULL);
  CHECK(p && p->constraint_matrix_num_nonzeros == 5);
  lp_problem_free(p);

  /* error paths: out-of-range COO index, NULL descriptor */
  const int bad[5] = {2, 0, 7, 0, 1};
  d.data.coo.row_ind = bad;
  CHECK(create_lp_problem(c, &d, cl, cu, vl, vu, NULL) == NULL);
  CHECK(create_lp_problem(c, NULL, cl, cu, vl, vu, NULL) == NULL);
  cupdlpx_result_free(NULL);
  lp_problem_free(NULL);
  return 0;
}

static int planners(void) {
  const int rules[3] = {0, 1, 2};
  for (int k = 0; k < 3; ++k) {
    dpdlp_gen_spec_t g;
    memset(&g, 0, sizeof(g));
    g.m = 700;
    g.n = 1500;
    g.nnz = rules[k] == 2 ? 0 : 9000;
    g.row_rule = rules[k];
    g.zipf_s = 1.5;
    g.zipf_max = 50000;
    g.poisson_lambda = 8.0;
    g.pm1_frac = 0.1;
    g.seed = 17 + (uint64_t)k;
    double opt = 0.0;
    lp_problem_t *p = dpdlp_generate(&g, &opt);
    CHECK(p && isfinite(opt));
    int m = p->num_constraints, n = p->num_variables, nnz = p->constraint_matrix_num_nonzeros;
    const int *rp = p->constraint_matrix_row_pointers, *ci = p->constraint_matrix_col_indices;
    int *trp = malloc((size_t)(n + 1) * sizeof(int)), *tci = malloc((size_t)nnz * sizeof(int));
    double *tv = malloc((size_t)nnz * sizeof(double));
    int *blk = malloc((size_t)(m + n + 2) * sizeof(int));
    CHECK(trp && tci && tv && blk);
    CHECK(dpdlp_host_transpose(m, n, rp, ci, p->constraint_matrix_values, trp, tci, tv) == 0);
    CHECK(trp[n] == nnz);
    const int caps[3] = {1024, 2048, 4096};
    for (int q = 0; q < 3; ++q) {
      int nb = dpdlp_plan_blocks(m, rp, caps[q], blk);
      CHECK(nb >= 1 && blk[0] == 0 && blk[nb] == m);
      nb = dpdlp_plan_blocks(n, trp, caps[q], blk);
      CHECK(nb >= 1 && blk[nb] == n);
    }
    for (int P = 2; P <= 8; P *= 2) {
      int nb = dpdlp_plan_blocks_panels(m, n, rp, ci, 1024, P, DPDLP_PANEL_RMAX, blk);
      CHECK(nb >= 1 && blk[nb] == m);
      int *toff = malloc(((size_t)nb * P + 1) * sizeof(int));
      unsigned short *rofs = malloc((size_t)P * ((size_t)m + nb) * sizeof(unsigned short));
      int *perm = malloc((size_t)nnz * sizeof(int)), *pci = malloc((size_t)nnz * sizeof(int));
      CHECK(toff && rofs && perm && pci);
      CHECK(dpdlp_build_panels(m, n, rp, ci, blk, nb, P, toff, rofs, perm, pci) == 0);
      CHECK(toff[(size_t)nb * P] == nnz);
      /* A''s panel plan from A (no transpose) equals the plan on A' */
      int *bt = malloc((size_t)(n + 2) * sizeof(int)), *bT = malloc((size_t)(n + 2) * sizeof(int));
      CHECK(bt && bT);
      int na = dpdlp_plan_blocks_panels(n, m, trp, tci, 1024, P, DPDLP_PANEL_RMAX, bt);
      int nT = dpdlp_plan_blocks_panels_T(m, n, rp, ci, 1024, P, DPDLP_PANEL_RMAX, bT);
      CHECK(na == nT && memcmp(bt, bT, (size_t)(na + 1) * sizeof(int)) == 0);
      free(bt);
      free(bT);
      free(toff);
      free(rofs);
      free(perm);
      free(pci);
    }
    /* the oracle solve of the small LP (a few hundred iterations) */
    pdhg_parameters_t prm;
    orc_default_parameters(&prm);
    prm.presolve = false;
    prm.termination_criteria.iteration_limit = 400;
    cupdlpx_result_t *r = orc_solve(p, &prm, NULL);
    CHECK(r && r->total_count <= 400);
    orc_result_free(r);
    free(trp);
    free(tci);
    free(tv);
    free(blk);
    lp_problem_free(p);
  }
  return 0;
}

int main(int argc, char **argv) {
  /* argv: fixtures expected to parse, then "--" and files expected to fail;
   * a first argument "--mps-only" reads the files and stops there */
  int expect_fail = 0, mps_only = argc > 1 && strcmp(argv[1], "--mps-only") == 0;
  for (int i = 1 + mps_only; i < argc; ++i) {
    if (strcmp(argv[i], "--") == 0) {
      expect_fail = 1;
      continue;
    }
    lp_problem_t *p = read_mps_file(argv[i]);
    if (expect_fail) {
      CHECK(p == NULL);
    } else {
      CHECK(p != NULL);
      lp_problem_free(p);
    }
  }
  if (!mps_only && (layouts() || planners())) return 1;
  printf("harness ok\n");
  return 0;
}
