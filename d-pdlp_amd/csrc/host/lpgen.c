/*
 * Seeded synthetic LPs, feasible by construction with a known optimum.
 *
 * Rows: uniform positions (rule 0), Zipf(s) lengths clipped to [1, zmax] and
 * rescaled so the total is exactly nnz (rule 1), or Poisson(lambda)+1 lengths
 * (rule 2).  Column indices are uniform and distinct within a row (sorted);
 * values are N(0,1) with a fraction replaced by +-1.
 *
 * Optimum: x*_j = 0 w.p. 0.3, = 10 w.p. 0.1 when u_j is finite, else U[0,10];
 * l_v = 0, u_v = 10 (+inf for 20% of columns); z*_j = |N| at x*_j = 0,
 * -|N| at x*_j = 10, else 0.  Rows: 30% equalities l = u = (Ax*)_i with
 * y*_i ~ N(0,1); 20% active >= rows l = (Ax*)_i, u = +inf, y*_i = |N|; 50%
 * slack >= rows l = (Ax*)_i - U[0,1], u = +inf, y*_i = 0.  c = A'y* + z*.
 * (x*, y*, z*) satisfies the KKT conditions, so c'x* is the optimal value.
 *
 * Random numbers: xoshiro256** (Blackman & Vigna) seeded with SplitMix64;
 * normals by the Marsaglia polar method.
 */
#include "cupdlpx.h"
#include "dpdlp_ext.h"
#include "dpdlp_host.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t s[4];
  int has;
  double saved;
} xr_t;

static uint64_t splitmix(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static void xr_seed(xr_t *r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix(&x);
  r->has = 0;
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t xr_next(xr_t *r) {
  uint64_t *s = r->s;
  uint64_t res = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return res;
}

static inline double xr_u01(xr_t *r) { return (double)(xr_next(r) >> 11) * 0x1p-53; }

static inline uint32_t xr_below(xr_t *r, uint32_t k) {
  return (uint32_t)(((xr_next(r) >> 32) * (uint64_t)k) >> 32);
}

static double xr_normal(xr_t *r) {
  if (r->has) {
    r->has = 0;
    return r->saved;
  }
  double u, v, q;
  do {
    u = 2.0 * xr_u01(r) - 1.0;
    v = 2.0 * xr_u01(r) - 1.0;
    q = u * u + v * v;
  } while (q >= 1.0 || q == 0.0);
  double f = sqrt(-2.0 * log(q) / q);
  r->saved = u * f;
  r->has = 1;
  return v * f;
}

static int xr_poisson(xr_t *r, double lam) {
  /* Knuth's product method; lambda here is small (<= ~30) */
  double L = exp(-lam), p = 1.0;
  int k = 0;
  do {
    ++k;
    p *= xr_u01(r);
  } while (p > L);
  return k - 1;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

typedef struct {
  double frac;
  int idx;
} fr_t;

static int cmp_frac_desc(const void *a, const void *b) {
  const fr_t *x = (const fr_t *)a, *y = (const fr_t *)b;
  if (x->frac != y->frac) return x->frac < y->frac ? 1 : -1;
  return x->idx - y->idx;
}

/* distinct sorted columns of one row into out[0..L) */
static void draw_row(xr_t *r, int n, int L, int *out, unsigned char *mark) {
  if (L <= 0) return;
  if ((long long)L * 4 > n) {
    /* dense-ish: partial selection via marks */
    int got = 0;
    while (got < L) {
      int j = (int)xr_below(r, (uint32_t)n);
      if (!mark[j]) {
        mark[j] = 1;
        out[got++] = j;
      }
    }
    for (int k = 0; k < L; ++k) mark[out[k]] = 0;
    qsort(out, (size_t)L, sizeof(int), cmp_int);
    return;
  }
  for (int k = 0; k < L; ++k) out[k] = (int)xr_below(r, (uint32_t)n);
  for (;;) {
    qsort(out, (size_t)L, sizeof(int), cmp_int);
    int u = 0;
    for (int k = 0; k < L; ++k)
      if (u == 0 || out[k] != out[u - 1]) out[u++] = out[k];
    if (u == L) return;
    for (int k = u; k < L; ++k) out[k] = (int)xr_below(r, (uint32_t)n);
  }
}

lp_problem_t *dpdlp_generate(const dpdlp_gen_spec_t *g, double *opt_obj) {
  if (!g || g->m <= 0 || g->n <= 0) return NULL;
  int m = g->m, n = g->n;
  xr_t R;
  xr_seed(&R, g->seed);
  long long *len = (long long *)calloc((size_t)m, sizeof(long long));
  if (!len) return NULL;
  long long total = 0;
  if (g->row_rule == 0) {
    for (long long k = 0; k < g->nnz; ++k) len[xr_below(&R, (uint32_t)m)]++;
  } else if (g->row_rule == 1) {
    int zmax = g->zipf_max > 0 ? g->zipf_max : 50000;
    double s = g->zipf_s > 0 ? g->zipf_s : 1.5;
    double *cdf = (double *)malloc((size_t)zmax * sizeof(double));
    double *raw = (double *)malloc((size_t)m * sizeof(double));
    if (!cdf || !raw) {
      free(cdf);
      free(raw);
      free(len);
      return NULL;
    }
    double acc = 0.0;
    for (int k = 1; k <= zmax; ++k) {
      acc += pow((double)k, -s);
      cdf[k - 1] = acc;
    }
    for (int i = 0; i < m; ++i) {
      double u = xr_u01(&R) * acc;
      int lo = 0, hi = zmax - 1;
      while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (cdf[mid] < u)
          lo = mid + 1;
        else
          hi = mid;
      }
      raw[i] = (double)(lo + 1);
    }
    /* largest f with sum max(1, floor(f L)) <= nnz, by bisection */
    double flo = 0.0, fhi = 1.0;
    for (;;) {
      long long t = 0;
      for (int i = 0; i < m; ++i) {
        long long v = (long long)floor(fhi * raw[i]);
        t += v < 1 ? 1 : v;
      }
      if (t > g->nnz || fhi > 1e9) break;
      fhi *= 2.0;
    }
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (flo + fhi);
      long long t = 0;
      for (int i = 0; i < m; ++i) {
        long long v = (long long)floor(mid * raw[i]);
        t += v < 1 ? 1 : v;
      }
      if (t <= g->nnz)
        flo = mid;
      else
        fhi = mid;
    }
    fr_t *fr = (fr_t *)malloc((size_t)m * sizeof(fr_t));
    if (!fr) {
      free(cdf);
      free(raw);
      free(len);
      return NULL;
    }
    for (int i = 0; i < m; ++i) {
      double v = flo * raw[i];
      long long fl = (long long)floor(v);
      len[i] = fl < 1 ? 1 : fl;
      total += len[i];
      fr[i].frac = fl < 1 ? -1.0 : v - (double)fl;
      fr[i].idx = i;
    }
    long long rem = g->nnz - total;
    if (rem > 0) {
      qsort(fr, (size_t)m, sizeof(fr_t), cmp_frac_desc);
      for (long long k = 0; k < rem && k < m; ++k) len[fr[k].idx]++;
    }
    free(fr);
    free(cdf);
    free(raw);
  } else {
    double lam = g->poisson_lambda > 0 ? g->poisson_lambda : 8.0;
    for (int i = 0; i < m; ++i) len[i] = xr_poisson(&R, lam) + 1;
  }
  total = 0;
  for (int i = 0; i < m; ++i) {
    if (len[i] > n) len[i] = n;
    total += len[i];
  }
  if (total > 2147483647LL) {
    free(len);
    return NULL;
  }
  int nnz = (int)total;
  lp_problem_t *p = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
  int *rp = (int *)malloc(((size_t)m + 1) * sizeof(int));
  int *ci = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
  double *v = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
  unsigned char *mark = (unsigned char *)calloc((size_t)n, 1);
  double *xs = (double *)malloc((size_t)n * sizeof(double));
  double *zs = (double *)malloc((size_t)n * sizeof(double));
  double *ys = (double *)malloc((size_t)m * sizeof(double));
  double *ax = (double *)malloc((size_t)m * sizeof(double));
  if (!p || !rp || !ci || !v || !mark || !xs || !zs || !ys || !ax) {
    free(p); free(rp); free(ci); free(v); free(mark); free(xs); free(zs); free(ys); free(ax);
    free(len);
    return NULL;
  }
  rp[0] = 0;
  for (int i = 0; i < m; ++i) rp[i + 1] = rp[i] + (int)len[i];
  free(len);
  for (int i = 0; i < m; ++i) {
    draw_row(&R, n, rp[i + 1] - rp[i], ci + rp[i], mark);
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (g->pm1_frac > 0 && xr_u01(&R) < g->pm1_frac)
        v[k] = (xr_u01(&R) < 0.5) ? -1.0 : 1.0;
      else
        v[k] = xr_normal(&R);
    }
  }
  free(mark);
  p->num_variables = n;
  p->num_constraints = m;
  p->constraint_matrix_row_pointers = rp;
  p->constraint_matrix_col_indices = ci;
  p->constraint_matrix_values = v;
  p->constraint_matrix_num_nonzeros = nnz;
  p->variable_lower_bound = (double *)malloc((size_t)n * sizeof(double));
  p->variable_upper_bound = (double *)malloc((size_t)n * sizeof(double));
  p->objective_vector = (double *)malloc((size_t)n * sizeof(double));
  p->constraint_lower_bound = (double *)malloc((size_t)m * sizeof(double));
  p->constraint_upper_bound = (double *)malloc((size_t)m * sizeof(double));
  if (!p->variable_lower_bound || !p->variable_upper_bound || !p->objective_vector ||
      !p->constraint_lower_bound || !p->constraint_upper_bound) {
    free(xs); free(zs); free(ys); free(ax);
    lp_problem_free(p);
    return NULL;
  }
  for (int j = 0; j < n; ++j) {
    int uinf = xr_u01(&R) < 0.2;
    double r = xr_u01(&R);
    p->variable_lower_bound[j] = 0.0;
    p->variable_upper_bound[j] = uinf ? INFINITY : 10.0;
    if (r < 0.3) {
      xs[j] = 0.0;
      zs[j] = fabs(xr_normal(&R));
    } else if (!uinf && r < 0.4) {
      xs[j] = 10.0;
      zs[j] = -fabs(xr_normal(&R));
    } else {
      xs[j] = 10.0 * xr_u01(&R);
      zs[j] = 0.0;
    }
  }
  for (int i = 0; i < m; ++i) {
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * xs[ci[k]];
    ax[i] = s;
  }
  for (int i = 0; i < m; ++i) {
    double r = xr_u01(&R);
    if (r < 0.3) {
      p->constraint_lower_bound[i] = ax[i];
      p->constraint_upper_bound[i] = ax[i];
      ys[i] = xr_normal(&R);
    } else if (r < 0.5) {
      p->constraint_lower_bound[i] = ax[i];
      p->constraint_upper_bound[i] = INFINITY;
      ys[i] = fabs(xr_normal(&R));
    } else {
      p->constraint_lower_bound[i] = ax[i] - xr_u01(&R);
      p->constraint_upper_bound[i] = INFINITY;
      ys[i] = 0.0;
    }
  }
  for (int j = 0; j < n; ++j) p->objective_vector[j] = zs[j];
  for (int i = 0; i < m; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) p->objective_vector[ci[k]] += v[k] * ys[i];
  p->objective_constant = 0.0;
  if (opt_obj) {
    double o = 0.0;
    for (int j = 0; j < n; ++j) o += p->objective_vector[j] * xs[j];
    *opt_obj = o;
  }
  free(xs);
  free(zs);
  free(ys);
  free(ax);
  return p;
}

/* Batch workload member k: m log-uniform in [m_lo, m_hi], n = m * U[1.5, 3],
 * Poisson(8)+1 nonzeros per row, LP seed base_seed + k.  The sizes come
 * from their own xoshiro stream seeded with (base_seed + k) ^ a constant, so
 * member k is the same LP whichever members are drawn with it. */
void dpdlp_batch_member_spec(int k, uint64_t base_seed, int m_lo, int m_hi,
                             dpdlp_gen_spec_t *out) {
  xr_t R;
  xr_seed(&R, (base_seed + (uint64_t)k) ^ 0x5eed5eed5eed5eedull);
  double lo = log((double)m_lo), hi = log((double)m_hi);
  int m = (int)llround(exp(lo + (hi - lo) * xr_u01(&R)));
  if (m < 1) m = 1;
  int n = (int)llround((double)m * (1.5 + 1.5 * xr_u01(&R)));
  if (n < 1) n = 1;
  memset(out, 0, sizeof(*out));
  out->m = m;
  out->n = n;
  out->nnz = 0;
  out->row_rule = 2;
  out->poisson_lambda = 8.0;
  out->zipf_s = 1.5;
  out->zipf_max = 50000;
  out->pm1_frac = 0.0;
  out->seed = base_seed + (uint64_t)k;
}
