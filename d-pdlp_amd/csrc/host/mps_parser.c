/*
 * Free-format MPS reader, read_mps_file() of include/mps_parser.h.
 *
 * Follows the reference's rules (src/mps_parser.c:365-777) so a file yields
 * the same lp_problem_t:
 *   - lines starting with '*', CR or LF are skipped; tokens split on space,
 *     tab, CR, LF; at most 6 tokens per line;
 *   - a line with one token starting with a letter is a section header
 *     (ROWS COLUMNS RHS RANGES BOUNDS OBJSENSE ENDATA; anything else ends the
 *     current section);
 *   - the first N row is the objective (the first row if there is no N row);
 *     other N rows are dropped; L -> [-inf,0], G -> [0,inf], E -> [0,0];
 *   - COLUMNS: odd token count starts a column, even continues it; 'MARKER'
 *     lines are skipped; entries of unknown rows are dropped; duplicates kept;
 *   - RHS / RANGES: token 0 is the set name; objective RHS gives c0 = -value;
 *   - BOUNDS: LO UP FX FR MI PL BV, unknown columns / types ignored;
 *     default bounds [0, +inf);
 *   - a MAX objective sense (only recognised on a data line of two or more
 *     tokens inside OBJSENSE, because a lone "MAX" is taken as a header, as in
 *     the reference) negates c and c0;
 *   - CSR keeps file order within each row; ".gz" names are read with zlib.
 * Differences: a file that cannot be opened returns NULL (the reference
 * dereferences NULL); lines longer than 4095 bytes are read whole instead of
 * ending the file.
 */
#define _POSIX_C_SOURCE 200809L
#include "mps_parser.h"
#include "cupdlpx.h"

#include <ctype.h>
#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

/* ---- string -> index map (open addressing, FNV-1a).  A 32-byte slot holds
 * the hash, the index and the key itself when it is shorter than 16 bytes
 * (MPS names usually are), so a lookup normally touches one cache line. ---- */
typedef struct {
  uint64_t h;
  int v;   /* -1: empty */
  int len;
  union {
    char in[16];
    char *out;
  } k;
} sslot_t;

typedef struct {
  sslot_t *slot;
  size_t cap, size;
} smap_t;

static uint64_t fnv1a_len(const char *s, int *len) {
  uint64_t h = 1469598103934665603ull;
  const char *p = s;
  for (; *p; ++p) {
    h ^= (unsigned char)*p;
    h *= 1099511628211ull;
  }
  *len = (int)(p - s);
  return h;
}

static sslot_t *slots_new(size_t cap) {
  sslot_t *s = (sslot_t *)malloc(cap * sizeof(sslot_t));
  if (s)
    for (size_t i = 0; i < cap; ++i) s[i].v = -1;
  return s;
}

static int smap_init(smap_t *m) {
  m->cap = 1024;
  m->size = 0;
  m->slot = slots_new(m->cap);
  return m->slot ? 0 : -1;
}

static void smap_free(smap_t *m) {
  if (m->slot)
    for (size_t i = 0; i < m->cap; ++i)
      if (m->slot[i].v >= 0 && m->slot[i].len >= 16) free(m->slot[i].k.out);
  free(m->slot);
  memset(m, 0, sizeof(*m));
}

static inline int slot_is(const sslot_t *e, const char *k, int len, uint64_t h) {
  if (e->h != h || e->len != len) return 0;
  return len < 16 ? memcmp(e->k.in, k, (size_t)len) == 0 : strcmp(e->k.out, k) == 0;
}

static int smap_get_h(const smap_t *m, const char *k, int len, uint64_t h) {
  size_t i = (size_t)(h & (m->cap - 1));
  while (m->slot[i].v >= 0) {
    if (slot_is(&m->slot[i], k, len, h)) return m->slot[i].v;
    i = (i + 1) & (m->cap - 1);
  }
  return -1;
}

static int smap_get(const smap_t *m, const char *k) {
  int len;
  uint64_t h = fnv1a_len(k, &len);
  return smap_get_h(m, k, len, h);
}

static int smap_grow(smap_t *m) {
  size_t nc = m->cap * 2;
  sslot_t *ns = slots_new(nc);
  if (!ns) return -1;
  for (size_t i = 0; i < m->cap; ++i) {
    if (m->slot[i].v < 0) continue;
    size_t j = (size_t)(m->slot[i].h & (nc - 1));
    while (ns[j].v >= 0) j = (j + 1) & (nc - 1);
    ns[j] = m->slot[i];
  }
  free(m->slot);
  m->slot = ns;
  m->cap = nc;
  return 0;
}

/* returns the existing index, or inserts with index = size */
static int smap_put(smap_t *m, const char *k) {
  int len;
  uint64_t h = fnv1a_len(k, &len);
  int v = smap_get_h(m, k, len, h);
  if (v >= 0) return v;
  if ((m->size + 1) * 2 > m->cap && smap_grow(m)) return -1;
  size_t i = (size_t)(h & (m->cap - 1));
  while (m->slot[i].v >= 0) i = (i + 1) & (m->cap - 1);
  sslot_t *e = &m->slot[i];
  if (len < 16) {
    memcpy(e->k.in, k, (size_t)len);
  } else {
    e->k.out = strdup(k);
    if (!e->k.out) return -1;
  }
  e->h = h;
  e->len = len;
  e->v = (int)m->size++;
  return e->v;
}

/* Decimal to double, as atof: numbers with at most 15 significant digits
 * and a decimal exponent within 10^±22 are converted exactly with one
 * correctly rounded multiply or divide (both operands are exact doubles,
 * Clinger's fast path); everything else goes to strtod. */
static double mps_atof(const char *s) {
  static const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                 1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  const char *p = s;
  int neg = 0;
  if (*p == '+' || *p == '-') neg = (*p++ == '-');
  uint64_t mant = 0;
  int digits = 0, scale = 0, any = 0;
  while (*p >= '0' && *p <= '9') {
    any = 1;
    if (mant || *p != '0') {
      if (++digits > 15) return atof(s);
      mant = mant * 10 + (uint64_t)(*p - '0');
    }
    ++p;
  }
  if (*p == '.') {
    ++p;
    while (*p >= '0' && *p <= '9') {
      any = 1;
      if (mant || *p != '0') {
        if (++digits > 15) return atof(s);
        mant = mant * 10 + (uint64_t)(*p - '0');
      }
      --scale;
      ++p;
    }
  }
  if (!any) return atof(s);
  if (*p == 'e' || *p == 'E') {
    ++p;
    int eneg = 0, e = 0, edig = 0;
    if (*p == '+' || *p == '-') eneg = (*p++ == '-');
    while (*p >= '0' && *p <= '9') {
      if (e < 10000) e = e * 10 + (*p - '0');
      ++p;
      ++edig;
    }
    if (!edig) return atof(s);
    scale += eneg ? -e : e;
  }
  if (*p != '\0') return atof(s);
  double v = (double)mant; /* exact: at most 15 digits */
  if (mant == 0) return neg ? -0.0 : 0.0;
  if (scale > 22 || scale < -22) return atof(s);
  v = scale >= 0 ? v * p10[scale] : v / p10[-scale];
  return neg ? -v : v;
}

/* ---- line source: plain stdio or zlib, read in 4 MiB chunks ---- */
typedef struct {
  gzFile gz;
  FILE *f;
  char *buf;
  size_t cap, beg, end; /* unread bytes are buf[beg, end) */
  int eof;
} src_t;

static size_t src_fill(src_t *s, char *dst, size_t n) {
  if (s->gz) {
    int r = gzread(s->gz, dst, (unsigned)(n > (1u << 30) ? (1u << 30) : n));
    return r > 0 ? (size_t)r : 0;
  }
  return fread(dst, 1, n, s->f);
}

/* next line, NUL-terminated in place without its newline; NULL at the end */
static char *src_line(src_t *s) {
  for (;;) {
    char *nl = s->beg < s->end ? (char *)memchr(s->buf + s->beg, '\n', s->end - s->beg) : NULL;
    if (nl) {
      char *line = s->buf + s->beg;
      *nl = '\0';
      s->beg = (size_t)(nl - s->buf) + 1;
      return line;
    }
    if (s->eof) {
      if (s->beg < s->end) { /* last line without a newline */
        char *line = s->buf + s->beg;
        s->buf[s->end] = '\0';
        s->beg = s->end;
        return line;
      }
      return NULL;
    }
    /* keep the partial line, make room, read more */
    size_t rest = s->end - s->beg;
    if (s->beg > 0) {
      memmove(s->buf, s->buf + s->beg, rest);
      s->beg = 0;
      s->end = rest;
    }
    if (s->cap - s->end < (1u << 20) + 1) {
      size_t nc = s->cap ? s->cap * 2 : (4u << 20);
      char *nb = (char *)realloc(s->buf, nc);
      if (!nb) return NULL;
      s->buf = nb;
      s->cap = nc;
    }
    size_t got = src_fill(s, s->buf + s->end, s->cap - s->end - 1);
    if (got == 0) s->eof = 1;
    s->end += got;
  }
}

typedef struct {
  smap_t rows, cols;
  char *types;       /* per constraint index */
  size_t types_cap;
  char **pend_names; /* ROWS section, before finalisation */
  char *pend_types;
  size_t npend, pend_cap;
  double *obj, *lb, *ub;
  size_t col_cap;
  double *clb, *cub;
  int *ci_r, *ci_c;
  double *cv;
  size_t nz, nz_cap;
  char *obj_row;
  char *cur_col;
  int cur_idx; /* index of cur_col (-1: not looked up yet) */
  double obj_const;
  bool maximize;
  int err;
} mps_t;

static int ensure_cols(mps_t *P) {
  if (P->cols.size < P->col_cap) return 0;
  size_t nc = P->col_cap ? P->col_cap * 2 : 256;
  double *o = (double *)realloc(P->obj, nc * sizeof(double));
  if (!o) return -1;
  P->obj = o;
  double *l = (double *)realloc(P->lb, nc * sizeof(double));
  if (!l) return -1;
  P->lb = l;
  double *u = (double *)realloc(P->ub, nc * sizeof(double));
  if (!u) return -1;
  P->ub = u;
  for (size_t i = P->col_cap; i < nc; ++i) {
    P->obj[i] = 0.0;
    P->lb[i] = 0.0;
    P->ub[i] = INFINITY;
  }
  P->col_cap = nc;
  return 0;
}

static int add_entry(mps_t *P, int r, int c, double v) {
  if (P->nz >= P->nz_cap) {
    size_t nc = P->nz_cap ? P->nz_cap * 2 : 1024;
    int *a = (int *)realloc(P->ci_r, nc * sizeof(int));
    if (!a) return -1;
    P->ci_r = a;
    int *b = (int *)realloc(P->ci_c, nc * sizeof(int));
    if (!b) return -1;
    P->ci_c = b;
    double *d = (double *)realloc(P->cv, nc * sizeof(double));
    if (!d) return -1;
    P->cv = d;
    P->nz_cap = nc;
  }
  P->ci_r[P->nz] = r;
  P->ci_c[P->nz] = c;
  P->cv[P->nz] = v;
  P->nz++;
  return 0;
}

/* mps_parser.c:511-594 */
static int finalize_rows(mps_t *P) {
  long obj = -1;
  for (size_t i = 0; i < P->npend; ++i)
    if (P->pend_types[i] == 'N') {
      obj = (long)i;
      break;
    }
  if (obj < 0 && P->npend > 0) obj = 0;
  if (obj >= 0) {
    P->obj_row = strdup(P->pend_names[obj]);
    if (!P->obj_row) return -1;
  }
  for (size_t i = 0; i < P->npend; ++i) {
    if ((long)i == obj) continue;
    char t = P->pend_types[i];
    if (t != 'E' && t != 'L' && t != 'G') continue;
    size_t at = P->rows.size;
    if (at >= P->types_cap) {
      size_t nc = P->types_cap ? P->types_cap * 2 : 64;
      char *nt = (char *)realloc(P->types, nc);
      if (!nt) return -1;
      P->types = nt;
      P->types_cap = nc;
    }
    if (smap_put(&P->rows, P->pend_names[i]) < 0) return -1;
    P->types[at] = t;
  }
  size_t m = P->rows.size;
  if (m > 0) {
    P->clb = (double *)malloc(m * sizeof(double));
    P->cub = (double *)malloc(m * sizeof(double));
    if (!P->clb || !P->cub) return -1;
    for (size_t i = 0; i < m; ++i) {
      char t = P->types[i];
      P->clb[i] = (t == 'L') ? -INFINITY : 0.0;
      P->cub[i] = (t == 'G') ? INFINITY : 0.0;
    }
  }
  return 0;
}

static int rows_line(mps_t *P, char **tk, int nt) {
  if (nt < 2) return 0;
  if (P->npend >= P->pend_cap) {
    size_t nc = P->pend_cap ? P->pend_cap * 2 : 64;
    char **nn = (char **)realloc(P->pend_names, nc * sizeof(char *));
    if (!nn) return -1;
    P->pend_names = nn;
    char *ty = (char *)realloc(P->pend_types, nc);
    if (!ty) return -1;
    P->pend_types = ty;
    P->pend_cap = nc;
  }
  P->pend_types[P->npend] = tk[0][0];
  P->pend_names[P->npend] = strdup(tk[1]);
  if (!P->pend_names[P->npend]) return -1;
  P->npend++;
  return 0;
}

/* mps_parser.c:596-650 */
static int columns_line(mps_t *P, char **tk, int nt) {
  if (nt < 2) return 0;
  if (strcmp(tk[1], "'MARKER'") == 0) return 0;
  int first;
  if (nt % 2 != 0) {
    if (!P->cur_col || strcmp(P->cur_col, tk[0]) != 0) {
      free(P->cur_col);
      P->cur_col = strdup(tk[0]);
      if (!P->cur_col) return -1;
      P->cur_idx = -1;
    }
    first = 1;
  } else {
    if (!P->cur_col) {
      fprintf(stderr, "ERROR: Column data found before any column name was defined.\n");
      return -1;
    }
    first = 0;
  }
  if (ensure_cols(P)) return -1;
  if (P->cur_idx < 0) {
    P->cur_idx = smap_put(&P->cols, P->cur_col);
    if (P->cur_idx < 0) return -1;
  }
  int col = P->cur_idx;
  for (int i = first; i + 1 < nt; i += 2) {
    double v = mps_atof(tk[i + 1]);
    if (P->obj_row && strcmp(tk[i], P->obj_row) == 0) {
      P->obj[col] += v;
    } else {
      int r = smap_get(&P->rows, tk[i]);
      if (r >= 0 && add_entry(P, r, col, v)) return -1;
    }
  }
  return 0;
}

/* mps_parser.c:652-678 */
static void rhs_line(mps_t *P, char **tk, int nt) {
  for (int i = 1; i + 1 < nt; i += 2) {
    double v = mps_atof(tk[i + 1]);
    if (P->obj_row && strcmp(tk[i], P->obj_row) == 0) {
      P->obj_const = -v;
      continue;
    }
    int r = smap_get(&P->rows, tk[i]);
    if (r < 0) continue;
    char t = P->types[r];
    if (t == 'L') {
      P->cub[r] = v;
    } else if (t == 'G') {
      P->clb[r] = v;
    } else {
      P->clb[r] = v;
      P->cub[r] = v;
    }
  }
}

/* mps_parser.c:680-707 */
static void ranges_line(mps_t *P, char **tk, int nt) {
  for (int i = 1; i + 1 < nt; i += 2) {
    double R = atof(tk[i + 1]);
    int r = smap_get(&P->rows, tk[i]);
    if (r < 0) continue;
    char t = P->types[r];
    double rhs = (t == 'L') ? P->cub[r] : P->clb[r];
    if (t == 'G') {
      P->cub[r] = rhs + fabs(R);
    } else if (t == 'L') {
      P->clb[r] = rhs - fabs(R);
    } else if (t == 'E') {
      if (R >= 0)
        P->cub[r] = rhs + R;
      else
        P->clb[r] = rhs + R;
    }
  }
}

/* mps_parser.c:709-742 */
static void bounds_line(mps_t *P, char **tk, int nt) {
  if (nt < 3) return;
  int col = smap_get(&P->cols, tk[2]);
  if (col < 0) return;
  double v = nt > 3 ? atof(tk[3]) : 0.0;
  const char *t = tk[0];
  if (!strcmp(t, "LO")) {
    P->lb[col] = v;
  } else if (!strcmp(t, "UP")) {
    P->ub[col] = v;
  } else if (!strcmp(t, "FX")) {
    P->lb[col] = v;
    P->ub[col] = v;
  } else if (!strcmp(t, "FR")) {
    P->lb[col] = -INFINITY;
    P->ub[col] = INFINITY;
  } else if (!strcmp(t, "MI")) {
    P->lb[col] = -INFINITY;
  } else if (!strcmp(t, "PL")) {
    P->ub[col] = INFINITY;
  } else if (!strcmp(t, "BV")) {
    P->lb[col] = 0.0;
    P->ub[col] = 1.0;
  }
}

static void mps_free(mps_t *P) {
  smap_free(&P->rows);
  smap_free(&P->cols);
  free(P->types);
  for (size_t i = 0; i < P->npend; ++i) free(P->pend_names[i]);
  free(P->pend_names);
  free(P->pend_types);
  free(P->obj);
  free(P->lb);
  free(P->ub);
  free(P->clb);
  free(P->cub);
  free(P->ci_r);
  free(P->ci_c);
  free(P->cv);
  free(P->obj_row);
  free(P->cur_col);
}

enum { S_NONE, S_ROWS, S_COLUMNS, S_RHS, S_RANGES, S_BOUNDS, S_OBJSENSE, S_END };

lp_problem_t *read_mps_file(const char *filename) {
  if (!filename) return NULL;
  src_t src = {0};
  size_t fl = strlen(filename);
  if (fl > 3 && strcmp(filename + fl - 3, ".gz") == 0)
    src.gz = gzopen(filename, "rb");
  else
    src.f = fopen(filename, "r");
  if (!src.gz && !src.f) {
    fprintf(stderr, "ERROR: Could not open file %s\n", filename);
    return NULL;
  }
  mps_t P;
  memset(&P, 0, sizeof(P));
  if (smap_init(&P.rows) || smap_init(&P.cols)) P.err = 1;
  int sec = S_NONE;
  bool rows_done = false;
  char *line;
  while (!P.err && (line = src_line(&src)) != NULL) {
    if (line[0] == '*' || line[0] == '\0' || line[0] == '\r') continue;
    char *tk[6];
    int nt = 0;
    char *save = NULL;
    for (char *t = strtok_r(line, " \t\n\r", &save); t && nt < 6;
         t = strtok_r(NULL, " \t\n\r", &save))
      tk[nt++] = t;
    if (nt == 0) continue;
    if (nt == 1 && isalpha((unsigned char)tk[0][0])) {
      int next = S_NONE;
      if (!strcmp(tk[0], "ROWS")) next = S_ROWS;
      else if (!strcmp(tk[0], "COLUMNS")) next = S_COLUMNS;
      else if (!strcmp(tk[0], "RHS")) next = S_RHS;
      else if (!strcmp(tk[0], "RANGES")) next = S_RANGES;
      else if (!strcmp(tk[0], "BOUNDS")) next = S_BOUNDS;
      else if (!strcmp(tk[0], "OBJSENSE")) next = S_OBJSENSE;
      else if (!strcmp(tk[0], "ENDATA")) next = S_END;
      if (sec == S_ROWS && next != S_ROWS && !rows_done) {
        if (finalize_rows(&P)) P.err = 1;
        rows_done = true;
      }
      sec = next;
      if (sec == S_END) break;
      continue;
    }
    switch (sec) {
    case S_OBJSENSE:
      if (!strcmp(tk[0], "MAX") || !strcmp(tk[0], "MAXIMIZE")) P.maximize = true;
      break;
    case S_ROWS:
      if (rows_line(&P, tk, nt)) P.err = 1;
      break;
    case S_COLUMNS:
      if (columns_line(&P, tk, nt)) P.err = 1;
      break;
    case S_RHS:
      rhs_line(&P, tk, nt);
      break;
    case S_RANGES:
      ranges_line(&P, tk, nt);
      break;
    case S_BOUNDS:
      bounds_line(&P, tk, nt);
      break;
    default:
      break;
    }
  }
  if (src.gz) gzclose(src.gz);
  if (src.f) fclose(src.f);
  free(src.buf);
  if (P.err) {
    fprintf(stderr, "ERROR: Failed to parse MPS file.\n");
    mps_free(&P);
    return NULL;
  }
  lp_problem_t *p = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
  if (!p) {
    mps_free(&P);
    return NULL;
  }
  int n = (int)P.cols.size, m = (int)P.rows.size;
  p->num_variables = n;
  p->num_constraints = m;
  p->constraint_matrix_num_nonzeros = (int)P.nz;
  p->objective_constant = P.maximize ? -P.obj_const : P.obj_const;
  p->objective_vector = P.obj;
  p->variable_lower_bound = P.lb;
  p->variable_upper_bound = P.ub;
  p->constraint_lower_bound = P.clb;
  p->constraint_upper_bound = P.cub;
  P.obj = P.lb = P.ub = P.clb = P.cub = NULL;
  if (P.maximize)
    for (int j = 0; j < n; ++j) p->objective_vector[j] *= -1.0;
  /* COO -> CSR keeping file order within rows (mps_parser.c:744-777) */
  p->constraint_matrix_row_pointers = (int *)calloc((size_t)m + 1, sizeof(int));
  p->constraint_matrix_col_indices = (int *)malloc((P.nz ? P.nz : 1) * sizeof(int));
  p->constraint_matrix_values = (double *)malloc((P.nz ? P.nz : 1) * sizeof(double));
  int *pos = (int *)malloc(((size_t)m + 1) * sizeof(int));
  if (!p->constraint_matrix_row_pointers || !p->constraint_matrix_col_indices ||
      !p->constraint_matrix_values || !pos) {
    free(pos);
    lp_problem_free(p);
    mps_free(&P);
    return NULL;
  }
  int *rp = p->constraint_matrix_row_pointers;
  for (size_t k = 0; k < P.nz; ++k) rp[P.ci_r[k] + 1]++;
  for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
  memcpy(pos, rp, ((size_t)m + 1) * sizeof(int));
  for (size_t k = 0; k < P.nz; ++k) {
    int d = pos[P.ci_r[k]]++;
    p->constraint_matrix_col_indices[d] = P.ci_c[k];
    p->constraint_matrix_values[d] = P.cv[k];
  }
  free(pos);
  mps_free(&P);
  return p;
}
