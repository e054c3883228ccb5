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
 *
 * A plain input is read whole into memory; a block-gzipped one (BGZF, every
 * member recording its compressed size) is inflated member by member on the
 * host worker threads; any other gzipped one is inflated by a producer
 * thread while the parser consumes it (see text_open).  The body of
 * a COLUMNS section (the bulk of a file) is parsed by host worker threads:
 * it is cut at line starts into chunks; each chunk's lines are split into
 * tokens, their numbers converted and their row names looked up (the row map
 * is complete and read only by then) into runs of entries, one run per
 * column name change; one thread then walks the chunks in file order,
 * numbering the columns and appending the entries, so the result is the one
 * a line-by-line read gives.
 * DPDLP_HOST_THREADS sets the thread count, DPDLP_MPS_CHUNK the smallest
 * chunk in bytes (default 1 MiB), DPDLP_MPS_WINDOW the window of a streamed
 * input (default 16 MiB); DPDLP_MPS_STREAM=0 reads a gzipped input whole
 * first.
 */
#define _GNU_SOURCE
#include "mps_parser.h"
#include "cupdlpx.h"
#include "dpdlp_host.h"

#include <ctype.h>
#include <math.h>
#include <pthread.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <zlib.h>

enum { S_NONE, S_ROWS, S_COLUMNS, S_RHS, S_RANGES, S_BOUNDS, S_OBJSENSE, S_END };

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

/* capacity for `count` keys at load factor 1/2 (rehashes at most once) */
static int smap_grow_to(smap_t *m, size_t nc) {
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

static int smap_grow(smap_t *m) { return smap_grow_to(m, m->cap * 2); }

static int smap_reserve(smap_t *m, size_t count) {
  size_t nc = m->cap;
  while ((count + 1) * 2 > nc) nc *= 2;
  return nc == m->cap ? 0 : smap_grow_to(m, nc);
}

/* returns the existing index, or inserts with index = size */
static int smap_put_h(smap_t *m, const char *k, int len, uint64_t h) {
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

static int smap_put(smap_t *m, const char *k) {
  int len;
  uint64_t h = fnv1a_len(k, &len);
  return smap_put_h(m, k, len, h);
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

/* ---- the input text.  A plain file is read whole before parsing.  A ".gz"
 * file is inflated by a producer thread into an address range reserved up
 * front (deflate expands at most 1032:1, so the compressed size bounds it),
 * publishing each piece under a mutex, while the parser consumes whole lines
 * behind it: the serial inflate then overlaps the threaded parsing instead
 * of preceding it.  The reader never looks past the published length, and
 * the producer never writes below it. ---- */
typedef struct {
  gzFile gz;
  char *buf;
  size_t cap, len; /* len: published bytes (under mu) */
  int eof, stop;   /* eof: all published, buf[len] = NUL; stop: reader gave up */
  int err;         /* zlib reported corrupt or truncated data (set with eof) */
  pthread_mutex_t mu;
  pthread_cond_t cv;
  pthread_t th;
} gzsrc_t;

#define GZ_PIECE (2u << 20)

static void *gz_produce(void *arg) {
  gzsrc_t *S = (gzsrc_t *)arg;
  size_t len = 0;
  for (;;) {
    pthread_mutex_lock(&S->mu);
    int stop = S->stop;
    pthread_mutex_unlock(&S->mu);
    size_t want = S->cap - 1 - len;
    if (want > GZ_PIECE) want = GZ_PIECE;
    int r = stop || want == 0 ? 0 : gzread(S->gz, S->buf + len, (unsigned)want);
    size_t got = r > 0 ? (size_t)r : 0;
    len += got;
    if (got == 0) S->buf[len] = '\0';
    pthread_mutex_lock(&S->mu);
    S->len = len;
    if (got == 0) S->eof = 1;
    if (r < 0) S->err = 1;
    pthread_cond_broadcast(&S->cv);
    pthread_mutex_unlock(&S->mu);
    if (got == 0) return NULL;
  }
}

typedef struct {
  char *buf;
  size_t len, pos; /* readable text is buf[0, len); unread from pos */
  int eof;         /* len is final */
  gzsrc_t *src;    /* streamed input, or NULL: the text is whole */
  size_t map_len;  /* the reservation of a streamed input */
} text_t;

/* waits until at least want bytes are readable or the input has ended */
static void text_wait(text_t *t, size_t want) {
  gzsrc_t *S = t->src;
  if (!S || t->eof) return;
  pthread_mutex_lock(&S->mu);
  while (S->len < want && !S->eof) pthread_cond_wait(&S->cv, &S->mu);
  t->len = S->len;
  t->eof = S->eof;
  pthread_mutex_unlock(&S->mu);
}

static void text_close(text_t *t) {
  if (t->src) {
    gzsrc_t *S = t->src;
    pthread_mutex_lock(&S->mu);
    S->stop = 1;
    pthread_mutex_unlock(&S->mu);
    pthread_join(S->th, NULL);
    gzclose(S->gz);
    pthread_mutex_destroy(&S->mu);
    pthread_cond_destroy(&S->cv);
    munmap(t->buf, t->map_len);
    free(S);
  } else {
    free(t->buf);
  }
  memset(t, 0, sizeof(*t));
}

/* a ".gz" input streamed by a producer thread; -1 if the reservation or the
 * thread cannot be had (the caller then reads the file whole) */
static int gz_stream_open(const char *filename, text_t *t) {
  struct stat st;
  if (stat(filename, &st) != 0 || st.st_size < 0) return -1;
  size_t cap = (size_t)st.st_size * 1040u + (16u << 20);
  gzsrc_t *S = (gzsrc_t *)calloc(1, sizeof(*S));
  if (!S) return -1;
  void *b = mmap(NULL, cap, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                 -1, 0);
  if (b == MAP_FAILED) {
    free(S);
    return -1;
  }
  S->gz = gzopen(filename, "rb");
  if (!S->gz) {
    munmap(b, cap);
    free(S);
    return -2;
  }
  gzbuffer(S->gz, 1u << 20);
  S->buf = (char *)b;
  S->cap = cap;
  pthread_mutex_init(&S->mu, NULL);
  pthread_cond_init(&S->cv, NULL);
  if (pthread_create(&S->th, NULL, gz_produce, S) != 0) {
    gzclose(S->gz);
    pthread_mutex_destroy(&S->mu);
    pthread_cond_destroy(&S->cv);
    munmap(b, cap);
    free(S);
    return -1;
  }
  t->buf = S->buf;
  t->len = t->pos = 0;
  t->eof = 0;
  t->src = S;
  t->map_len = cap;
  return 0;
}

/* the whole input, NUL-terminated: plain stdio or zlib */
static char *load_text(const char *filename, size_t *len_out) {
  size_t fl = strlen(filename), len = 0, cap = 4u << 20;
  int corrupt = 0;
  gzFile gz = NULL;
  FILE *f = NULL;
  if (fl > 3 && strcmp(filename + fl - 3, ".gz") == 0) {
    gz = gzopen(filename, "rb");
  } else {
    f = fopen(filename, "rb");
    if (f && fseek(f, 0, SEEK_END) == 0) {
      long sz = ftell(f);
      if (sz >= 0) cap = (size_t)sz + 2;
      if (fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        f = NULL;
      }
    }
  }
  *len_out = 0;
  if (!gz && !f) return NULL;
  char *buf = (char *)malloc(cap);
  for (;;) {
    if (!buf) break;
    if (cap - len < 2) {
      char *nb = (char *)realloc(buf, cap * 2);
      if (!nb) {
        free(buf);
        buf = NULL;
        break;
      }
      buf = nb;
      cap *= 2;
    }
    size_t want = cap - len - 1, got;
    if (gz) {
      int r = gzread(gz, buf + len, (unsigned)(want > (1u << 30) ? (1u << 30) : want));
      got = r > 0 ? (size_t)r : 0;
      if (r < 0) { /* corrupt or truncated data: a read failure, not the end */
        fprintf(stderr, "ERROR: corrupt or truncated gzip data in %s\n", filename);
        free(buf);
        buf = NULL;
        corrupt = 1;
        break;
      }
    } else {
      got = fread(buf + len, 1, want, f);
    }
    len += got;
    if (got == 0) break;
  }
  if (gz) gzclose(gz);
  if (f) fclose(f);
  if (!buf) {
    if (!corrupt) fprintf(stderr, "ERROR: Out of memory reading %s\n", filename);
    *len_out = 1; /* opened, not read */
    return NULL;
  }
  buf[len] = '\0';
  *len_out = len;
  return buf;
}

/* ---- block-gzipped input (BGZF): every gzip member (RFC 1952) carries its
 * own compressed size in a "BC" extra subfield, so the members can be found
 * without inflating anything and inflated in parallel on the host worker
 * threads, each into its place in the text (its uncompressed size is the
 * member's ISIZE field).  Every member's CRC-32 and size are checked.  Any
 * other gzip layout (one stream, or members without that subfield) is
 * inflated serially by the streamed reader. ---- */
typedef struct {
  size_t in_off, in_len, out_off;
  uint32_t out_len, crc;
} gzmember_t;

static uint32_t le32(const unsigned char *p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

/* total size of the gzip member starting at z (n bytes of it readable) from
 * its "BC" subfield, or 0 if it has none (or is not a gzip member); *hdr: its
 * header length.  The size may run past the n bytes given. */
static size_t bgzf_member_size(const unsigned char *z, size_t n, size_t *hdr) {
  if (n < 18 || z[0] != 0x1f || z[1] != 0x8b || z[2] != 8 || z[3] != 4) return 0;
  const size_t xlen = (size_t)z[10] | (size_t)z[11] << 8;
  if (12 + xlen > n) return 0;
  for (size_t q = 12; q + 4 <= 12 + xlen;) {
    const size_t slen = (size_t)z[q + 2] | (size_t)z[q + 3] << 8;
    if (q + 4 + slen > 12 + xlen) return 0;
    if (z[q] == 'B' && z[q + 1] == 'C' && slen == 2) {
      const size_t size = ((size_t)z[q + 4] | (size_t)z[q + 5] << 8) + 1;
      *hdr = 12 + xlen;
      return size >= *hdr + 8 ? size : 0;
    }
    q += 4 + slen;
  }
  return 0;
}

typedef struct {
  const unsigned char *z;
  const gzmember_t *mem;
  int nmem, ntask;
  char *out;
  int err;
} bgzf_job_t;

static void bgzf_task(void *arg, int t) {
  bgzf_job_t *J = (bgzf_job_t *)arg;
  const int k0 = (int)((long long)J->nmem * t / J->ntask);
  const int k1 = (int)((long long)J->nmem * (t + 1) / J->ntask);
  for (int k = k0; k < k1; ++k) {
    const gzmember_t *M = &J->mem[k];
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit2(&zs, -15) != Z_OK) {
      __atomic_store_n(&J->err, 1, __ATOMIC_RELAXED);
      return;
    }
    zs.next_in = (Bytef *)(J->z + M->in_off);
    zs.avail_in = (uInt)M->in_len;
    zs.next_out = (Bytef *)(J->out + M->out_off);
    zs.avail_out = M->out_len;
    const int r = M->out_len ? inflate(&zs, Z_FINISH) : Z_STREAM_END;
    const int ok = (r == Z_STREAM_END || (M->out_len == 0 && r == Z_BUF_ERROR)) &&
                   zs.total_out == M->out_len &&
                   crc32(0L, (const Bytef *)(J->out + M->out_off), M->out_len) == M->crc;
    inflateEnd(&zs);
    if (!ok) __atomic_store_n(&J->err, 1, __ATOMIC_RELAXED);
  }
}

/* 0: the text is whole in t; 1: not block-gzipped (t untouched); -1: the
 * file cannot be opened; -2: it cannot be read or a member is corrupt */
static int bgzf_open(const char *filename, text_t *t) {
  FILE *f = fopen(filename, "rb");
  if (!f) return -1;
  unsigned char head[64];
  const size_t h = fread(head, 1, sizeof(head), f);
  size_t hdr = 0;
  if (bgzf_member_size(head, h, &hdr) == 0) {
    fclose(f);
    return 1;
  }
  struct stat st;
  if (fstat(fileno(f), &st) != 0 || st.st_size <= 0 || fseek(f, 0, SEEK_SET) != 0) {
    fclose(f);
    return -2;
  }
  const size_t n = (size_t)st.st_size;
  unsigned char *z = (unsigned char *)malloc(n);
  if (!z || fread(z, 1, n, f) != n) {
    free(z);
    fclose(f);
    return -2;
  }
  fclose(f);
  size_t cap = 1024, nmem = 0, total = 0;
  gzmember_t *mem = (gzmember_t *)malloc(cap * sizeof(gzmember_t));
  int rc = mem ? 0 : -2;
  for (size_t pos = 0; rc == 0 && pos < n;) {
    const size_t size = bgzf_member_size(z + pos, n - pos, &hdr);
    if (size > n - pos) {
      fprintf(stderr, "ERROR: truncated gzip member in %s\n", filename);
      rc = -2;
      break;
    }
    if (size == 0) {
      rc = 1; /* a member without its size: the serial reader takes the file */
      break;
    }
    if (nmem == cap) {
      gzmember_t *nm = (gzmember_t *)realloc(mem, 2 * cap * sizeof(gzmember_t));
      if (!nm) {
        rc = -2;
        break;
      }
      mem = nm;
      cap *= 2;
    }
    gzmember_t *M = &mem[nmem++];
    M->in_off = pos + hdr;
    M->in_len = size - hdr - 8;
    M->crc = le32(z + pos + size - 8);
    M->out_len = le32(z + pos + size - 4);
    M->out_off = total;
    total += M->out_len;
    pos += size;
  }
  char *out = NULL;
  if (rc == 0 && !(out = (char *)malloc(total + 1))) rc = -2;
  if (rc == 0) {
    int nt = 4 * dpdlp_host_threads();
    if ((size_t)nt > nmem) nt = (int)nmem;
    bgzf_job_t J = {z, mem, (int)nmem, nt > 0 ? nt : 1, out, 0};
    if (nmem > 0) dpdlp_par_for(J.ntask, bgzf_task, &J);
    if (J.err) {
      fprintf(stderr, "ERROR: corrupt gzip member in %s\n", filename);
      rc = -2;
    }
  }
  free(z);
  free(mem);
  if (rc != 0) {
    free(out);
    return rc;
  }
  out[total] = '\0';
  t->buf = out;
  t->len = total;
  t->pos = 0;
  t->eof = 1;
  return 0;
}

/* opens the input: 0, or -1 if it cannot be opened, -2 if it cannot be read */
static int text_open(const char *filename, text_t *t) {
  memset(t, 0, sizeof(*t));
  size_t fl = strlen(filename);
  const char *env = getenv("DPDLP_MPS_STREAM");
  int stream = !(env && env[0] == '0');
  if (fl > 3 && strcmp(filename + fl - 3, ".gz") == 0) {
    const int rc = bgzf_open(filename, t);
    if (rc == 0) return 0;
    if (rc < 0) return rc;
  }
  if (stream && fl > 3 && strcmp(filename + fl - 3, ".gz") == 0) {
    int rc = gz_stream_open(filename, t);
    if (rc == 0) return 0;
    if (rc == -2) return -1;
  }
  t->buf = load_text(filename, &t->len);
  if (!t->buf) return t->len == 0 ? -1 : -2;
  t->eof = 1;
  return 0;
}

/* next line, NUL-terminated in place without its newline; NULL at the end */
static char *next_line(text_t *t) {
  for (;;) {
    if (t->pos < t->len) {
      char *line = t->buf + t->pos;
      char *nl = (char *)memchr(line, '\n', t->len - t->pos);
      if (nl) {
        *nl = '\0';
        t->pos = (size_t)(nl - t->buf) + 1;
        return line;
      }
      if (t->eof) {
        t->pos = t->len; /* the last line has no newline; buf[len] is NUL */
        return line;
      }
    } else if (t->eof) {
      return NULL;
    }
    text_wait(t, t->len + 1);
  }
}

static inline int is_delim(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

/* the tokens of a line as strtok_r(" \t\n\r") splits them, at most 6,
 * each NUL-terminated in place */
static int tokenize(char *p, char **tk) {
  int nt = 0;
  for (;;) {
    while (*p && is_delim(*p)) ++p;
    if (!*p || nt == 6) break;
    tk[nt++] = p;
    while (*p && !is_delim(*p)) ++p;
    if (!*p) break;
    *p++ = '\0';
  }
  return nt;
}

/* the line [p, e) (e: its newline or the end of the text) is a section
 * header: not a comment or blank line, exactly one token, a letter first */
static int is_header(const char *p, const char *e) {
  if (p == e || *p == '*' || *p == '\r' || *p == '\0') return 0;
  while (p < e && *p && is_delim(*p)) ++p;
  if (p == e || !*p || !isalpha((unsigned char)*p)) return 0;
  while (p < e && *p && !is_delim(*p)) ++p;
  while (p < e && *p && is_delim(*p)) ++p;
  return p == e || !*p;
}

typedef struct {
  int *er;    /* constraint (row) index */
  int *ec;    /* column index */
  double *ev; /* value */
  size_t ne;
} eblk_t;

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
  eblk_t *blk; /* entries, one block per parsed COLUMNS chunk, in file order */
  size_t nblk, blk_cap;
  size_t nz;
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

/* mps_parser.c:511-594 */
static int finalize_rows(mps_t *P) {
  long obj = -1;
  for (size_t i = 0; i < P->npend; ++i)
    if (P->pend_types[i] == 'N') {
      obj = (long)i;
      break;
    }
  if (obj < 0 && P->npend > 0) obj = 0;
  if (smap_reserve(&P->rows, P->rows.size + P->npend)) return -1;
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

/* ---- COLUMNS (mps_parser.c:596-650) ---- */

typedef struct {
  const char *name; /* column name (NUL-terminated in the text); NULL: the
                       chunk's leading lines that continue the previous column */
  uint64_t h;
  int len;
  size_t beg; /* first entry of the run */
  int col;    /* column index, set when the chunks are merged */
} crun_t;

typedef struct {
  char *beg, *end; /* whole lines */
  char *hdr;       /* header scan: first header line in [beg, end), or NULL */
  crun_t *run;
  size_t nrun, run_cap;
  int *er; /* constraint entries: row index, value, column (set by the merge) */
  double *ev;
  int *ec;
  size_t ne, e_cap;
  size_t *orun; /* objective entries: run, value */
  double *ov;
  size_t no, o_cap;
  struct srec *rec; /* RHS, RANGES, BOUNDS: one record per name/value */
  size_t nrec, rec_cap;
  int err;
} cchunk_t;

typedef struct {
  const mps_t *P; /* the maps and the objective row name: read only here */
  cchunk_t *ch;
  int sec;
} cjob_t;

static int chunk_run(cchunk_t *c, const char *name) {
  if (c->nrun == c->run_cap) {
    size_t nc = c->run_cap ? c->run_cap * 2 : 256;
    crun_t *nr = (crun_t *)realloc(c->run, nc * sizeof(crun_t));
    if (!nr) return -1;
    c->run = nr;
    c->run_cap = nc;
  }
  crun_t *r = &c->run[c->nrun++];
  r->name = name;
  r->len = 0;
  r->h = name ? fnv1a_len(name, &r->len) : 0;
  r->beg = c->ne;
  return 0;
}

static int chunk_obj(cchunk_t *c, double v) {
  if (c->no == c->o_cap) {
    size_t nc = c->o_cap ? c->o_cap * 2 : 1024;
    size_t *a = (size_t *)realloc(c->orun, nc * sizeof(size_t));
    if (!a) return -1;
    c->orun = a;
    double *b = (double *)realloc(c->ov, nc * sizeof(double));
    if (!b) return -1;
    c->ov = b;
    c->o_cap = nc;
  }
  c->orun[c->no] = c->nrun - 1;
  c->ov[c->no] = v;
  c->no++;
  return 0;
}

static int chunk_entry(cchunk_t *c, int row, double v) {
  if (c->ne == c->e_cap) {
    size_t nc = c->e_cap ? c->e_cap * 2 : 4096;
    int *a = (int *)realloc(c->er, nc * sizeof(int));
    if (!a) return -1;
    c->er = a;
    double *b = (double *)realloc(c->ev, nc * sizeof(double));
    if (!b) return -1;
    c->ev = b;
    c->e_cap = nc;
  }
  c->er[c->ne] = row;
  c->ev[c->ne] = v;
  c->ne++;
  return 0;
}

/* worker: the first header line of a chunk */
static void chunk_find_header(void *arg, int t) {
  cchunk_t *c = &((cjob_t *)arg)->ch[t];
  c->hdr = NULL;
  for (char *p = c->beg; p < c->end;) {
    char *nl = (char *)memchr(p, '\n', (size_t)(c->end - p));
    char *e = nl ? nl : c->end;
    if (is_header(p, e)) {
      c->hdr = p;
      return;
    }
    p = nl ? nl + 1 : c->end;
  }
}

/* worker: the lines of one chunk of a COLUMNS body (no header among them).
 * An odd token count names the column (a change of name starts a run), an
 * even count continues the current one; 'MARKER' lines and lines of fewer
 * than two tokens are skipped; entries of unknown rows are dropped. */
static void chunk_parse(void *arg, int t) {
  const cjob_t *J = (const cjob_t *)arg;
  const mps_t *P = J->P;
  cchunk_t *c = &J->ch[t];
  const char *cur = NULL;
  for (char *p = c->beg; p < c->end && !c->err;) {
    char *line = p;
    char *nl = (char *)memchr(p, '\n', (size_t)(c->end - p));
    if (nl) {
      *nl = '\0';
      p = nl + 1;
    } else {
      p = c->end; /* the text's last line: NUL-terminated at the end */
    }
    if (line[0] == '*' || line[0] == '\0' || line[0] == '\r') continue;
    char *tk[6];
    int nt = tokenize(line, tk);
    if (nt < 2 || strcmp(tk[1], "'MARKER'") == 0) continue;
    int first = nt % 2;
    if (first) {
      if (!cur || strcmp(cur, tk[0]) != 0) {
        if (chunk_run(c, tk[0])) c->err = 1;
        cur = tk[0];
      }
    } else if (c->nrun == 0) {
      if (chunk_run(c, NULL)) c->err = 1;
    }
    for (int i = first; i + 1 < nt && !c->err; i += 2) {
      double v = mps_atof(tk[i + 1]);
      if (P->obj_row && strcmp(tk[i], P->obj_row) == 0) {
        if (chunk_obj(c, v)) c->err = 1;
      } else {
        int r = smap_get(&P->rows, tk[i]);
        if (r >= 0 && chunk_entry(c, r, v)) c->err = 1;
      }
    }
  }
}

/* worker: the column of every constraint entry of a chunk */
static void chunk_cols(void *arg, int t) {
  cchunk_t *c = &((cjob_t *)arg)->ch[t];
  if (c->ne == 0) return;
  c->ec = (int *)malloc(c->ne * sizeof(int));
  if (!c->ec) {
    c->err = 1;
    return;
  }
  for (size_t k = 0; k < c->nrun; ++k) {
    size_t end = k + 1 < c->nrun ? c->run[k + 1].beg : c->ne;
    for (size_t e = c->run[k].beg; e < end; ++e) c->ec[e] = c->run[k].col;
  }
}

/* in file order: number the columns (first appearance) and add the
 * objective entries to c; then hand the chunks' entries to the problem */
static int chunks_merge(mps_t *P, cjob_t *job, int nch) {
  cchunk_t *ch = job->ch;
  for (int t = 0; t < nch; ++t)
    if (ch[t].err) return -1;
  size_t runs = 0;
  for (int t = 0; t < nch; ++t) runs += ch[t].nrun;
  if (smap_reserve(&P->cols, P->cols.size + runs)) return -1;
  const char *cur = P->cur_col;
  int col = P->cur_idx;
  for (int t = 0; t < nch; ++t) {
    cchunk_t *c = &ch[t];
    for (size_t k = 0; k < c->nrun; ++k) {
      crun_t *r = &c->run[k];
      if (!r->name) {
        if (!cur) {
          fprintf(stderr, "ERROR: Column data found before any column name was defined.\n");
          return -1;
        }
      } else if (!cur || strcmp(cur, r->name) != 0) {
        cur = r->name;
        col = -1;
      }
      if (col < 0) {
        if (ensure_cols(P)) return -1;
        int clen = r->len;
        uint64_t h = r->h; /* hashed by the worker */
        if (!r->name || cur != r->name) h = fnv1a_len(cur, &clen);
        col = smap_put_h(&P->cols, cur, clen, h);
        if (col < 0) return -1;
      }
      r->col = col;
    }
    for (size_t o = 0; o < c->no; ++o) P->obj[c->run[c->orun[o]].col] += c->ov[o];
  }
  if (cur != P->cur_col) { /* the text is freed after parsing: keep a copy */
    char *own = strdup(cur);
    if (!own) return -1;
    free(P->cur_col);
    P->cur_col = own;
  }
  P->cur_idx = col;
  dpdlp_par_for(nch, chunk_cols, job);
  if (P->nblk + (size_t)nch > P->blk_cap) {
    size_t nc = P->nblk + (size_t)nch + 16;
    eblk_t *nb = (eblk_t *)realloc(P->blk, nc * sizeof(eblk_t));
    if (!nb) return -1;
    P->blk = nb;
    P->blk_cap = nc;
  }
  for (int t = 0; t < nch; ++t) {
    cchunk_t *c = &ch[t];
    if (c->err) return -1;
    if (c->ne == 0) continue;
    P->blk[P->nblk++] = (eblk_t){c->er, c->ec, c->ev, c->ne};
    P->nz += c->ne;
    c->er = c->ec = NULL;
    c->ev = NULL;
  }
  return 0;
}

/* cut [b, e) at line starts into nch chunks */
static void cut_chunks(char *b, char *e, int nch, cchunk_t *ch) {
  size_t len = (size_t)(e - b);
  char *prev = b;
  for (int k = 0; k < nch; ++k) {
    char *s = k == 0 ? b : b + (size_t)((double)len * k / nch);
    if (s < prev) s = prev;
    if (k > 0 && s > b && s < e && s[-1] != '\n') {
      char *nl = (char *)memchr(s, '\n', (size_t)(e - s));
      s = nl ? nl + 1 : e;
    }
    ch[k].beg = s;
    if (k > 0) ch[k - 1].end = s;
    prev = s;
  }
  ch[nch - 1].end = e;
}

/* ---- RHS, RANGES, BOUNDS (mps_parser.c:652-742): worker threads turn the
 * lines of a section body into records (numbers converted, names looked up;
 * unknown names and bound types dropped) that one thread applies in file
 * order ---- */
typedef struct srec {
  int kind, idx;
  double v;
} srec_t;

enum { R_OBJ = 1, R_ROW, B_LO, B_UP, B_FX, B_FR, B_MI, B_PL, B_BV };

static int chunk_rec(cchunk_t *c, int kind, int idx, double v) {
  if (c->nrec == c->rec_cap) {
    size_t nc = c->rec_cap ? c->rec_cap * 2 : 1024;
    srec_t *nr = (srec_t *)realloc(c->rec, nc * sizeof(srec_t));
    if (!nr) return -1;
    c->rec = nr;
    c->rec_cap = nc;
  }
  c->rec[c->nrec++] = (srec_t){kind, idx, v};
  return 0;
}

static int bound_kind(const char *t) {
  static const char *names[] = {"LO", "UP", "FX", "FR", "MI", "PL", "BV"};
  for (int k = 0; k < 7; ++k)
    if (!strcmp(t, names[k])) return B_LO + k;
  return 0;
}

static void chunk_records(void *arg, int t) {
  const cjob_t *J = (const cjob_t *)arg;
  const mps_t *P = J->P;
  cchunk_t *c = &J->ch[t];
  for (char *p = c->beg; p < c->end && !c->err;) {
    char *line = p;
    char *nl = (char *)memchr(p, '\n', (size_t)(c->end - p));
    if (nl) {
      *nl = '\0';
      p = nl + 1;
    } else {
      p = c->end;
    }
    if (line[0] == '*' || line[0] == '\0' || line[0] == '\r') continue;
    char *tk[6];
    int nt = tokenize(line, tk);
    if (J->sec == S_BOUNDS) {
      if (nt < 3) continue;
      int col = smap_get(&P->cols, tk[2]);
      int kind = bound_kind(tk[0]);
      if (col < 0 || !kind) continue;
      if (chunk_rec(c, kind, col, nt > 3 ? atof(tk[3]) : 0.0)) c->err = 1;
      continue;
    }
    for (int i = 1; i + 1 < nt && !c->err; i += 2) {
      if (J->sec == S_RHS) {
        double v = mps_atof(tk[i + 1]);
        if (P->obj_row && strcmp(tk[i], P->obj_row) == 0) {
          if (chunk_rec(c, R_OBJ, 0, v)) c->err = 1;
          continue;
        }
        int r = smap_get(&P->rows, tk[i]);
        if (r >= 0 && chunk_rec(c, R_ROW, r, v)) c->err = 1;
      } else { /* RANGES */
        double R = atof(tk[i + 1]);
        int r = smap_get(&P->rows, tk[i]);
        if (r >= 0 && chunk_rec(c, R_ROW, r, R)) c->err = 1;
      }
    }
  }
}

/* mps_parser.c:652-678 */
static void apply_rhs(mps_t *P, const srec_t *q) {
  if (q->kind == R_OBJ) {
    P->obj_const = -q->v;
    return;
  }
  int r = q->idx;
  char t = P->types[r];
  if (t == 'L') {
    P->cub[r] = q->v;
  } else if (t == 'G') {
    P->clb[r] = q->v;
  } else {
    P->clb[r] = q->v;
    P->cub[r] = q->v;
  }
}

/* mps_parser.c:680-707 */
static void apply_range(mps_t *P, const srec_t *q) {
  int r = q->idx;
  double R = q->v;
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

/* mps_parser.c:709-742 */
static void apply_bound(mps_t *P, const srec_t *q) {
  int col = q->idx;
  double v = q->v;
  switch (q->kind) {
  case B_LO: P->lb[col] = v; break;
  case B_UP: P->ub[col] = v; break;
  case B_FX: P->lb[col] = v; P->ub[col] = v; break;
  case B_FR: P->lb[col] = -INFINITY; P->ub[col] = INFINITY; break;
  case B_MI: P->lb[col] = -INFINITY; break;
  case B_PL: P->ub[col] = INFINITY; break;
  case B_BV: P->lb[col] = 0.0; P->ub[col] = 1.0; break;
  default: break;
  }
}

/* One window [b, end) of whole lines of a section body: the part before
 * the first header line in it is cut into chunks, parsed by the worker
 * threads and merged or applied in file order.  *stop is set to that
 * header (or end if there is none) and *found to whether there was one. */
static int body_window(mps_t *P, char *b, char *end, int sec, char **stop, int *found) {
  const char *env = getenv("DPDLP_MPS_CHUNK");
  long long min_chunk = env ? atoll(env) : (1LL << 20);
  if (min_chunk < 1) min_chunk = 1;
  int nth = dpdlp_host_threads();
  long long want = (long long)(end - b) / min_chunk;
  int nch = want < 1 ? 1 : (want > 8LL * nth ? 8 * nth : (int)want);
  cchunk_t *ch = (cchunk_t *)calloc((size_t)nch, sizeof(cchunk_t));
  if (!ch) return -1;
  cjob_t job = {P, ch, sec};
  cut_chunks(b, end, nch, ch);
  dpdlp_par_for(nch, chunk_find_header, &job);
  char *e = end;
  *found = 0;
  for (int k = 0; k < nch; ++k)
    if (ch[k].hdr) {
      e = ch[k].hdr;
      *found = 1;
      break;
    }
  want = (long long)(e - b) / min_chunk;
  if (want < nch) nch = want < 1 ? 1 : (int)want;
  cut_chunks(b, e, nch, ch);
  int rc = 0;
  if (sec == S_COLUMNS) {
    dpdlp_par_for(nch, chunk_parse, &job);
    rc = chunks_merge(P, &job, nch);
  } else {
    dpdlp_par_for(nch, chunk_records, &job);
    for (int k = 0; k < nch && !rc; ++k) {
      if (ch[k].err) rc = -1;
      for (size_t i = 0; i < ch[k].nrec && !rc; ++i) {
        const srec_t *q = &ch[k].rec[i];
        if (sec == S_RHS)
          apply_rhs(P, q);
        else if (sec == S_RANGES)
          apply_range(P, q);
        else
          apply_bound(P, q);
      }
    }
  }
  for (int k = 0; k < nch; ++k) {
    free(ch[k].run);
    free(ch[k].er);
    free(ch[k].ec);
    free(ch[k].ev);
    free(ch[k].orun);
    free(ch[k].ov);
    free(ch[k].rec);
  }
  free(ch);
  *stop = e;
  return rc;
}

/* The body of the section whose header was just read (it runs from the
 * text's read position to the next header line, or the end).  A whole text
 * is one window; a streamed one is taken in windows of whole lines as they
 * arrive (DPDLP_MPS_WINDOW bytes, default 16 MiB), each parsed while the
 * producer inflates the next.  The read position is left at the header that
 * ends the body. */
static int section_body(mps_t *P, text_t *T, int sec) {
  const char *env = getenv("DPDLP_MPS_WINDOW");
  long long wv = env ? atoll(env) : (16LL << 20);
  size_t win = wv < 1 ? 1 : (size_t)wv;
  size_t want = T->pos + win;
  for (;;) {
    text_wait(T, want);
    char *b = T->buf + T->pos, *end = T->buf + T->len;
    if (!T->eof) {
      char *ln = b < end ? (char *)memrchr(b, '\n', (size_t)(end - b)) : NULL;
      if (!ln) { /* not one whole line yet */
        want = T->len + win;
        continue;
      }
      end = ln + 1;
    }
    if (b == end) return 0; /* the end of the text */
    char *e;
    int found;
    if (body_window(P, b, end, sec, &e, &found)) return -1;
    T->pos = (size_t)(e - T->buf);
    if (found || (T->eof && T->pos >= T->len)) return 0;
    want = T->pos + win;
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
  for (size_t b = 0; b < P->nblk; ++b) {
    free(P->blk[b].er);
    free(P->blk[b].ec);
    free(P->blk[b].ev);
  }
  free(P->blk);
  free(P->obj_row);
  free(P->cur_col);
}

/* ---- entries -> CSR: a stable counting sort by row over G groups of
 * consecutive entry blocks; each group counts its rows, the counts become
 * per-group write positions (row by row, groups in file order), and each
 * group scatters its entries, so every row keeps file order. ---- */
typedef struct {
  const mps_t *P;
  int m, G, R;
  size_t *gb;  /* group g holds blocks [gb[g], gb[g + 1]) */
  int *cnt;    /* cnt[g * (m + 1) + r] */
  int *rp, *ci;
  double *v;
} csr_job_t;

static void csr_count(void *arg, int g) {
  csr_job_t *J = (csr_job_t *)arg;
  int *c = J->cnt + (size_t)g * (J->m + 1);
  for (size_t b = J->gb[g]; b < J->gb[g + 1]; ++b) {
    const eblk_t *B = &J->P->blk[b];
    for (size_t e = 0; e < B->ne; ++e) c[B->er[e]]++;
  }
}

static void csr_rows(int m, int R, int t, int *r0, int *r1) {
  *r0 = (int)((long long)m * t / R);
  *r1 = (int)((long long)m * (t + 1) / R);
}

static void csr_totals(void *arg, int t) {
  csr_job_t *J = (csr_job_t *)arg;
  int r0, r1;
  csr_rows(J->m, J->R, t, &r0, &r1);
  for (int r = r0; r < r1; ++r) {
    int s = 0;
    for (int g = 0; g < J->G; ++g) s += J->cnt[(size_t)g * (J->m + 1) + r];
    J->rp[r + 1] = s;
  }
}

static void csr_offsets(void *arg, int t) {
  csr_job_t *J = (csr_job_t *)arg;
  int r0, r1;
  csr_rows(J->m, J->R, t, &r0, &r1);
  for (int r = r0; r < r1; ++r) {
    int run = J->rp[r];
    for (int g = 0; g < J->G; ++g) {
      int *c = &J->cnt[(size_t)g * (J->m + 1) + r];
      int k = *c;
      *c = run;
      run += k;
    }
  }
}

static void csr_scatter(void *arg, int g) {
  csr_job_t *J = (csr_job_t *)arg;
  int *pos = J->cnt + (size_t)g * (J->m + 1);
  for (size_t b = J->gb[g]; b < J->gb[g + 1]; ++b) {
    const eblk_t *B = &J->P->blk[b];
    for (size_t e = 0; e < B->ne; ++e) {
      int d = pos[B->er[e]]++;
      J->ci[d] = B->ec[e];
      J->v[d] = B->ev[e];
    }
  }
}

static int build_csr(const mps_t *P, int m, int *rp, int *ci, double *v) {
  rp[0] = 0;
  if (m == 0 || P->nz == 0) {
    for (int i = 0; i < m; ++i) rp[i + 1] = 0;
    return 0;
  }
  int nth = dpdlp_host_threads();
  long long gmax = (32LL << 20) / ((long long)m + 1); /* counts stay within 128 MB */
  int G = (int)P->nblk < nth ? (int)P->nblk : nth;
  if (G > gmax) G = gmax < 1 ? 1 : (int)gmax;
  size_t *gb = (size_t *)malloc(((size_t)G + 1) * sizeof(size_t));
  int *cnt = (int *)calloc((size_t)G * ((size_t)m + 1), sizeof(int));
  if (!gb || !cnt) {
    free(gb);
    free(cnt);
    return -1;
  }
  /* groups of consecutive blocks with about nz / G entries each */
  gb[0] = 0;
  size_t acc = 0, b = 0;
  for (int g = 1; g < G; ++g) {
    size_t target = P->nz / (size_t)G * (size_t)g;
    while (b < P->nblk && acc < target) acc += P->blk[b++].ne;
    gb[g] = b;
  }
  gb[G] = P->nblk;
  int R = m < 4096 ? 1 : 4 * nth;
  csr_job_t J = {P, m, G, R, gb, cnt, rp, ci, v};
  dpdlp_par_for(G, csr_count, &J);
  dpdlp_par_for(R, csr_totals, &J);
  for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
  dpdlp_par_for(R, csr_offsets, &J);
  dpdlp_par_for(G, csr_scatter, &J);
  free(gb);
  free(cnt);
  return 0;
}

lp_problem_t *read_mps_file(const char *filename) {
  if (!filename) return NULL;
  text_t T;
  int orc = text_open(filename, &T);
  if (orc) {
    if (orc == -1) fprintf(stderr, "ERROR: Could not open file %s\n", filename);
    return NULL;
  }
  mps_t P;
  memset(&P, 0, sizeof(P));
  P.cur_idx = -1;
  if (smap_init(&P.rows) || smap_init(&P.cols)) P.err = 1;
  int sec = S_NONE;
  bool rows_done = false;
  char *line;
  while (!P.err && (line = next_line(&T)) != NULL) {
    if (line[0] == '*' || line[0] == '\0' || line[0] == '\r') continue;
    char *tk[6];
    int nt = tokenize(line, tk);
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
      if ((sec == S_COLUMNS || sec == S_RHS || sec == S_RANGES || sec == S_BOUNDS) && !P.err &&
          section_body(&P, &T, sec))
        P.err = 1;
      continue;
    }
    switch (sec) {
    case S_OBJSENSE:
      if (!strcmp(tk[0], "MAX") || !strcmp(tk[0], "MAXIMIZE")) P.maximize = true;
      break;
    case S_ROWS:
      if (rows_line(&P, tk, nt)) P.err = 1;
      break;
    default:
      break;
    }
  }
  if (T.src && !P.err) {
    /* the inflate of a streamed input runs to the end of the file (the
     * gzip trailer's CRC-32 and size are checked there); corrupt or
     * truncated data fails the read */
    text_wait(&T, (size_t)-1);
    pthread_mutex_lock(&T.src->mu);
    const int bad = T.src->err;
    pthread_mutex_unlock(&T.src->mu);
    if (bad) {
      fprintf(stderr, "ERROR: corrupt or truncated gzip data in %s\n", filename);
      P.err = 1;
    }
  }
  text_close(&T);
  if (P.err) {
    fprintf(stderr, "ERROR: Failed to parse MPS file.\n");
    mps_free(&P);
    return NULL;
  }
  if (P.nz > 0x7fffffff) {
    fprintf(stderr, "ERROR: More than 2^31-1 nonzeros.\n");
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
  /* entries -> CSR keeping file order within rows (mps_parser.c:744-777) */
  p->constraint_matrix_row_pointers = (int *)calloc((size_t)m + 1, sizeof(int));
  p->constraint_matrix_col_indices = (int *)malloc((P.nz ? P.nz : 1) * sizeof(int));
  p->constraint_matrix_values = (double *)malloc((P.nz ? P.nz : 1) * sizeof(double));
  if (!p->constraint_matrix_row_pointers || !p->constraint_matrix_col_indices ||
      !p->constraint_matrix_values ||
      build_csr(&P, m, p->constraint_matrix_row_pointers, p->constraint_matrix_col_indices,
                p->constraint_matrix_values)) {
    lp_problem_free(p);
    mps_free(&P);
    return NULL;
  }
  mps_free(&P);
  return p;
}
