/*
 * Starting vector of the power iteration: the reference draws it from a
 * process-global std::mt19937 gen(1) with std::normal_distribution<double>
 * (src/utils.cu:29-30,80-82).  This library reseeds per solve, which
 * reproduces the vector of the reference's first solve in a process and makes
 * every solve independent of call order.
 *
 * The draws are the standard library's, bit for bit, computed faster.
 * std::normal_distribution<double> is Marsaglia's polar method over
 * generate_canonical<double, 53>: every attempt takes four 32-bit engine
 * outputs (two uniforms of two words each, low word first), and an accepted
 * attempt (0 < r2 <= 1) yields two values, y*mult first, then x*mult.  The
 * engine runs once, sequentially, and keeps the accepted (x, y, r2) triples;
 * the logarithms and square roots, which dominate the cost, then run on the
 * host worker threads over slices of the triples.
 */
#include <cmath>
#include <random>
#include <stdint.h>
#include <stdlib.h>

extern "C" void dpdlp_par_for(int ntask, void (*fn)(void *ctx, int task), void *ctx);

namespace {

/* MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998), the engine
 * std::mt19937 specifies: state of 624 words, seeded by the recurrence
 * x[i] = 1812433253 (x[i-1] ^ (x[i-1] >> 30)) + i, regenerated 624 words at a
 * time and tempered.  Regenerating and tempering a whole block in plain loops
 * lets the compiler vectorise them. */
struct mt_t {
  uint32_t x[624], out[624];
  int pos;
};

void mt_seed(mt_t &s, uint32_t seed) {
  s.x[0] = seed;
  for (uint32_t i = 1; i < 624; ++i) s.x[i] = 1812433253u * (s.x[i - 1] ^ (s.x[i - 1] >> 30)) + i;
  s.pos = 624;
}

inline uint32_t mt_mix(uint32_t hi, uint32_t lo, uint32_t far) {
  const uint32_t y = (hi & 0x80000000u) | (lo & 0x7fffffffu);
  return far ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
}

void mt_refill(mt_t &s) {
  uint32_t *x = s.x;
  for (int k = 0; k < 227; ++k) x[k] = mt_mix(x[k], x[k + 1], x[k + 397]);
  for (int k = 227; k < 623; ++k) x[k] = mt_mix(x[k], x[k + 1], x[k - 227]);
  x[623] = mt_mix(x[623], x[0], x[396]);
  for (int k = 0; k < 624; ++k) {
    uint32_t z = x[k];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    s.out[k] = z;
  }
  s.pos = 0;
}

inline uint32_t mt_next(mt_t &s) {
  if (s.pos == 624) mt_refill(s);
  return s.out[s.pos++];
}

/* generate_canonical<double, 53>(mt19937): sum = w0 + w1 * 2^32 rounded to
 * double, divided by 2^64; a sum that rounds up to 2^64 gives the largest
 * double below 1 */
inline double canonical(mt_t &g) {
  const double w0 = (double)mt_next(g);
  const double w1 = (double)mt_next(g);
  double r = (w0 + w1 * 4294967296.0) / 18446744073709551616.0;
  if (r >= 1.0) r = std::nextafter(1.0, 0.0);
  return r;
}

struct pair_job_t {
  const double *xyr; /* accepted attempts: x, y, r2 */
  double *out;
  long long count;   /* values to write */
  long long npairs;
  int ntask;
};

void pair_task(void *arg, int t) {
  const pair_job_t *J = (const pair_job_t *)arg;
  const long long p0 = J->npairs * t / J->ntask, p1 = J->npairs * (t + 1) / J->ntask;
  for (long long p = p0; p < p1; ++p) {
    const double x = J->xyr[3 * p], y = J->xyr[3 * p + 1], r2 = J->xyr[3 * p + 2];
    const double mult = std::sqrt(-2 * std::log(r2) / r2);
    /* mean 0, standard deviation 1 applied as the library does (maps -0 to +0) */
    J->out[2 * p] = y * mult * 1.0 + 0.0;
    if (2 * p + 1 < J->count) J->out[2 * p + 1] = x * mult * 1.0 + 0.0;
  }
}

} // namespace

extern "C" void dpdlp_normal_fill(uint32_t seed, int count, double *out) {
  if (count <= 0) return;
  const long long npairs = ((long long)count + 1) / 2;
  double *xyr = (double *)malloc((size_t)npairs * 3 * sizeof(double));
  if (!xyr) { /* no memory for the staging: the standard library's own loop */
    std::mt19937 g(seed);
    std::normal_distribution<double> dist(0.0, 1.0);
    for (int i = 0; i < count; ++i) out[i] = dist(g);
    return;
  }
  mt_t gen;
  mt_seed(gen, seed);
  for (long long p = 0; p < npairs; ++p) {
    double x, y, r2;
    do {
      x = 2.0 * canonical(gen) - 1.0;
      y = 2.0 * canonical(gen) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    xyr[3 * p] = x;
    xyr[3 * p + 1] = y;
    xyr[3 * p + 2] = r2;
  }
  pair_job_t J = {xyr, out, count, npairs, (int)(npairs / 16384 + 1)};
  if (J.ntask > 64) J.ntask = 64;
  dpdlp_par_for(J.ntask, pair_task, &J);
  free(xyr);
}
