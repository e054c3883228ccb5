/*
 * Starting vector of the power iteration: the reference draws it from a
 * process-global std::mt19937 gen(1) with std::normal_distribution<double>
 * (src/utils.cu:29-30,80-82).  This library reseeds per solve, which
 * reproduces the vector of the reference's first solve in a process and makes
 * every solve independent of call order.
 */
#include <random>
#include <stdint.h>

extern "C" void dpdlp_normal_fill(uint32_t seed, int count, double *out) {
  std::mt19937 gen(seed);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (int i = 0; i < count; ++i) out[i] = dist(gen);
}
