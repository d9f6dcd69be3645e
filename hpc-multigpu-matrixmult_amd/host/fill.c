#include "fill.h"

#include <string.h>

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int phpc_host_fill(double *x, size_t n, const char *dist, uint64_t seed) {
  const int ramp = strcmp(dist, "ramp") == 0, uniform = strcmp(dist, "uniform") == 0;
  if (!ramp && !uniform && strcmp(dist, "wide") != 0) return -1;
  for (size_t i = 0; i < n; ++i) {
    if (ramp) {
      x[i] = (double)i;
      continue;
    }
    const uint64_t z = mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
    if (uniform) {
      x[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
    } else {
      int e = (int)(((z >> 32) * 41ULL) >> 32) - 20;
      double v = 1.0; /* 2^e by exact doublings / halvings (no libm needed) */
      for (; e > 0; --e) v *= 2.0;
      for (; e < 0; ++e) v *= 0.5;
      x[i] = (z & 1) ? -v : v;
    }
  }
  return 0;
}
