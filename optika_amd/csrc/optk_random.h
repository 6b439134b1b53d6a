// optk_random.h -- counter-based random numbers and the draws of the
// detector exposure (optk_expose), shared by the kernel and the host test
// harness.
//
// Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011): 10 rounds of two 32x32->64
// multiplications with the published multipliers and Weyl key increments.
// Poisson draws: multiplication of uniforms below a mean of 10 (Knuth,
// TAOCP vol. 2, 3.4.1), W. Hormann's transformed rejection with squeeze
// (PTRS, "The transformed rejection method for generating Poisson random
// variables", Insurance: Mathematics and Economics 12, 1993) above.
// Gaussian read noise: Box-Muller.
#pragma once

#include <math.h>
#include <stdint.h>

#include "optk_physics.h"

namespace optk {

OPTK_HD void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

// four 32-bit words for (pixel, draw) under the key
OPTK_HD void philox(uint64_t pixel, uint32_t draw, uint64_t seed, uint32_t out[4]) {
  uint32_t c[4] = {(uint32_t)pixel, (uint32_t)(pixel >> 32), draw, 0x6f70746bu};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k);
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
  for (int j = 0; j < 4; ++j) out[j] = c[j];
}

// uniform in (0, 1): 53 bits, never 0
OPTK_HD double unit53(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

struct UniformStream {
  uint64_t pixel, seed;
  uint32_t draw = 0;
  uint32_t words[4];
  int used = 4;
  OPTK_HD double next() {
    if (used >= 4) {
      philox(pixel, draw++, seed, words);
      used = 0;
    }
    const double u = unit53(words[used], words[used + 1]);
    used += 2;
    return u;
  }
};

// Both loops are bounded: a draw that has not finished after POISSON_MAX_STEPS
// steps (probability far below 1e-300 for finite means) returns the rounded
// mean, so every thread ends.
constexpr int POISSON_MAX_STEPS = 1000;

OPTK_HD double poisson_draw(double mean, UniformStream &rng) {
  if (!(mean > 0.0)) return mean == 0.0 ? 0.0 : NAN;
  if (!(mean < 1e300)) return mean;  // infinite mean
  if (mean < 10.0) {
    // multiplication of uniforms until the product falls below exp(-mean)
    const double limit = exp(-mean);
    double prod = rng.next();
    double k = 0.0;
    for (int step = 0; prod > limit && step < POISSON_MAX_STEPS; ++step) {
      prod *= rng.next();
      k += 1.0;
    }
    return k;
  }
  // transformed rejection with squeeze (Hormann 1993, PTRS)
  const double slam = sqrt(mean), loglam = log(mean);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int step = 0; step < POISSON_MAX_STEPS; ++step) {
    const double U = rng.next() - 0.5;
    const double V = rng.next();
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + mean + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -mean + k * loglam - lgamma(k + 1.0))
      return k;
  }
  return rint(mean);
}


// one zero-mean, unit-variance Gaussian from two uniforms
OPTK_HD double gaussian_draw(UniformStream &rng) {
  const double u1 = rng.next(), u2 = rng.next();
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

}  // namespace optk
