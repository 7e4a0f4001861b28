/* Checks that the kernels' division by the radius (three operations from
 * rinv = RN(1 / radius): q0 = RN(a rinv), e = fma(-q0, radius, a),
 * q = fma(e, rinv, q0); a zero a keeps its sign) equals IEEE a / radius for
 * every float a with 2^-100 <= |a| <= 2^100, and for +0 and -0.
 *
 * usage: over_radius_check radius...   prints "ok" or the first mismatch.
 * Built with -ffp-contract=off, so only the explicit fmaf calls fuse. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static uint32_t bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

static float over_radius(float a, float radius, float rinv) {
  const float q0 = a * rinv;
  const float residual = fmaf(-q0, radius, a);
  const float q = fmaf(residual, rinv, q0);
  return a == 0.0f ? q0 : q;
}

int main(int argc, char** argv) {
  const uint32_t lo = bits(0x1p-100f), hi = bits(0x1p100f);
  for (int k = 1; k < argc; ++k) {
    const float radius = strtof(argv[k], NULL);
    const float rinv = 1.0f / radius;
    long bad = -1;
#pragma omp parallel for schedule(static) reduction(max : bad)
    for (long u = lo; u <= (long)hi; ++u) {
      for (int sign = 0; sign < 2; ++sign) {
        const float a = from_bits((uint32_t)u | (sign ? 0x80000000u : 0u));
        if (bits(over_radius(a, radius, rinv)) != bits(a / radius) && u > bad) bad = u;
      }
    }
    const float zeros[2] = {0.0f, -0.0f};
    for (int z = 0; z < 2; ++z)
      if (bits(over_radius(zeros[z], radius, rinv)) != bits(zeros[z] / radius)) bad = 0;
    if (bad >= 0) {
      printf("mismatch: radius %.9g, a bits 0x%08lx\n", radius, bad);
      return 1;
    }
  }
  printf("ok\n");
  return 0;
}
