/* The step kernel's free-flight draw (flight and flight_nz in
 * attribute-aware-rbfs_amd/csrc/render_common.h) computes 1 - xi as
 * fma((float)s, -2^-32, 1) instead of 1.0f - (float)s * 2^-32, the
 * reference's xi = lcg() followed by 1 - xi (rng.h, deviceCodeRBF.hlsl).
 * The product (float)s * 2^-32 is exact (a power-of-two scaling of a value
 * >= 1 or of 0), so both round the same real number once; this program
 * checks the two against each other for every 32-bit LCG state (equal bits
 * in, equal aa_flight out).
 *
 *   gcc -O2 -ffp-contract=off -I include tests/cpp/flight_identity.c -lm
 *
 * Exits 0 when all 2^32 states agree. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "aarbf_detmath.h"

static uint32_t bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

int main(void) {
  uint64_t bad = 0;
  for (uint64_t k = 0; k < (1ull << 32); ++k) {
    uint32_t s = (uint32_t)k;
    const float xi = aa_lcg_next(&s);
    const float ref = 1.0f - xi;
    const float fused = fmaf((float)s, -2.3283064365386963e-10f, 1.0f);
    if (bits(ref) != bits(fused)) {
      if (bad < 5) printf("state %u: 1 - xi = %a, fused %a\n", (uint32_t)k, ref, fused);
      ++bad;
    }
  }
  printf("%llu of 4294967296 states differ\n", (unsigned long long)bad);
  return bad != 0;
}
