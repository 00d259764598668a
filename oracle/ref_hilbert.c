/* Harness that exports the reference's hilbert.h (compiled unmodified from
 * the reference tree) for building golden vectors.  Test infrastructure. */
#include "hilbert.h"

__attribute__((visibility("default"))) unsigned long long ref_hilbert_c2i(const unsigned long long coord[3]) {
  return hilbert_c2i(3, 16, coord);
}

__attribute__((visibility("default"))) void ref_hilbert_i2c(unsigned long long index, unsigned long long coord[3]) {
  hilbert_i2c(3, 16, index, coord);
}

/* The quantisation lambda of hostCode.cpp:200-207, restated with plain
 * floats (linalg's float3 is not available outside GPRT). */
__attribute__((visibility("default"))) void ref_hilbert_codes(const float *xyzw, unsigned long long n,
                                                              const float aabb[6], unsigned long long *codes) {
  for (unsigned long long i = 0; i < n; ++i) {
    float t[3];
    for (int d = 0; d < 3; ++d) {
      t[d] = (xyzw[4 * i + d] - aabb[d]) / (aabb[3 + d] - aabb[d]);
      t[d] = t[d] * (float)(1 << 16);
    }
    const bitmask_t coord[3] = {(bitmask_t)t[0], (bitmask_t)t[1], (bitmask_t)t[2]};
    codes[i] = hilbert_c2i(3, 16, coord);
  }
}
