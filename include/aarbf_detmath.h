/*
 * aarbf_detmath.h -- the renderer's portable, deterministic FP32 math.
 *
 * Stochastic tracking makes a discrete decision at every step (the
 * free-flight distance t -= ln(1-xi)*unit, the accept test xi < density,
 * the coverage test |P-x| < r).  One ULP of difference in a logarithm
 * between two implementations is enough to flip such a decision and send
 * the whole path elsewhere, so "same pixel for the same RNG stream" is
 * only possible if host and device evaluate the transcendental functions
 * with identical arithmetic.  The reference leaves these ULPs to the
 * Vulkan driver (deviceCodeRBF.hlsl:75,153,111,132; sharedCode.h:139), so
 * this header *defines* them for this renderer:
 *
 *   - only IEEE-754 binary32 operations that are correctly rounded on both
 *     x86-64 (SSE) and gfx950 (VALU): +, -, *, /, sqrt, fma, floor, rint,
 *     min/max and bit manipulation;
 *   - every multiply-add that should be fused is written as an explicit
 *     fmaf, and everything else must be compiled with -ffp-contract=off;
 *   - polynomial coefficients are this project's own least-squares fits
 *     (64 Chebyshev nodes, 60-digit decimal arithmetic), rounded to
 *     binary32; tests/golden/fit_detmath.py reproduces them.
 *
 * The header is part of the C ABI contract: the HIP kernels and any host
 * restatement of the renderer include it unchanged.  Accuracy against a
 * correctly rounded libm is <= 2 ulp for aa_logf/aa_expf on the ranges the
 * renderer uses (checked by the test suite).
 */
#ifndef AARBF_DETMATH_H
#define AARBF_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define AA_FN static inline __host__ __device__
#define AA_FMAF(a, b, c) __builtin_fmaf((a), (b), (c))
#define AA_SQRTF(a) __builtin_sqrtf(a)
#define AA_FLOORF(a) __builtin_floorf(a)
#define AA_RINTF(a) __builtin_rintf(a)
#define AA_FMINF(a, b) __builtin_fminf((a), (b))
#define AA_FMAXF(a, b) __builtin_fmaxf((a), (b))
#define AA_FABSF(a) __builtin_fabsf(a)
#else
#include <math.h>
#define AA_FN static inline
#define AA_FMAF(a, b, c) fmaf((a), (b), (c))
#define AA_SQRTF(a) sqrtf(a)
#define AA_FLOORF(a) floorf(a)
#define AA_RINTF(a) rintf(a)
#define AA_FMINF(a, b) fminf((a), (b))
#define AA_FMAXF(a, b) fmaxf((a), (b))
#define AA_FABSF(a) fabsf(a)
#endif

AA_FN uint32_t aa_f2u(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
AA_FN float aa_u2f(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

#define AA_INF_F (aa_u2f(0x7f800000u))
#define AA_NAN_F (aa_u2f(0x7fc00000u))

/* ln(x).  x is split as m * 2^e with m in [sqrt(1/2), sqrt(2)); with
 * f = m - 1 (exact by Sterbenz), ln(1+f) = f + f^2 * Q(f), Q a degree-8
 * fit of (ln(1+f) - f) / f^2 on [-0.2929, 0.4143].  ln2 is split into a
 * 16-bit head and a binary32 tail so e*ln2_hi is exact. */
AA_FN float aa_logf(float x) {
  uint32_t u = aa_f2u(x);
  if (u == 0u || u == 0x80000000u) return -AA_INF_F;
  if (u >> 31) return AA_NAN_F;             /* x < 0 */
  if (u >= 0x7f800000u) return x;           /* +inf, NaN */
  int e = 0;
  if (u < 0x00800000u) {                    /* subnormal: scale by 2^25 */
    u = aa_f2u(x * 33554432.0f);
    e = -25;
  }
  e += (int)(u >> 23) - 127;
  uint32_t mb = (u & 0x007fffffu) | 0x3f800000u;
  if (mb > 0x3fb504f3u) { /* m > sqrt(2) */
    mb -= 0x00800000u;
    e += 1;
  }
  float f = aa_u2f(mb) - 1.0f;
  float q = -0.07866667956113815f;
  q = AA_FMAF(q, f, 0.12821748852729797f);
  q = AA_FMAF(q, f, -0.131016343832016f);
  q = AA_FMAF(q, f, 0.1418900042772293f);
  q = AA_FMAF(q, f, -0.1662759929895401f);
  q = AA_FMAF(q, f, 0.2000196874141693f);
  q = AA_FMAF(q, f, -0.2500077188014984f);
  q = AA_FMAF(q, f, 0.33333322405815125f);
  q = AA_FMAF(q, f, -0.4999999701976776f);
  float f2 = f * f;
  float ef = (float)e;
  float lo = AA_FMAF(ef, 1.4286068203094172e-06f, f2 * q);
  float r = lo + f;
  return AA_FMAF(ef, 0.693145751953125f, r);
}

/* ln(x) for the arguments a tracker step can produce, 1 - xi with xi an
 * LCG draw in [0, 1]: a normal number in [2^-24, 1] (x = 0 is
 * aa_log_step's).  The reduction and polynomial of aa_logf without its
 * branches, and a shorter tail: e is in [-24, 0] here, so e * ln2 in one
 * fused multiply-add with binary32 ln2 (error at most 24 * 1.9e-9, under
 * 0.03 ulp of the result) replaces the two-part product, and f + f^2 q
 * is one fused multiply-add.  Within 0.873 ulp of ln(x) on the whole
 * domain (checked exhaustively against the double-precision logarithm). */
AA_FN float aa_log_step_nz(float x) {
  uint32_t u = aa_f2u(x);
  /* The same m and e as aa_logf's "m > sqrt(2)" adjustment, from one
   * integer subtraction: u - bits(0x3f3504f4) has the exponent of
   * x / m in its top bits (arithmetic shift; m in [0x3f3504f4, 0x3fb504f3]
   * as a float), and removing that exponent from u leaves m.  x = 0 gives
   * e = -127, m = 1 as aa_logf does, before e is replaced by -inf. */
  const uint32_t tmp = u - 0x3f3504f4u;
  int e = (int)((int32_t)tmp >> 23);
  uint32_t mb = u - (tmp & 0xff800000u);
  float f = aa_u2f(mb) - 1.0f;
  float q = -0.07866667956113815f;
  q = AA_FMAF(q, f, 0.12821748852729797f);
  q = AA_FMAF(q, f, -0.131016343832016f);
  q = AA_FMAF(q, f, 0.1418900042772293f);
  q = AA_FMAF(q, f, -0.1662759929895401f);
  q = AA_FMAF(q, f, 0.2000196874141693f);
  q = AA_FMAF(q, f, -0.2500077188014984f);
  q = AA_FMAF(q, f, 0.33333322405815125f);
  q = AA_FMAF(q, f, -0.4999999701976776f);
  float f2 = f * f;
  float r = AA_FMAF(f2, q, f);
  return AA_FMAF((float)e, 0.693147182464599609375f, r);
}

/* The tracker-step logarithm on its whole domain: 0 (ln 0 = -inf) and
 * aa_log_step_nz's [2^-24, 1]. */
AA_FN float aa_log_step(float x) { return x == 0.0f ? -AA_INF_F : aa_log_step_nz(x); }

/* A delta-tracking free flight, t - ln(x) * unit (x = 1 - xi) with the
 * product and the difference rounded once (one fused multiply-add, as a
 * shader compiler may contract "t - (log(1 - xi) / majorant) * unit" with
 * majorant 1, deviceCodeRBF.hlsl:153).  x = 0 gives +inf. */
AA_FN float aa_flight(float t, float x, float unit) { return AA_FMAF(aa_log_step(x), -unit, t); }

/* e^x.  k = rint(x / ln2), r = x - k*ln2 (two-part, fused), e^r = 1 + r +
 * r^2 * P(r) with P a degree-5 fit on |r| <= ln2/2, then scaled by 2^k
 * through the exponent field (two factors when 2^k is outside the normal
 * range). */
AA_FN float aa_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return AA_INF_F;
  if (x < -103.97208404541016f) return 0.0f;
  float k = AA_RINTF(x * 1.4426950216293335f);
  float r = AA_FMAF(k, -0.693145751953125f, x);
  r = AA_FMAF(k, -1.4286068203094172e-06f, r);
  float p = 0.00019899275503121316f;
  p = AA_FMAF(p, r, 0.0013933643931522965f);
  p = AA_FMAF(p, r, 0.0083332983776927f);
  p = AA_FMAF(p, r, 0.04166646674275398f);
  p = AA_FMAF(p, r, 0.1666666716337204f);
  p = AA_FMAF(p, r, 0.5f);
  float y = AA_FMAF(r * r, p, r) + 1.0f;
  int ki = (int)k;
  if (ki > 127) { /* x just below the overflow threshold */
    y *= 2.0f;
    ki -= 1;
  }
  if (ki < -126) { /* gradual underflow: scale in two steps */
    y *= aa_u2f((uint32_t)(ki + 126 + 127) << 23);
    return y * aa_u2f(0x00800000u); /* 2^-126 */
  }
  return y * aa_u2f((uint32_t)(ki + 127) << 23);
}

/* x^y for x >= 0 (the renderer only raises colours and alphas, which are
 * non-negative).  x == 0 gives 0 for y > 0, x == 1 gives 1; otherwise
 * e^(y ln x).  Negative or NaN x gives NaN. */
AA_FN float aa_powf(float x, float y) {
  if (x != x || y != y) return AA_NAN_F;
  if (x < 0.0f) return AA_NAN_F;
  if (x == 1.0f || y == 0.0f) return 1.0f;
  if (x == 0.0f) return (y > 0.0f) ? 0.0f : AA_INF_F;
  return aa_expf(y * aa_logf(x));
}

/* ln(p) for p in [0, 1], the products of tracker steps below: 0 gives
 * -inf, a subnormal p is scaled by 2^25 first; the reduction, polynomial
 * and tail of aa_log_step_nz, whose binary32 ln2 in e * ln2 errs by at
 * most 126 * 1.9e-9, under 0.04 ulp of the result for e <= -1. */
AA_FN float aa_log_unit(float p) {
  uint32_t u = aa_f2u(p);
  int eo = 0;
  if (u < 0x00800000u) { /* 0 or subnormal */
    u = aa_f2u(p * 33554432.0f);
    eo = -25;
  }
  const uint32_t tmp = u - 0x3f3504f4u;
  const int e = (int)((int32_t)tmp >> 23) + eo;
  const uint32_t mb = u - (tmp & 0xff800000u);
  const float f = aa_u2f(mb) - 1.0f;
  float q = -0.07866667956113815f;
  q = AA_FMAF(q, f, 0.12821748852729797f);
  q = AA_FMAF(q, f, -0.131016343832016f);
  q = AA_FMAF(q, f, 0.1418900042772293f);
  q = AA_FMAF(q, f, -0.1662759929895401f);
  q = AA_FMAF(q, f, 0.2000196874141693f);
  q = AA_FMAF(q, f, -0.2500077188014984f);
  q = AA_FMAF(q, f, 0.33333322405815125f);
  q = AA_FMAF(q, f, -0.4999999701976776f);
  const float f2 = f * f;
  const float r = AA_FMAF(f2, q, f);
  const float l = AA_FMAF((float)e, 0.693147182464599609375f, r);
  return p == 0.0f ? -AA_INF_F : l;
}

/* Delta-tracking free flights (deviceCodeRBF.hlsl:153, majorant 1): the
 * position after a step is t = t_base - ln(P) * unit, P the product of
 * the steps' 1 - xi since the ray's last anchor: the ray's start, and
 * every step after which its step count i is a multiple of
 * AA_FLIGHT_BLOCK (i counts the steps of the primary and its shadow ray
 * together).  At an anchor t_base = t and P = 1.  ln(P) is the sum of the
 * steps' ln(1 - xi), so the positions are those of the reference's
 * statement, t -= ln(1 - xi) * unit per step, rounded differently: one
 * product per step and one fused multiply-add per position, which lets a
 * step in empty space be a multiplication and a compare (below). */
#define AA_FLIGHT_BLOCK 32u

/* The position after a step: fma(ln P, -unit, t_base), one rounding. */
AA_FN float aa_flight_pos(float tb, float p, float unit) { return AA_FMAF(aa_log_unit(p), -unit, tb); }

/* Cube used for alpha^3 (deviceCodeRBF.hlsl:375-376). */
AA_FN float aa_cube(float a) { return (a * a) * a; }

/* |a - b| with the summation order x, y, z (HLSL distance()). */
AA_FN float aa_dist3(float ax, float ay, float az, float bx, float by, float bz) {
  float dx = ax - bx, dy = ay - by, dz = az - bz;
  return AA_SQRTF((dx * dx + dy * dy) + dz * dz);
}

/* |a - b|^2 with the summation order of aa_dist3. */
AA_FN float aa_dist2(float ax, float ay, float az, float bx, float by, float bz) {
  float dx = ax - bx, dy = ay - by, dz = az - bz;
  return (dx * dx + dy * dy) + dz * dz;
}

/* The renderer's form of the RBF of sharedCode.h:137-140 (sigma = 3):
 * exp(-0.5 (3 d / r)^2) = exp(d^2 m) with the per-particle coefficient
 * m = -4.5 / r^2, and the particle covers a point where (3 d / r)^2 < 9,
 * i.e. d^2 m > -4.5 (the reference's distance(X, P) < r).  The point
 * queries evaluate e = aa_dist2(...) * m, test e > -4.5 and weight exp(e):
 * no square root and no division per test.  m = -inf for r = 0 (never
 * covers). */
AA_FN float aa_rbf_coef(float r) { return -4.5f / (r * r); }

/* Gaussian RBF of sharedCode.h:137-140 with sigma = 3:
 * exp(-0.5 * ((d * 3) / r)^2) (the splat path's form). */
AA_FN float aa_rbf(float d, float r) {
  float q = (d * 3.0f) / r;
  return aa_expf(-0.5f * (q * q));
}

/* aa_expf for -87 <= x <= 87 (no NaN, overflow or underflow handling):
 * the same operations on that range, so the same bits.  The renderer's
 * RBF weights take it (their argument lies in (-4.5, 0]). */
AA_FN float aa_expf_mid(float x) {
  float k = AA_RINTF(x * 1.4426950216293335f);
  float r = AA_FMAF(k, -0.693145751953125f, x);
  r = AA_FMAF(k, -1.4286068203094172e-06f, r);
  float p = 0.00019899275503121316f;
  p = AA_FMAF(p, r, 0.0013933643931522965f);
  p = AA_FMAF(p, r, 0.0083332983776927f);
  p = AA_FMAF(p, r, 0.04166646674275398f);
  p = AA_FMAF(p, r, 0.1666666716337204f);
  p = AA_FMAF(p, r, 0.5f);
  float y = AA_FMAF(r * r, p, r) + 1.0f;
  return y * aa_u2f((uint32_t)((int)k + 127) << 23);
}

/* A product bound for empty-space steps: any binary32 p > the result
 * places aa_flight_pos(tb, p, unit) below lim (inv_unit = 1 / unit,
 * correctly rounded).  With c = fl(fl(tb - lim) * inv_unit), the position
 * is below lim - |lim| 2^-22 (so it rounds below lim) once
 * ln p > (tb - lim + |lim| 2^-22) / unit + 4 ulp of the logarithm; the
 * three roundings in c, the ulp of lim and the 4 ulp together stay under
 * (|c| + |lim| inv_unit) 2^-19, which is added, and exp of the sum is
 * rounded up by (1 + 2^-20) to cover aa_expf's ulp.  The sum is clamped
 * to [-87, 0] (a larger bound stays a bound; above 0 no p <= 1 passes).
 * lim = +inf gives 0 (every step is in empty space), a NaN lim NaN (no
 * step is). */
AA_FN float aa_flight_bound(float tb, float lim, float inv_unit) {
  const float c = (tb - lim) * inv_unit;
  const float m = (AA_FABSF(c) + AA_FABSF(lim) * inv_unit) * 1.9073486328125e-06f; /* 2^-19 */
  const float b = aa_expf_mid(AA_FMINF(AA_FMAXF(c + m, -87.0f), 0.0f)) * 1.0000009536743164f; /* 1 + 2^-20 */
  return lim == AA_INF_F ? 0.0f : (lim == lim ? b : AA_NAN_F);
}


/* over (sharedCode.h:142-149): a composited over b. */
AA_FN void aa_over(const float a[4], const float b[4], float out[4]) {
  float ra = a[3] + b[3] * (1.0f - a[3]);
  if (ra == 0.0f) {
    out[0] = a[0];
    out[1] = a[1];
    out[2] = a[2];
    out[3] = a[3];
    return;
  }
  float r0 = (a[0] * a[3] + (b[0] * b[3]) * (1.0f - a[3])) / ra;
  float r1 = (a[1] * a[3] + (b[1] * b[3]) * (1.0f - a[3])) / ra;
  float r2 = (a[2] * a[3] + (b[2] * b[3]) * (1.0f - a[3])) / ra;
  out[0] = r0;
  out[1] = r1;
  out[2] = r2;
  out[3] = ra;
}

/* Whether the segment [tmin, tmax] of the ray o + t*d meets the box
 * [l, h] (slab test with the reciprocal direction i = 1/d; an empty box,
 * l.x > h.x, is never met).  This stands for the box test of a ray-tracing
 * traversal, which reports the primitives of every box the ray segment
 * meets. */
AA_FN int aa_seg_box(float ox, float oy, float oz, float ix, float iy, float iz, float tmin, float tmax,
                     float lx, float ly, float lz, float hx, float hy, float hz) {
  if (!(lx <= hx)) return 0;
  float x1 = (lx - ox) * ix, x2 = (hx - ox) * ix;
  float y1 = (ly - oy) * iy, y2 = (hy - oy) * iy;
  float z1 = (lz - oz) * iz, z2 = (hz - oz) * iz;
  float tn = AA_FMAXF(AA_FMAXF(AA_FMINF(x1, x2), AA_FMINF(y1, y2)), AA_FMINF(z1, z2));
  float tf = AA_FMINF(AA_FMINF(AA_FMAXF(x1, x2), AA_FMAXF(y1, y2)), AA_FMAXF(z1, z2));
  tn = AA_FMAXF(tn, tmin);
  tf = AA_FMINF(tf, tmax);
  return tn <= tf;
}

/* ParticleSplatIntersection's test (deviceCodeSplat.hlsl:164-181) for a
 * particle at c with radius r: the sample sits at t = |c - o| along the
 * ray, and hits when it lies within r of c and tmin <= t < tmax. */
AA_FN int aa_splat_hit(float cx, float cy, float cz, float r, float ox, float oy, float oz, float dx, float dy,
                       float dz, float tmin, float tmax, float *t_out) {
  float t = aa_dist3(cx, cy, cz, ox, oy, oz);
  float fx = cx - (ox + t * dx), fy = cy - (oy + t * dy), fz = cz - (oz + t * dz);
  *t_out = t;
  return ((fx * fx + fy * fy) + fz * fz) < r * r && tmin <= t && t < tmax;
}

/* IEEE binary16 <-> binary32 for the FP16 particle payload.  aa_h2f is
 * exact; aa_f2h_rn rounds to nearest even, aa_f2h_rz towards zero; both take
 * finite non-negative inputs below 65504 (colours, alphas, radii). */
AA_FN float aa_h2f(uint32_t h) {
  uint32_t e = (h >> 10) & 31u, m = h & 1023u;
  if (e == 0u) return (float)m * 5.9604644775390625e-08f; /* m * 2^-24, exact */
  return aa_u2f(((e + 112u) << 23) | (m << 13));
}
AA_FN uint32_t aa_f2h_rn(float x) {
  uint32_t u = aa_f2u(x);
  int e = (int)((u >> 23) & 255u) - 127;
  if (e < -14) { /* subnormal half: round x * 2^24 to an integer */
    float m = AA_RINTF(x * 16777216.0f);
    return (uint32_t)m; /* 1024 is the smallest normal */
  }
  uint32_t mant = u & 0x7fffffu, h = ((uint32_t)(e + 15) << 10) + (mant >> 13), rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
  return h;
}
AA_FN uint32_t aa_f2h_rz(float x) {
  uint32_t u = aa_f2u(x);
  int e = (int)((u >> 23) & 255u) - 127;
  if (e < -14) return (uint32_t)AA_FLOORF(x * 16777216.0f);
  return ((uint32_t)(e + 15) << 10) + ((u & 0x7fffffu) >> 13);
}

/* One LCG step (rng.h:52-57) and its [0,1] float (rng.h:59-61): the uint32
 * is converted with round-to-nearest-even, so states >= 0xFFFFFF80 give
 * exactly 1.0f. */
AA_FN float aa_lcg_next(uint32_t *state) {
  *state = *state * 1664525u + 1013904223u;
  return (float)(*state) * 2.3283064365386963e-10f; /* 2^-32 */
}

AA_FN uint32_t aa_murmur_mix(uint32_t h, uint32_t k) {
  k *= 0xcc9e2d51u;
  k = (k << 15) | (k >> 17);
  k *= 0x1b873593u;
  h ^= k;
  h = (h << 13) | (h >> 19);
  return h * 5u + 0xe6546b64u;
}

AA_FN uint32_t aa_murmur_fmix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

/* Per-pixel seed (rng.h:63-70): two murmur3 block mixes of the linear
 * pixel index and the frame id, then the murmur3 finaliser. */
AA_FN uint32_t aa_rng_seed(uint32_t frame_id, uint32_t px, uint32_t py, uint32_t width) {
  uint32_t h = aa_murmur_mix(0u, px + py * width);
  h = aa_murmur_mix(h, frame_id);
  return aa_murmur_fmix(h);
}

/* 1D linearly filtered lookup into an RGBA float table of n texels with
 * clamp-to-edge addressing, matching a normalised-coordinate texture fetch
 * (texel i centred at (i + 0.5) / n).  Returns component c of the sample;
 * the interpolation is a + f * (b - a), so constant tables are returned
 * exactly.  NaN coordinates read texel 0. */
/* aa_tf_coords with nf = (float)n supplied (a kernel keeps it uniform). */
AA_FN void aa_tf_coords_nf(float u, uint32_t n, float nf, uint32_t *j0, uint32_t *j1, float *f) {
  float x = u * nf - 0.5f;
  x = AA_FMAXF(x, -1.0f); /* also maps NaN to -1 */
  x = AA_FMINF(x, nf);
  float fl = AA_FLOORF(x);
  int i0 = (int)fl;
  *f = x - fl;
  int a = i0 < 0 ? 0 : (i0 > (int)n - 1 ? (int)n - 1 : i0);
  int b = i0 + 1 < 0 ? 0 : (i0 + 1 > (int)n - 1 ? (int)n - 1 : i0 + 1);
  *j0 = (uint32_t)a;
  *j1 = (uint32_t)b;
}

AA_FN void aa_tf_coords(float u, uint32_t n, uint32_t *j0, uint32_t *j1, float *f) {
  aa_tf_coords_nf(u, n, (float)n, j0, j1, f);
}

AA_FN float aa_lerp_tex(float a, float b, float f) { return a + f * (b - a); }

#endif /* AARBF_DETMATH_H */
