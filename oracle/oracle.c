/*
 * oracle.c -- CPU restatement of the reference render path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP
 * renderer: only the test suite, the smoke check and the benchmark's CPU
 * baseline load it.  Nothing in the product links it.
 *
 * It restates, as scalar C, the reference functions on the path:
 *   hilbert.h:16-88,124-225           Hilbert index <-> coordinates
 *   hostCode.cpp:183-250               space-time AABB, quantisation,
 *                                      sort, attribute normalisation
 *   deviceCodeBounds.hlsl:29-82        GenRBFBounds
 *   deviceCodeRBF.hlsl:58-203          ParticleTracker (marching, tracking)
 *   deviceCodeRBF.hlsl:206-336         ParticleRBFRayGen
 *   deviceCodeRBF.hlsl:338-389         intersection + any-hit accumulation
 *   sharedCode.h:127-185               light direction, RBF, over, slab test
 *   rng.h:20-70                        murmur3 seeding and LCG
 *   deviceCodeSplat.hlsl:29-194        splatting
 *   deviceCodeCommon.hlsl:27-43,67-115 TAA resolve
 *
 * Choices the reference leaves to the driver are fixed the same way the
 * renderer fixes them: transcendental functions from aarbf_detmath.h, the
 * density sum taken in ascending particle index, TF textures sampled with
 * linear filtering and clamp-to-edge addressing, normalize(v) = v/|v|.
 *
 * Point queries use a uniform grid whose cell lists are in ascending
 * particle order; the set of contributing particles is by definition
 * {i : |P_i - x| < r_i, colormap(w_i).a != 0}, which does not depend on
 * the leaf grouping or on the acceleration structure.
 *
 * Build: see oracle/Makefile (-O2 -ffp-contract=off, no SIMD intrinsics).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/aarbf_detmath.h"

#define ORC_EXPORT __attribute__((visibility("default")))
#define ORC_MAX_DEPTH 10000 /* sharedCode.h:27 */

/* ======================================================================
 * Hilbert curve, 3 axes x 16 bits (hilbert.h:181-225 and 124-166)
 * ====================================================================== */

static inline uint32_t rotr3(uint32_t v, uint32_t r) { return ((v >> r) | (v << (3 - r))) & 7u; }
static inline uint32_t rotl3(uint32_t v, uint32_t r) { return ((v << r) | (v >> (3 - r))) & 7u; }

/* Rotation update of hilbert.h:16-24 for 3 axes: advance by one plus the
 * position of the lowest set bit among the two low bits of `digit`. */
static inline uint32_t next_rotation(uint32_t rotation, uint32_t digit) {
  uint32_t low = digit & (0u - digit) & 3u;
  uint32_t shift = (low == 1u) ? 1u : (low == 2u ? 2u : 0u);
  uint32_t r = rotation + shift + 1u;
  return r >= 3u ? r - 3u : r;
}

/* hilbert_c2i(3, 16, coord) (hilbert.h:181-225).  The coordinates are
 * packed into one 48-bit word by shift-and-OR (hilbert.h:189-193), so a
 * 17-bit value 65536 leaks its top bit into bit 0 of the next axis (and out
 * of the word for z); the restatement keeps that packing. */
ORC_EXPORT uint64_t orc_hilbert_c2i(const uint64_t coord[3]) {
  uint64_t packed = ((coord[2] << 16 | coord[1]) << 16) | coord[0];
  uint32_t field[3];
  for (int d = 0; d < 3; ++d) field[d] = (uint32_t)((packed >> (16 * d)) & 0xffffu);
  /* level b's 3-bit digit holds bit b of axis d at bit d; each digit is
   * XORed with the digit of the next-coarser level (hilbert.h:203-204) */
  uint32_t digit[17];
  digit[16] = 0;
  for (int b = 0; b < 16; ++b)
    digit[b] = ((field[0] >> b) & 1u) | (((field[1] >> b) & 1u) << 1) | (((field[2] >> b) & 1u) << 2);
  uint64_t index = 0;
  uint32_t rotation = 0, flip = 0;
  for (int b = 15; b >= 0; --b) {
    uint32_t g = digit[b] ^ digit[b + 1];
    uint32_t h = rotr3(flip ^ g, rotation);
    index = (index << 3) | h;
    flip = 1u << rotation;
    rotation = next_rotation(rotation, h);
  }
  index ^= 0x124924924924ull; /* ones(48) / 7 >> 1 */
  for (int s = 1; s < 48; s *= 2) index ^= index >> s;
  return index;
}

/* hilbert_i2c(3, 16, index) (hilbert.h:124-166). */
ORC_EXPORT void orc_hilbert_i2c(uint64_t index, uint64_t coord[3]) {
  index ^= (index ^ 0x249249249249ull) >> 1;
  uint64_t inter = 0;
  uint32_t rotation = 0, flip = 0;
  for (int b = 15; b >= 0; --b) {
    uint32_t g = (uint32_t)(index >> (3 * b)) & 7u;
    inter = (inter << 3) | (rotl3(g, rotation) ^ flip);
    flip = 1u << rotation;
    rotation = next_rotation(rotation, g);
  }
  for (int s = 3; s < 48; s *= 2) inter ^= inter >> s;
  for (int d = 0; d < 3; ++d) {
    uint64_t c = 0;
    for (int b = 0; b < 16; ++b) c |= ((inter >> (3 * b + d)) & 1ull) << b;
    coord[d] = c;
  }
}

/* Quantisation of hostCode.cpp:200-207: ((p - min) / (max - min)) * 2^16,
 * truncated.  A zero-extent axis gives 0/0 in the reference (undefined on
 * conversion); here any NaN or negative value maps to 0. */
static inline uint64_t quantise16(float p, float lo, float hi) {
  float t = (p - lo) / (hi - lo);
  t = t * 65536.0f;
  if (!(t >= 0.0f)) return 0;
  return (uint64_t)t;
}

ORC_EXPORT void orc_hilbert_codes(const float *xyzw, uint64_t n, const float aabb[6], uint64_t *codes) {
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t c[3];
    for (int d = 0; d < 3; ++d) c[d] = quantise16(xyzw[4 * i + d], aabb[d], aabb[3 + d]);
    codes[i] = orc_hilbert_c2i(c);
  }
}

/* ======================================================================
 * Pre-process (hostCode.cpp:183-250)
 * ====================================================================== */

/* linalg::min / max per component (hostCode.cpp:187-192). */
ORC_EXPORT void orc_spacetime_aabb(const float *xyzw, uint64_t n, float aabb[6]) {
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      float v = xyzw[4 * i + d];
      aabb[d] = (aabb[d] < v) ? aabb[d] : v;         /* linalg::min: a < b ? a : b */
      aabb[3 + d] = (aabb[3 + d] < v) ? v : aabb[3 + d]; /* linalg::max: a < b ? b : a */
    }
}

/* Monotone key of a float: orders like '<' except that -0 sorts before +0
 * (std::sort leaves such ties in an unspecified order). */
static inline uint32_t float_key(float f) {
  uint32_t u = aa_f2u(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

typedef struct {
  uint64_t code;
  uint32_t k[4];
  uint32_t idx;
} sort_item;

static int cmp_items(const void *pa, const void *pb) {
  const sort_item *a = (const sort_item *)pa, *b = (const sort_item *)pb;
  if (a->code != b->code) return a->code < b->code ? -1 : 1;
  for (int j = 0; j < 4; ++j)
    if (a->k[j] != b->k[j]) return a->k[j] < b->k[j] ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* std::sort of pair<code, float4> (hostCode.cpp:133,212-221): by code, then
 * lexicographically by x, y, z, w; equal elements keep input order. */
ORC_EXPORT void orc_sort_particles(float *xyzw, uint64_t *codes, uint64_t n, uint32_t *perm) {
  sort_item *it = (sort_item *)malloc(sizeof(sort_item) * (n ? n : 1));
  float *tmp = (float *)malloc(sizeof(float) * 4 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {
    it[i].code = codes[i];
    for (int j = 0; j < 4; ++j) it[i].k[j] = float_key(xyzw[4 * i + j]);
    it[i].idx = (uint32_t)i;
  }
  qsort(it, n, sizeof(sort_item), cmp_items);
  memcpy(tmp, xyzw, sizeof(float) * 4 * n);
  for (uint64_t i = 0; i < n; ++i) {
    memcpy(xyzw + 4 * i, tmp + 4 * (uint64_t)it[i].idx, 16);
    codes[i] = it[i].code;
    if (perm) perm[i] = it[i].idx;
  }
  free(tmp);
  free(it);
}

/* std::min / std::max folds of hostCode.cpp:227-237. */
ORC_EXPORT void orc_attribute_range(const float *xyzw, uint64_t n, float range[2]) {
  for (uint64_t i = 0; i < n; ++i) {
    float w = xyzw[4 * i + 3];
    range[0] = (range[0] < w) ? range[0] : w; /* std::min(w, min) */
    range[1] = (range[1] < w) ? w : range[1]; /* std::max(w, max) */
  }
}

ORC_EXPORT void orc_normalize_attributes(float *xyzw, uint64_t n, const float range[2]) {
  if (!(range[1] > range[0])) return;
  float den = range[1] - range[0];
  for (uint64_t i = 0; i < n; ++i) xyzw[4 * i + 3] = (xyzw[4 * i + 3] - range[0]) / den;
}

/* ======================================================================
 * Transfer functions (hostCode.cpp:355-365,524-572)
 * ====================================================================== */

ORC_EXPORT void orc_tf_sample(const float *table, uint32_t n, float u, float out[4]) {
  uint32_t j0, j1;
  float f;
  aa_tf_coords(u, n, &j0, &j1, &f);
  for (int c = 0; c < 4; ++c) out[c] = aa_lerp_tex(table[4 * j0 + c], table[4 * j1 + c], f);
}

static inline float tf1(const float *table, uint32_t n, float u, int c) {
  uint32_t j0, j1;
  float f;
  aa_tf_coords(u, n, &j0, &j1, &f);
  return aa_lerp_tex(table[4 * j0 + c], table[4 * j1 + c], f);
}

/* ======================================================================
 * GenRBFBounds (deviceCodeBounds.hlsl:29-82)
 * ====================================================================== */

ORC_EXPORT void orc_gen_bounds(const float *xyzw, uint32_t numParticles, uint32_t ppl,
                               float rbfRadius, const float *radiusmap, uint32_t rm_n,
                               const float *radii, float *aabbs, uint32_t numAABBs) {
  const float nanv = AA_NAN_F;
  for (uint32_t c = 0; c < numAABBs; ++c) {
    float lo[3] = {AA_INF_F, AA_INF_F, AA_INF_F}, hi[3] = {-AA_INF_F, -AA_INF_F, -AA_INF_F};
    if ((uint64_t)c * ppl > numParticles) { /* ':45' uses '>' */
      for (int d = 0; d < 3; ++d) aabbs[6 * (uint64_t)c + d] = aabbs[6 * (uint64_t)c + 3 + d] = nanv;
      continue;
    }
    for (uint32_t i = 0; i < ppl; ++i) {
      uint64_t p = (uint64_t)c * ppl + i;
      if (p >= numParticles) break;
      const float *q = xyzw + 4 * p;
      float rp = tf1(radiusmap, rm_n, q[3], 0);
      if (rp == 0.0f) continue;
      float radius = radii ? radii[p] : rbfRadius;
      float e = radius * rp;
      for (int d = 0; d < 3; ++d) {
        float a = q[d] - e, b = q[d] + e;
        if (i == 0) {
          lo[d] = a;
          hi[d] = b;
        } else {
          lo[d] = fminf(lo[d], a);
          hi[d] = fmaxf(hi[d], b);
        }
      }
    }
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2])
      for (int d = 0; d < 3; ++d) lo[d] = hi[d] = nanv;
    for (int d = 0; d < 3; ++d) {
      aabbs[6 * (uint64_t)c + d] = lo[d];
      aabbs[6 * (uint64_t)c + 3 + d] = hi[d];
    }
  }
}

/* ======================================================================
 * Hilbert-packed BVH (checker for the device builder).  Level 0 holds one
 * box per leaf, widened per axis by 2^-20 of its largest magnitude so that
 * the box test never rejects a point whose distance test succeeds; NaN
 * (inactive) leaves become empty boxes (lo = +inf, hi = -inf).  Node j of
 * level l+1 is the exact union of nodes [jK, jK+K) of level l.  Boxes are
 * 8 floats: lo.xyz, 0, hi.xyz, 0.
 * ====================================================================== */

ORC_EXPORT uint32_t orc_bvh_num_levels(uint32_t num_leaves, uint32_t K) {
  uint32_t levels = 1, n = num_leaves;
  while (n > 1 || levels < 2) {
    n = (n + K - 1) / K;
    ++levels;
  }
  return levels;
}

/* child_is_leaf: `child` holds n_child reference leaf boxes (6 floats) and
 * the output is level 0; otherwise `child` is level l (8 floats per node)
 * and the output is level l+1. */
ORC_EXPORT void orc_bvh_build_level(const float *child, uint32_t child_is_leaf, uint32_t n_child,
                                    uint32_t K, float *out) {
  if (child_is_leaf) {
    for (uint32_t j = 0; j < n_child; ++j) {
      const float *b = child + 6 * (uint64_t)j;
      float *o = out + 8 * (uint64_t)j;
      int nan = 0;
      for (int d = 0; d < 6; ++d) nan |= (b[d] != b[d]);
      for (int d = 0; d < 3; ++d) {
        if (!nan && b[d] <= b[3 + d]) {
          float m = fmaxf(fabsf(b[d]), fabsf(b[3 + d]));
          o[d] = b[d] - m * 9.5367431640625e-07f; /* 2^-20 */
          o[4 + d] = b[3 + d] + m * 9.5367431640625e-07f;
        } else {
          o[d] = AA_INF_F;
          o[4 + d] = -AA_INF_F;
        }
      }
      o[3] = o[7] = 0.0f;
    }
    return;
  }
  uint32_t n = (n_child + K - 1) / K;
  for (uint32_t j = 0; j < n; ++j) {
    float lo[3] = {AA_INF_F, AA_INF_F, AA_INF_F}, hi[3] = {-AA_INF_F, -AA_INF_F, -AA_INF_F};
    for (uint32_t c = j * K; c < j * K + K && c < n_child; ++c) {
      const float *b = child + 8 * (uint64_t)c;
      for (int d = 0; d < 3; ++d) {
        lo[d] = fminf(lo[d], b[d]);
        hi[d] = fmaxf(hi[d], b[4 + d]);
      }
    }
    float *o = out + 8 * (uint64_t)j;
    for (int d = 0; d < 3; ++d) {
      o[d] = lo[d];
      o[4 + d] = hi[d];
    }
    o[3] = o[7] = 0.0f;
  }
}

/* ======================================================================
 * Renderer (deviceCodeRBF.hlsl)
 * ====================================================================== */

typedef struct orc_params {
  /* PushConstants (sharedCode.h:33-73) */
  uint32_t accumID, frameID, enableBlueNoise, showNoise, visualizeAttributes;
  float light_azimuth, light_elevation, light_ambient;
  float cam_pos[3], cam_d00[3], cam_du[3], cam_dv[3];
  float rbfRadius, unit;
  /* RayGenData (sharedCode.h:108-117) */
  int32_t width, height;
  float aabb_min[3], aabb_max[3];
  /* scene */
  const float *xyzw;      /* numParticles float4, sorted */
  const float *radii;     /* NULL = rbfRadius */
  uint32_t numParticles;
  const float *colormap, *radiusmap, *densitymap;
  uint32_t cm_n, rm_n, dm_n;
  const uint8_t *stbn;    /* RGBA8 atlas or NULL */
  int32_t stbn_w, stbn_h;
  uint32_t max_depth;     /* 0 = ORC_MAX_DEPTH */
  const float *baked;     /* NULL, or 8 floats per particle (x, y, z, r, alpha^3, rgb^2.2) used as is */
} orc_params;

typedef struct {
  float x, y, z, r;  /* position and effective radius (0 = never covers) */
  float a3;          /* colormap(w).a ^ 3 */
  float c22[3];      /* colormap(w).rgb ^ 2.2 */
  float m;           /* RBF coefficient -4.5 / r^2 (aa_rbf_coef) */
} baked_particle;

typedef struct {
  const orc_params *p;
  baked_particle *bp;
  float light[3];
  /* grid */
  float glo[3], ginv;
  int gn[3];
  uint32_t *cell_start; /* gn0*gn1*gn2 + 1 */
  uint32_t *cell_items;
  int empty;
} orc_scene;

/* getLightDirection (sharedCode.h:127-132), evaluated on the host with the
 * C library's sinf/cosf. */
ORC_EXPORT void orc_light_direction(float azimuth, float elevation, float out[3]) {
  float a = elevation * 3.14f - 0.5f * 3.14f;
  float b = azimuth * 2.0f * 3.14f;
  out[0] = sinf(a) * cosf(b);
  out[1] = sinf(a) * sinf(b);
  out[2] = cosf(a);
}

/* Per-particle values of the intersection/any-hit programs that depend
 * only on the particle (deviceCodeRBF.hlsl:350-351,366-376). */
static void bake(const orc_params *p, baked_particle *bp) {
  if (p->baked) { /* a payload given in decoded form */
    for (uint32_t i = 0; i < p->numParticles; ++i) {
      const float *b = p->baked + 8 * (uint64_t)i;
      bp[i].x = b[0];
      bp[i].y = b[1];
      bp[i].z = b[2];
      bp[i].r = b[3];
      bp[i].a3 = b[4];
      for (int k = 0; k < 3; ++k) bp[i].c22[k] = b[5 + k];
      bp[i].m = aa_rbf_coef(b[3]);
    }
    return;
  }
  for (uint32_t i = 0; i < p->numParticles; ++i) {
    const float *q = p->xyzw + 4 * (uint64_t)i;
    float rp = tf1(p->radiusmap, p->rm_n, q[3], 0);
    float radius = (p->radii ? p->radii[i] : p->rbfRadius) * rp;
    float c[4];
    orc_tf_sample(p->colormap, p->cm_n, q[3], c);
    baked_particle *b = bp + i;
    b->x = q[0];
    b->y = q[1];
    b->z = q[2];
    b->r = (c[3] == 0.0f) ? 0.0f : radius; /* transparent: ignoreHit ':369-372' */
    b->a3 = aa_cube(c[3]);
    for (int k = 0; k < 3; ++k) b->c22[k] = aa_powf(c[k], 2.2f);
    b->m = aa_rbf_coef(b->r);
  }
}

static int build_grid(orc_scene *s) {
  const orc_params *p = s->p;
  float lo[3] = {AA_INF_F, AA_INF_F, AA_INF_F}, hi[3] = {-AA_INF_F, -AA_INF_F, -AA_INF_F};
  float rmax = 0.0f;
  uint64_t live = 0;
  for (uint32_t i = 0; i < p->numParticles; ++i) {
    baked_particle *b = s->bp + i;
    if (!(b->r > 0.0f)) continue;
    float c[3] = {b->x, b->y, b->z};
    for (int d = 0; d < 3; ++d) {
      lo[d] = fminf(lo[d], c[d] - b->r);
      hi[d] = fmaxf(hi[d], c[d] + b->r);
    }
    rmax = fmaxf(rmax, b->r);
    ++live;
  }
  if (live == 0) {
    s->empty = 1;
    return 0;
  }
  float ext = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  float margin = ext * 1e-5f + 1e-30f;
  for (int d = 0; d < 3; ++d) {
    lo[d] -= margin;
    hi[d] += margin;
  }
  double vol = ((double)hi[0] - lo[0]) * ((double)hi[1] - lo[1]) * ((double)hi[2] - lo[2]);
  double h = cbrt(vol / (double)live) * 1.5;
  if (h < 2.0 * rmax) h = 2.0 * rmax;
  for (int d = 0; d < 3; ++d) {
    double cells = ((double)hi[d] - lo[d]) / h;
    if (cells > 512) cells = 512;
    s->gn[d] = (int)cells + 1;
  }
  double hmax = 0;
  for (int d = 0; d < 3; ++d) {
    double hd = ((double)hi[d] - lo[d]) / s->gn[d];
    if (hd > hmax) hmax = hd;
  }
  s->ginv = (float)(1.0 / hmax);
  for (int d = 0; d < 3; ++d) s->glo[d] = lo[d];
  uint64_t ncell = (uint64_t)s->gn[0] * s->gn[1] * s->gn[2];
  s->cell_start = (uint32_t *)calloc(ncell + 1, sizeof(uint32_t));
  if (!s->cell_start) return -1;
  /* pass 1: count, pass 2: fill in ascending particle order */
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      uint64_t acc = 0;
      for (uint64_t c = 0; c <= ncell; ++c) {
        uint32_t v = s->cell_start[c];
        s->cell_start[c] = (uint32_t)acc;
        acc += v;
      }
      s->cell_items = (uint32_t *)malloc(sizeof(uint32_t) * (acc ? acc : 1));
      if (!s->cell_items) return -1;
    }
    uint32_t *fill = NULL;
    if (pass == 1) {
      fill = (uint32_t *)malloc(sizeof(uint32_t) * (ncell ? ncell : 1));
      for (uint64_t c = 0; c < ncell; ++c) fill[c] = s->cell_start[c];
    }
    for (uint32_t i = 0; i < p->numParticles; ++i) {
      baked_particle *b = s->bp + i;
      if (!(b->r > 0.0f)) continue;
      float c[3] = {b->x, b->y, b->z};
      int c0[3], c1[3];
      for (int d = 0; d < 3; ++d) {
        float m = b->r * 1.0001f + margin;
        int a = (int)floorf((c[d] - m - s->glo[d]) * s->ginv);
        int e = (int)floorf((c[d] + m - s->glo[d]) * s->ginv);
        c0[d] = a < 0 ? 0 : (a >= s->gn[d] ? s->gn[d] - 1 : a);
        c1[d] = e < 0 ? 0 : (e >= s->gn[d] ? s->gn[d] - 1 : e);
      }
      for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
          for (int x = c0[0]; x <= c1[0]; ++x) {
            uint64_t cell = ((uint64_t)z * s->gn[1] + y) * s->gn[0] + x;
            if (pass == 0)
              s->cell_start[cell]++;
            else
              s->cell_items[fill[cell]++] = i;
          }
    }
    free(fill);
  }
  return 0;
}

typedef struct {
  uint32_t count;
  float density;
  float color[3];
} orc_payload;

/* TraceRay point query (deviceCodeRBF.hlsl:90-107) with the intersection
 * (':338-360') and any-hit (':362-389') programs inlined; contributions are
 * summed in ascending particle order. */
static void point_query(const orc_scene *s, const float x[3], orc_payload *pl) {
  pl->count = 0;
  pl->density = 0.0f;
  pl->color[0] = pl->color[1] = pl->color[2] = 0.0f;
  if (s->empty) return;
  int c[3];
  for (int d = 0; d < 3; ++d) {
    float g = (x[d] - s->glo[d]) * s->ginv;
    if (!(g >= 0.0f) || g >= (float)s->gn[d]) return;
    c[d] = (int)g;
    if (c[d] >= s->gn[d]) c[d] = s->gn[d] - 1;
  }
  uint64_t cell = ((uint64_t)c[2] * s->gn[1] + c[1]) * s->gn[0] + c[0];
  int clamp1 = !s->p->visualizeAttributes;
  for (uint32_t k = s->cell_start[cell]; k < s->cell_start[cell + 1]; ++k) {
    const baked_particle *b = s->bp + s->cell_items[k];
    /* distance(X, P) < r and the RBF in the form of aa_rbf_coef */
    float e = aa_dist2(b->x, b->y, b->z, x[0], x[1], x[2]) * b->m;
    if (!(e > -4.5f)) continue;
    float rbf = aa_expf(e);
    float wa = rbf * b->a3;
    pl->count += 1;
    pl->density += wa;
    for (int j = 0; j < 3; ++j) pl->color[j] += b->c22[j] * wa;
    if (clamp1) pl->density = fminf(pl->density, 1.0f); /* ':379-387' */
  }
}

/* aabbIntersection (sharedCode.h:151-185). */
static int aabb_intersection(const float o[3], const float dir[3], float tmin, float tmax,
                             const float lb[3], const float rt[3], float *tenter, float *texit) {
  float inv[3], t1[3], t2[3];
  for (int d = 0; d < 3; ++d) inv[d] = 1.0f / dir[d];
  for (int d = 0; d < 3; ++d) {
    t2[d] = (rt[d] - o[d]) * inv[d];
    t1[d] = (lb[d] - o[d]) * inv[d];
  }
  float te = fmaxf(fmaxf(fminf(t1[0], t2[0]), fminf(t1[1], t2[1])), fminf(t1[2], t2[2]));
  float tx = fminf(fminf(fmaxf(t1[0], t2[0]), fmaxf(t1[1], t2[1])), fmaxf(t1[2], t2[2]));
  te = fmaxf(te, tmin);
  tx = fminf(tx, tmax);
  *tenter = te;
  *texit = tx;
  int hit = 1;
  if (tx < 0.0f) hit = 0;
  if (te >= tx) hit = 0;
  return hit;
}

typedef struct {
  uint32_t i;        /* step counter shared by the primary and shadow rays */
  float random;      /* random.x */
  uint32_t rng;      /* tracking RNG */
  float albedo[4];
  float t;
  float LHS;
  int shadow;
  int marching;
  uint32_t max_depth;
  uint64_t steps;
} orc_tracker;

static void colour_of(const orc_scene *s, const orc_payload *pl, float rgb[3]) {
  const orc_params *p = s->p;
  for (int j = 0; j < 3; ++j) rgb[j] = aa_powf(pl->color[j] / pl->density, 1.0f / 2.2f);
  if (!p->visualizeAttributes)
    for (int j = 0; j < 3; ++j) rgb[j] = tf1(p->colormap, p->cm_n, pl->density, j);
}

/* stochasticMarching (deviceCodeRBF.hlsl:67-139). */
static void track_marching(const orc_scene *s, orc_tracker *tr, const float o[3], const float dir[3],
                           float tmax) {
  const orc_params *p = s->p;
  float RHS = -aa_logf(1.0f - tr->random * 0.99f);
  for (; tr->i < tr->max_depth; ++tr->i) {
    tr->steps++;
    tr->t = tr->t + p->unit;
    if (tr->t >= tmax) return;
    float x[3];
    for (int d = 0; d < 3; ++d) x[d] = o[d] + tr->t * dir[d];
    orc_payload pl;
    point_query(s, x, &pl);
    if (!(pl.count > 0 && pl.density > 0.0f)) continue;
    float density = tf1(p->densitymap, p->dm_n, pl.density, 0);
    tr->LHS += density;
    if (tr->LHS > RHS && !tr->shadow) {
      float rgb[3];
      colour_of(s, &pl, rgb);
      tr->albedo[0] = rgb[0];
      tr->albedo[1] = rgb[1];
      tr->albedo[2] = rgb[2];
      tr->albedo[3] = 1.0f;
      return;
    }
    if (tr->shadow) {
      tr->albedo[3] += 1.0f - aa_expf(-density);
      if (tr->albedo[3] >= 1.0f) {
        tr->albedo[3] = 1.0f;
        return;
      }
    }
  }
}

/* stochasticTracking (deviceCodeRBF.hlsl:141-203); majorant 1, so the
 * divisions by it are exact and omitted.  t restarts at TMin (':150').
 * The free flight of ':153' in the product form of aarbf_detmath.h
 * (aa_flight_pos): P collects the steps' 1 - xi since the anchor tb, the
 * ray's start and every step after which i is a multiple of
 * AA_FLIGHT_BLOCK. */
static void track_delta(const orc_scene *s, orc_tracker *tr, const float o[3], const float dir[3],
                        float tmax) {
  const orc_params *p = s->p;
  tr->t = 0.0f;
  float tb = 0.0f, P = 1.0f;
  for (; tr->i < tr->max_depth; ++tr->i) {
    tr->steps++;
    float xi = aa_lcg_next(&tr->rng);
    P = P * (1.0f - xi);
    tr->t = aa_flight_pos(tb, P, p->unit);
    if ((tr->i + 1u) % AA_FLIGHT_BLOCK == 0u) {
      tb = tr->t;
      P = 1.0f;
    }
    if (tr->t >= tmax) return;
    float x[3];
    for (int d = 0; d < 3; ++d) x[d] = o[d] + tr->t * dir[d];
    orc_payload pl;
    point_query(s, x, &pl);
    if (pl.count == 0) continue;
    float density = tf1(p->densitymap, p->dm_n, pl.density, 0);
    float xi2 = aa_lcg_next(&tr->rng);
    if (xi2 < density) {
      float rgb[3];
      colour_of(s, &pl, rgb);
      tr->albedo[0] = rgb[0];
      tr->albedo[1] = rgb[1];
      tr->albedo[2] = rgb[2];
      tr->albedo[3] = 1.0f;
      return;
    }
  }
}

static void track(const orc_scene *s, orc_tracker *tr, const float o[3], const float d[3], float tmax) {
  if (tr->marching)
    track_marching(s, tr, o, d, tmax);
  else
    track_delta(s, tr, o, d, tmax);
}

/* over (sharedCode.h:142-149). */
static void over(const float a[4], const float b[4], float out[4]) {
  float ra = a[3] + b[3] * (1.0f - a[3]);
  if (ra == 0.0f) {
    memcpy(out, a, 16);
    return;
  }
  for (int j = 0; j < 3; ++j) out[j] = (a[j] * a[3] + (b[j] * b[3]) * (1.0f - a[3])) / ra;
  out[3] = ra;
}

typedef struct {
  uint64_t steps, exhausted;
} orc_counts;

/* ParticleRBFRayGen (deviceCodeRBF.hlsl:206-336) for one pixel and one
 * (frameID, accumID); reads and updates accum[], writes image[]. */
static void raygen(const orc_scene *s, int px, int py, uint32_t frameID, uint32_t accumID,
                   float *image, float *accum, orc_counts *cnt) {
  const orc_params *p = s->p;
  const int W = p->width, H = p->height;
  const uint64_t ofs = (uint64_t)px + (uint64_t)W * py;
  uint32_t rng = aa_rng_seed(frameID, (uint32_t)px, (uint32_t)py, (uint32_t)W);
  float sx = ((float)px + 0.5f) / (float)W;
  float sy = ((float)py + 0.5f) / (float)H;
  float rt[3], lb[3], dir[3];
  for (int d = 0; d < 3; ++d) {
    rt[d] = p->aabb_max[d] + 2.0f * p->rbfRadius;
    lb[d] = p->aabb_min[d] - 2.0f * p->rbfRadius;
    dir[d] = (p->cam_d00[d] + sx * p->cam_du[d]) + sy * p->cam_dv[d];
  }
  float len = sqrtf((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
  for (int d = 0; d < 3; ++d) dir[d] = dir[d] / len;
  float tenter, texit;
  aabb_intersection(p->cam_pos, dir, 0.0f, 10000.0f, lb, rt, &tenter, &texit);

  orc_tracker tr;
  memset(&tr, 0, sizeof(tr));
  tr.max_depth = p->max_depth ? p->max_depth : ORC_MAX_DEPTH;
  float random[3];
  if (p->enableBlueNoise) {
    int gx = (int)(frameID % 8u), gy = (int)((frameID / 8u) % 16u);
    int cx = gx * 256 + px % 256, cy = gy * 256 + py % 256;
    uint8_t v = 0;
    if (p->stbn && cx < p->stbn_w && cy < p->stbn_h) v = p->stbn[4 * ((uint64_t)cy * p->stbn_w + cx)];
    random[0] = (float)v / 255.0f;
    tr.marching = 1;
  } else {
    random[0] = aa_lcg_next(&rng);
    tr.marching = 0;
    tr.rng = rng;
  }
  random[1] = aa_lcg_next(&rng);
  random[2] = aa_lcg_next(&rng);
  tr.random = random[0];
  float color[4] = {0, 0, 0, 0};
  if (tenter < texit) {
    tr.t = tenter + p->unit * random[1];
    track(s, &tr, p->cam_pos, dir, texit);
    float albedo[4];
    memcpy(albedo, tr.albedo, 16);
    float t = tr.t;
    float visibility = 1.0f;
    if (albedo[3] > 0.0f && p->light_ambient != 1.0f) {
      float so[3], ste, stx;
      for (int d = 0; d < 3; ++d) so[d] = p->cam_pos[d] + dir[d] * t;
      aabb_intersection(so, s->light, 0.0f, 10000.0f, lb, rt, &ste, &stx);
      memset(tr.albedo, 0, 16);
      tr.LHS = 0.0f;
      tr.shadow = 1;
      tr.t = p->enableBlueNoise ? -p->unit * random[2] : 0.0f;
      track(s, &tr, so, s->light, stx);
      visibility = 1.0f - tr.albedo[3];
    }
    if (albedo[3] == 1.0f) {
      for (int j = 0; j < 3; ++j)
        color[j] = (albedo[j] * visibility) * (1.0f - p->light_ambient) + albedo[j] * p->light_ambient;
      color[3] = 1.0f;
    }
  }
  if (tr.i >= tr.max_depth) cnt->exhausted++;
  cnt->steps += tr.steps;
  float bg[4] = {0, 0, 0, 1}, c2[4];
  over(color, bg, c2);
  float w1 = 1.0f / (float)accumID;
  float w0 = (float)((int)accumID - 1) / (float)accumID;
  float fin[4];
  for (int j = 0; j < 4; ++j) fin[j] = w1 * c2[j] + w0 * accum[4 * ofs + j];
  memcpy(accum + 4 * ofs, fin, 16);
  if (p->showNoise) {
    fin[0] = fin[1] = fin[2] = random[0];
    fin[3] = 1.0f;
  }
  memcpy(image + 4 * ofs, fin, 16);
}

static int scene_init_grid(orc_scene *s, const orc_params *p, int grid) {
  memset(s, 0, sizeof(*s));
  s->p = p;
  s->bp = (baked_particle *)malloc(sizeof(baked_particle) * (p->numParticles ? p->numParticles : 1));
  if (!s->bp) return -1;
  bake(p, s->bp);
  orc_light_direction(p->light_azimuth, p->light_elevation, s->light);
  return grid ? build_grid(s) : 0;
}

static int scene_init(orc_scene *s, const orc_params *p) { return scene_init_grid(s, p, 1); }

static void scene_free(orc_scene *s) {
  free(s->bp);
  free(s->cell_start);
  free(s->cell_items);
}

typedef struct {
  const orc_scene *s;
  const int32_t *pixels; /* pairs (x, y), or NULL = whole image row-major */
  uint64_t npix;
  uint32_t spp;
  float *image, *accum;
  uint64_t next;
  pthread_mutex_t lock;
  orc_counts cnt;
} orc_job;

static void *worker(void *arg) {
  orc_job *job = (orc_job *)arg;
  const orc_params *p = job->s->p;
  orc_counts local = {0, 0};
  for (;;) {
    pthread_mutex_lock(&job->lock);
    uint64_t begin = job->next;
    job->next += 64;
    pthread_mutex_unlock(&job->lock);
    if (begin >= job->npix) break;
    uint64_t end = begin + 64 < job->npix ? begin + 64 : job->npix;
    for (uint64_t k = begin; k < end; ++k) {
      int px, py;
      if (job->pixels) {
        px = job->pixels[2 * k];
        py = job->pixels[2 * k + 1];
      } else {
        px = (int)(k % (uint64_t)p->width);
        py = (int)(k / (uint64_t)p->width);
      }
      for (uint32_t sidx = 0; sidx < job->spp; ++sidx)
        raygen(job->s, px, py, (p->frameID + sidx) & 0xffffu, (p->accumID + sidx) & 0xffffu,
               job->image, job->accum, &local);
    }
  }
  pthread_mutex_lock(&job->lock);
  job->cnt.steps += local.steps;
  job->cnt.exhausted += local.exhausted;
  pthread_mutex_unlock(&job->lock);
  return NULL;
}

static int run_job(const orc_scene *s, const int32_t *pixels, uint64_t npix, uint32_t spp, float *image,
                   float *accum, int threads, uint64_t counts[2]) {
  const orc_params *p = s->p;
  orc_job job;
  memset(&job, 0, sizeof(job));
  job.s = s;
  job.pixels = pixels;
  job.npix = pixels ? npix : (uint64_t)p->width * p->height;
  job.spp = spp;
  job.image = image;
  job.accum = accum;
  pthread_mutex_init(&job.lock, NULL);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, &job);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&job.lock);
  if (counts) {
    counts[0] = job.cnt.steps;
    counts[1] = job.cnt.exhausted;
  }
  return 0;
}

/* Renders spp consecutive frames (frameID + s, accumID + s) for the listed
 * pixels (or the whole image) with `threads` host threads.  image/accum are
 * full W*H float4 images.  counts (optional) receives total tracker steps
 * and the number of samples that exhausted the step budget. */
ORC_EXPORT int orc_render(const orc_params *p, const int32_t *pixels, uint64_t npix, uint32_t spp,
                          float *image, float *accum, int threads, uint64_t counts[2]) {
  orc_scene s;
  if (scene_init(&s, p) != 0) {
    scene_free(&s);
    return -1;
  }
  run_job(&s, pixels, npix, spp, image, accum, threads, counts);
  scene_free(&s);
  return 0;
}

/* The same split in set-up (bake + grid) and rendering, so that the two can
 * be timed separately.  The params struct must outlive the scene. */
ORC_EXPORT void *orc_scene_new(const orc_params *p) {
  orc_scene *s = (orc_scene *)malloc(sizeof(orc_scene));
  if (!s) return NULL;
  if (scene_init(s, p) != 0) {
    scene_free(s);
    free(s);
    return NULL;
  }
  return s;
}

ORC_EXPORT int orc_scene_render(void *scene, const int32_t *pixels, uint64_t npix, uint32_t spp, float *image,
                                float *accum, int threads, uint64_t counts[2]) {
  return run_job((const orc_scene *)scene, pixels, npix, spp, image, accum, threads, counts);
}

ORC_EXPORT void orc_scene_free(void *scene) {
  if (!scene) return;
  scene_free((orc_scene *)scene);
  free(scene);
}

/* Point query exposed for tests: count, density, rgb sums at each point. */
ORC_EXPORT int orc_point_queries(const orc_params *p, const float *pts, uint64_t n, float *out5) {
  orc_scene s;
  if (scene_init(&s, p) != 0) {
    scene_free(&s);
    return -1;
  }
  for (uint64_t i = 0; i < n; ++i) {
    orc_payload pl;
    point_query(&s, pts + 3 * i, &pl);
    out5[5 * i + 0] = (float)pl.count;
    out5[5 * i + 1] = pl.density;
    out5[5 * i + 2] = pl.color[0];
    out5[5 * i + 3] = pl.color[1];
    out5[5 * i + 4] = pl.color[2];
  }
  scene_free(&s);
  return 0;
}

/* ======================================================================
 * Splatting (deviceCodeSplat.hlsl:29-194)
 *
 * ParticleSplatRayGen marches slabs of 15/8 = 1 unit (integer division)
 * from t = 0; for every slab that reaches past tenter it traces the ray
 * segment [max(tenter, tslab), min(texit, tslab + unit)], keeps the first
 * 15 hits the any-hit program sees, sorts them by t and composites them
 * front to back.  The traversal reports the particles of every leaf whose
 * box the segment meets (aa_seg_box on the leaf boxes widened by 2^-20
 * relative, as the BVH stores them) that pass the intersection test; the
 * hit order is ascending particle index (the reference's is the driver's
 * traversal order, which only matters when a slab has more than 15 hits).
 * The sphere radius is rbfRadius (the radius map is not used, ':166') or
 * the per-particle radius when radii are given.
 * ====================================================================== */

typedef struct {
  uint32_t p;
  float t;
} splat_hit;

static int cmp_hit_t(const void *a, const void *b) {
  const splat_hit *x = (const splat_hit *)a, *y = (const splat_hit *)b;
  if (x->t < y->t) return -1;
  if (x->t > y->t) return 1;
  return (x->p > y->p) - (x->p < y->p);
}

static int cmp_hit_p(const void *a, const void *b) {
  const splat_hit *x = (const splat_hit *)a, *y = (const splat_hit *)b;
  return (x->p > y->p) - (x->p < y->p);
}

/* hits: NULL, or receives the largest hit count of a slab and of a run of
 * 16 consecutive slabs among the slabs the pixel visits (before the first
 * 15 are kept), so tests can see which overflow cases a scene reaches. */
static void splat_pixel(const orc_params *p, const float *leaves /* 8 floats per widened leaf */, uint32_t nleaf,
                        uint32_t ppl, int px, int py, splat_hit *cand, splat_hit *slab, float out[4],
                        uint32_t hits[2]) {
  const int W = p->width, H = p->height;
  float sx = ((float)px + 0.5f) / (float)W, sy = ((float)py + 0.5f) / (float)H;
  float lb[3], rt[3], dir[3], inv[3];
  for (int d = 0; d < 3; ++d) {
    rt[d] = p->aabb_max[d] + p->rbfRadius; /* +- rbfRadius, not 2x (':66-67') */
    lb[d] = p->aabb_min[d] - p->rbfRadius;
    dir[d] = (p->cam_d00[d] + sx * p->cam_du[d]) + sy * p->cam_dv[d];
  }
  float len = sqrtf((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
  for (int d = 0; d < 3; ++d) {
    dir[d] = dir[d] / len;
    inv[d] = 1.0f / dir[d];
  }
  const float *o = p->cam_pos;
  float tenter, texit;
  aabb_intersection(o, dir, 0.0f, 10000.0f, lb, rt, &tenter, &texit);
  float res[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (tenter < texit) {
    /* every particle any slab can report: its leaf box meets [tenter, texit]
     * and its sample lies in the sphere with tenter <= t < texit */
    uint32_t nc = 0;
    for (uint32_t l = 0; l < nleaf; ++l) {
      const float *b = leaves + 8 * (uint64_t)l;
      if (!aa_seg_box(o[0], o[1], o[2], inv[0], inv[1], inv[2], tenter, texit, b[0], b[1], b[2], b[4], b[5], b[6]))
        continue;
      for (uint32_t q = l * ppl; q < (l + 1) * ppl && q < p->numParticles; ++q) {
        const float *c = p->xyzw + 4 * (uint64_t)q;
        float r = p->radii ? p->radii[q] : p->rbfRadius, t;
        if (aa_splat_hit(c[0], c[1], c[2], r, o[0], o[1], o[2], dir[0], dir[1], dir[2], tenter, texit, &t)) {
          cand[nc].p = q;
          cand[nc].t = t;
          ++nc;
        }
      }
    }
    qsort(cand, nc, sizeof(splat_hit), cmp_hit_t);
    const float pps = (float)(15 / 8);
    const float sp = pps * p->unit;
    const float fade = aa_expf(-p->unit * pps);
    float tslab = 0.0f;
    uint32_t lo = 0;
    uint32_t window[16] = {0}, wsum = 0, nslab = 0;
    while (tslab < texit) {
      if (res[3] > 0.99f) break;
      if (!(tslab + sp > tslab)) break; /* unit below half an ulp of t: stop */
      float tmin = fmaxf(tenter, tslab), tmax = fminf(texit, tslab + sp);
      if (tmax > tenter) {
        while (lo < nc && cand[lo].t < tmin) ++lo;
        uint32_t ns = 0;
        for (uint32_t k = lo; k < nc && cand[k].t < tmax; ++k) {
          const float *b = leaves + 8 * (uint64_t)(cand[k].p / ppl);
          if (aa_seg_box(o[0], o[1], o[2], inv[0], inv[1], inv[2], tmin, tmax, b[0], b[1], b[2], b[4], b[5], b[6]))
            slab[ns++] = cand[k];
        }
        if (hits) {
          wsum += ns - window[nslab & 15u];
          window[nslab & 15u] = ns;
          ++nslab;
          if (ns > hits[0]) hits[0] = ns;
          if (wsum > hits[1]) hits[1] = wsum;
        }
        /* the payload keeps the first 15 hits in traversal order, then
         * SplatPayload::sort orders them by t (stable bubble sort) */
        qsort(slab, ns, sizeof(splat_hit), cmp_hit_p);
        if (ns > 15) ns = 15;
        for (uint32_t a = 1; a < ns; ++a) {
          splat_hit h = slab[a];
          uint32_t b = a;
          while (b > 0 && slab[b - 1].t > h.t) {
            slab[b] = slab[b - 1];
            --b;
          }
          slab[b] = h;
        }
        for (uint32_t k = 0; k < ns; ++k) {
          const float *c = p->xyzw + 4 * (uint64_t)slab[k].p;
          float r = p->radii ? p->radii[slab[k].p] : p->rbfRadius;
          float t = slab[k].t;
          float x = o[0] + dir[0] * t, y = o[1] + dir[1] * t, z = o[2] + dir[2] * t;
          float drbf = aa_rbf(aa_dist3(x, y, z, c[0], c[1], c[2]), r);
          float col[4];
          if (p->visualizeAttributes) {
            orc_tf_sample(p->colormap, p->cm_n, c[3], col);
            col[3] = col[3] * tf1(p->densitymap, p->dm_n, drbf, 0);
          } else {
            float cm[4];
            orc_tf_sample(p->colormap, p->cm_n, drbf, cm);
            col[0] = cm[0];
            col[1] = cm[1];
            col[2] = cm[2];
            col[3] = tf1(p->densitymap, p->dm_n, drbf, 0);
          }
          col[3] = col[3] * fade;
          float nr[4];
          aa_over(res, col, nr);
          memcpy(res, nr, 16);
        }
      }
      tslab = tslab + sp;
    }
  }
  const float black[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  aa_over(res, black, out);
}

/* Renders the given pixels (all when pixels == NULL) into image (W*H float4).
 * leaf_aabbs: the GenRBFBounds output (6 floats per leaf). */
/* hits: NULL, or 2 counts per rendered pixel (see splat_pixel). */
ORC_EXPORT int orc_splat(const orc_params *p, const float *leaf_aabbs, uint32_t num_leaves, uint32_t ppl,
                         const int32_t *pixels, uint64_t npix, float *image, uint32_t *hits) {
  if (ppl == 0) return -1;
  float *leaves = (float *)malloc(sizeof(float) * 8 * ((uint64_t)num_leaves + 1));
  splat_hit *cand = (splat_hit *)malloc(sizeof(splat_hit) * ((uint64_t)p->numParticles + 1));
  splat_hit *slab = (splat_hit *)malloc(sizeof(splat_hit) * ((uint64_t)p->numParticles + 1));
  if (!leaves || !cand || !slab) {
    free(leaves);
    free(cand);
    free(slab);
    return -1;
  }
  orc_bvh_build_level(leaf_aabbs, 1, num_leaves, 2, leaves);
  const uint64_t n = pixels ? npix : (uint64_t)p->width * p->height;
  for (uint64_t i = 0; i < n; ++i) {
    int px = pixels ? pixels[2 * i] : (int)(i % (uint64_t)p->width);
    int py = pixels ? pixels[2 * i + 1] : (int)(i / (uint64_t)p->width);
    uint32_t *h = hits ? hits + 2 * i : NULL;
    if (h) h[0] = h[1] = 0;
    splat_pixel(p, leaves, num_leaves, ppl, px, py, cand, slab, image + 4 * ((uint64_t)px + (uint64_t)p->width * py),
                h);
  }
  free(leaves);
  free(cand);
  free(slab);
  return 0;
}

/* ======================================================================
 * FP64 quadrature of what delta tracking converges to
 *
 * ParticleRBFRayGen with delta tracking (majorant 1 per unit, accept with
 * probability D = densitymap(rho) wherever some particle covers the point)
 * samples a collision at t with density sigma(t) T(t), where sigma = D/unit
 * (0 where nothing covers) and T(t) = exp(-int_0^t sigma).  A collision
 * shades with the colour of the point times (amb + (1 - amb) T_L), T_L the
 * transmittance of the shadow ray to the light; a ray that leaves the box
 * is black.  The expected pixel is therefore
 *     E = int_0^texit sigma(t) T(t) albedo(t) (amb + (1 - amb) T_L(t)) dt,
 * which this section evaluates in double precision: the rays are split at
 * every sphere entry and exit (the covering set, and so the integrand, is
 * smooth in between), each piece into panels no longer than r_min / 8, and
 * every panel integrated with 8-point Gauss-Legendre (the optical depth to
 * each node by a nested 8-point rule).  The step budget (MAX_DEPTH) is
 * ignored: its exhaustion probability is reported by the renderer.
 * ====================================================================== */

static const double gl8_x[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                                -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                                0.7966664774136267,  0.9602898564975363};
static const double gl8_w[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                                0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};

static double tf1d(const float *table, uint32_t n, double u, int c) {
  double x = u * (double)n - 0.5;
  if (!(x > -1.0)) x = -1.0;
  if (x > (double)n) x = (double)n;
  double fl = floor(x);
  int i0 = (int)fl;
  double f = x - fl;
  int a = i0 < 0 ? 0 : (i0 > (int)n - 1 ? (int)n - 1 : i0);
  int b = i0 + 1 < 0 ? 0 : (i0 + 1 > (int)n - 1 ? (int)n - 1 : i0 + 1);
  double ta = table[4 * a + c], tb = table[4 * b + c];
  return ta + f * (tb - ta);
}

typedef struct {
  const orc_scene *s;
  double o[3], d[3];
  uint32_t *act; /* particles covering some point of the current piece */
  uint32_t nact;
} q_ray;

/* density sum, colour sum and count at o + t d over the active particles */
static void q_payload(const q_ray *r, double t, double *rho, double col[3], uint32_t *count) {
  const orc_scene *s = r->s;
  const orc_params *p = s->p;
  double x[3] = {r->o[0] + t * r->d[0], r->o[1] + t * r->d[1], r->o[2] + t * r->d[2]};
  *rho = 0.0;
  col[0] = col[1] = col[2] = 0.0;
  *count = 0;
  for (uint32_t k = 0; k < r->nact; ++k) {
    const baked_particle *b = s->bp + r->act[k];
    double dx = b->x - x[0], dy = b->y - x[1], dz = b->z - x[2];
    double dd = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(dd < b->r)) continue;
    double q = dd * 3.0 / b->r;
    double wa = exp(-0.5 * q * q) * b->a3;
    *count += 1;
    *rho += wa;
    for (int j = 0; j < 3; ++j) col[j] += b->c22[j] * wa;
    if (!p->visualizeAttributes && *rho > 1.0) *rho = 1.0;
  }
}

static double q_sigma(const q_ray *r, double t) {
  const orc_params *p = r->s->p;
  double rho, col[3];
  uint32_t count;
  q_payload(r, t, &rho, col, &count);
  if (count == 0) return 0.0;
  double D = tf1d(p->densitymap, p->dm_n, rho, 0);
  D = D < 0.0 ? 0.0 : (D > 1.0 ? 1.0 : D);
  return D / (double)p->unit;
}

typedef struct {
  double t0, t1;
  uint32_t p;
} q_span;

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* per-ray work arrays, grown on demand */
typedef struct {
  q_span *sp;
  double *cuts;
  uint32_t *act;
  uint64_t cap;
  int oom;
} q_scratch;

static int q_reserve(q_scratch *w, uint64_t n) {
  if (n <= w->cap) return 1;
  uint64_t c = w->cap ? w->cap : 1024;
  while (c < n) c *= 2;
  q_span *sp = (q_span *)realloc(w->sp, sizeof(q_span) * c);
  if (sp) w->sp = sp;
  double *cuts = (double *)realloc(w->cuts, sizeof(double) * 2 * c);
  if (cuts) w->cuts = cuts;
  uint32_t *act = (uint32_t *)realloc(w->act, sizeof(uint32_t) * c);
  if (act) w->act = act;
  if (!sp || !cuts || !act) {
    w->oom = 1;
    return 0;
  }
  w->cap = c;
  return 1;
}

static void q_scratch_free(q_scratch *w) {
  free(w->sp);
  free(w->cuts);
  free(w->act);
}

/* sphere spans of all particles along the ray over [0, tmax]; returns
 * count (0 with ws->oom set when the arrays cannot grow) */
static uint32_t q_spans(const orc_scene *s, const double o[3], const double d[3], double tmax, q_scratch *ws) {
  const orc_params *p = s->p;
  uint32_t n = 0;
  for (uint32_t i = 0; i < p->numParticles; ++i) {
    const baked_particle *b = s->bp + i;
    if (!(b->r > 0.0f)) continue;
    double v[3] = {b->x - o[0], b->y - o[1], b->z - o[2]};
    double tc = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];
    double h2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - tc * tc;
    double w2 = (double)b->r * b->r - h2;
    if (!(w2 > 0.0)) continue;
    double w = sqrt(w2);
    double a = tc - w, c = tc + w;
    if (c < 0.0 || a > tmax) continue;
    if (!q_reserve(ws, (uint64_t)n + 1)) return 0;
    ws->sp[n].t0 = a < 0.0 ? 0.0 : a;
    ws->sp[n].t1 = c > tmax ? tmax : c;
    ws->sp[n].p = i;
    ++n;
  }
  return n;
}

/* visits the pieces of [0, tmax] between sphere boundaries that have a
 * non-empty covering set, calling f(piece r, t0, t1, panel length) */
typedef void (*q_piece_fn)(q_ray *r, double a, double b, double h, void *user);

static void q_pieces(q_ray *r, double tmax, q_scratch *w, q_piece_fn f, void *user) {
  uint32_t ns = q_spans(r->s, r->o, r->d, tmax, w);
  q_span *sp = w->sp;
  double *cuts = w->cuts;
  uint32_t *act = w->act;
  uint32_t nc = 0;
  for (uint32_t k = 0; k < ns; ++k) {
    cuts[nc++] = sp[k].t0;
    cuts[nc++] = sp[k].t1;
  }
  qsort(cuts, nc, sizeof(double), cmp_double);
  r->act = act;
  for (uint32_t c = 0; c + 1 < nc; ++c) {
    double a = cuts[c], b = cuts[c + 1];
    if (!(b > a)) continue;
    double m = 0.5 * (a + b), rmin = 1e300;
    r->nact = 0;
    for (uint32_t k = 0; k < ns; ++k)
      if (sp[k].t0 <= m && m <= sp[k].t1) {
        act[r->nact++] = sp[k].p;
        double rr = r->s->bp[sp[k].p].r;
        if (rr < rmin) rmin = rr;
      }
    if (r->nact == 0) continue;
    f(r, a, b, rmin / 8.0, user);
  }
}

typedef struct {
  double tau;
} q_depth;

static void q_depth_piece(q_ray *r, double a, double b, double h, void *user) {
  q_depth *u = (q_depth *)user;
  int m = (int)ceil((b - a) / h);
  if (m < 1) m = 1;
  double step = (b - a) / m;
  for (int k = 0; k < m; ++k) {
    double lo = a + k * step, half = 0.5 * step, mid = lo + half;
    for (int j = 0; j < 8; ++j) u->tau += gl8_w[j] * half * q_sigma(r, mid + half * gl8_x[j]);
  }
}

/* optical depth from x along dir to where it leaves the clip box */
static double q_shadow_depth(const orc_scene *s, const double x[3], q_scratch *w) {
  const orc_params *p = s->p;
  float so[3] = {(float)x[0], (float)x[1], (float)x[2]}, lb[3], rt[3], te, tx;
  for (int d = 0; d < 3; ++d) {
    rt[d] = p->aabb_max[d] + 2.0f * p->rbfRadius;
    lb[d] = p->aabb_min[d] - 2.0f * p->rbfRadius;
  }
  aabb_intersection(so, s->light, 0.0f, 10000.0f, lb, rt, &te, &tx);
  if (!(tx > 0.0f)) return 0.0;
  q_ray r;
  r.s = s;
  for (int d = 0; d < 3; ++d) {
    r.o[d] = x[d];
    r.d[d] = s->light[d];
  }
  q_depth u = {0.0};
  q_pieces(&r, (double)tx, w, q_depth_piece, &u);
  return u.tau;
}

typedef struct {
  double T;      /* transmittance at the start of the next piece */
  double e[3];   /* expected colour */
  q_scratch *shadow;
} q_primary;

static void q_primary_piece(q_ray *r, double a, double b, double h, void *user) {
  q_primary *u = (q_primary *)user;
  const orc_scene *s = r->s;
  const orc_params *p = s->p;
  const double amb = p->light_ambient;
  int m = (int)ceil((b - a) / h);
  if (m < 1) m = 1;
  double step = (b - a) / m;
  for (int k = 0; k < m; ++k) {
    double lo = a + k * step, half = 0.5 * step, mid = lo + half;
    double panel_tau = 0.0;
    for (int j = 0; j < 8; ++j) {
      double t = mid + half * gl8_x[j];
      /* optical depth from the panel start to the node */
      double hh = 0.5 * (t - lo), mm = lo + hh, tau = 0.0;
      for (int i = 0; i < 8; ++i) tau += gl8_w[i] * hh * q_sigma(r, mm + hh * gl8_x[i]);
      double rho, col[3];
      uint32_t count;
      q_payload(r, t, &rho, col, &count);
      if (count == 0) continue;
      double D = tf1d(p->densitymap, p->dm_n, rho, 0);
      D = D < 0.0 ? 0.0 : (D > 1.0 ? 1.0 : D);
      double sig = D / (double)p->unit;
      if (sig == 0.0) continue;
      double alb[3];
      for (int c = 0; c < 3; ++c)
        alb[c] = p->visualizeAttributes ? pow(col[c] / rho, 1.0 / 2.2) : tf1d(p->colormap, p->cm_n, rho, c);
      double shade = amb;
      if (amb != 1.0) {
        double x[3] = {r->o[0] + t * r->d[0], r->o[1] + t * r->d[1], r->o[2] + t * r->d[2]};
        shade += (1.0 - amb) * exp(-q_shadow_depth(s, x, u->shadow));
      }
      double wgt = gl8_w[j] * half * sig * u->T * exp(-tau);
      for (int c = 0; c < 3; ++c) u->e[c] += wgt * alb[c] * shade;
    }
    for (int j = 0; j < 8; ++j) panel_tau += gl8_w[j] * half * q_sigma(r, mid + half * gl8_x[j]);
    u->T *= exp(-panel_tau);
  }
}

typedef struct {
  const orc_scene *s;
  const int32_t *pixels;
  uint64_t n;
  double *out;
  uint64_t next;
  int oom;
  pthread_mutex_t lock;
} q_job;

static void q_pixel(const orc_scene *s, int px, int py, q_scratch *w, q_scratch *w2, double *e) {
  const orc_params *p = s->p;
  float sx = ((float)px + 0.5f) / (float)p->width, sy = ((float)py + 0.5f) / (float)p->height;
  float rt[3], lb[3], dir[3];
  for (int d = 0; d < 3; ++d) {
    rt[d] = p->aabb_max[d] + 2.0f * p->rbfRadius;
    lb[d] = p->aabb_min[d] - 2.0f * p->rbfRadius;
    dir[d] = (p->cam_d00[d] + sx * p->cam_du[d]) + sy * p->cam_dv[d];
  }
  float len = sqrtf((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
  for (int d = 0; d < 3; ++d) dir[d] = dir[d] / len;
  float tenter, texit;
  aabb_intersection(p->cam_pos, dir, 0.0f, 10000.0f, lb, rt, &tenter, &texit);
  e[0] = e[1] = e[2] = 0.0;
  if (!(tenter < texit)) return;
  q_ray r;
  r.s = s;
  for (int d = 0; d < 3; ++d) {
    r.o[d] = p->cam_pos[d];
    r.d[d] = dir[d];
  }
  q_primary u = {1.0, {0.0, 0.0, 0.0}, w2};
  q_pieces(&r, (double)texit, w, q_primary_piece, &u);
  for (int c = 0; c < 3; ++c) e[c] = u.e[c];
}

static void *q_worker(void *arg) {
  q_job *job = (q_job *)arg;
  const orc_params *p = job->s->p;
  q_scratch w = {0}, w2 = {0};
  for (;;) {
    pthread_mutex_lock(&job->lock);
    uint64_t i = job->next++;
    pthread_mutex_unlock(&job->lock);
    if (i >= job->n) break;
    int px = job->pixels ? job->pixels[2 * i] : (int)(i % (uint64_t)p->width);
    int py = job->pixels ? job->pixels[2 * i + 1] : (int)(i / (uint64_t)p->width);
    q_pixel(job->s, px, py, &w, &w2, job->out + 3 * i);
    if (w.oom || w2.oom) break;
  }
  pthread_mutex_lock(&job->lock);
  if (w.oom || w2.oom) job->oom = 1;
  pthread_mutex_unlock(&job->lock);
  q_scratch_free(&w);
  q_scratch_free(&w2);
  return NULL;
}

/* Expected RGB of the given pixels (all when pixels == NULL) for delta
 * tracking (enableBlueNoise = 0); out = 3 doubles per pixel; pixels are
 * shared among `threads` workers (each pixel's sum is sequential, so the
 * result does not depend on the thread count). */
ORC_EXPORT int orc_expected(const orc_params *p, const int32_t *pixels, uint64_t npix, double *out, int threads) {
  if (p->enableBlueNoise) return -1;
  orc_scene s;
  if (scene_init_grid(&s, p, 0) != 0) return -1;
  q_job job;
  job.s = &s;
  job.pixels = pixels;
  job.n = pixels ? npix : (uint64_t)p->width * p->height;
  job.out = out;
  job.next = 0;
  job.oom = 0;
  pthread_mutex_init(&job.lock, NULL);
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  int started = 0;
  for (int k = 1; k < threads; ++k)
    if (pthread_create(&th[started], NULL, q_worker, &job) == 0) ++started;
  q_worker(&job);
  for (int k = 0; k < started; ++k) pthread_join(th[k], NULL);
  pthread_mutex_destroy(&job.lock);
  scene_free(&s);
  return job.oom ? -1 : 0;
}

/* ======================================================================
 * TAA resolve (deviceCodeCommon.hlsl:27-43,67-115); border texels are
 * clamped to the edge.  Writes taa[] and copies it to taa_prev[]
 * (hostCode.cpp:829-830).
 * ====================================================================== */

static void rgb2ycocg(const float c[3], float o[3]) {
  float co = c[0] - c[2];
  float t = c[2] + co / 2.0f;
  float cg = c[1] - t;
  float y = t + cg / 2.0f;
  o[0] = y;
  o[1] = co;
  o[2] = cg;
}

static void ycocg2rgb(const float c[3], float o[3]) {
  float t = c[0] - c[2] / 2.0f;
  float g = c[2] + t;
  float b = t - c[1] / 2.0f;
  float r = c[1] + b;
  o[0] = r;
  o[1] = g;
  o[2] = b;
}

/* sRGB electro-optical transfer function of an R8G8B8A8_SRGB texel
 * channel (the GUI layer, hostCode.cpp:328-330). */
static float srgb_decode8(uint32_t v) {
  float c = (float)v / 255.0f;
  return c <= 0.04045f ? c / 12.92f : aa_powf((c + 0.055f) / 1.055f, 2.4f);
}

static uint32_t to_unorm8(float v) {
  v = fminf(fmaxf(v, 0.0f), 1.0f); /* NaN -> 0 */
  return (uint32_t)(v * 255.0f + 0.5f);
}

/* CompositeGui (deviceCodeCommon.hlsl:67-115) for every pixel: the TAA
 * resolve into taa[] (unless disabled), the GUI layer composited on top
 * with over(), and the BGRA8 frame-buffer word (B | G<<8 | R<<16 | A<<24,
 * channels clamped to [0, 1] and rounded half up); then the host's history
 * copy taa -> taa_prev (hostCode.cpp:829-830).  gui (RGBA8 words, sRGB
 * RGB) and frame may be NULL. */
ORC_EXPORT void orc_taa_resolve(const float *image, float *taa, float *taa_prev, int W, int H,
                                int disable, const uint32_t *gui, uint32_t *frame) {
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      uint64_t ofs = (uint64_t)px + (uint64_t)W * py;
      const float *center = image + 4 * ofs;
      float out[4];
      if (!disable) {
        float mn[3], mx[3];
        rgb2ycocg(center, mn);
        memcpy(mx, mn, 12);
        for (int iy = -1; iy <= 1; ++iy)
          for (int ix = -1; ix <= 1; ++ix) {
            if (ix == 0 && iy == 0) continue;
            int qx = px + ix, qy = py + iy;
            qx = qx < 0 ? 0 : (qx >= W ? W - 1 : qx);
            qy = qy < 0 ? 0 : (qy >= H ? H - 1 : qy);
            float c[3];
            rgb2ycocg(image + 4 * ((uint64_t)qx + (uint64_t)W * qy), c);
            for (int j = 0; j < 3; ++j) {
              mn[j] = fminf(mn[j], c[j]);
              mx[j] = fmaxf(mx[j], c[j]);
            }
          }
        float old[4], y[3];
        memcpy(old, taa_prev + 4 * ofs, 16);
        rgb2ycocg(old, y);
        for (int j = 0; j < 3; ++j) y[j] = fmaxf(y[j], mn[j]);
        ycocg2rgb(y, old);
        rgb2ycocg(old, y);
        for (int j = 0; j < 3; ++j) y[j] = fminf(y[j], mx[j]);
        ycocg2rgb(y, old);
        /* lerp(old, center, .2f) */
        for (int j = 0; j < 4; ++j) out[j] = old[j] + 0.2f * (center[j] - old[j]);
        memcpy(taa + 4 * ofs, out, 16);
      } else {
        memcpy(out, center, 16);
      }
      if (frame) {
        float g[4] = {0.0f, 0.0f, 0.0f, 0.0f}, c[4];
        if (gui) {
          uint32_t v = gui[ofs];
          g[0] = srgb_decode8(v & 255u);
          g[1] = srgb_decode8((v >> 8) & 255u);
          g[2] = srgb_decode8((v >> 16) & 255u);
          g[3] = (float)(v >> 24) / 255.0f;
        }
        over(g, out, c);
        frame[ofs] = to_unorm8(c[2]) | (to_unorm8(c[1]) << 8) | (to_unorm8(c[0]) << 16) |
                     (to_unorm8(c[3]) << 24);
      }
    }
  /* the host's history copy, after the pass */
  memcpy(taa_prev, taa, 16 * (size_t)W * (size_t)H);
}

/* The shadow ray of ParticleRBFRayGen's next-event estimation under
 * stochastic marching (deviceCodeRBF.hlsl:287-312 with the tracker's
 * shadow branch, :131-137), from `origin` toward the light, starting at
 * t = t0 (the raygen passes -unit * random.z) with a fresh step budget.
 * Writes the visibility 1 - albedo.w; returns 0, or -1 when out of memory. */
ORC_EXPORT int orc_shadow_march(const orc_params *p, const float origin[3], float t0, float *visibility) {
  orc_scene s;
  if (scene_init(&s, p) != 0) return -1;
  float lb[3], rt[3], ste, stx;
  for (int d = 0; d < 3; ++d) {
    rt[d] = p->aabb_max[d] + 2.0f * p->rbfRadius;
    lb[d] = p->aabb_min[d] - 2.0f * p->rbfRadius;
  }
  aabb_intersection(origin, s.light, 0.0f, 10000.0f, lb, rt, &ste, &stx);
  orc_tracker tr;
  memset(&tr, 0, sizeof(tr));
  tr.max_depth = p->max_depth ? p->max_depth : ORC_MAX_DEPTH;
  tr.marching = 1;
  tr.shadow = 1;
  tr.t = t0;
  track(&s, &tr, origin, s.light, stx);
  scene_free(&s);
  *visibility = 1.0f - tr.albedo[3];
  return 0;
}

/* Deterministic math exposed for its accuracy tests: which = 0 log, 1 exp,
 * 2 pow(x, y), 3 rbf(d = x, r = y), 4 lcg draw from state bits of x,
 * 5 tracker-step log, 6 flight bound for anchor x, limit y, unit 1/2500. */
ORC_EXPORT void orc_math(int which, const float *x, const float *y, float *out, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    switch (which) {
    case 0: out[i] = aa_logf(x[i]); break;
    case 1: out[i] = aa_expf(x[i]); break;
    case 2: out[i] = aa_powf(x[i], y[i]); break;
    case 3: out[i] = aa_rbf(x[i], y[i]); break;
    case 5: out[i] = aa_log_step(x[i]); break;
    case 6: out[i] = aa_flight_bound(x[i], y[i], 2500.0f); break;
    default: {
      uint32_t s = aa_f2u(x[i]);
      out[i] = aa_lcg_next(&s);
    }
    }
  }
}

/* Number of binary32 x in [-87, 87] where aa_expf_mid and aa_expf differ
 * bitwise (every 13th binary32 in [-4.5, 4.5], every 1009th beyond). */
ORC_EXPORT uint64_t orc_expf_mid_mismatches(void) {
  uint64_t bad = 0;
  for (uint32_t u = 0u; u <= 0x42ae0000u; u += (u < 0x40900000u ? 13u : 1009u)) { /* [0, 87] */
    float x = aa_u2f(u), y = -x;
    if (aa_f2u(aa_expf_mid(x)) != aa_f2u(aa_expf(x))) ++bad;
    if (aa_f2u(aa_expf_mid(y)) != aa_f2u(aa_expf(y))) ++bad;
  }
  return bad;
}

/* Largest error of aa_log_step over its domain (every binary32 in
 * [2^-24, 1]) in ulps of the binary32 result, against the double-precision
 * logarithm; x where it occurs in *at.  ln(1) = 0 must be exact (-1 if
 * not). */
ORC_EXPORT double orc_log_step_max_ulp(float *at) {
  double mx = 0.0;
  if (aa_log_step(1.0f) != 0.0f) return -1.0;
  for (uint32_t u = 0x33800000u; u < 0x3f800000u; ++u) {
    const float x = aa_u2f(u);
    const double ref = log((double)x);
    const float r32 = fabsf((float)ref);
    const double ulp = (double)(nextafterf(r32, AA_INF_F) - r32);
    const double err = fabs((double)aa_log_step(x) - ref) / ulp;
    if (err > mx) {
      mx = err;
      *at = x;
    }
  }
  return mx;
}

/* Largest error of aa_log_unit over every stride-th binary32 in (0, 1)
 * (subnormals included), in ulps of the binary32 result, against the
 * double-precision logarithm; x where it occurs in *at.  ln 0 = -inf and
 * ln 1 = 0 must be exact (-1 if not). */
ORC_EXPORT double orc_log_unit_max_ulp(uint32_t stride, float *at) {
  double mx = 0.0;
  if (aa_log_unit(1.0f) != 0.0f || aa_log_unit(0.0f) != -AA_INF_F) return -1.0;
  for (uint32_t u = 1u; u < 0x3f800000u; u += stride) {
    const float x = aa_u2f(u);
    const double ref = log((double)x);
    const float r32 = fabsf((float)ref);
    const double ulp = (double)(nextafterf(r32, AA_INF_F) - r32);
    const double err = fabs((double)aa_log_unit(x) - ref) / ulp;
    if (err > mx) {
      mx = err;
      *at = x;
    }
  }
  return mx;
}

/* Soundness of aa_flight_bound: in n seeded trials (anchor tb, limit lim,
 * unit), every product p above the bound, the 64 binary32 values above it
 * and 64 random ones in (bound, 1], must place aa_flight_pos(tb, p, unit)
 * below lim.  Returns the number of such p that do not.  naive = 1 uses
 * exp((tb - lim) / unit) without the margins instead (a check that the
 * trials reach the boundary). */
ORC_EXPORT uint64_t orc_flight_bound_violations(uint64_t n, uint32_t seed, int naive) {
  uint64_t bad = 0;
  uint32_t s = seed;
  for (uint64_t k = 0; k < n; ++k) {
    const float unit = (k & 1) ? 0.0004f : 1e-5f + 1e-2f * aa_lcg_next(&s);
    const float tb = (k & 2) ? 5.0f * aa_lcg_next(&s) : 1e-3f * aa_lcg_next(&s);
    float lim;
    switch (k % 5) {
    case 0: lim = tb + unit * 40.0f * aa_lcg_next(&s); break;  /* within 40 mean steps */
    case 1: lim = tb + unit * aa_lcg_next(&s); break;          /* within one */
    case 2: { /* a few ulps above tb */
      lim = tb;
      for (uint32_t j = (uint32_t)(8.0f * aa_lcg_next(&s)); j; --j) lim = nextafterf(lim, AA_INF_F);
      break;
    }
    case 3: lim = tb + unit * 100.0f * aa_lcg_next(&s); break;
    default: lim = 2.0f * tb + 1e-6f; break;
    }
    const float inv = 1.0f / unit;
    const float b = naive ? expf((tb - lim) * inv) : aa_flight_bound(tb, lim, inv);
    float p = b;
    for (int j = 0; j < 128; ++j) {
      if (j < 64) p = nextafterf(p, AA_INF_F);
      else p = b + (1.0f - b) * aa_lcg_next(&s);
      if (!(p > b) || p > 1.0f) continue;
      if (!(aa_flight_pos(tb, p, unit) < lim)) ++bad;
    }
  }
  return bad;
}

/* get_rng(frame, pixel, dims).state (rng.h:63-70) and the first k draws. */
ORC_EXPORT uint32_t orc_rng_stream(uint32_t frame, uint32_t px, uint32_t py, uint32_t width, float *draws,
                                   uint32_t k) {
  uint32_t s = aa_rng_seed(frame, px, py, width);
  uint32_t s0 = s;
  for (uint32_t i = 0; i < k; ++i) draws[i] = aa_lcg_next(&s);
  return s0;
}

ORC_EXPORT uint32_t orc_murmur_mix(uint32_t h, uint32_t k) { return aa_murmur_mix(h, k); }
ORC_EXPORT uint32_t orc_murmur_fmix(uint32_t h) { return aa_murmur_fmix(h); }
