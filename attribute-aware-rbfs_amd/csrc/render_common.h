// render_common.h -- device code shared by the render kernels (render.hip,
// wavefront.hip): kernel arguments, transfer-function sampling, the BVH
// point query and segment walks, slab test, colour resolve.
#pragma once

#include <algorithm>
#include <cmath>

#include "aarbf_detmath.h"
#include "internal.h"

namespace aarbf {

// One view of a render call (the wavefront kernels read these from a
// device table, so a batch of views of uneven size shares one schedule).
// Its pixels are the tiles [tile0, tile0 + tiles_x * tiles_y) of the
// combined 8x8-tile pixel space.
struct ReqArgs {
  float3 pos, d00, du, dv, light;
  float amb;
  int W, H;
  uint32_t tiles_x, tile0;
  uint32_t frameID, accumID, spp;
  uint32_t blue, showNoise;
  const uint8_t *stbn;
  int stbn_w, stbn_h;
  float4 *image, *accum;
  uint64_t samples; // W * H * spp (completion count of the latency record)
};

// Kernel arguments of the RBF render kernels (one launch).
struct RenderArgs {
  uint32_t frameID, accumID, spp;
  uint32_t blue, showNoise, visAttr;
  uint32_t max_depth;
  int W, H;
  float amb, unit;
  float inv_unit; // 1 / unit, correctly rounded (aa_flight_bound)
  float3 light, pos, d00, du, dv, lb, rt;
  const float4 *pr; // baked x, y, z, radius (0 = never covers)
  const float4 *pc; // baked alpha^3, rgb^2.2
  const float *pm;  // baked RBF coefficient -4.5 / r^2 (both payload formats)
  // FP16 payload (q16 != 0): per particle 16 bytes -- x, y, z and radius
  // as 16-bit fixed point in the frame of its chunk of 32 particles, alpha
  // as 16-bit unorm, rgb^2.2 as binary16; per chunk qc[2c] = origin (xyz,
  // radius base), qc[2c+1] = per-axis step (xyz, radius step)
  const uint4 *qp;
  const float4 *qc;
  uint32_t q16;
  uint32_t numParticles, ppl;
  const float4 *cmap;
  const float4 *dmap;
  uint32_t cm_n, dm_n;
  float dm_nf; // (float)dm_n
  const uint8_t *stbn;
  int stbn_w, stbn_h;
  float4 *image, *accum;
  uint64_t *stats;
  BVHView bvh;
  // wavefront schedule: the views of a batch, the view of every tile, and
  // per view (samples finished, wall clock at the last) with [2 * nreq] =
  // the wall clock at the start, or null
  ReqArgs v0; // the view of a single-view call
  const ReqArgs *req;
  const uint32_t *tile_req;
  uint32_t nreq;
  unsigned long long *lat;
  // step kernel: a wave leaves its loop when fewer than this many of its
  // lanes are still stepping (0 = never), so that the next round packs the
  // rays it held into full waves
  uint32_t min_live;
};

// Default schedule (wavefront.hip): rounds of gather and step kernels.
int render_wavefront(aarbf_context *ctx, const RenderArgs &A, const ReqArgs *views, uint32_t nviews,
                     bool batch);

constexpr uint32_t kQChunkLog2 = 5; // particles per FP16 quantisation chunk: 32

// binary16 -> binary32 by the hardware conversion (exact, subnormal halves
// included: HIP kernels run with FP16 denormals enabled).
__device__ inline float h16(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }

// 16-bit fixed point of a chunk: origin + q * step with one rounding (the
// encoder evaluates the same expression, so it knows every decoded value).
__device__ inline float fx16(uint32_t q, float step, float origin) { return __builtin_fmaf((float)q, step, origin); }
// radius code 0 = radius 0 (never covers)
__device__ inline float fx16_r(uint32_t q, float step, float origin) { return q ? fx16(q, step, origin) : 0.0f; }
// alpha as 16-bit unorm
__device__ inline float un16(uint32_t q) { return (float)q * (1.0f / 65535.0f); }

// One free-flight draw of the delta tracker: advances the LCG state s and
// returns 1 - xi, bit for bit 1.0f - aa_lcg_next(&s), as one fused
// multiply-add: xi = (float)s * 2^-32 is exact, so fma((float)s, -2^-32, 1)
// rounds the same value as 1 - xi.  tests/cpp/flight_identity.c checks
// this over all 2^32 states.
__device__ inline float flight_x(uint32_t &s) {
  s = s * 1664525u + 1013904223u;
  return __builtin_fmaf((float)s, -2.3283064365386963e-10f, 1.0f);
}

// Baked particle p: position and radius, and (alpha^3, rgb^2.2).  Q: 0 =
// the payload A holds, 1 = FP32, 2 = FP16 (kernels instantiated per
// payload; A.q16 must match).
template <int Q = 0>
__device__ inline float4 ld_pr(const RenderArgs &A, uint32_t p) {
  if (Q == 1 || (Q == 0 && !A.q16)) return A.pr[p];
  const uint4 w = A.qp[p];
  const float4 o = A.qc[2 * (p >> kQChunkLog2)], s = A.qc[2 * (p >> kQChunkLog2) + 1];
  return make_float4(fx16(w.x & 0xffffu, s.x, o.x), fx16(w.x >> 16, s.y, o.y), fx16(w.y & 0xffffu, s.z, o.z),
                     fx16_r(w.y >> 16, s.w, o.w));
}
__device__ inline float4 ld_pc(const RenderArgs &A, uint32_t p) {
  if (!A.q16) return A.pc[p];
  const uint4 w = A.qp[p];
  return make_float4(aa_cube(un16(w.z & 0xffffu)), h16(w.z >> 16), h16(w.w & 0xffffu), h16(w.w >> 16));
}
// The FP16 chunk frame of particle p (origin, step), for ld_pr_in: the
// children of one BVH group (at most 16 consecutive leaves, starting at a
// multiple of the branching factor) lie in one 32-particle chunk, so a
// leaf batch loads the frame once.
__device__ inline void ld_chunk(const RenderArgs &A, uint32_t p, float4 &o, float4 &s) {
  o = A.qc[2 * (p >> kQChunkLog2)];
  s = A.qc[2 * (p >> kQChunkLog2) + 1];
}
// ld_pr with particle p's chunk frame already loaded (FP16 payload)
template <int Q = 0>
__device__ inline float4 ld_pr_in(const RenderArgs &A, uint32_t p, const float4 &o, const float4 &s) {
  if (Q == 1 || (Q == 0 && !A.q16)) return A.pr[p];
  const uint4 w = A.qp[p];
  return make_float4(fx16(w.x & 0xffffu, s.x, o.x), fx16(w.x >> 16, s.y, o.y), fx16(w.y & 0xffffu, s.z, o.z),
                     fx16_r(w.y >> 16, s.w, o.w));
}
// Everything a candidate test reads of particle p -- position and radius,
// colour payload and RBF coefficient -- with every load issued before any
// value is used (one memory round trip; the same values as ld_pr, ld_pc and
// A.pm[p]).  Q: 0 = the payload A holds, 1 = FP32, 2 = FP16 (for kernels
// instantiated per payload, which A.q16 must match).
template <int Q = 0>
__device__ inline void ld_cand(const RenderArgs &A, uint32_t p, float4 &pr, float4 &pc, float &m) {
  if (Q == 1 || (Q == 0 && !A.q16)) {
    pr = A.pr[p];
    pc = A.pc[p];
    m = A.pm[p];
    return;
  }
  const uint4 w = A.qp[p];
  const float4 o = A.qc[2 * (p >> kQChunkLog2)], s = A.qc[2 * (p >> kQChunkLog2) + 1];
  m = A.pm[p];
  pr = make_float4(fx16(w.x & 0xffffu, s.x, o.x), fx16(w.x >> 16, s.y, o.y), fx16(w.y & 0xffffu, s.z, o.z),
                   fx16_r(w.y >> 16, s.w, o.w));
  pc = make_float4(aa_cube(un16(w.z & 0xffffu)), h16(w.z >> 16), h16(w.w & 0xffffu), h16(w.w >> 16));
}

namespace {

// Per-block LDS copy of the BVH level table (offsets and node counts), so
// that the walk can index it by level without copying the kernel-argument
// array to scratch memory.
__shared__ uint32_t g_lvl_ofs[kMaxLevels];
__shared__ uint32_t g_lvl_cnt[kMaxLevels];
__shared__ uint32_t g_lvl_qofs[kMaxLevels];

__device__ inline void load_levels(const RenderArgs &A) {
  if (threadIdx.x < kMaxLevels) {
    g_lvl_ofs[threadIdx.x] = A.bvh.level_ofs[threadIdx.x];
    g_lvl_cnt[threadIdx.x] = A.bvh.counts[threadIdx.x];
    g_lvl_qofs[threadIdx.x] = A.bvh.group_ofs[threadIdx.x];
  }
  __syncthreads();
}

struct Counters {
  uint32_t samples, steps, queries, nodes, leaves, particles, covering, exhausted;
};

struct Payload {
  uint32_t count;
  float dens, r, g, b;
};

__device__ inline float4 tf4(const float4 *__restrict__ t, uint32_t n, float u) {
  uint32_t j0, j1;
  float f;
  aa_tf_coords(u, n, &j0, &j1, &f);
  float4 a = t[j0], b = t[j1];
  return make_float4(aa_lerp_tex(a.x, b.x, f), aa_lerp_tex(a.y, b.y, f), aa_lerp_tex(a.z, b.z, f),
                     aa_lerp_tex(a.w, b.w, f));
}

__device__ inline float tf_r(const float4 *__restrict__ t, uint32_t n, float u) {
  uint32_t j0, j1;
  float f;
  aa_tf_coords(u, n, &j0, &j1, &f);
  return aa_lerp_tex(t[j0].x, t[j1].x, f);
}

// Per-particle constants of the intersection/any-hit programs

__device__ inline bool inside(const float4 &lo, const float4 &hi, float x, float y, float z) {
  return (x >= lo.x) & (x <= hi.x) & (y >= lo.y) & (y <= hi.y) & (z >= lo.z) & (z <= hi.z);
}

// TraceRay point query (deviceCodeRBF.hlsl:90-107) + intersection
// (':338-360') + any-hit (':362-389').
template <bool STATS>
__device__ inline void point_query(const RenderArgs &A, float x, float y, float z, Payload &pl, Counters &c) {
  pl.count = 0;
  pl.dens = 0.0f;
  pl.r = pl.g = pl.b = 0.0f;
  const float4 *__restrict__ N = A.bvh.nodes;
  const uint32_t log2K = A.bvh.log2K, kmask = (1u << log2K) - 1u, top = A.bvh.top;
  const uint32_t *lofs = g_lvl_ofs, *lcnt = g_lvl_cnt;
  uint32_t level = top, idx = 0;
  if (STATS) c.queries++;
  for (;;) {
    const float4 *lv = N + lofs[level];
    float4 lo = lv[2 * idx], hi = lv[2 * idx + 1];
    if (STATS) c.nodes++;
    if (inside(lo, hi, x, y, z)) {
      if (level != 1) {
        --level;
        idx <<= log2K;
        continue;
      }
      uint32_t l0 = idx << log2K, l1 = min(l0 + kmask + 1u, lcnt[0]);
      for (uint32_t leaf = l0; leaf < l1; ++leaf) {
        if (A.ppl > 1) {
          const float4 *lb = N + lofs[0];
          if (STATS) c.leaves++;
          if (!inside(lb[2 * leaf], lb[2 * leaf + 1], x, y, z)) continue;
        }
        uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
        for (uint32_t p = p0; p < p1; ++p) {
          float4 q = ld_pr(A, p);
          if (STATS) c.particles++;
          const float e = aa_dist2(q.x, q.y, q.z, x, y, z) * A.pm[p];
          if (e > -4.5f) {
            float4 k = ld_pc(A, p);
            float wa = aa_expf_mid(e) * k.x;
            if (STATS) c.covering++;
            pl.count += 1;
            pl.dens += wa;
            pl.r += k.y * wa;
            pl.g += k.z * wa;
            pl.b += k.w * wa;
            if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
          }
        }
      }
    }
    // next node in depth-first order
    for (;;) {
      if (level == top) return;
      if ((idx & kmask) != kmask && idx + 1 < lcnt[level]) {
        ++idx;
        break;
      }
      ++level;
      idx >>= log2K;
    }
  }
}


// aabbIntersection (sharedCode.h:151-185).
__device__ inline void aabb_intersection(float3 o, float3 d, float tmin, float tmax, float3 lb, float3 rt,
                                         float &te, float &tx) {
  float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
  float t2 = (rt.x - o.x) * ix, t4 = (rt.y - o.y) * iy, t6 = (rt.z - o.z) * iz;
  float t1 = (lb.x - o.x) * ix, t3 = (lb.y - o.y) * iy, t5 = (lb.z - o.z) * iz;
  te = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
  tx = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
  te = fmaxf(te, tmin);
  tx = fminf(tx, tmax);
}

__device__ inline float3 colour_of(const RenderArgs &A, const Payload &pl) {
  float3 c;
  if (A.visAttr) {
    c.x = aa_powf(pl.r / pl.dens, 1.0f / 2.2f);
    c.y = aa_powf(pl.g / pl.dens, 1.0f / 2.2f);
    c.z = aa_powf(pl.b / pl.dens, 1.0f / 2.2f);
  } else {
    float4 t = tf4(A.cmap, A.cm_n, pl.dens);
    c = make_float3(t.x, t.y, t.z);
  }
  return c;
}

// ParticleTracker::track (deviceCodeRBF.hlsl:58-203).  `i` is the step
__device__ inline uint64_t wave_sum(uint32_t v) {
  uint64_t s = v;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// One lane per pixel, spp samples in sequence; 256-thread blocks cover
// 16x16 pixels, each wave an 8x8 tile.

struct SegQuery {
  float ax, ay, az; // segment start
  float dx, dy, dz; // unit direction
  float ix, iy, iz; // 1/d with zero components replaced by +-1e30
  float len;        // segment length
  float m;          // conservative margin for rounding of o + t*d and |P-x|
};

__device__ inline float safe_inv(float d) { return 1.0f / (fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d); }

// Slab test that also returns where the segment enters the (widened) box,
// relative to the segment start.
__device__ inline bool seg_box_t(const SegQuery &s, float4 lo, float4 hi, float &tnear) {
  float t1 = (lo.x - s.m - s.ax) * s.ix, t2 = (hi.x + s.m - s.ax) * s.ix;
  float nr = fminf(t1, t2), fr = fmaxf(t1, t2);
  t1 = (lo.y - s.m - s.ay) * s.iy;
  t2 = (hi.y + s.m - s.ay) * s.iy;
  nr = fmaxf(nr, fminf(t1, t2));
  fr = fminf(fr, fmaxf(t1, t2));
  t1 = (lo.z - s.m - s.az) * s.iz;
  t2 = (hi.z + s.m - s.az) * s.iz;
  nr = fmaxf(nr, fminf(t1, t2));
  fr = fminf(fr, fmaxf(t1, t2));
  tnear = fmaxf(nr, 0.0f);
  return tnear <= fminf(fr, s.len);
}

__device__ inline bool seg_box(const SegQuery &s, float4 lo, float4 hi) {
  if (!(lo.x <= hi.x)) return false; // empty node
  float t1 = (lo.x - s.m - s.ax) * s.ix, t2 = (hi.x + s.m - s.ax) * s.ix;
  float nr = fminf(t1, t2), fr = fmaxf(t1, t2);
  t1 = (lo.y - s.m - s.ay) * s.iy;
  t2 = (hi.y + s.m - s.ay) * s.iy;
  nr = fmaxf(nr, fminf(t1, t2));
  fr = fminf(fr, fmaxf(t1, t2));
  t1 = (lo.z - s.m - s.az) * s.iz;
  t2 = (hi.z + s.m - s.az) * s.iz;
  nr = fmaxf(nr, fminf(t1, t2));
  fr = fminf(fr, fmaxf(t1, t2));
  return fmaxf(nr, 0.0f) <= fminf(fr, s.len);
}


// ---- batched walk ------------------------------------------------------------
// Depth-first like the stackless walk above, but a hit node tests all K of
// its children with independent loads (one memory round trip per level
// instead of one per child).  The hit mask of each level still to be
// visited is kept in a per-lane LDS column (stk[level * STRIDE]); children
// are taken lowest bit first, so leaves are still visited in ascending
// order.  Level-1 hits hand their K leaves to `leaves(base, m)`.

constexpr uint32_t kStackLevels = 16;

template <int STRIDE, bool STATS, class BoxTest, class Leaves>
__device__ inline void walk_batched(const RenderArgs &A, uint16_t *stk, BoxTest test, Leaves leaves, Counters &c) {
  const float4 *__restrict__ N = A.bvh.nodes;
  const uint32_t log2K = A.bvh.log2K, K = 1u << log2K, top = A.bvh.top;
  const float4 *root = N + g_lvl_ofs[top];
  if (STATS) c.nodes++;
  if (!test(root[0], root[1])) return;
  if (top == 1) {
    leaves(0u, min(K, g_lvl_cnt[0]));
    return;
  }
  auto group = [&](uint32_t lvl, uint32_t base) -> uint32_t {
    const float4 *lv = N + g_lvl_ofs[lvl] + 2 * (uint64_t)base;
    uint32_t m = min(K, g_lvl_cnt[lvl] - base), mask = 0;
    if (STATS) c.nodes += m;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k)
      if (k < m && test(lv[2 * k], lv[2 * k + 1])) mask |= 1u << k;
    for (uint32_t k = 8; k < m; ++k) // K = 16
      if (test(lv[2 * k], lv[2 * k + 1])) mask |= 1u << k;
    return mask;
  };
  uint32_t lvl = top - 1, base = 0;
  uint32_t mask = group(lvl, 0);
  for (;;) {
    if (mask == 0) {
      if (lvl == top - 1) return;
      ++lvl;
      mask = stk[lvl * STRIDE];
      base = ((base >> log2K) >> log2K) << log2K;
      continue;
    }
    uint32_t bit = (uint32_t)__ffs(mask) - 1u;
    mask &= mask - 1u;
    uint32_t node = base + bit;
    if (lvl == 1) {
      uint32_t lb = node << log2K;
      if (leaves(lb, min(K, g_lvl_cnt[0] - lb))) return;
      continue;
    }
    stk[lvl * STRIDE] = (uint16_t)mask;
    --lvl;
    base = node << log2K;
    mask = group(lvl, base);
  }
}

// ---- quantised walk ------------------------------------------------------------
// The same depth-first order over the QGroup copy of the tree (K <= 8): a
// visited node costs one 80-byte group load, and its children's boxes are
// dequantised (lo + q * step, rounded outwards at build time) before the
// box test.  Level-0 hits are handed to `leaves(base, mask)` all at once.

__device__ inline uint32_t qbyte(const QGroup &g, uint32_t field, uint32_t k) {
  return (g.q[2 * field + (k >> 2)] >> ((k & 3u) * 8u)) & 255u;
}

template <int STRIDE, bool STATS, class BoxTest, class Leaves, class StkT>
__device__ inline void walk_q(const RenderArgs &A, StkT *stk, BoxTest test, Leaves leaves, Counters &c) {
  const QGroup *__restrict__ Q = A.bvh.groups;
  const uint32_t log2K = A.bvh.log2K, K = 1u << log2K, top = A.bvh.top;
  const float4 *root = A.bvh.nodes + g_lvl_ofs[top];
  if (STATS) c.nodes++;
  if (!test(root[0], root[1])) return;
  auto group = [&](uint32_t lvl, uint32_t node) -> uint32_t {
    const QGroup g = Q[g_lvl_qofs[lvl] + node];
    uint32_t first = node << log2K;
    uint32_t m = min(K, g_lvl_cnt[lvl - 1] - first), mask = 0;
    if (STATS) c.nodes += m;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      uint32_t lx = qbyte(g, 0, k), hx = qbyte(g, 3, k);
      if (k < m && lx <= hx) {
        float4 lo, hi;
        lo.x = g.a.x + (float)lx * g.a.w;
        lo.y = g.a.y + (float)qbyte(g, 1, k) * g.b.x;
        lo.z = g.a.z + (float)qbyte(g, 2, k) * g.b.y;
        hi.x = g.a.x + (float)hx * g.a.w;
        hi.y = g.a.y + (float)qbyte(g, 4, k) * g.b.x;
        hi.z = g.a.z + (float)qbyte(g, 5, k) * g.b.y;
        lo.w = hi.w = 0.0f;
        if (test(lo, hi)) mask |= 1u << k;
      }
    }
    return mask;
  };
  uint32_t lvl = top - 1, base = 0;
  uint32_t mask = group(top, 0);
  for (;;) {
    if (mask == 0) {
      if (lvl == top - 1) return;
      ++lvl;
      mask = stk[lvl * STRIDE];
      base = ((base >> log2K) >> log2K) << log2K;
      continue;
    }
    if (lvl == 0) {
      if (leaves(base, mask)) return;
      mask = 0;
      continue;
    }
    uint32_t bit = (uint32_t)__ffs(mask) - 1u;
    mask &= mask - 1u;
    uint32_t node = base + bit;
    stk[lvl * STRIDE] = (uint16_t)mask;
    base = node << log2K;
    mask = group(lvl, node);
    --lvl;
  }
}

// Child ordering keys: the entry distance (>= 0, so its bits order as
// unsigned integers; the sign bit is cleared for -0) with its low three bits
// replaced by the child number, so one unsigned min / max orders
// (distance, child) pairs; distances within 8 ulps order by child.  A miss
// sorts after every hit.
constexpr uint32_t kMissKey = 0xfffffff8u;
__device__ inline uint32_t entry_key(float t, uint32_t k) { return ((__float_as_uint(t) & 0x7ffffff8u) | k); }

// Batcher's odd-even merge sort network for 8 keys (fixed indices: the keys
// stay in registers)
__device__ inline void sort8(uint32_t key[8]) {
#define AARBF_KSWAP(i, j)                                                                                   \
  {                                                                                                          \
    const uint32_t a_ = key[i], b_ = key[j];                                                                 \
    key[i] = min(a_, b_);                                                                                    \
    key[j] = max(a_, b_);                                                                                    \
  }
  AARBF_KSWAP(0, 1) AARBF_KSWAP(2, 3) AARBF_KSWAP(4, 5) AARBF_KSWAP(6, 7)
  AARBF_KSWAP(0, 2) AARBF_KSWAP(1, 3) AARBF_KSWAP(4, 6) AARBF_KSWAP(5, 7)
  AARBF_KSWAP(1, 2) AARBF_KSWAP(5, 6) AARBF_KSWAP(0, 4) AARBF_KSWAP(3, 7)
  AARBF_KSWAP(1, 5) AARBF_KSWAP(2, 6) AARBF_KSWAP(1, 4) AARBF_KSWAP(3, 6)
  AARBF_KSWAP(2, 4) AARBF_KSWAP(3, 5) AARBF_KSWAP(3, 4)
#undef AARBF_KSWAP
}

// Children of quantised node `node` at level `lvl` that the segment q
// enters, in order of where it enters their boxes: nibbles (child + 1),
// first child in the lowest nibble (K <= 8).  The front-to-back walks use
// it to reach the nearest particles first, so their t_cut prunes the rest.
__device__ inline uint32_t ftb_children(const RenderArgs &A, const SegQuery &q, uint32_t lvl, uint32_t node) {
  const QGroup g = A.bvh.groups[g_lvl_qofs[lvl] + node];
  const uint32_t log2K = A.bvh.log2K, K = 1u << log2K;
  const uint32_t first = node << log2K;
  const uint32_t mcnt = min(K, g_lvl_cnt[lvl - 1] - first);
  // The slab parameters of a child bound b = origin + code * step are
  // (b -+ m - start) / d = code * (step / d) + (origin -+ m - start) / d:
  // per group a scale and two offsets per axis, then one fused
  // multiply-add per bound.  Rounding moves a parameter by ulps, far
  // inside the margin m, so the test stays conservative.
  const float sx = g.a.w * q.ix, sy = g.b.x * q.iy, sz = g.b.y * q.iz;
  const float lx0 = (g.a.x - q.m - q.ax) * q.ix, hx0 = (g.a.x + q.m - q.ax) * q.ix;
  const float ly0 = (g.a.y - q.m - q.ay) * q.iy, hy0 = (g.a.y + q.m - q.ay) * q.iy;
  const float lz0 = (g.a.z - q.m - q.az) * q.iz, hz0 = (g.a.z + q.m - q.az) * q.iz;
  uint32_t key[8];
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    key[k] = kMissKey | k;
    const uint32_t lx = qbyte(g, 0, k), hx = qbyte(g, 3, k);
    if (k < mcnt && lx <= hx) {
      float t1 = __builtin_fmaf((float)lx, sx, lx0), t2 = __builtin_fmaf((float)hx, sx, hx0);
      float nr = fminf(t1, t2), fr = fmaxf(t1, t2);
      t1 = __builtin_fmaf((float)qbyte(g, 1, k), sy, ly0);
      t2 = __builtin_fmaf((float)qbyte(g, 4, k), sy, hy0);
      nr = fmaxf(nr, fminf(t1, t2));
      fr = fminf(fr, fmaxf(t1, t2));
      t1 = __builtin_fmaf((float)qbyte(g, 2, k), sz, lz0);
      t2 = __builtin_fmaf((float)qbyte(g, 5, k), sz, hz0);
      nr = fmaxf(nr, fminf(t1, t2));
      fr = fminf(fr, fmaxf(t1, t2));
      const float tn = fmaxf(nr, 0.0f);
      if (tn <= fminf(fr, q.len)) key[k] = entry_key(tn, k);
    }
  }
  sort8(key);
  uint32_t ord = 0;
#pragma unroll
  for (uint32_t r = 0; r < 8; ++r)
    if (key[r] < kMissKey) ord |= ((key[r] & 7u) + 1u) << (4u * r);
  return ord;
}

// Point query over the quantised tree: leaves whose (dequantised) box holds
// the point get the exact test in ascending order, as in point_query.
template <int STRIDE, bool STATS>
__device__ inline void point_query_q(const RenderArgs &A, uint16_t *stk, float x, float y, float z, Payload &pl,
                                     Counters &c) {
  pl.count = 0;
  pl.dens = 0.0f;
  pl.r = pl.g = pl.b = 0.0f;
  if (STATS) c.queries++;
  auto test = [&](const float4 &lo, const float4 &hi) { return inside(lo, hi, x, y, z); };
  auto particle = [&](uint32_t p, const float4 &q) {
    const float e = aa_dist2(q.x, q.y, q.z, x, y, z) * A.pm[p];
    if (e > -4.5f) {
      float4 k = ld_pc(A, p);
      float wa = aa_expf_mid(e) * k.x;
      if (STATS) c.covering++;
      pl.count += 1;
      pl.dens += wa;
      pl.r += k.y * wa;
      pl.g += k.z * wa;
      pl.b += k.w * wa;
      if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
    }
  };
  auto leaves = [&](uint32_t lb, uint32_t mask) -> bool {
    if (A.ppl == 1) {
      float4 q[8], co, cs;
      if (A.q16) ld_chunk(A, lb, co, cs);
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if ((mask >> k) & 1u) q[k] = ld_pr_in(A, lb + k, co, cs);
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if ((mask >> k) & 1u) {
          if (STATS) c.particles++;
          particle(lb + k, q[k]);
        }
      return false;
    }
    while (mask) {
      uint32_t leaf = lb + (uint32_t)__ffs(mask) - 1u;
      mask &= mask - 1u;
      uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
      for (uint32_t p = p0; p < p1; ++p) {
        if (STATS) c.particles++;
        particle(p, ld_pr(A, p));
      }
    }
    return false;
  };
  walk_q<STRIDE, STATS>(A, stk, test, leaves, c);
}

// Point query with the batched walk; same covering set, order and
// arithmetic as point_query.
template <int STRIDE, bool STATS>
__device__ inline void point_query_b(const RenderArgs &A, uint16_t *stk, float x, float y, float z, Payload &pl,
                                     Counters &c) {
  pl.count = 0;
  pl.dens = 0.0f;
  pl.r = pl.g = pl.b = 0.0f;
  if (STATS) c.queries++;
  auto test = [&](const float4 &lo, const float4 &hi) { return inside(lo, hi, x, y, z); };
  auto particle = [&](uint32_t p, const float4 &q) {
    const float e = aa_dist2(q.x, q.y, q.z, x, y, z) * A.pm[p];
    if (e > -4.5f) {
      float4 k = ld_pc(A, p);
      float wa = aa_expf_mid(e) * k.x;
      if (STATS) c.covering++;
      pl.count += 1;
      pl.dens += wa;
      pl.r += k.y * wa;
      pl.g += k.z * wa;
      pl.b += k.w * wa;
      if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
    }
  };
  auto leaves = [&](uint32_t lb, uint32_t m) -> bool {
    if (A.ppl == 1) {
      float4 q[8], co, cs;
      if (A.q16) ld_chunk(A, lb, co, cs);
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (k < m) q[k] = ld_pr_in(A, lb + k, co, cs);
      if (STATS) c.particles += m;
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (k < m) particle(lb + k, q[k]);
      for (uint32_t k = 8; k < m; ++k) particle(lb + k, ld_pr(A, lb + k));
      return false;
    }
    const float4 *L0 = A.bvh.nodes + g_lvl_ofs[0];
    for (uint32_t k = 0; k < m; ++k) {
      uint32_t leaf = lb + k;
      if (STATS) c.leaves++;
      if (!inside(L0[2 * leaf], L0[2 * leaf + 1], x, y, z)) continue;
      uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
      for (uint32_t p = p0; p < p1; ++p) {
        if (STATS) c.particles++;
        particle(p, ld_pr(A, p));
      }
    }
    return false;
  };
  walk_batched<STRIDE, STATS>(A, stk, test, leaves, c);
}

} // namespace
} // namespace aarbf
