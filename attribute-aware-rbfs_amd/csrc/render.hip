// render.hip -- ParticleRBFRayGen with its tracker, point query,
// intersection and any-hit programs (deviceCodeRBF.hlsl:29-389) on gfx950.
//
// The reference issues one TraceRay point query per tracking step and lets
// the RT cores find every leaf box that contains the point.  CDNA4 has no
// ray-tracing units, so the query is a software walk of the implicit
// Hilbert-packed BVH (bvh.hip): depth-first, children in ascending order,
// stackless (the next node is found arithmetically from the level and index),
// which visits leaves -- and therefore sums particle contributions -- in
// ascending particle order.  Everything that only depends on a particle's
// attribute (radius, alpha^3, colour^2.2) is baked once per launch.
//
// Arithmetic follows aarbf_detmath.h and is compiled with -ffp-contract=off,
// so every accept/reject decision matches a host evaluation of the same
// definitions bit for bit.
#include <algorithm>
#include <cmath>

#include "aarbf_detmath.h"
#include "internal.h"

namespace aarbf {
namespace {

struct RenderArgs {
  uint32_t frameID, accumID, spp;
  uint32_t blue, showNoise, visAttr;
  uint32_t max_depth;
  int W, H;
  float amb, unit;
  float3 light, pos, d00, du, dv, lb, rt;
  const float4 *pr; // baked x, y, z, radius (0 = never covers)
  const float4 *pc; // baked alpha^3, rgb^2.2
  uint32_t numParticles, ppl;
  const float4 *cmap;
  const float4 *dmap;
  uint32_t cm_n, dm_n;
  const uint8_t *stbn;
  int stbn_w, stbn_h;
  float4 *image, *accum;
  uint64_t *stats;
  BVHView bvh;
};

// Per-block LDS copy of the BVH level table (offsets and node counts), so
// that the walk can index it by level without copying the kernel-argument
// array to scratch memory.
__shared__ uint32_t g_lvl_ofs[kMaxLevels];
__shared__ uint32_t g_lvl_cnt[kMaxLevels];

__device__ inline void load_levels(const RenderArgs &A) {
  if (threadIdx.x < kMaxLevels) {
    g_lvl_ofs[threadIdx.x] = A.bvh.level_ofs[threadIdx.x];
    g_lvl_cnt[threadIdx.x] = A.bvh.counts[threadIdx.x];
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
// (deviceCodeRBF.hlsl:350-351,366-376).
__global__ __launch_bounds__(256) void bake_kernel(const float4 *__restrict__ particles,
                                                   const float *__restrict__ radii, uint32_t n,
                                                   float rbfRadius, const float4 *__restrict__ rmap,
                                                   uint32_t rm_n, const float4 *__restrict__ cmap,
                                                   uint32_t cm_n, float4 *__restrict__ pr,
                                                   float4 *__restrict__ pc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 q = particles[i];
  float rp = tf_r(rmap, rm_n, q.w);
  float radius = (radii ? radii[i] : rbfRadius) * rp;
  float4 c = tf4(cmap, cm_n, q.w);
  pr[i] = make_float4(q.x, q.y, q.z, c.w == 0.0f ? 0.0f : radius);
  pc[i] = make_float4(aa_cube(c.w), aa_powf(c.x, 2.2f), aa_powf(c.y, 2.2f), aa_powf(c.z, 2.2f));
}

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
          float4 q = A.pr[p];
          if (STATS) c.particles++;
          float d = aa_dist3(q.x, q.y, q.z, x, y, z);
          if (d < q.w) {
            float4 k = A.pc[p];
            float wa = aa_rbf(d, q.w) * k.x;
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
// counter shared by the primary and the shadow ray.
template <bool STATS>
__device__ void track(const RenderArgs &A, bool marching, bool shadow, uint32_t &i, float &t, uint32_t &rng,
                      float random0, float3 o, float3 d, float tmax, float4 &albedo, Counters &c) {
  Payload pl;
  if (marching) { // stochasticMarching ':67-139'
    float RHS = -aa_logf(1.0f - random0 * 0.99f);
    float LHS = 0.0f;
    for (; i < A.max_depth; ++i) {
      if (STATS) c.steps++;
      t = t + A.unit;
      if (t >= tmax) return;
      point_query<STATS>(A, o.x + t * d.x, o.y + t * d.y, o.z + t * d.z, pl, c);
      if (!(pl.count > 0 && pl.dens > 0.0f)) continue;
      float D = tf_r(A.dmap, A.dm_n, pl.dens);
      LHS += D;
      if (LHS > RHS && !shadow) {
        float3 col = colour_of(A, pl);
        albedo = make_float4(col.x, col.y, col.z, 1.0f);
        return;
      }
      if (shadow) {
        albedo.w += 1.0f - aa_expf(-D);
        if (albedo.w >= 1.0f) {
          albedo.w = 1.0f;
          return;
        }
      }
    }
  } else { // stochasticTracking ':141-203', majorant 1
    t = 0.0f;
    for (; i < A.max_depth; ++i) {
      if (STATS) c.steps++;
      float xi = aa_lcg_next(&rng);
      t = t - aa_logf(1.0f - xi) * A.unit;
      if (t >= tmax) return;
      point_query<STATS>(A, o.x + t * d.x, o.y + t * d.y, o.z + t * d.z, pl, c);
      if (pl.count == 0) continue;
      float D = tf_r(A.dmap, A.dm_n, pl.dens);
      float xi2 = aa_lcg_next(&rng);
      if (xi2 < D) {
        float3 col = colour_of(A, pl);
        albedo = make_float4(col.x, col.y, col.z, 1.0f);
        return;
      }
    }
  }
}

// ParticleRBFRayGen (deviceCodeRBF.hlsl:206-324) for one sample; returns the
// colour composited over black, and random.x in rx.
template <bool STATS>
__device__ float4 raygen_sample(const RenderArgs &A, int px, int py, uint32_t frameID, float &rx, Counters &c) {
  uint32_t rng = aa_rng_seed(frameID, (uint32_t)px, (uint32_t)py, (uint32_t)A.W);
  float sx = ((float)px + 0.5f) / (float)A.W;
  float sy = ((float)py + 0.5f) / (float)A.H;
  float3 dir;
  dir.x = (A.d00.x + sx * A.du.x) + sy * A.dv.x;
  dir.y = (A.d00.y + sx * A.du.y) + sy * A.dv.y;
  dir.z = (A.d00.z + sx * A.du.z) + sy * A.dv.z;
  float len = sqrtf((dir.x * dir.x + dir.y * dir.y) + dir.z * dir.z);
  dir.x = dir.x / len;
  dir.y = dir.y / len;
  dir.z = dir.z / len;
  float tenter, texit;
  aabb_intersection(A.pos, dir, 0.0f, 10000.0f, A.lb, A.rt, tenter, texit);
  bool marching = A.blue != 0;
  float r0;
  uint32_t trng = 0;
  if (marching) {
    int gx = (int)(frameID % 8u), gy = (int)((frameID / 8u) % 16u);
    int cx = gx * 256 + px % 256, cy = gy * 256 + py % 256;
    uint8_t v = 0;
    if (cx < A.stbn_w && cy < A.stbn_h) v = A.stbn[4 * ((uint64_t)cy * A.stbn_w + cx)];
    r0 = (float)v / 255.0f;
  } else {
    r0 = aa_lcg_next(&rng);
    trng = rng;
  }
  float r1 = aa_lcg_next(&rng);
  float r2 = aa_lcg_next(&rng);
  rx = r0;
  float4 color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t i = 0;
  if (tenter < texit) {
    float t = tenter + A.unit * r1;
    float4 albedo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    track<STATS>(A, marching, false, i, t, trng, r0, A.pos, dir, texit, albedo, c);
    float vis = 1.0f;
    if (albedo.w > 0.0f && A.amb != 1.0f) {
      float3 so = make_float3(A.pos.x + dir.x * t, A.pos.y + dir.y * t, A.pos.z + dir.z * t);
      float ste, stx;
      aabb_intersection(so, A.light, 0.0f, 10000.0f, A.lb, A.rt, ste, stx);
      float4 sh = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      float ts = marching ? -A.unit * r2 : 0.0f;
      track<STATS>(A, marching, true, i, ts, trng, r0, so, A.light, stx, sh, c);
      vis = 1.0f - sh.w;
    }
    if (albedo.w == 1.0f) {
      color.x = (albedo.x * vis) * (1.0f - A.amb) + albedo.x * A.amb;
      color.y = (albedo.y * vis) * (1.0f - A.amb) + albedo.y * A.amb;
      color.z = (albedo.z * vis) * (1.0f - A.amb) + albedo.z * A.amb;
      color.w = 1.0f;
    }
  }
  if (STATS) {
    c.samples++;
    if (i >= A.max_depth) c.exhausted++;
  }
  // over(color, black) (sharedCode.h:142-149)
  float ra = color.w + 1.0f * (1.0f - color.w);
  if (ra == 0.0f) return color;
  float4 o;
  o.x = (color.x * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
  o.y = (color.y * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
  o.z = (color.z * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
  o.w = ra;
  return o;
}

__device__ inline uint64_t wave_sum(uint32_t v) {
  uint64_t s = v;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// One lane per pixel, spp samples in sequence; 256-thread blocks cover
// 16x16 pixels, each wave an 8x8 tile.
template <bool STATS>
__global__ __launch_bounds__(256) void render_kernel(RenderArgs A) {
  load_levels(A);
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int px = blockIdx.x * 16 + (w & 1) * 8 + (l & 7);
  const int py = blockIdx.y * 16 + (w >> 1) * 8 + (l >> 3);
  Counters c = {};
  if (px < A.W && py < A.H) {
    const uint64_t ofs = (uint64_t)px + (uint64_t)A.W * py;
    float4 acc = A.accum[ofs], fin = acc;
    float rx = 0.0f;
    for (uint32_t s = 0; s < A.spp; ++s) {
      uint32_t fid = (A.frameID + s) & 0xffffu, aid = (A.accumID + s) & 0xffffu;
      float4 col = raygen_sample<STATS>(A, px, py, fid, rx, c);
      // running mean (deviceCodeRBF.hlsl:326-328)
      float w1 = 1.0f / (float)aid;
      float w0 = (float)((int)aid - 1) / (float)aid;
      fin.x = w1 * col.x + w0 * acc.x;
      fin.y = w1 * col.y + w0 * acc.y;
      fin.z = w1 * col.z + w0 * acc.z;
      fin.w = w1 * col.w + w0 * acc.w;
      acc = fin;
    }
    A.accum[ofs] = acc;
    if (A.showNoise) fin = make_float4(rx, rx, rx, 1.0f);
    A.image[ofs] = fin;
  }
  if (STATS) {
    uint32_t v[8] = {c.samples, c.steps, c.queries, c.nodes, c.leaves, c.particles, c.covering, c.exhausted};
    for (int k = 0; k < 8; ++k) {
      uint64_t s = wave_sum(v[k]);
      if (l == 0 && s) atomicAdd((unsigned long long *)&A.stats[k], (unsigned long long)s);
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent, segment-gather renderer.
//
// Bit-for-bit the same samples as render_kernel, scheduled differently:
//  * persistent waves pull pixels from a global counter (8x8 tiles in
//    row-major order), so a lane that finishes a pixel refills at once and
//    no wave idles behind its longest path;
//  * each lane runs a flat state machine (new sample -> primary steps ->
//    shadow steps -> accumulate) in which one loop iteration is one tracker
//    step for every lane, so lanes stay converged on the step code;
//  * instead of a BVH walk per step, a lane walks the BVH once for a whole
//    ray segment [t0, t1] and keeps, in LDS, the ascending list of particles
//    whose sphere (widened by a rounding margin) meets the segment.  Every
//    step in the segment only tests those candidates with the exact
//    |P - x| < r test, in ascending index order, so each step sees the same
//    covering set summed in the same order as a full point query.  The
//    segment length adapts to the candidate count; where more than kCand
//    particles crowd a short segment the lane falls back to point queries.
// ---------------------------------------------------------------------------

constexpr uint32_t kCand = 24;
constexpr uint32_t kDense = 0xffffffffu;
constexpr int kPBlock = 256;

struct SegQuery {
  float ax, ay, az; // segment start
  float dx, dy, dz; // unit direction
  float ix, iy, iz; // 1/d with zero components replaced by +-1e30
  float len;        // segment length
  float m;          // conservative margin for rounding of o + t*d and |P-x|
};

__device__ inline float safe_inv(float d) { return 1.0f / (fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d); }

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

__device__ inline bool seg_particle(const SegQuery &s, float4 q) {
  if (!(q.w > 0.0f)) return false;
  float vx = q.x - s.ax, vy = q.y - s.ay, vz = q.z - s.az;
  float p = vx * s.dx + vy * s.dy + vz * s.dz;
  p = fminf(fmaxf(p, 0.0f), s.len);
  float wx = vx - p * s.dx, wy = vy - p * s.dy, wz = vz - p * s.dz;
  float rr = q.w + s.m;
  return (wx * wx + wy * wy + wz * wz) <= rr * rr;
}

// Candidates of one segment, appended to the lane's LDS column in ascending
// particle order; kDense when more than kCand particles qualify.
template <bool STATS>
__device__ uint32_t gather(const RenderArgs &A, const SegQuery &s, uint32_t *my, Counters &c) {
  const float4 *__restrict__ N = A.bvh.nodes;
  const uint32_t log2K = A.bvh.log2K, kmask = (1u << log2K) - 1u, top = A.bvh.top;
  const uint32_t *lofs = g_lvl_ofs, *lcnt = g_lvl_cnt;
  uint32_t level = top, idx = 0, n = 0;
  if (STATS) c.queries++;
  for (;;) {
    const float4 *lv = N + lofs[level];
    if (STATS) c.nodes++;
    if (seg_box(s, lv[2 * idx], lv[2 * idx + 1])) {
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
          if (!seg_box(s, lb[2 * leaf], lb[2 * leaf + 1])) continue;
        }
        uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
        for (uint32_t p = p0; p < p1; ++p) {
          if (STATS) c.particles++;
          if (seg_particle(s, A.pr[p])) {
            if (n == kCand) return kDense;
            my[n * kPBlock] = p;
            ++n;
          }
        }
      }
    }
    for (;;) {
      if (level == top) return n;
      if ((idx & kmask) != kmask && idx + 1 < lcnt[level]) {
        ++idx;
        break;
      }
      ++level;
      idx >>= log2K;
    }
  }
}

// ---- batched walk ------------------------------------------------------------
// Depth-first like the stackless walk above, but a hit node tests all K of
// its children with independent loads (one memory round trip per level
// instead of one per child).  The hit mask of each level still to be
// visited is kept in a per-lane LDS column (stk[level * kPBlock]); children
// are taken lowest bit first, so leaves are still visited in ascending
// order.  Level-1 hits hand their K leaves to `leaves(base, m)`.

constexpr uint32_t kStackLevels = 16;

template <bool STATS, class BoxTest, class Leaves>
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
      mask = stk[lvl * kPBlock];
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
    stk[lvl * kPBlock] = (uint16_t)mask;
    --lvl;
    base = node << log2K;
    mask = group(lvl, base);
  }
}

// Point query with the batched walk; same covering set, order and
// arithmetic as point_query.
template <bool STATS>
__device__ inline void point_query_b(const RenderArgs &A, uint16_t *stk, float x, float y, float z, Payload &pl,
                                     Counters &c) {
  pl.count = 0;
  pl.dens = 0.0f;
  pl.r = pl.g = pl.b = 0.0f;
  if (STATS) c.queries++;
  auto test = [&](const float4 &lo, const float4 &hi) { return inside(lo, hi, x, y, z); };
  auto particle = [&](uint32_t p, const float4 &q) {
    float d = aa_dist3(q.x, q.y, q.z, x, y, z);
    if (d < q.w) {
      float4 k = A.pc[p];
      float wa = aa_rbf(d, q.w) * k.x;
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
      float4 q[8];
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (k < m) q[k] = A.pr[lb + k];
      if (STATS) c.particles += m;
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (k < m) particle(lb + k, q[k]);
      for (uint32_t k = 8; k < m; ++k) particle(lb + k, A.pr[lb + k]);
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
        particle(p, A.pr[p]);
      }
    }
    return false;
  };
  walk_batched<STATS>(A, stk, test, leaves, c);
}

// Segment gather with the batched walk (kDense on overflow).
template <bool STATS>
__device__ inline uint32_t gather_b(const RenderArgs &A, uint16_t *stk, const SegQuery &s, uint32_t *my,
                                    Counters &c) {
  uint32_t n = 0;
  bool overflow = false;
  if (STATS) c.queries++;
  auto test = [&](const float4 &lo, const float4 &hi) { return seg_box(s, lo, hi); };
  auto take = [&](uint32_t p, const float4 &q) -> bool {
    if (seg_particle(s, q)) {
      if (n == kCand) {
        overflow = true;
        return true;
      }
      my[n * kPBlock] = p;
      ++n;
    }
    return false;
  };
  auto leaves = [&](uint32_t lb, uint32_t m) -> bool {
    if (A.ppl == 1) {
      float4 q[8];
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (k < m) q[k] = A.pr[lb + k];
      if (STATS) c.particles += m;
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k)
        if (k < m && take(lb + k, q[k])) return true;
      for (uint32_t k = 8; k < m; ++k)
        if (take(lb + k, A.pr[lb + k])) return true;
      return false;
    }
    const float4 *L0 = A.bvh.nodes + g_lvl_ofs[0];
    for (uint32_t k = 0; k < m; ++k) {
      uint32_t leaf = lb + k;
      if (STATS) c.leaves++;
      if (!seg_box(s, L0[2 * leaf], L0[2 * leaf + 1])) continue;
      uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
      for (uint32_t p = p0; p < p1; ++p) {
        if (STATS) c.particles++;
        if (take(p, A.pr[p])) return true;
      }
    }
    return false;
  };
  walk_batched<STATS>(A, stk, test, leaves, c);
  return overflow ? kDense : n;
}

// Point query restricted to the segment's candidates (same arithmetic and
// order as point_query's leaf loop).
template <bool STATS>
__device__ inline void scan(const RenderArgs &A, const uint32_t *my, uint32_t n, float x, float y, float z,
                            Payload &pl, Counters &c) {
  pl.count = 0;
  pl.dens = 0.0f;
  pl.r = pl.g = pl.b = 0.0f;
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t p = my[k * kPBlock];
    float4 q = A.pr[p];
    if (STATS) c.particles++;
    float d = aa_dist3(q.x, q.y, q.z, x, y, z);
    if (d < q.w) {
      float4 kk = A.pc[p];
      float wa = aa_rbf(d, q.w) * kk.x;
      if (STATS) c.covering++;
      pl.count += 1;
      pl.dens += wa;
      pl.r += kk.y * wa;
      pl.g += kk.z * wa;
      pl.b += kk.w * wa;
      if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
    }
  }
}

enum : int { ST_FETCH = 0, ST_INIT = 1, ST_TRACK = 2, ST_WRITE = 3, ST_EXIT = 4 };

__global__ void zero_counter(uint32_t *c) { *c = 0; }

template <bool STATS>
__global__ __launch_bounds__(kPBlock) void render_persistent(RenderArgs A, uint32_t *__restrict__ counter) {
  __shared__ uint32_t cand[kCand * kPBlock];
  __shared__ uint16_t stack_mem[kStackLevels * kPBlock];
  load_levels(A);
  uint32_t *my = cand + threadIdx.x;
  uint16_t *stk = stack_mem + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t tilesX = (uint32_t)(A.W + 7) / 8, tilesY = (uint32_t)(A.H + 7) / 8;
  const uint32_t total = tilesX * tilesY * 64u;
  const bool marching = A.blue != 0;
  const float max_seg = 1024.0f * A.unit;
  Counters c = {};

  int state = ST_FETCH;
  int px = 0, py = 0;
  uint64_t ofs = 0;
  uint32_t s = 0, i = 0, trng = 0;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f), fin = acc, albedo = acc;
  float rx = 0.0f, r2 = 0.0f, RHS = 0.0f, LHS = 0.0f, shw = 0.0f;
  float3 dir = make_float3(0.0f, 0.0f, 0.0f), o = dir, rd = dir;
  float t = 0.0f, tmax = 0.0f;
  bool shadow = false;
  float seg_t0 = 0.0f, seg_t1 = -AA_INF_F, seg_len = 64.0f * A.unit;
  uint32_t seg_n = 0;

  for (;;) {
    // ---- refill: lanes without a pixel take the next ones, one atomic per wave
    bool need = state == ST_FETCH;
    uint64_t mask = __ballot(need);
    if (mask) {
      int leader = __ffsll((unsigned long long)mask) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
      base = __shfl(base, leader, 64);
      if (need) {
        uint32_t idx = base + (uint32_t)__popcll(mask & lt);
        if (idx >= total) {
          state = ST_EXIT;
        } else {
          uint32_t tile = idx >> 6, w = idx & 63u;
          px = (int)((tile % tilesX) * 8u + (w & 7u));
          py = (int)((tile / tilesX) * 8u + (w >> 3));
          if (px < A.W && py < A.H) {
            ofs = (uint64_t)px + (uint64_t)A.W * py;
            acc = A.accum[ofs];
            fin = acc;
            s = 0;
            state = ST_INIT;
          }
        }
      }
    }
    if (__ballot(state != ST_EXIT) == 0) break;

    // ---- new sample: ParticleRBFRayGen set-up (deviceCodeRBF.hlsl:207-280)
    bool finish = false;
    float vis = 1.0f;
    if (state == ST_INIT) {
      uint32_t fid = (A.frameID + s) & 0xffffu;
      uint32_t rng = aa_rng_seed(fid, (uint32_t)px, (uint32_t)py, (uint32_t)A.W);
      float sx = ((float)px + 0.5f) / (float)A.W;
      float sy = ((float)py + 0.5f) / (float)A.H;
      dir.x = (A.d00.x + sx * A.du.x) + sy * A.dv.x;
      dir.y = (A.d00.y + sx * A.du.y) + sy * A.dv.y;
      dir.z = (A.d00.z + sx * A.du.z) + sy * A.dv.z;
      float len = sqrtf((dir.x * dir.x + dir.y * dir.y) + dir.z * dir.z);
      dir.x = dir.x / len;
      dir.y = dir.y / len;
      dir.z = dir.z / len;
      float tenter, texit;
      aabb_intersection(A.pos, dir, 0.0f, 10000.0f, A.lb, A.rt, tenter, texit);
      float r0;
      if (marching) {
        int gx = (int)(fid % 8u), gy = (int)((fid / 8u) % 16u);
        int cx = gx * 256 + px % 256, cy = gy * 256 + py % 256;
        uint8_t v = 0;
        if (cx < A.stbn_w && cy < A.stbn_h) v = A.stbn[4 * ((uint64_t)cy * A.stbn_w + cx)];
        r0 = (float)v / 255.0f;
      } else {
        r0 = aa_lcg_next(&rng);
        trng = rng;
      }
      float r1 = aa_lcg_next(&rng);
      r2 = aa_lcg_next(&rng);
      rx = r0;
      i = 0;
      albedo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (tenter < texit) {
        o = A.pos;
        rd = dir;
        tmax = texit;
        shadow = false;
        t = marching ? tenter + A.unit * r1 : 0.0f;
        LHS = 0.0f;
        RHS = marching ? -aa_logf(1.0f - r0 * 0.99f) : 0.0f;
        seg_t1 = -AA_INF_F;
        state = ST_TRACK;
      } else {
        finish = true;
      }
    }

    // ---- one tracker step (deviceCodeRBF.hlsl:78-138 / 151-202)
    if (state == ST_TRACK) {
      bool end = false;
      bool stepping = i < A.max_depth;
      Payload pl;
      pl.count = 0;
      if (!stepping) {
        end = true;
      } else {
        if (STATS) c.steps++;
        if (marching) {
          t = t + A.unit;
        } else {
          float xi = aa_lcg_next(&trng);
          t = t - aa_logf(1.0f - xi) * A.unit;
        }
        if (t >= tmax) end = true;
      }
      bool live = stepping && !end;
      // segment refresh, batched: when any lane must gather, lanes past the
      // middle of their segment gather too
      bool must = live && !(t <= seg_t1);
      bool any = __ballot(must) != 0;
      if (any && live && (must || t > 0.5f * (seg_t0 + seg_t1))) {
        float L = seg_len;
        for (;;) {
          float tb = fminf(t + L, tmax);
          SegQuery q;
          q.ax = o.x + t * rd.x;
          q.ay = o.y + t * rd.y;
          q.az = o.z + t * rd.z;
          q.dx = rd.x;
          q.dy = rd.y;
          q.dz = rd.z;
          q.ix = safe_inv(rd.x);
          q.iy = safe_inv(rd.y);
          q.iz = safe_inv(rd.z);
          q.len = tb - t;
          q.m = 1e-5f * (1.0f + fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) + fabsf(tb));
          uint32_t n = gather_b<STATS>(A, stk, q, my, c);
          if (n != kDense) {
            seg_t0 = t;
            seg_t1 = tb;
            seg_n = n;
            seg_len = (n <= kCand / 4) ? fminf(2.0f * L, max_seg) : (n > kCand / 2 ? 0.5f * L : L);
            break;
          }
          L *= 0.25f;
          if (L < A.unit) {
            seg_t0 = t;
            seg_t1 = fminf(t + 4.0f * A.unit, tmax);
            seg_n = kDense;
            seg_len = 4.0f * A.unit;
            break;
          }
        }
      }
      if (live) {
        float x = o.x + t * rd.x, y = o.y + t * rd.y, z = o.z + t * rd.z;
        if (seg_n == kDense)
          point_query_b<STATS>(A, stk, x, y, z, pl, c);
        else
          scan<STATS>(A, my, seg_n, x, y, z, pl, c);
        if (marching) {
          if (pl.count > 0 && pl.dens > 0.0f) {
            float D = tf_r(A.dmap, A.dm_n, pl.dens);
            LHS += D;
            if (LHS > RHS && !shadow) {
              float3 col = colour_of(A, pl);
              albedo = make_float4(col.x, col.y, col.z, 1.0f);
              end = true;
            } else if (shadow) {
              shw += 1.0f - aa_expf(-D);
              if (shw >= 1.0f) {
                shw = 1.0f;
                end = true;
              }
            }
          }
        } else if (pl.count > 0) {
          float D = tf_r(A.dmap, A.dm_n, pl.dens);
          float xi2 = aa_lcg_next(&trng);
          if (xi2 < D) {
            if (shadow) {
              shw = 1.0f;
            } else {
              float3 col = colour_of(A, pl);
              albedo = make_float4(col.x, col.y, col.z, 1.0f);
            }
            end = true;
          }
        }
        if (!end) ++i;
      }
      if (end) {
        if (!shadow && albedo.w > 0.0f && A.amb != 1.0f) {
          // NEE shadow ray (':288-311'); the step counter is not reset
          float3 so = make_float3(A.pos.x + dir.x * t, A.pos.y + dir.y * t, A.pos.z + dir.z * t);
          float ste, stx;
          aabb_intersection(so, A.light, 0.0f, 10000.0f, A.lb, A.rt, ste, stx);
          o = so;
          rd = A.light;
          tmax = stx;
          shadow = true;
          shw = 0.0f;
          LHS = 0.0f;
          t = marching ? -A.unit * r2 : 0.0f;
          seg_t1 = -AA_INF_F;
        } else {
          vis = shadow ? 1.0f - shw : 1.0f;
          finish = true;
        }
      }
    }

    // ---- shade, composite and accumulate (':316-328')
    if (finish) {
      float4 color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (albedo.w == 1.0f) {
        color.x = (albedo.x * vis) * (1.0f - A.amb) + albedo.x * A.amb;
        color.y = (albedo.y * vis) * (1.0f - A.amb) + albedo.y * A.amb;
        color.z = (albedo.z * vis) * (1.0f - A.amb) + albedo.z * A.amb;
        color.w = 1.0f;
      }
      if (STATS) {
        c.samples++;
        if (i >= A.max_depth) c.exhausted++;
      }
      float ra = color.w + 1.0f * (1.0f - color.w);
      if (ra != 0.0f) {
        color.x = (color.x * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
        color.y = (color.y * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
        color.z = (color.z * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
        color.w = ra;
      }
      uint32_t aid = (A.accumID + s) & 0xffffu;
      float w1 = 1.0f / (float)aid;
      float w0 = (float)((int)aid - 1) / (float)aid;
      fin.x = w1 * color.x + w0 * acc.x;
      fin.y = w1 * color.y + w0 * acc.y;
      fin.z = w1 * color.z + w0 * acc.z;
      fin.w = w1 * color.w + w0 * acc.w;
      acc = fin;
      ++s;
      state = (s == A.spp) ? ST_WRITE : ST_INIT;
    }
    if (state == ST_WRITE) {
      A.accum[ofs] = acc;
      A.image[ofs] = A.showNoise ? make_float4(rx, rx, rx, 1.0f) : fin;
      state = ST_FETCH;
    }
  }
  if (STATS) {
    uint32_t v[8] = {c.samples, c.steps, c.queries, c.nodes, c.leaves, c.particles, c.covering, c.exhausted};
    for (int k = 0; k < 8; ++k) {
      uint64_t sum = wave_sum(v[k]);
      if (lane == 0 && sum) atomicAdd((unsigned long long *)&A.stats[k], (unsigned long long)sum);
    }
  }
}

} // namespace
} // namespace aarbf

using namespace aarbf;

// getLightDirection (sharedCode.h:127-132), evaluated once on the host.
static float3 light_direction(float azimuth, float elevation) {
  float a = elevation * 3.14f - 0.5f * 3.14f;
  float b = azimuth * 2.0f * 3.14f;
  return make_float3(sinf(a) * cosf(b), sinf(a) * sinf(b), cosf(a));
}

extern "C" {

int aarbf_render_rbf_spp(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W,
                         int H, uint32_t spp) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, rg && pc, "aarbf_render_rbf: null struct");
  AARBF_REQUIRE(ctx, W > 0 && H > 0 && spp >= 1, "aarbf_render_rbf: empty launch");
  AARBF_REQUIRE(ctx, rg->imageBuffer && rg->accumBuffer, "aarbf_render_rbf: null image/accum buffer");
  BVH *bvh = get_bvh(ctx, pc->world);
  AARBF_REQUIRE(ctx, bvh, "aarbf_render_rbf: world is not a BVH handle");
  TF *cm = get_tf(ctx, pc->colormap), *rm = get_tf(ctx, pc->radiusmap), *dm = get_tf(ctx, pc->densitymap);
  AARBF_REQUIRE(ctx, cm && rm && dm, "aarbf_render_rbf: colormap/radiusmap/densitymap are not TF handles");
  AARBF_REQUIRE(ctx, pc->particlesPerLeaf >= 1, "aarbf_render_rbf: particlesPerLeaf must be >= 1");
  AARBF_REQUIRE(ctx, pc->numParticles == 0 || pc->particles, "aarbf_render_rbf: null particles");
  AARBF_REQUIRE(ctx, (uint64_t)bvh->counts[0] * pc->particlesPerLeaf >= pc->numParticles,
                "aarbf_render_rbf: the BVH has fewer leaves than numParticles / particlesPerLeaf");
  STBN *stbn = nullptr;
  if (pc->enableBlueNoise) {
    stbn = get_stbn(ctx, rg->stbnTexture);
    AARBF_REQUIRE(ctx, stbn, "aarbf_render_rbf: blue noise enabled without an STBN atlas");
  }
  hipSetDevice(ctx->device);
  uint32_t n = pc->numParticles;
  size_t need = 2 * (size_t)(n ? n : 1);
  if (ctx->baked_count < need) {
    if (ctx->baked) {
      hipStreamSynchronize(ctx->stream);
      hipFree(ctx->baked);
      ctx->baked = nullptr;
      ctx->baked_count = 0;
    }
    AARBF_HIP(ctx, hipMalloc(&ctx->baked, sizeof(float4) * need));
    ctx->baked_count = need;
  }
  float4 *pr = ctx->baked, *pcol = ctx->baked + (n ? n : 1);
  if (n)
    bake_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const float4 *)pc->particles, (const float *)pc->radii,
                                                          n, pc->rbfRadius, rm->table, rm->n, cm->table, cm->n, pr,
                                                          pcol);
  RenderArgs A{};
  A.frameID = pc->frameID;
  A.accumID = pc->accumID;
  A.spp = spp;
  A.blue = pc->enableBlueNoise ? 1 : 0;
  A.showNoise = pc->showNoise ? 1 : 0;
  A.visAttr = pc->visualizeAttributes ? 1 : 0;
  A.max_depth = AARBF_MAX_DEPTH;
  A.W = W;
  A.H = H;
  A.amb = pc->light.ambient;
  A.unit = pc->unit;
  A.light = light_direction(pc->light.azimuth, pc->light.elevation);
  A.pos = make_float3(pc->camera.pos.x, pc->camera.pos.y, pc->camera.pos.z);
  A.d00 = make_float3(pc->camera.dir_00.x, pc->camera.dir_00.y, pc->camera.dir_00.z);
  A.du = make_float3(pc->camera.dir_du.x, pc->camera.dir_du.y, pc->camera.dir_du.z);
  A.dv = make_float3(pc->camera.dir_dv.x, pc->camera.dir_dv.y, pc->camera.dir_dv.z);
  float m2 = 2.0f * pc->rbfRadius;
  A.rt = make_float3(rg->globalAABBMax.x + m2, rg->globalAABBMax.y + m2, rg->globalAABBMax.z + m2);
  A.lb = make_float3(rg->globalAABBMin.x - m2, rg->globalAABBMin.y - m2, rg->globalAABBMin.z - m2);
  A.pr = pr;
  A.pc = pcol;
  A.numParticles = n;
  A.ppl = pc->particlesPerLeaf;
  A.cmap = cm->table;
  A.cm_n = cm->n;
  A.dmap = dm->table;
  A.dm_n = dm->n;
  A.stbn = stbn ? stbn->rgba : nullptr;
  A.stbn_w = stbn ? stbn->width : 0;
  A.stbn_h = stbn ? stbn->height : 0;
  A.image = (float4 *)rg->imageBuffer;
  A.accum = (float4 *)rg->accumBuffer;
  A.stats = ctx->d_stats;
  A.bvh = make_view(*bvh);
  if (ctx->render_mode == 1 || bvh->num_levels > kStackLevels) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    if (ctx->stats)
      render_kernel<true><<<grid, 256, 0, ctx->stream>>>(A);
    else
      render_kernel<false><<<grid, 256, 0, ctx->stream>>>(A);
    AARBF_CHECK_LAUNCH(ctx, "render_kernel");
    return AARBF_OK;
  }
  // persistent: as many blocks as fit on the device at once, capped by the work
  int per_cu = 0;
  hipError_t e = ctx->stats ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, render_persistent<true>, kPBlock, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, render_persistent<false>, kPBlock, 0);
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  uint64_t tiles = (uint64_t)((W + 7) / 8) * (uint64_t)((H + 7) / 8);
  uint64_t want = (tiles * 64 + kPBlock - 1) / kPBlock;
  uint32_t blocks = (uint32_t)std::min<uint64_t>(want, (uint64_t)per_cu * (uint64_t)(ctx->cus > 0 ? ctx->cus : 1));
  zero_counter<<<1, 1, 0, ctx->stream>>>(ctx->d_counter);
  if (ctx->stats)
    render_persistent<true><<<blocks, kPBlock, 0, ctx->stream>>>(A, ctx->d_counter);
  else
    render_persistent<false><<<blocks, kPBlock, 0, ctx->stream>>>(A, ctx->d_counter);
  AARBF_CHECK_LAUNCH(ctx, "render_persistent");
  return AARBF_OK;
}

int aarbf_render_rbf(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W,
                     int H) {
  return aarbf_render_rbf_spp(ctx, rg, pc, W, H, 1);
}

} // extern "C"
