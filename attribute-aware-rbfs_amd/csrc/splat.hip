// splat.hip -- ParticleSplatRayGen (deviceCodeSplat.hlsl:29-194): Wald-style
// slab splatting, one lane per pixel.
//
// The reference marches slabs of one unit (15/8 = 1 particle per slab,
// integer division) and traces every slab segment: the any-hit program
// keeps the first 15 hits, which are sorted by t and composited front to
// back.  Tracing 1-unit segments one by one would cost a BVH walk per slab,
// so this kernel walks the tree once per run of slabs (a chunk): the walk
// collects, in ascending particle order, every particle whose leaf box the
// chunk's segment meets and whose splat sample lies in the chunk.  Each
// slab then takes from that list the hits whose t falls in the slab and
// whose leaf box its own segment meets -- the same set, in the same order,
// that a walk of the slab alone reports.  A walk of several slabs stops at
// the first hit its per-lane list cannot hold, and the chunk shrinks to a
// quarter (down to a single slab, which is walked to the end and where only
// the first 15 hits matter).  The walk opens a node by testing all its
// children's boxes, loaded together, and keeps the children still to visit
// as one bit mask per level in LDS.
//
// Box tests (aa_seg_box) use the leaf boxes the BVH stores (GenRBFBounds
// boxes widened by 2^-20 relative); hits are taken in ascending particle
// index (the reference's order is the RT traversal's, visible only when a
// slab has more than 15 hits).
#include "render_common.h"

namespace aarbf {
namespace {

constexpr int kSplatBlock = 64;       // one wave, an 8x8 pixel tile
constexpr uint32_t kSplatCap = 32;    // hits one chunk may collect per lane
constexpr uint32_t kPayload = 15;     // PARTICLE_BUFFER_SIZE (deviceCodeSplat.hlsl:37)
constexpr uint32_t kChunkMax = 512;   // slabs per chunk

struct SplatArgs {
  int W, H;
  uint32_t visAttr;
  float unit, rbf;
  float3 pos, d00, du, dv, lb, rt;
  const float4 *particles; // raw x, y, z, attribute
  const float *radii;      // per-particle radius or null (rbf)
  uint32_t n, ppl;
  const float4 *cmap, *dmap;
  uint32_t cm_n, dm_n;
  float4 *image;
  BVHView bvh;
};

__device__ inline void over4(float4 &res, const float4 &c) {
  float a[4] = {res.x, res.y, res.z, res.w}, b[4] = {c.x, c.y, c.z, c.w}, o[4];
  aa_over(a, b, o);
  res = make_float4(o[0], o[1], o[2], o[3]);
}

__global__ __launch_bounds__(kSplatBlock) void splat_kernel(SplatArgs A) {
  __shared__ uint32_t s_id[kSplatCap * kSplatBlock];
  __shared__ float s_t[kSplatCap * kSplatBlock];
  __shared__ uint8_t s_pe[kPayload * kSplatBlock];    // a slab's payload: positions in the hit list
  __shared__ uint16_t s_mask[kMaxLevels * kSplatBlock]; // per level: siblings still to visit
  if (threadIdx.x < kMaxLevels) {
    g_lvl_ofs[threadIdx.x] = A.bvh.level_ofs[threadIdx.x];
    g_lvl_cnt[threadIdx.x] = A.bvh.counts[threadIdx.x];
  }
  __syncthreads();
  const int px = (int)(blockIdx.x * 8u + (threadIdx.x & 7u)), py = (int)(blockIdx.y * 8u + (threadIdx.x >> 3));
  if (px >= A.W || py >= A.H) return;
  uint32_t *hid = s_id + threadIdx.x;
  float *ht = s_t + threadIdx.x;
  uint8_t *pe = s_pe + threadIdx.x;
  uint16_t *smask = s_mask + threadIdx.x;

  const float sx = ((float)px + 0.5f) / (float)A.W, sy = ((float)py + 0.5f) / (float)A.H;
  float3 d;
  d.x = (A.d00.x + sx * A.du.x) + sy * A.dv.x;
  d.y = (A.d00.y + sx * A.du.y) + sy * A.dv.y;
  d.z = (A.d00.z + sx * A.du.z) + sy * A.dv.z;
  const float len = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
  d.x = d.x / len;
  d.y = d.y / len;
  d.z = d.z / len;
  const float3 o = A.pos, iv = make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  float tenter, texit;
  aabb_intersection(o, d, 0.0f, 10000.0f, A.lb, A.rt, tenter, texit);
  float4 res = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

  const float4 *__restrict__ N = A.bvh.nodes;
  const uint32_t log2K = A.bvh.log2K, kmask = (1u << log2K) - 1u, top = A.bvh.top;
  auto leaf_meets = [&](uint32_t leaf, float t0, float t1) {
    const float4 *lb = N + g_lvl_ofs[0];
    const float4 lo = lb[2 * leaf], hi = lb[2 * leaf + 1];
    return aa_seg_box(o.x, o.y, o.z, iv.x, iv.y, iv.z, t0, t1, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z) != 0;
  };
  // Nodes first .. first + K - 1 of `level` (clipped to the level's count)
  // against [t0, t1]: bit j set when node first + j meets the segment.
  // The boxes are loaded eight at a time before any is tested, so a group
  // costs one memory round trip instead of one per node (lanes past the
  // group's end reload its last box and are masked off).
  auto group_mask = [&](uint32_t level, uint32_t first, float t0, float t1) -> uint32_t {
    const float4 *lv = N + g_lvl_ofs[level] + 2u * first;
    const uint32_t n = min(kmask + 1u, g_lvl_cnt[level] - first);
    uint32_t m = 0;
    for (uint32_t h = 0; h < n; h += 8u) {
      float4 lo[8], hi[8];
#pragma unroll
      for (uint32_t j = 0; j < 8u; ++j) {
        const uint32_t c = min(h + j, n - 1u);
        lo[j] = lv[2u * c];
        hi[j] = lv[2u * c + 1u];
      }
#pragma unroll
      for (uint32_t j = 0; j < 8u; ++j)
        if (h + j < n && aa_seg_box(o.x, o.y, o.z, iv.x, iv.y, iv.z, t0, t1, lo[j].x, lo[j].y, lo[j].z, hi[j].x,
                                    hi[j].y, hi[j].z))
          m |= 1u << (h + j);
    }
    return m;
  };
  // every particle of a leaf met by [t0, t1] whose sample lies in [t0, t1),
  // in ascending order; returns how many there are (only kSplatCap kept)
  // and stops at `limit` hits (the list is then incomplete).  The
  // walk is depth first in ascending node order: `m` holds the current
  // group's children still to visit, smask[level] those of the groups above.
  auto collect = [&](float t0, float t1, uint32_t limit) -> uint32_t {
    uint32_t cnt = 0;
    if (!group_mask(top, 0u, t0, t1)) return 0; // the root
    uint32_t level = top - 1u, base = 0;
    uint32_t m = group_mask(level, 0u, t0, t1);
    for (;;) {
      if (level != 0 && m) { // open the next child
        const uint32_t node = base + (uint32_t)__builtin_ctz(m);
        smask[level * kSplatBlock] = (uint16_t)(m & (m - 1u));
        --level;
        base = node << log2K;
        m = group_mask(level, base, t0, t1);
        continue;
      }
      if (level == 0 && A.ppl == 1u && A.n != 0) {
        // one particle per leaf: the particles of up to eight met leaves
        // are loaded together, then tested in ascending order
        while (m) {
          uint32_t id[8];
          float4 pc[8];
          float pr[8];
#pragma unroll
          for (uint32_t j = 0; j < 8u; ++j) {
            id[j] = m ? base + (uint32_t)__builtin_ctz(m) : ~0u;
            m &= m - 1u;
          }
#pragma unroll
          for (uint32_t j = 0; j < 8u; ++j) {
            const uint32_t q = min(id[j], A.n - 1u); // unused slots (~0u) and leaves past n load particle n - 1
            pc[j] = A.particles[q];
            pr[j] = A.radii ? A.radii[q] : A.rbf;
          }
#pragma unroll
          for (uint32_t j = 0; j < 8u; ++j) {
            float t;
            if (id[j] < A.n &&
                aa_splat_hit(pc[j].x, pc[j].y, pc[j].z, pr[j], o.x, o.y, o.z, d.x, d.y, d.z, t0, t1, &t)) {
              if (cnt < kSplatCap) {
                hid[cnt * kSplatBlock] = id[j];
                ht[cnt * kSplatBlock] = t;
              }
              if (++cnt >= limit) return cnt;
            }
          }
        }
      }
      for (; level == 0 && m; m &= m - 1u) {
        const uint32_t leaf = base + (uint32_t)__builtin_ctz(m);
        {
          const uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.n);
          for (uint32_t p = p0; p < p1; ++p) {
            const float4 c = A.particles[p];
            const float r = A.radii ? A.radii[p] : A.rbf;
            float t;
            if (aa_splat_hit(c.x, c.y, c.z, r, o.x, o.y, o.z, d.x, d.y, d.z, t0, t1, &t)) {
              if (cnt < kSplatCap) {
                hid[cnt * kSplatBlock] = p;
                ht[cnt * kSplatBlock] = t;
              }
              if (++cnt >= limit) return cnt;
            }
          }
        }
      }
      // this group is done: back to the parent's group
      if (level == top - 1u) return cnt;
      ++level;
      base = (base >> log2K) & ~kmask;
      m = smask[level * kSplatBlock];
    }
  };

  if (tenter < texit) {
    const float pps = (float)(15 / 8);
    const float sp = pps * A.unit;
    const float fade = aa_expf(-A.unit * pps);
    float tslab = 0.0f;
    // A slab that no longer moves tslab (unit below half an ulp of t)
    // would loop for ever; such a pixel stops there.
    // slabs that end before the box hold nothing (':100')
    while (tslab < texit && !(fminf(texit, tslab + sp) > tenter) && tslab + sp > tslab) tslab = tslab + sp;
    uint32_t S = 16; // slabs per chunk
    bool done = !(tslab + sp > tslab);
    while (!done && tslab < texit) {
      if (res.w > 0.99f) break;
      // the chunk: up to S slabs from tslab
      const float c0 = fmaxf(tenter, tslab);
      float c1 = c0, tl = tslab;
      uint32_t ns = 0;
      for (; ns < S && tl < texit; ++ns) {
        c1 = fminf(texit, tl + sp);
        tl = tl + sp;
      }
      // A chunk of several slabs whose hits overflow the list is walked
      // no further than its first hit past the list (a chunk that has grown
      // long in a sparse stretch and runs into a dense region would
      // otherwise visit every leaf its whole segment meets); it composites
      // nothing and shrinks to a quarter.  A one-slab chunk walks to the
      // end and keeps the first kSplatCap hits, which hold its first 15.
      const uint32_t cnt = collect(c0, c1, S > 1 ? kSplatCap + 1u : ~0u);
      if (S > 1 && cnt > kSplatCap) {
        S = max(1u, S / 4u);
        continue;
      }
      const uint32_t m = min(cnt, kSplatCap);
      uint32_t first = 0; // list entries before the current slab
      for (uint32_t k = 0; k < ns; ++k) {
        if (res.w > 0.99f) {
          done = true;
          break;
        }
        const float tmin = fmaxf(tenter, tslab), tmax = fminf(texit, tslab + sp);
        // payload: the first 15 hits of this slab in traversal order
        uint32_t tail = 0;
        for (uint32_t e = first; e < m && tail < kPayload; ++e) {
          const float t = ht[e * kSplatBlock];
          if (!(tmin <= t && t < tmax)) continue;
          const uint32_t p = hid[e * kSplatBlock];
          if (!leaf_meets(p / A.ppl, tmin, tmax)) continue;
          pe[tail * kSplatBlock] = (uint8_t)e;
          ++tail;
        }
        // SplatPayload::sort (':44-55'): stable by t
        for (uint32_t a = 1; a < tail; ++a) {
          const uint32_t ea = pe[a * kSplatBlock];
          const float ta = ht[ea * kSplatBlock];
          uint32_t b = a;
          while (b > 0 && ht[pe[(b - 1) * kSplatBlock] * kSplatBlock] > ta) {
            pe[b * kSplatBlock] = pe[(b - 1) * kSplatBlock];
            --b;
          }
          pe[b * kSplatBlock] = (uint8_t)ea;
        }
        // front-to-back integration (':117-140')
        for (uint32_t e = 0; e < tail; ++e) {
          const uint32_t le = pe[e * kSplatBlock];
          const uint32_t p = hid[le * kSplatBlock];
          const float t = ht[le * kSplatBlock];
          const float4 c = A.particles[p];
          const float r = A.radii ? A.radii[p] : A.rbf;
          const float x = o.x + d.x * t, y = o.y + d.y * t, z = o.z + d.z * t;
          float drbf = aa_rbf(aa_dist3(x, y, z, c.x, c.y, c.z), r);
          float4 col;
          if (A.visAttr) {
            col = tf4(A.cmap, A.cm_n, c.w);
            col.w = col.w * tf_r(A.dmap, A.dm_n, drbf);
          } else {
            const float4 cm = tf4(A.cmap, A.cm_n, drbf);
            col = make_float4(cm.x, cm.y, cm.z, tf_r(A.dmap, A.dm_n, drbf));
          }
          col.w = col.w * fade;
          over4(res, col);
        }
        if (!(tslab + sp > tslab)) {
          done = true;
          break;
        }
        tslab = tslab + sp;
        // entries before this slab's end are spent (the list is in index
        // order, not t order, so only a prefix can be skipped)
        while (first < m && ht[first * kSplatBlock] < tmax) ++first;
      }
      if (cnt <= kSplatCap / 2) S = min(2u * S, kChunkMax);
    }
  }
  over4(res, make_float4(0.0f, 0.0f, 0.0f, 1.0f));
  A.image[(uint64_t)px + (uint64_t)A.W * py] = res;
}

} // namespace
} // namespace aarbf

using namespace aarbf;

extern "C" int aarbf_render_splat(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc,
                                  int W, int H) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, rg && pc, "aarbf_render_splat: null struct");
  AARBF_REQUIRE(ctx, W > 0 && H > 0, "aarbf_render_splat: empty launch");
  AARBF_REQUIRE(ctx, rg->imageBuffer, "aarbf_render_splat: null image buffer");
  BVH *bvh = get_bvh(ctx, pc->world);
  AARBF_REQUIRE(ctx, bvh, "aarbf_render_splat: world is not a BVH handle");
  TF *cm = get_tf(ctx, pc->colormap), *dm = get_tf(ctx, pc->densitymap);
  AARBF_REQUIRE(ctx, cm && dm, "aarbf_render_splat: colormap/densitymap are not TF handles");
  AARBF_REQUIRE(ctx, pc->particlesPerLeaf >= 1, "aarbf_render_splat: particlesPerLeaf must be >= 1");
  AARBF_REQUIRE(ctx, pc->numParticles == 0 || pc->particles, "aarbf_render_splat: null particles");
  AARBF_REQUIRE(ctx, (uint64_t)bvh->counts[0] * pc->particlesPerLeaf >= pc->numParticles,
                "aarbf_render_splat: the BVH has fewer leaves than numParticles / particlesPerLeaf");
  AARBF_REQUIRE(ctx, bvh->num_levels <= kMaxLevels, "aarbf_render_splat: BVH too deep");
  AARBF_REQUIRE(ctx, pc->unit > 0.0f, "aarbf_render_splat: unit must be positive");
  hipSetDevice(ctx->device);
  SplatArgs A{};
  A.W = W;
  A.H = H;
  A.visAttr = pc->visualizeAttributes ? 1u : 0u;
  A.unit = pc->unit;
  A.rbf = pc->rbfRadius;
  A.pos = make_float3(pc->camera.pos.x, pc->camera.pos.y, pc->camera.pos.z);
  A.d00 = make_float3(pc->camera.dir_00.x, pc->camera.dir_00.y, pc->camera.dir_00.z);
  A.du = make_float3(pc->camera.dir_du.x, pc->camera.dir_du.y, pc->camera.dir_du.z);
  A.dv = make_float3(pc->camera.dir_dv.x, pc->camera.dir_dv.y, pc->camera.dir_dv.z);
  const float m = pc->rbfRadius; // the splat box grows by one radius (':66-67')
  A.rt = make_float3(rg->globalAABBMax.x + m, rg->globalAABBMax.y + m, rg->globalAABBMax.z + m);
  A.lb = make_float3(rg->globalAABBMin.x - m, rg->globalAABBMin.y - m, rg->globalAABBMin.z - m);
  A.particles = (const float4 *)pc->particles;
  A.radii = (const float *)pc->radii;
  A.n = pc->numParticles;
  A.ppl = pc->particlesPerLeaf;
  A.cmap = cm->table;
  A.cm_n = cm->n;
  A.dmap = dm->table;
  A.dm_n = dm->n;
  A.image = (float4 *)rg->imageBuffer;
  A.bvh = make_view(*bvh);
  dim3 grid((W + 7) / 8, (H + 7) / 8);
  splat_kernel<<<grid, kSplatBlock, 0, ctx->stream>>>(A);
  AARBF_CHECK_LAUNCH(ctx, "splat_kernel");
  return AARBF_OK;
}
