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
#include "render_common.h"

#include <cstring>
#include <mutex>
#include <vector>

namespace aarbf {
namespace {

// (deviceCodeRBF.hlsl:350-351,366-376).
__global__ __launch_bounds__(256) void bake_kernel(const float4 *__restrict__ particles,
                                                   const float *__restrict__ radii, uint32_t n,
                                                   float rbfRadius, const float4 *__restrict__ rmap,
                                                   uint32_t rm_n, const float4 *__restrict__ cmap,
                                                   uint32_t cm_n, float4 *__restrict__ pr,
                                                   float4 *__restrict__ pc, float *__restrict__ pm) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 q = particles[i];
  float rp = tf_r(rmap, rm_n, q.w);
  float radius = (radii ? radii[i] : rbfRadius) * rp;
  float4 c = tf4(cmap, cm_n, q.w);
  const float r = c.w == 0.0f ? 0.0f : radius;
  pr[i] = make_float4(q.x, q.y, q.z, r);
  pm[i] = aa_rbf_coef(r);
  pc[i] = make_float4(aa_cube(c.w), aa_powf(c.x, 2.2f), aa_powf(c.y, 2.2f), aa_powf(c.z, 2.2f));
}


// FP16 payload, one thread per particle, 32 lanes per chunk of 32
// Hilbert-consecutive (so spatially compact) particles:
//  * positions: the chunk's centre box gives an origin and a per-axis step
//    ((max - min) / 65535, in double, stored as float); each coordinate
//    rounds to the nearest grid point;
//  * radius: shrunk by the particle's rounding shift (so the decoded sphere
//    lies inside the original one, and so inside its FP32 leaf box: the BVH
//    stays valid), then coded against the chunk's range of shrunk radii
//    (code 1 = the smallest, 65535 = the largest; 0 = radius 0), rounded
//    down -- the encoder evaluates the decoder's own expression;
//  * alpha (cubed when decoded) as 16-bit unorm, rounded to nearest, a
//    positive alpha at least one step (a zero alpha would leave points with
//    count > 0 and density 0, whose colour divides 0 by 0);
//  * rgb^2.2 as binary16, rounded to nearest (subnormal halves included:
//    dark channels keep an absolute precision of 2^-25).
__device__ inline float chunk_min(float v) {
  for (int m = 16; m > 0; m >>= 1) v = fminf(v, __shfl_xor(v, m, 32));
  return v;
}
__device__ inline float chunk_max(float v) {
  for (int m = 16; m > 0; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 32));
  return v;
}
__device__ inline double chunk_min(double v) {
  for (int m = 16; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m, 32));
  return v;
}
__device__ inline double chunk_max(double v) {
  for (int m = 16; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 32));
  return v;
}

__global__ __launch_bounds__(256) void bake_q(const float4 *__restrict__ particles, const float *__restrict__ radii,
                                              uint32_t n, float rbfRadius, const float4 *__restrict__ rmap,
                                              uint32_t rm_n, const float4 *__restrict__ cmap, uint32_t cm_n,
                                              float4 *__restrict__ qc, uint4 *__restrict__ qp,
                                              float *__restrict__ pm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n; // every lane of a chunk takes part in its reductions
  const float4 q = live ? particles[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float v[3] = {q.x, q.y, q.z};
  float ov[3], sv[3];
  for (int d = 0; d < 3; ++d) {
    ov[d] = chunk_min(live ? v[d] : AA_INF_F);
    const float hi = chunk_max(live ? v[d] : -AA_INF_F);
    sv[d] = (float)(((double)hi - (double)ov[d]) / 65535.0);
  }
  uint32_t qa[3];
  double shift2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double t = sv[d] > 0.0f ? rint(((double)v[d] - (double)ov[d]) / (double)sv[d]) : 0.0;
    qa[d] = (uint32_t)fmin(fmax(t, 0.0), 65535.0);
    const double e = (double)fx16(qa[d], sv[d], ov[d]) - (double)v[d];
    shift2 += e * e;
  }
  float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  double rr = 0.0;
  if (live) {
    const float rp = tf_r(rmap, rm_n, q.w);
    const float radius = (radii ? radii[i] : rbfRadius) * rp;
    c = tf4(cmap, cm_n, q.w);
    const float r = c.w == 0.0f ? 0.0f : radius;
    if (r > 0.0f) rr = ((double)r - sqrt(shift2)) * (1.0 - 0x1p-30);
  }
  const bool has_r = rr > 0.0;
  // chunk radius range: base + code * step, code 1 .. 65535
  const double rmin = chunk_min(has_r ? rr : (double)AA_INF_F);
  const double rmax = chunk_max(has_r ? rr : -(double)AA_INF_F);
  float r_step = 0.0f, r_base = 0.0f;
  if (rmin < (double)AA_INF_F) {
    r_step = (float)((rmax - rmin) / 65534.0);
    r_base = (float)(rmin - (double)r_step);
    // code 1 must not decode above the smallest shrunk radius
    while ((double)fx16(1u, r_step, r_base) > rmin) r_base = nextafterf(r_base, -AA_INF_F);
  }
  uint32_t qr = 0;
  if (has_r) {
    double t = r_step > 0.0f ? floor((rr - (double)r_base) / (double)r_step) : 1.0;
    qr = (uint32_t)fmin(fmax(t, 1.0), 65535.0);
    while (qr > 1u && (double)fx16(qr, r_step, r_base) > rr) --qr;
    if ((double)fx16(qr, r_step, r_base) > rr || !(fx16(qr, r_step, r_base) > 0.0f)) qr = 0u; // cannot shrink: drop
  }
  if ((i & 31u) == 0u && live) {
    qc[2 * (i >> kQChunkLog2)] = make_float4(ov[0], ov[1], ov[2], r_base);
    qc[2 * (i >> kQChunkLog2) + 1] = make_float4(sv[0], sv[1], sv[2], r_step);
  }
  if (!live) return;
  uint32_t au = (uint32_t)fminf(rintf(c.w * 65535.0f), 65535.0f);
  if (au == 0u && c.w > 0.0f) au = 1u;
  pm[i] = aa_rbf_coef(fx16_r(qr, r_step, r_base)); // of the decoded radius
  qp[i] = make_uint4(qa[0] | (qa[1] << 16), qa[2] | (qr << 16), au | (aa_f2h_rn(aa_powf(c.x, 2.2f)) << 16),
                     aa_f2h_rn(aa_powf(c.y, 2.2f)) | (aa_f2h_rn(aa_powf(c.z, 2.2f)) << 16));
}

// The decoded payload (x, y, z, r, alpha^3, rgb^2.2 per particle) of either format.
__global__ __launch_bounds__(256) void payload_decode(RenderArgs A, float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.numParticles) return;
  const float4 a = ld_pr(A, i), b = ld_pc(A, i);
  float *o = out + 8 * (uint64_t)i;
  o[0] = a.x;
  o[1] = a.y;
  o[2] = a.z;
  o[3] = a.w;
  o[4] = b.x;
  o[5] = b.y;
  o[6] = b.z;
  o[7] = b.w;
}

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
  } else { // stochasticTracking ':141-203', majorant 1; free flights in product form (aa_flight_pos)
    t = 0.0f;
    float tb = 0.0f, P = 1.0f;
    for (; i < A.max_depth; ++i) {
      if (STATS) c.steps++;
      float xi = aa_lcg_next(&rng);
      P = P * (1.0f - xi);
      t = aa_flight_pos(tb, P, A.unit);
      if ((i + 1u) % AA_FLIGHT_BLOCK == 0u) {
        tb = t;
        P = 1.0f;
      }
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

} // namespace
} // namespace aarbf

using namespace aarbf;

// getLightDirection (sharedCode.h:127-132), evaluated once on the host
// (camera.cpp).
static float3 light_direction(float azimuth, float elevation) {
  float l[3];
  aarbf_light_direction(azimuth, elevation, l);
  return make_float3(l[0], l[1], l[2]);
}

// Per-view fields of a render call (camera, counters, light, buffers).
static ReqArgs view_args(const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W, int H, uint32_t spp,
                         const STBN *stbn) {
  ReqArgs R{};
  R.pos = make_float3(pc->camera.pos.x, pc->camera.pos.y, pc->camera.pos.z);
  R.d00 = make_float3(pc->camera.dir_00.x, pc->camera.dir_00.y, pc->camera.dir_00.z);
  R.du = make_float3(pc->camera.dir_du.x, pc->camera.dir_du.y, pc->camera.dir_du.z);
  R.dv = make_float3(pc->camera.dir_dv.x, pc->camera.dir_dv.y, pc->camera.dir_dv.z);
  R.light = light_direction(pc->light.azimuth, pc->light.elevation);
  R.amb = pc->light.ambient;
  R.W = W;
  R.H = H;
  R.tiles_x = (uint32_t)(W + 7) / 8;
  R.frameID = pc->frameID;
  R.accumID = pc->accumID;
  R.spp = spp;
  R.blue = pc->enableBlueNoise ? 1u : 0u;
  R.showNoise = pc->showNoise ? 1u : 0u;
  R.stbn = stbn ? stbn->rgba : nullptr;
  R.stbn_w = stbn ? stbn->width : 0;
  R.stbn_h = stbn ? stbn->height : 0;
  R.image = (float4 *)rg->imageBuffer;
  R.accum = (float4 *)rg->accumBuffer;
  R.samples = (uint64_t)W * (uint64_t)H * spp;
  return R;
}

// Size of the baked payload of n particles in float4 units: FP32 32 B per
// particle, FP16 16 B per particle then 32 B per chunk, whichever format is
// larger, followed by the RBF coefficients (one float per particle).
static size_t payload_float4s(uint32_t n, size_t *pm_off) {
  size_t need = 2 * (size_t)(n ? n : 1);
  const size_t need16 = (size_t)(n ? n : 1) + 2 * (size_t)((n + 31u) >> kQChunkLog2);
  if (need16 > need) need = need16;
  *pm_off = need;
  return need + ((size_t)(n ? n : 1) + 3) / 4;
}

// Validates a render call, bakes the particle payload and fills the
// kernel arguments.  *stbn_out: the call's blue-noise atlas (or null).
// fixed_payload: a device buffer of at least payload_float4s(n) float4 for
// the baked payload (checked by the caller), or null for the context's
// grow-only buffer.
static int render_setup(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W,
                        int H, uint32_t spp, RenderArgs &A, BVH *&bvh_out, STBN *&stbn_out,
                        float4 *fixed_payload = nullptr) {
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
  bvh_out = bvh;
  stbn_out = stbn;
  hipSetDevice(ctx->device);
  uint32_t n = pc->numParticles;
  size_t pm_off = 0;
  const size_t need = payload_float4s(n, &pm_off);
  float4 *buf = fixed_payload;
  if (!buf) {
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
    buf = ctx->baked;
  }
  float4 *pr = buf, *pcol = buf + (n ? n : 1);
  float *pm = (float *)(buf + pm_off);
  uint4 *qp = (uint4 *)buf;
  float4 *qc = buf + (n ? n : 1);
  const bool q16 = ctx->payload == 1;
  if (n && !q16)
    bake_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const float4 *)pc->particles, (const float *)pc->radii,
                                                          n, pc->rbfRadius, rm->table, rm->n, cm->table, cm->n, pr,
                                                          pcol, pm);
  if (n && q16)
    bake_q<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const float4 *)pc->particles, (const float *)pc->radii, n,
                                                     pc->rbfRadius, rm->table, rm->n, cm->table, cm->n, qc, qp, pm);
  AARBF_CHECK_LAUNCH(ctx, "bake");
  if (!fixed_payload) { // what aarbf_payload_read decodes
    ctx->baked_fmt = q16 ? 1 : 0;
    ctx->baked_n = n;
  }
  A = RenderArgs{};
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
  A.inv_unit = 1.0f / pc->unit;
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
  A.pm = pm;
  A.qp = qp;
  A.qc = qc;
  A.q16 = q16 ? 1u : 0u;
  A.numParticles = n;
  A.ppl = pc->particlesPerLeaf;
  A.cmap = cm->table;
  A.cm_n = cm->n;
  A.dmap = dm->table;
  A.dm_n = dm->n;
  A.dm_nf = (float)dm->n;
  A.stbn = stbn ? stbn->rgba : nullptr;
  A.stbn_w = stbn ? stbn->width : 0;
  A.stbn_h = stbn ? stbn->height : 0;
  A.image = (float4 *)rg->imageBuffer;
  A.accum = (float4 *)rg->accumBuffer;
  A.stats = ctx->d_stats;
  A.bvh = make_view(*bvh);
  return AARBF_OK;
}

// The wavefront schedule needs the quantised tree within its LDS stack.
static bool wavefront_ok(const BVH *bvh) { return bvh->num_levels <= kStackLevels && bvh->groups; }

extern "C" {

int aarbf_render_rbf_spp(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W,
                         int H, uint32_t spp) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  RenderArgs A;
  BVH *bvh = nullptr;
  STBN *stbn = nullptr;
  int rc = render_setup(ctx, rg, pc, W, H, spp, A, bvh, stbn);
  if (rc != AARBF_OK) return rc;
  if (ctx->render_mode == 1 || !wavefront_ok(bvh)) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    if (ctx->stats)
      render_kernel<true><<<grid, 256, 0, ctx->stream>>>(A);
    else
      render_kernel<false><<<grid, 256, 0, ctx->stream>>>(A);
    AARBF_CHECK_LAUNCH(ctx, "render_kernel");
    return AARBF_OK;
  }
  const ReqArgs R = view_args(rg, pc, W, H, spp, stbn);
  return render_wavefront(ctx, A, &R, 1, false);
}

int aarbf_render_batch(aarbf_context *ctx, const aarbf_request *reqs, uint32_t n) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, reqs && n >= 1, "aarbf_render_batch: no requests");
  const aarbf_push_constants &p0 = reqs[0].pc;
  const aarbf_raygen_data &g0 = reqs[0].rg;
  uint32_t max_spp = 0;
  std::vector<ReqArgs> R(n);
  std::vector<STBN *> atlas(n, nullptr);
  for (uint32_t i = 0; i < n; ++i) {
    const aarbf_request &q = reqs[i];
    AARBF_REQUIRE(ctx, q.width > 0 && q.height > 0 && q.spp >= 1, "aarbf_render_batch: empty request");
    AARBF_REQUIRE(ctx, q.rg.imageBuffer && q.rg.accumBuffer, "aarbf_render_batch: null image/accum buffer");
    const aarbf_push_constants &p = q.pc;
    // the scene is shared: every field that is not per view must match request 0
    AARBF_REQUIRE(ctx,
                  p.world == p0.world && p.particles == p0.particles && p.numParticles == p0.numParticles &&
                      p.particlesPerLeaf == p0.particlesPerLeaf && p.colormap == p0.colormap &&
                      p.radiusmap == p0.radiusmap && p.densitymap == p0.densitymap && p.radii == p0.radii &&
                      p.rbfRadius == p0.rbfRadius && p.unit == p0.unit &&
                      p.visualizeAttributes == p0.visualizeAttributes &&
                      !memcmp(&q.rg.globalAABBMin, &g0.globalAABBMin, sizeof(g0.globalAABBMin)) &&
                      !memcmp(&q.rg.globalAABBMax, &g0.globalAABBMax, sizeof(g0.globalAABBMax)),
                  "aarbf_render_batch: requests must share the scene (world, particles, radii, transfer "
                  "functions, rbfRadius, unit, visualizeAttributes, global AABB)");
    if (p.enableBlueNoise) {
      atlas[i] = get_stbn(ctx, q.rg.stbnTexture);
      AARBF_REQUIRE(ctx, atlas[i], "aarbf_render_batch: blue noise enabled without an STBN atlas");
    }
    max_spp = std::max(max_spp, q.spp);
  }
  RenderArgs A;
  BVH *bvh = nullptr;
  STBN *stbn = nullptr;
  int rc = render_setup(ctx, &g0, &p0, reqs[0].width, reqs[0].height, max_spp, A, bvh, stbn);
  if (rc != AARBF_OK) return rc;
  AARBF_REQUIRE(ctx, wavefront_ok(bvh), "aarbf_render_batch: the BVH is too deep for the wavefront schedule");
  A.blue = 0;
  for (uint32_t i = 0; i < n; ++i) {
    R[i] = view_args(&reqs[i].rg, &reqs[i].pc, reqs[i].width, reqs[i].height, reqs[i].spp, atlas[i]);
    A.blue |= R[i].blue;
  }
  return render_wavefront(ctx, A, R.data(), n, true);
}

int aarbf_set_payload(aarbf_context *ctx, int format) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, format == 0 || format == 1, "aarbf_set_payload: format must be 0 (FP32) or 1 (FP16)");
  ctx->payload = format;
  return AARBF_OK;
}

int aarbf_payload_read(aarbf_context *ctx, uint64_t out, uint32_t n) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, out && ctx->baked && n == ctx->baked_n, "aarbf_payload_read: no payload of that size");
  RenderArgs A{};
  const uint32_t m = n ? n : 1;
  A.pr = ctx->baked;
  A.pc = ctx->baked + m;
  A.qp = (const uint4 *)ctx->baked;
  A.qc = ctx->baked + m;
  A.q16 = ctx->baked_fmt ? 1u : 0u;
  A.numParticles = n;
  if (n) payload_decode<<<(n + 255) / 256, 256, 0, ctx->stream>>>(A, (float *)out);
  AARBF_CHECK_LAUNCH(ctx, "payload_decode");
  return AARBF_OK;
}

int aarbf_render_rbf(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W,
                     int H) {
  return aarbf_render_rbf_spp(ctx, rg, pc, W, H, 1);
}

} // extern "C"

// ---- the allocation-free render entry point ----------------------------------
//
// aarbf_render_rbf_spp_noalloc bakes the particle payload into this
// file-scope array instead of the context's grow-only buffer, so the call
// allocates nothing.  Each device holds its own copy of the array (for the
// life of the process) at its own address; host code uses that address,
// taken once per device with hipGetSymbolAddress when a context is created
// on the device, never the array's name, which on the host is only a
// stand-in for the device symbol.
static constexpr size_t kNoallocFloat4s =
    2 * (size_t)AARBF_NOALLOC_MAX_PARTICLES + ((size_t)AARBF_NOALLOC_MAX_PARTICLES + 3) / 4;
static __device__ __align__(256) float4 noalloc_scratch[kNoallocFloat4s];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_lock;
float4 *g_noalloc_ptr[kMaxDevices] = {};
} // namespace

namespace aarbf {
// Called by aarbf_context_create with `device` current.
int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return AARBF_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> g(g_noalloc_lock);
  if (g_noalloc_ptr[device]) return AARBF_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)) != hipSuccess || !p) return AARBF_ERR_HIP;
  g_noalloc_ptr[device] = (float4 *)p;
  return AARBF_OK;
}
} // namespace aarbf

extern "C" int aarbf_render_rbf_spp_noalloc(aarbf_context *ctx, const aarbf_raygen_data *rg,
                                            const aarbf_push_constants *pc, int W, int H, uint32_t spp) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, pc, "aarbf_render_rbf_spp_noalloc: null struct");
  AARBF_REQUIRE(ctx, pc->numParticles <= AARBF_NOALLOC_MAX_PARTICLES,
                "aarbf_render_rbf_spp_noalloc: numParticles exceeds AARBF_NOALLOC_MAX_PARTICLES (1,048,576); "
                "use aarbf_render_rbf_spp");
  float4 *buf = nullptr;
  {
    std::lock_guard<std::mutex> g(g_noalloc_lock);
    buf = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_noalloc_ptr[ctx->device] : nullptr;
  }
  AARBF_REQUIRE(ctx, buf, "aarbf_render_rbf_spp_noalloc: the device's scratch array was not set up");
  size_t pm_off = 0;
  AARBF_REQUIRE(ctx, payload_float4s(pc->numParticles, &pm_off) <= kNoallocFloat4s,
                "aarbf_render_rbf_spp_noalloc: payload larger than the scratch array");
  RenderArgs A;
  BVH *bvh = nullptr;
  STBN *stbn = nullptr;
  int rc = render_setup(ctx, rg, pc, W, H, spp, A, bvh, stbn, buf);
  if (rc != AARBF_OK) return rc;
  dim3 grid((W + 15) / 16, (H + 15) / 16);
  if (ctx->stats)
    render_kernel<true><<<grid, 256, 0, ctx->stream>>>(A);
  else
    render_kernel<false><<<grid, 256, 0, ctx->stream>>>(A);
  AARBF_CHECK_LAUNCH(ctx, "render_kernel");
  return AARBF_OK;
}
