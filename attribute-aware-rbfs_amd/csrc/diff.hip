// diff.hip -- the differentiable renderer: the expected stochastic-marching
// image and its adjoint with respect to particle positions, radii,
// attributes and the colormap / densitymap / radiusmap texels.  No
// counterpart in the reference; its forward is the expectation of
// ParticleTracker::stochasticMarching (deviceCodeRBF.hlsl:67-139, primary
// ray, ambient 1) over the blue-noise threshold.
//
// A sample marches t_k = tenter + unit * (r1 + k) exactly as the marching
// tracker does; at the points some particle covers it adds D_k =
// densitymap(density) to L and stops there with probability
// F(L_k) - F(L_{k-1}), F(L) = min(1, (1 - e^-L) / 0.99) (the chance that
// the threshold -ln(1 - 0.99 xi) lies below L).  The sample's colour is
// C = sum_k (F(L_k) - F(L_{k-1})) c_k with c_k the attribute-weighted colour
// of point k, a pixel the mean of its samples.
//
// The adjoint replays each sample's march instead of storing its steps:
// dC/dD_j = sum_{k>=j} F'(L_k) <g, c_k - c_{k+1}> is that sum over the whole
// path minus the prefix so far.  The forward keeps the whole-path vector
// sum_k F'(L_k) (c_k - c_{k+1}) per sample (12 bytes), so the adjoint needs
// one replay; without it a first replay recomputes it.  The replay scatters
// the chain rule through the densitymap, the colour mean, each covering
// particle's RBF weight and its colormap / radiusmap lookups.  Particle
// gradients go to global memory with atomics; texel gradients are summed
// per block in LDS first.  Jumps of the truncated RBF
// where a sphere boundary crosses a march point are not differentiated.
#include "render_common.h"

namespace aarbf {
namespace {

constexpr int kDBlock = 64;
constexpr uint32_t kTexMax = 64;  // texels per table with LDS gradient sums (larger tables: global atomics)
constexpr uint32_t kCache = 4;    // per-lane particle adjoint cache entries
constexpr uint32_t kSlot = 9;     // particle index + 8 adjoint sums
constexpr uint32_t kRec = 16;     // covering particles a march point keeps for the adjoint

struct DiffArgs {
  RenderArgs A;
  const float4 *particles; // raw x, y, z, attribute
  const float *radii;
  const float4 *rmap;
  uint32_t rm_n;
  uint32_t spp;
  float4 *out;             // image (forward)
  float4 *path;            // per sample: sum_k F'(L_k) (c_k - c_{k+1}) (forward writes, adjoint reads), or null
  const float4 *gimg;      // dLoss/dimage (adjoint)
  float *g_pos, *g_rad, *g_attr, *g_cm, *g_dm, *g_rm;
};

__global__ __launch_bounds__(256) void diff_bake(const float4 *__restrict__ particles, const float *__restrict__ radii,
                                                 uint32_t n, const float4 *__restrict__ rmap, uint32_t rm_n,
                                                 const float4 *__restrict__ cmap, uint32_t cm_n,
                                                 float4 *__restrict__ pr, float4 *__restrict__ pc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 q = particles[i];
  float4 c = tf4(cmap, cm_n, q.w);
  float radius = radii[i] * tf_r(rmap, rm_n, q.w);
  pr[i] = make_float4(q.x, q.y, q.z, c.w == 0.0f ? 0.0f : radius);
  pc[i] = make_float4(aa_cube(c.w), aa_powf(c.x, 2.2f), aa_powf(c.y, 2.2f), aa_powf(c.z, 2.2f));
}

// value, slope and texel weights of a linearly filtered table channel
__device__ inline float tf_grad(const float4 *t, uint32_t n, int c, float u, float &slope, uint32_t &j0, uint32_t &j1,
                                float &f) {
  aa_tf_coords(u, n, &j0, &j1, &f);
  const float *a = (const float *)(t + j0), *b = (const float *)(t + j1);
  slope = j0 == j1 ? 0.0f : (float)n * (b[c] - a[c]);
  return aa_lerp_tex(a[c], b[c], f);
}

__device__ inline float Fcdf(float L, float &dF) {
  float e = aa_expf(-L);
  float v = (1.0f - e) / 0.99f;
  if (v >= 1.0f) {
    dF = 0.0f;
    return 1.0f;
  }
  dF = e / 0.99f;
  return v;
}

// every particle covering x, in ascending order (the stackless walk of
// point_query with a callback)
template <class Fn>
__device__ inline void for_covering(const RenderArgs &A, float x, float y, float z, Fn fn) {
  const float4 *__restrict__ N = A.bvh.nodes;
  const uint32_t log2K = A.bvh.log2K, kmask = (1u << log2K) - 1u, top = A.bvh.top;
  uint32_t level = top, idx = 0;
  for (;;) {
    const float4 *lv = N + g_lvl_ofs[level];
    float4 lo = lv[2 * idx], hi = lv[2 * idx + 1];
    if (inside(lo, hi, x, y, z)) {
      if (level != 1) {
        --level;
        idx <<= log2K;
        continue;
      }
      uint32_t l0 = idx << log2K, l1 = min(l0 + kmask + 1u, g_lvl_cnt[0]);
      for (uint32_t leaf = l0; leaf < l1; ++leaf) {
        const float4 *lb = N + g_lvl_ofs[0];
        if (!inside(lb[2 * leaf], lb[2 * leaf + 1], x, y, z)) continue;
        uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
        for (uint32_t p = p0; p < p1; ++p) {
          float4 q = A.pr[p];
          float d = aa_dist3(q.x, q.y, q.z, x, y, z);
          if (d < q.w) fn(p, q, d);
        }
      }
    }
    for (;;) {
      if (level == top) return;
      if ((idx & kmask) != kmask && idx + 1 < g_lvl_cnt[level]) {
        ++idx;
        break;
      }
      ++level;
      idx >>= log2K;
    }
  }
}

struct DPay {
  uint32_t count;
  float rho, S[3];
};

__device__ inline void payload(const RenderArgs &A, float x, float y, float z, DPay &pl) {
  pl.count = 0;
  pl.rho = pl.S[0] = pl.S[1] = pl.S[2] = 0.0f;
  for_covering(A, x, y, z, [&](uint32_t p, const float4 &q, float d) {
    float4 k = A.pc[p];
    float w = aa_rbf(d, q.w) * k.x;
    pl.count += 1;
    pl.rho += w;
    pl.S[0] += k.y * w;
    pl.S[1] += k.z * w;
    pl.S[2] += k.w * w;
  });
}

// the march of sample s of pixel (px, py): fn(x, y, z, payload) at every
// point some particle covers.  With rec != null the covering particles of
// the point (index, distance) are also kept there, up to kRec of them,
// stride kDBlock.
template <class Fn>
__device__ inline void march(const DiffArgs &D, int px, int py, uint32_t s, Fn fn, uint32_t *rec = nullptr) {
  const RenderArgs &A = D.A;
  uint32_t fid = (A.frameID + s) & 0xffffu;
  uint32_t rng = aa_rng_seed(fid, (uint32_t)px, (uint32_t)py, (uint32_t)A.W);
  float r1 = aa_lcg_next(&rng);
  float sx = ((float)px + 0.5f) / (float)A.W, sy = ((float)py + 0.5f) / (float)A.H;
  float3 d;
  d.x = (A.d00.x + sx * A.du.x) + sy * A.dv.x;
  d.y = (A.d00.y + sx * A.du.y) + sy * A.dv.y;
  d.z = (A.d00.z + sx * A.du.z) + sy * A.dv.z;
  float len = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
  d.x = d.x / len;
  d.y = d.y / len;
  d.z = d.z / len;
  float te, tx;
  aabb_intersection(A.pos, d, 0.0f, 10000.0f, A.lb, A.rt, te, tx);
  if (!(te < tx)) return;
  float t = te + A.unit * r1;
  for (uint32_t i = 0; i < A.max_depth; ++i) {
    t = t + A.unit;
    if (t >= tx) return;
    const float x = A.pos.x + t * d.x, y = A.pos.y + t * d.y, z = A.pos.z + t * d.z;
    DPay pl;
    if (rec) {
      pl.count = 0;
      pl.rho = pl.S[0] = pl.S[1] = pl.S[2] = 0.0f;
      for_covering(A, x, y, z, [&](uint32_t p, const float4 &q, float dd) {
        float4 k = A.pc[p];
        float w = aa_rbf(dd, q.w) * k.x;
        if (pl.count < kRec) {
          rec[(2 * pl.count) * kDBlock] = p;
          rec[(2 * pl.count + 1) * kDBlock] = __float_as_uint(dd);
        }
        pl.count += 1;
        pl.rho += w;
        pl.S[0] += k.y * w;
        pl.S[1] += k.z * w;
        pl.S[2] += k.w * w;
      });
    } else {
      payload(A, x, y, z, pl);
    }
    if (pl.count > 0 && pl.rho > 0.0f) fn(x, y, z, pl);
  }
}

__device__ inline void colour(const DPay &pl, float c[3]) {
  for (int k = 0; k < 3; ++k) c[k] = aa_powf(pl.S[k] / pl.rho, 1.0f / 2.2f);
}

__global__ __launch_bounds__(kDBlock) void diff_forward(DiffArgs D) {
  const RenderArgs &A = D.A;
  if (threadIdx.x < kMaxLevels) {
    g_lvl_ofs[threadIdx.x] = A.bvh.level_ofs[threadIdx.x];
    g_lvl_cnt[threadIdx.x] = A.bvh.counts[threadIdx.x];
  }
  __syncthreads();
  const int px = (int)(blockIdx.x * 8u + (threadIdx.x & 7u)), py = (int)(blockIdx.y * 8u + (threadIdx.x >> 3));
  if (px >= A.W || py >= A.H) return;
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t s = 0; s < D.spp; ++s) {
    float L = 0.0f, C[3] = {0.0f, 0.0f, 0.0f}, V[3] = {0.0f, 0.0f, 0.0f}, prev[3] = {0.0f, 0.0f, 0.0f};
    bool have = false;
    march(D, px, py, s, [&](float, float, float, const DPay &pl) {
      float sl, f, dF;
      uint32_t j0, j1;
      float Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, sl, j0, j1, f);
      float c[3];
      colour(pl, c);
      float F0 = Fcdf(L, dF);
      // the previous point's dC/dL term, now that c_k is known
      if (have)
        for (int k = 0; k < 3; ++k) V[k] += dF * (prev[k] - c[k]);
      have = true;
      L += Dv;
      float F1 = Fcdf(L, dF);
      for (int k = 0; k < 3; ++k) {
        C[k] += (F1 - F0) * c[k];
        prev[k] = c[k];
      }
    });
    {
      float dF;
      Fcdf(L, dF);
      for (int k = 0; k < 3; ++k) V[k] += dF * prev[k];
    }
    if (D.path)
      D.path[((uint64_t)py * A.W + px) * D.spp + s] = make_float4(V[0], V[1], V[2], 0.0f);
    for (int k = 0; k < 3; ++k) acc[k] += C[k];
  }
  const float inv = 1.0f / (float)D.spp;
  D.out[(uint64_t)px + (uint64_t)A.W * py] = make_float4(acc[0] * inv, acc[1] * inv, acc[2] * inv, 1.0f);
}

__global__ __launch_bounds__(kDBlock) void diff_adjoint(DiffArgs D) {
  const RenderArgs &A = D.A;
  __shared__ float s_gcm[4 * kTexMax], s_gdm[kTexMax], s_grm[kTexMax];
  __shared__ uint32_t s_cache[kCache * kSlot * kDBlock];
  __shared__ uint32_t s_rec[2 * kRec * kDBlock];
  uint32_t *rec = s_rec + threadIdx.x;
  if (threadIdx.x < kMaxLevels) {
    g_lvl_ofs[threadIdx.x] = A.bvh.level_ofs[threadIdx.x];
    g_lvl_cnt[threadIdx.x] = A.bvh.counts[threadIdx.x];
  }
  for (uint32_t j = threadIdx.x; j < 4 * kTexMax; j += kDBlock) s_gcm[j] = 0.0f;
  for (uint32_t j = threadIdx.x; j < kTexMax; j += kDBlock) s_gdm[j] = s_grm[j] = 0.0f;
  __syncthreads();
  const bool lds_cm = A.cm_n <= kTexMax, lds_dm = A.dm_n <= kTexMax, lds_rm = D.rm_n <= kTexMax;
  auto add_cm = [&](uint32_t j, float v) {
    if (lds_cm) atomicAdd(&s_gcm[j], v);
    else atomicAdd(&D.g_cm[j], v);
  };
  auto add_dm = [&](uint32_t j, float v) {
    if (lds_dm) atomicAdd(&s_gdm[j], v);
    else atomicAdd(&D.g_dm[j], v);
  };
  auto add_rm = [&](uint32_t j, float v) {
    if (lds_rm) atomicAdd(&s_grm[j], v);
    else atomicAdd(&D.g_rm[j], v);
  };
  // A particle covers runs of consecutive march points, so its per-point
  // adjoints (position, radius, alpha^3, colour^2.2) are summed in a small
  // direct-mapped per-lane cache and pushed through its radiusmap /
  // colormap lookups and to global memory once per run.
  uint32_t *cache = s_cache + threadIdx.x; // [kCache][kSlot] words, stride kDBlock
  auto word = [&](uint32_t e, uint32_t w) -> float & { return *(float *)&cache[(e * kSlot + w) * kDBlock]; };
  auto flush = [&](uint32_t e) {
    const uint32_t p = cache[e * kSlot * kDBlock];
    if (p == 0xffffffffu) return;
    cache[e * kSlot * kDBlock] = 0xffffffffu;
    atomicAdd(&D.g_pos[3 * p], word(e, 1));
    atomicAdd(&D.g_pos[3 * p + 1], word(e, 2));
    atomicAdd(&D.g_pos[3 * p + 2], word(e, 3));
    const float a_r = word(e, 4), a_a3 = word(e, 5);
    const float4 raw = D.particles[p];
    const float4 k = A.pc[p];
    float sl, f, a_attr = 0.0f;
    uint32_t k0, k1;
    // r = radius * radiusmap(attribute)
    const float rmv = tf_grad(D.rmap, D.rm_n, 0, raw.w, sl, k0, k1, f);
    atomicAdd(&D.g_rad[p], a_r * rmv);
    const float a_rm = a_r * D.radii[p];
    add_rm(k0, a_rm * (1.0f - f));
    add_rm(k1, a_rm * f);
    a_attr += a_rm * sl;
    // alpha^3 and colour^2.2 from colormap(attribute)
    for (int c2 = 0; c2 < 4; ++c2) {
      const float cv = tf_grad(A.cmap, A.cm_n, c2, raw.w, sl, k0, k1, f);
      float a_c;
      if (c2 == 3) {
        a_c = a_a3 * 3.0f * cv * cv;
      } else {
        const float c22 = c2 == 0 ? k.y : (c2 == 1 ? k.z : k.w);
        a_c = cv > 0.0f ? word(e, 6 + c2) * 2.2f * (c22 / cv) : 0.0f; // d(c^2.2)/dc = 2.2 c^2.2 / c
      }
      add_cm(4 * k0 + c2, a_c * (1.0f - f));
      add_cm(4 * k1 + c2, a_c * f);
      a_attr += a_c * sl;
    }
    atomicAdd(&D.g_attr[p], a_attr);
  };
  // the cache entry of particle p (words strided by kDBlock: use word())
  auto slot = [&](uint32_t p) -> uint32_t {
    const uint32_t e = p & (kCache - 1u);
    if (cache[e * kSlot * kDBlock] != p) {
      flush(e);
      cache[e * kSlot * kDBlock] = p;
      for (uint32_t w = 1; w < kSlot; ++w) word(e, w) = 0.0f;
    }
    return e;
  };
  const int px = (int)(blockIdx.x * 8u + (threadIdx.x & 7u)), py = (int)(blockIdx.y * 8u + (threadIdx.x >> 3));
  if (px < A.W && py < A.H) {
    const float4 gi = D.gimg[(uint64_t)px + (uint64_t)A.W * py];
    const float inv = 1.0f / (float)D.spp;
    const float g[3] = {gi.x * inv, gi.y * inv, gi.z * inv};
    for (uint32_t s = 0; s < D.spp; ++s) {
      for (uint32_t e = 0; e < kCache; ++e) cache[e * kSlot * kDBlock] = 0xffffffffu;
      // pass 1: the total of dC/dL_k over the path (kept by the forward when
      // it had a path buffer)
      float L = 0.0f, total = 0.0f, prev[3] = {0, 0, 0};
      bool have = false;
      if (D.path) {
        const float4 V = D.path[((uint64_t)py * A.W + px) * D.spp + s];
        total = g[0] * V.x + g[1] * V.y + g[2] * V.z;
        have = true;
      } else march(D, px, py, s, [&](float, float, float, const DPay &pl) {
        float sl, f, dF;
        uint32_t j0, j1;
        float Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, sl, j0, j1, f);
        float c[3];
        colour(pl, c);
        if (have) {
          Fcdf(L, dF);
          total += dF * (g[0] * (prev[0] - c[0]) + g[1] * (prev[1] - c[1]) + g[2] * (prev[2] - c[2]));
        }
        L += Dv;
        prev[0] = c[0];
        prev[1] = c[1];
        prev[2] = c[2];
        have = true;
      });
      if (!have) continue;
      if (!D.path) {
        float dF;
        Fcdf(L, dF);
        total += dF * (g[0] * prev[0] + g[1] * prev[1] + g[2] * prev[2]);
      }
      // pass 2: dC/dD_j = total - prefix, and the chain rule at every point
      float Lp = 0.0f, prefix = 0.0f;
      have = false;
      march(D, px, py, s, [&](float x, float y, float z, const DPay &pl) {
        float sD, fD, dF;
        uint32_t j0, j1;
        float Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, sD, j0, j1, fD);
        float c[3];
        colour(pl, c);
        if (have) {
          Fcdf(Lp, dF);
          prefix += dF * (g[0] * (prev[0] - c[0]) + g[1] * (prev[1] - c[1]) + g[2] * (prev[2] - c[2]));
        }
        const float aD = total - prefix;
        float F0 = Fcdf(Lp, dF);
        Lp += Dv;
        float F1 = Fcdf(Lp, dF);
        const float pj = F1 - F0;
        add_dm(j0, aD * (1.0f - fD));
        add_dm(j1, aD * fD);
        float a_rho = aD * sD, a_S[3];
        for (int k = 0; k < 3; ++k) {
          const float gc = g[k] * pj, m = pl.S[k] / pl.rho;
          const float dcdm = m > 0.0f ? (1.0f / 2.2f) * aa_powf(m, 1.0f / 2.2f - 1.0f) : 0.0f;
          a_S[k] = gc * dcdm / pl.rho;
          a_rho -= gc * dcdm * m / pl.rho;
        }
        auto scatter = [&](uint32_t p, const float4 &q, float d) {
          const float4 k = A.pc[p];
          const float qq = (d * 3.0f) / q.w;
          const float rbf = aa_expf(-0.5f * (qq * qq)), w = rbf * k.x;
          const float a_w = a_rho + a_S[0] * k.y + a_S[1] * k.z + a_S[2] * k.w;
          const float a_rbf = a_w * k.x;
          const float kk = d > 0.0f ? a_rbf * (-9.0f / (q.w * q.w) * rbf) : 0.0f; // d(rbf)/dd / d
          const uint32_t e = slot(p);
          word(e, 1) += kk * (q.x - x);
          word(e, 2) += kk * (q.y - y);
          word(e, 3) += kk * (q.z - z);
          word(e, 4) += a_rbf * (9.0f * d * d / (q.w * q.w * q.w) * rbf); // a_r
          word(e, 5) += a_w * rbf;                                         // a_a3
          word(e, 6) += a_S[0] * w;                                        // a_c22
          word(e, 7) += a_S[1] * w;
          word(e, 8) += a_S[2] * w;
        };
        if (pl.count <= kRec) { // the covering particles the march kept
          for (uint32_t r = 0; r < pl.count; ++r) {
            const uint32_t p = rec[(2 * r) * kDBlock];
            scatter(p, A.pr[p], __uint_as_float(rec[(2 * r + 1) * kDBlock]));
          }
        } else {
          for_covering(A, x, y, z, scatter);
        }
        prev[0] = c[0];
        prev[1] = c[1];
        prev[2] = c[2];
        have = true;
      }, rec);
      for (uint32_t e = 0; e < kCache; ++e) flush(e);
    }
  }
  __syncthreads();
  if (lds_cm)
    for (uint32_t j = threadIdx.x; j < 4 * A.cm_n; j += kDBlock)
      if (s_gcm[j] != 0.0f) atomicAdd(&D.g_cm[j], s_gcm[j]);
  if (lds_dm)
    for (uint32_t j = threadIdx.x; j < A.dm_n; j += kDBlock)
      if (s_gdm[j] != 0.0f) atomicAdd(&D.g_dm[j], s_gdm[j]);
  if (lds_rm)
    for (uint32_t j = threadIdx.x; j < D.rm_n; j += kDBlock)
      if (s_grm[j] != 0.0f) atomicAdd(&D.g_rm[j], s_grm[j]);
}

} // namespace
} // namespace aarbf

using namespace aarbf;

static int diff_setup(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W, int H,
                      uint32_t spp, DiffArgs &D) {
  AARBF_REQUIRE(ctx, rg && pc, "aarbf_diff: null struct");
  AARBF_REQUIRE(ctx, W > 0 && H > 0 && spp >= 1, "aarbf_diff: empty launch");
  AARBF_REQUIRE(ctx, pc->radii, "aarbf_diff: per-particle radii are required");
  AARBF_REQUIRE(ctx, pc->visualizeAttributes, "aarbf_diff: only the attribute-colour mode is differentiable");
  BVH *bvh = get_bvh(ctx, pc->world);
  AARBF_REQUIRE(ctx, bvh, "aarbf_diff: world is not a BVH handle");
  AARBF_REQUIRE(ctx, bvh->num_levels <= kMaxLevels, "aarbf_diff: BVH too deep");
  TF *cm = get_tf(ctx, pc->colormap), *rm = get_tf(ctx, pc->radiusmap), *dm = get_tf(ctx, pc->densitymap);
  AARBF_REQUIRE(ctx, cm && rm && dm, "aarbf_diff: colormap/radiusmap/densitymap are not TF handles");
  AARBF_REQUIRE(ctx, pc->particlesPerLeaf >= 1, "aarbf_diff: particlesPerLeaf must be >= 1");
  AARBF_REQUIRE(ctx, pc->numParticles == 0 || pc->particles, "aarbf_diff: null particles");
  AARBF_REQUIRE(ctx, (uint64_t)bvh->counts[0] * pc->particlesPerLeaf >= pc->numParticles,
                "aarbf_diff: the BVH has fewer leaves than numParticles / particlesPerLeaf");
  hipSetDevice(ctx->device);
  const uint32_t n = pc->numParticles;
  float4 *pr = (float4 *)scratch(ctx, 32 * (size_t)(n ? n : 1));
  if (!pr) return set_error(ctx, AARBF_ERR_OUT_OF_MEMORY, "aarbf_diff: payload");
  float4 *pcol = pr + (n ? n : 1);
  if (n)
    diff_bake<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const float4 *)pc->particles, (const float *)pc->radii, n,
                                                         rm->table, rm->n, cm->table, cm->n, pr, pcol);
  AARBF_CHECK_LAUNCH(ctx, "diff_bake");
  RenderArgs &A = D.A;
  A = RenderArgs{};
  A.frameID = pc->frameID;
  A.max_depth = AARBF_MAX_DEPTH;
  A.W = W;
  A.H = H;
  A.unit = pc->unit;
  A.pos = make_float3(pc->camera.pos.x, pc->camera.pos.y, pc->camera.pos.z);
  A.d00 = make_float3(pc->camera.dir_00.x, pc->camera.dir_00.y, pc->camera.dir_00.z);
  A.du = make_float3(pc->camera.dir_du.x, pc->camera.dir_du.y, pc->camera.dir_du.z);
  A.dv = make_float3(pc->camera.dir_dv.x, pc->camera.dir_dv.y, pc->camera.dir_dv.z);
  const float m2 = 2.0f * pc->rbfRadius;
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
  A.bvh = make_view(*bvh);
  D.particles = (const float4 *)pc->particles;
  D.radii = (const float *)pc->radii;
  D.rmap = rm->table;
  D.rm_n = rm->n;
  D.spp = spp;
  return AARBF_OK;
}

extern "C" {

int aarbf_diff_forward(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W, int H,
                       uint32_t spp, uint64_t image, uint64_t path) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, image, "aarbf_diff_forward: null image");
  DiffArgs D{};
  int rc = diff_setup(ctx, rg, pc, W, H, spp, D);
  if (rc != AARBF_OK) return rc;
  D.out = (float4 *)image;
  D.path = (float4 *)path;
  dim3 grid((W + 7) / 8, (H + 7) / 8);
  diff_forward<<<grid, kDBlock, 0, ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_forward");
  return AARBF_OK;
}

int aarbf_diff_backward(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W, int H,
                        uint32_t spp, uint64_t path, uint64_t grad_image, uint64_t g_pos, uint64_t g_radii,
                        uint64_t g_attr, uint64_t g_colormap, uint64_t g_densitymap, uint64_t g_radiusmap) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, grad_image && g_pos && g_radii && g_attr && g_colormap && g_densitymap && g_radiusmap,
                "aarbf_diff_backward: null buffer");
  DiffArgs D{};
  int rc = diff_setup(ctx, rg, pc, W, H, spp, D);
  if (rc != AARBF_OK) return rc;
  D.gimg = (const float4 *)grad_image;
  D.path = (float4 *)path;
  D.g_pos = (float *)g_pos;
  D.g_rad = (float *)g_radii;
  D.g_attr = (float *)g_attr;
  D.g_cm = (float *)g_colormap;
  D.g_dm = (float *)g_densitymap;
  D.g_rm = (float *)g_radiusmap;
  dim3 grid((W + 7) / 8, (H + 7) / 8);
  diff_adjoint<<<grid, kDBlock, 0, ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_adjoint");
  return AARBF_OK;
}

} // extern "C"
