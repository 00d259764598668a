// diff.hip -- the differentiable renderer: an expected image and its
// adjoint with respect to particle positions, radii, attributes and the
// colormap / densitymap / radiusmap texels.  No counterpart in the
// reference.  Estimator 0 is the expectation of
// ParticleTracker::stochasticMarching (deviceCodeRBF.hlsl:67-139, primary
// ray, ambient 1) over the blue-noise threshold; estimator 1 is ratio
// tracking, the expectation of ParticleTracker::stochasticTracking
// (':141-203') over its accept draws at the tracker's tentative points,
// C = sum_i T_{i-1} D_i c_i with T_i = T_{i-1} (1 - D_i).
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

#include <algorithm>

namespace aarbf {
namespace {

constexpr int kDBlock = 64;
constexpr uint32_t kTexMax = 256; // texels per table with LDS gradient sums (larger tables: global atomics)
#ifndef AARBF_DIFF_REGCACHE
#define AARBF_DIFF_REGCACHE 1 // a four-entry adjoint cache in registers (0: the LDS cache of kCache entries)
#endif
#ifndef AARBF_DIFF_CACHE
#define AARBF_DIFF_CACHE 4
#endif
constexpr uint32_t kCache = AARBF_DIFF_CACHE; // per-lane particle adjoint cache entries (a power of 2; 1, 2 and 4
                                              // measured within 3 % of each other)
constexpr uint32_t kSlot = 9;     // particle index + 8 adjoint sums
#ifndef AARBF_DIFF_CAND
#define AARBF_DIFF_CAND 12
#endif
constexpr uint32_t kDC = AARBF_DIFF_CAND; // candidates per march segment (<= 32: the active set is a bitmask)
constexpr uint32_t kDense = 32;   // points queried one by one where candidates crowd
#ifndef AARBF_DIFF_MINSEG
#define AARBF_DIFF_MINSEG 8
#endif
constexpr uint32_t kMinSeg = AARBF_DIFF_MINSEG; // shorter segments (in points) count as crowded

struct DiffArgs {
  RenderArgs A;
  const float4 *particles; // raw x, y, z, attribute
  const float *radii;
  const float4 *rmap;
  uint32_t rm_n;
  uint32_t spp;
  float4 *out;             // image (forward)
  float4 *colours;         // per sample: its colour (forward), averaged into `out` by diff_fold
  float4 *path;            // per sample: sum_k F'(L_k) (c_k - c_{k+1}) (forward writes, adjoint reads), or null
  const float4 *gimg;      // dLoss/dimage (adjoint)
  float *g_pos, *g_rad, *g_attr, *g_cm, *g_dm, *g_rm;
  float *acc5;             // adjoint: per particle (r, alpha^3, rgb^2.2) sums, before the TF chain rule
  const float2 *pm;        // per particle: m = -4.5 / r^2 and 1 / r (0, 0 for r = 0)
  uint32_t ratio;          // estimator: 0 expected marching, 1 ratio tracking
};

// Ratio tracking ends a sample after the first point with T < 2^-20.
constexpr float kRatioTmin = 9.5367431640625e-07f;

// The expected image and its adjoint are compared with the FP64 oracle to a
// tolerance, not bit for bit, so the march's transcendentals use the
// hardware's exp2 / log2 (about 1 ulp) instead of the renderer's
// deterministic polynomials.  The RBF weight of a particle at a point is
// exp(e), e = |x - P|^2 m with the baked m = -4.5 / r^2 (the particle covers
// the point where e > -4.5): no square root and no division per particle
// and point; its derivatives are d exp(e) / dP = 2 m exp(e) (P - x) and
// d exp(e) / dr = -2 e / r exp(e).
__device__ inline float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ inline float fpow(float m, float y) { return m > 0.0f ? __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(m)) : 0.0f; }

__global__ __launch_bounds__(256) void diff_bake(const float4 *__restrict__ particles, const float *__restrict__ radii,
                                                 uint32_t n, const float4 *__restrict__ rmap, uint32_t rm_n,
                                                 const float4 *__restrict__ cmap, uint32_t cm_n,
                                                 float4 *__restrict__ pr, float4 *__restrict__ pc,
                                                 float2 *__restrict__ pm) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 q = particles[i];
  float4 c = tf4(cmap, cm_n, q.w);
  float radius = c.w == 0.0f ? 0.0f : radii[i] * tf_r(rmap, rm_n, q.w);
  pr[i] = make_float4(q.x, q.y, q.z, radius);
  pc[i] = make_float4(aa_cube(c.w), aa_powf(c.x, 2.2f), aa_powf(c.y, 2.2f), aa_powf(c.z, 2.2f));
  pm[i] = radius > 0.0f ? make_float2(-4.5f / (radius * radius), 1.0f / radius) : make_float2(0.0f, 0.0f);
}

// e = |x - P|^2 m of particle p at x, and whether p covers x (e > -4.5;
// a radius of 0 never covers)
__device__ inline bool rbf_exponent(const DiffArgs &D, uint32_t p, const float4 &q, float x, float y, float z,
                                    float &e) {
  const float m = D.pm[p].x;
  e = aa_dist2(q.x, q.y, q.z, x, y, z) * m;
  return m < 0.0f && e > -4.5f;
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
  constexpr float inv = 1.0f / 0.99f;
  float e = fexp(-L);
  float v = (1.0f - e) * inv;
  if (v >= 1.0f) {
    dF = 0.0f;
    return 1.0f;
  }
  dF = e * inv;
  return v;
}

// every particle covering x, in ascending order (the stackless walk of
// point_query with a callback): fn(p, baked position and radius, e)
template <class Fn>
__device__ inline void for_covering(const DiffArgs &D, float x, float y, float z, Fn fn) {
  const RenderArgs &A = D.A;
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
          float e;
          if (rbf_exponent(D, p, q, x, y, z, e)) fn(p, q, e);
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


// Per-lane LDS columns of the march (stride kDBlock): the current
// segment's candidates and the walk stack.
struct DSeg {
  uint32_t *idx;
  float *tin, *tout;
  uint32_t *stk;  // per-level words of the walks (visit order / hit masks)
  uint32_t *list; // the current point's covering particles (kList words), or null
};

// Covering particles a point's payload walk records for the adjoint's
// scatter, which then reads them back instead of walking the tree again
// (a point covered by more walks the tree again).
constexpr uint32_t kList = 24;

// The march of sample s of pixel (px, py): fn(x, y, z, payload, cover) at
// every point some particle covers, where cover(g) calls g(p, q, e) for
// each covering particle p (baked q, RBF exponent e) in the order the
// payload summed them; fn returns false to end the march.  With W.list set,
// cover() reads the particles the payload walk recorded there.
//
// Points are t_k = t_{k-1} + unit (float recurrence from tenter + unit r1).
// Which particles can cover them comes from a segment gather: one walk of
// the quantised tree collects the kDC particles whose sphere (widened by a
// rounding margin) the ray enters first, with entry / exit parameters; the
// segment ends where the first particle left out would enter, so every
// particle covering a point before that is a candidate.  Candidates enter
// and leave an active set at their parameters; a point with an empty set
// costs only the recurrence, the others test the active candidates
// exactly (|P - x| < r).  Where more than kDC candidates crowd the segment
// start the next kDense points use full point queries instead.
template <class Fn>
__device__ inline void march(const DiffArgs &D, const DSeg &W, int px, int py, uint32_t s, Fn fn) {
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
  const float3 o = A.pos;
  const float m = 2e-5f * (1.0f + fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) + fmaxf(fabsf(te), fabsf(tx)));
  const bool seg_ok = A.bvh.groups != nullptr;
  uint32_t n = 0, act = 0, eptr = 0;
  float seg_end = -AA_INF_F, t_enter = AA_INF_F, t_exit = AA_INF_F, dense_end = -AA_INF_F;
  // the candidates of the segment from t0 on, sorted by entry
  auto gather = [&](float t0) {
    SegQuery q;
    q.ax = o.x + t0 * d.x;
    q.ay = o.y + t0 * d.y;
    q.az = o.z + t0 * d.z;
    q.dx = d.x;
    q.dy = d.y;
    q.dz = d.z;
    q.ix = safe_inv(d.x);
    q.iy = safe_inv(d.y);
    q.iz = safe_inv(d.z);
    q.m = m;
    q.len = tx - t0 + 4.0f * m;
    n = 0;
    float maxtin = -AA_INF_F, t_cut = AA_INF_F;
    uint32_t maxpos = 0;
    auto take = [&](uint32_t p) {
      const float4 P = A.pr[p];
      if (!(P.w > 0.0f)) return;
      const float vx = P.x - o.x, vy = P.y - o.y, vz = P.z - o.z;
      const float tc = vx * d.x + vy * d.y + vz * d.z;
      const float hx = vx - tc * d.x, hy = vy - tc * d.y, hz = vz - tc * d.z;
      const float rr = P.w + m;
      const float w2 = rr * rr - (hx * hx + hy * hy + hz * hz);
      if (!(w2 >= 0.0f)) return;
      const float w = sqrtf(w2);
      const float tin = tc - w - m, tout = tc + w + m;
      if (!(tout >= t0) || !(tin <= tx) || !(tin < t_cut)) {
        if (tout >= t0 && tin <= tx) t_cut = fminf(t_cut, tin);
        return;
      }
      if (n < kDC) {
        W.idx[n * kDBlock] = p;
        W.tin[n * kDBlock] = tin;
        W.tout[n * kDBlock] = tout;
        if (tin > maxtin) {
          maxtin = tin;
          maxpos = n;
        }
        ++n;
      } else if (tin < maxtin) { // evict the latest-entering candidate
        t_cut = fminf(t_cut, maxtin);
        W.idx[maxpos * kDBlock] = p;
        W.tin[maxpos * kDBlock] = tin;
        W.tout[maxpos * kDBlock] = tout;
        maxtin = -AA_INF_F;
        for (uint32_t k = 0; k < kDC; ++k) {
          const float v = W.tin[k * kDBlock];
          if (v > maxtin) {
            maxtin = v;
            maxpos = k;
          }
        }
      } else {
        t_cut = fminf(t_cut, tin);
      }
      q.len = fminf(tx, t_cut + 4.0f * m) - t0;
    };
    // front-to-back walk: the children of a hit node in order of entry, so
    // the nearest particles come first and t_cut prunes the rest early
    {
      const uint32_t log2K = A.bvh.log2K, top = A.bvh.top;
      const float4 *root = A.bvh.nodes + g_lvl_ofs[top];
      if (top >= 1 && seg_box(q, root[0], root[1])) {
        uint32_t lvl = top - 1, base = 0;
        uint32_t ord = ftb_children(A, q, top, 0);
        for (;;) {
          if (lvl == 0) { // a whole leaf group, in entry order
            for (; ord; ord >>= 4) {
              const uint32_t leaf = base + (ord & 15u) - 1u;
              const uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
              for (uint32_t p = p0; p < p1; ++p) take(p);
            }
          }
          const uint32_t nib = ord & 15u;
          if (nib == 0) {
            if (lvl == top - 1) break;
            ++lvl;
            ord = W.stk[(lvl - 1u) * kDBlock];
            base = ((base >> log2K) >> log2K) << log2K;
            continue;
          }
          ord >>= 4;
          const uint32_t node = base + nib - 1u;
          W.stk[(lvl - 1u) * kDBlock] = ord;
          base = node << log2K;
          ord = ftb_children(A, q, lvl, node);
          --lvl;
        }
      }
    }
    for (uint32_t a = 1; a < n; ++a) { // insertion sort by entry
      const uint32_t pi = W.idx[a * kDBlock];
      const float ti = W.tin[a * kDBlock], to = W.tout[a * kDBlock];
      uint32_t b = a;
      for (; b > 0 && W.tin[(b - 1) * kDBlock] > ti; --b) {
        W.idx[b * kDBlock] = W.idx[(b - 1) * kDBlock];
        W.tin[b * kDBlock] = W.tin[(b - 1) * kDBlock];
        W.tout[b * kDBlock] = W.tout[(b - 1) * kDBlock];
      }
      W.idx[b * kDBlock] = pi;
      W.tin[b * kDBlock] = ti;
      W.tout[b * kDBlock] = to;
    }
    seg_end = t_cut;
    act = 0;
    eptr = 0;
    t_exit = AA_INF_F;
    t_enter = n ? W.tin[0] : AA_INF_F;
  };
  // marching: t_k = t_{k-1} + unit from tenter + unit r1; ratio tracking:
  // the delta tracker's free flights from t = 0, one LCG draw each after
  // the sample's three set-up draws (r0, r1, r2)
  const bool ratio = D.ratio != 0;
  float t = te + A.unit * r1;
  if (ratio) {
    aa_lcg_next(&rng);
    aa_lcg_next(&rng);
    t = 0.0f;
  }
  for (uint32_t i = 0; i < A.max_depth; ++i) {
    if (ratio) {
      const float xi = aa_lcg_next(&rng);
      t = t - aa_log_step(1.0f - xi) * A.unit;
    } else {
      t = t + A.unit;
    }
    if (t >= tx) return;
    bool dense = !seg_ok || t < dense_end;
    if (!dense && !(t < seg_end)) {
      gather(t);
      // crowded (the list is full within a few points): point queries for
      // a while instead of a gather every few points
      if (!(t + (float)kMinSeg * A.unit < seg_end)) {
        dense = true;
        dense_end = t + (float)kDense * A.unit;
        seg_end = -AA_INF_F;
      }
    }
    if (!dense) {
      if (t >= t_enter) {
        while (eptr < n && W.tin[eptr * kDBlock] <= t) {
          const float to = W.tout[eptr * kDBlock];
          if (t <= to) {
            act |= 1u << eptr;
            t_exit = fminf(t_exit, to);
          }
          ++eptr;
        }
        t_enter = eptr < n ? W.tin[eptr * kDBlock] : AA_INF_F;
      }
      if (t > t_exit) {
        t_exit = AA_INF_F;
        for (uint32_t b = act; b; b &= b - 1u) {
          const uint32_t k = (uint32_t)__ffs(b) - 1u;
          const float to = W.tout[k * kDBlock];
          if (to < t) act &= ~(1u << k);
          else t_exit = fminf(t_exit, to);
        }
      }
      if (act == 0) continue;
    }
    const float x = o.x + t * d.x, y = o.y + t * d.y, z = o.z + t * d.z;
    // the covering particles, in one fixed order for the payload and cover()
    auto cover = [&](auto g) {
      if (dense) {
        if (!seg_ok) {
          for_covering(D, x, y, z, g);
          return;
        }
        // point query over the quantised tree, leaves in ascending order
        Counters c = {};
        walk_q<kDBlock, false>(
            A, W.stk, [&](const float4 &lo, const float4 &hi) { return inside(lo, hi, x, y, z); },
            [&](uint32_t base, uint32_t mask) {
              for (; mask; mask &= mask - 1u) {
                const uint32_t leaf = base + (uint32_t)__ffs(mask) - 1u;
                const uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
                for (uint32_t p = p0; p < p1; ++p) {
                  const float4 q = A.pr[p];
                  float e;
                  if (rbf_exponent(D, p, q, x, y, z, e)) g(p, q, e);
                }
              }
              return false;
            },
            c);
        return;
      }
      for (uint32_t b = act; b; b &= b - 1u) {
        const uint32_t p = W.idx[((uint32_t)__ffs(b) - 1u) * kDBlock];
        const float4 q = A.pr[p];
        float e;
        if (rbf_exponent(D, p, q, x, y, z, e)) g(p, q, e);
      }
    };
    DPay pl;
    pl.count = 0;
    pl.rho = pl.S[0] = pl.S[1] = pl.S[2] = 0.0f;
    cover([&](uint32_t p, const float4 &, float e) {
      const float4 k = A.pc[p];
      const float w = fexp(e) * k.x;
      if (W.list && pl.count < kList) W.list[pl.count * kDBlock] = p;
      pl.count += 1;
      pl.rho += w;
      pl.S[0] += k.y * w;
      pl.S[1] += k.z * w;
      pl.S[2] += k.w * w;
    });
    if (pl.count > 0 && pl.rho > 0.0f) {
      auto recorded = [&](auto g) {
        if (pl.count > kList) {
          cover(g);
          return;
        }
        for (uint32_t i = 0; i < pl.count; ++i) {
          const uint32_t p = W.list[i * kDBlock];
          const float4 q = A.pr[p];
          g(p, q, aa_dist2(q.x, q.y, q.z, x, y, z) * D.pm[p].x);
        }
      };
      if (W.list) {
        if (!fn(x, y, z, pl, recorded)) return;
      } else if (!fn(x, y, z, pl, cover)) {
        return;
      }
    }
  }
}

// the point's colour (S / rho)^(1/2.2); rho > 0 at every point the march
// reports
__device__ inline void colour(const DPay &pl, float c[3]) {
  const float inv_rho = __builtin_amdgcn_rcpf(pl.rho);
  for (int k = 0; k < 3; ++k) c[k] = fpow(pl.S[k] * inv_rho, 1.0f / 2.2f);
}

#define AARBF_DSEG_LDS(list)                                                                                \
  __shared__ uint32_t s_cidx[kDC * kDBlock];                                                                  \
  __shared__ float s_ctin[kDC * kDBlock], s_ctout[kDC * kDBlock];                                             \
  extern __shared__ uint32_t s_stk[]; /* walk stacks, max(top, 2) words per lane (sized at launch) */         \
  const DSeg W{s_cidx + threadIdx.x, s_ctin + threadIdx.x, s_ctout + threadIdx.x, s_stk + threadIdx.x, list};

// One lane per (pixel, sample): the samples of a pixel sit on neighbouring
// lanes (they march the same ray from different offsets, so the wave stays
// coherent), and a pixel whose ray crosses a dense cluster spreads its cost
// over spp lanes instead of running spp samples in a row.  Pixels are
// visited in 8 x 8 tiles.  Returns false for lanes past the image.
__device__ inline bool lane_sample(const DiffArgs &D, int &px, int &py, uint32_t &s) {
  const RenderArgs &A = D.A;
  const uint64_t g = (uint64_t)blockIdx.x * kDBlock + threadIdx.x;
  s = (uint32_t)(g % D.spp);
  const uint64_t p = g / D.spp;
  const uint32_t tiles_x = ((uint32_t)A.W + 7u) / 8u;
  const uint64_t tile = p >> 6;
  const uint32_t in = (uint32_t)(p & 63u);
  px = (int)((uint32_t)(tile % tiles_x) * 8u + (in & 7u));
  py = (int)((uint32_t)(tile / tiles_x) * 8u + (in >> 3));
  return px < A.W && py < A.H;
}

__global__ __launch_bounds__(kDBlock) void diff_forward(DiffArgs D) {
  const RenderArgs &A = D.A;
  AARBF_DSEG_LDS(nullptr)
  load_levels(A);
  int px, py;
  uint32_t s;
  if (!lane_sample(D, px, py, s)) return;
  {
    float L = 0.0f, C[3] = {0.0f, 0.0f, 0.0f}, V[3] = {0.0f, 0.0f, 0.0f}, prev[3] = {0.0f, 0.0f, 0.0f};
    float T = 1.0f;
    bool have = false;
    march(D, W, px, py, s, [&](float, float, float, const DPay &pl, auto) {
      float sl, f, dF;
      uint32_t j0, j1;
      float Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, sl, j0, j1, f);
      float c[3];
      colour(pl, c);
      if (D.ratio) { // C = sum_i T_{i-1} D_i c_i
        for (int k = 0; k < 3; ++k) C[k] += (T * Dv) * c[k];
        T = T * (1.0f - Dv);
        return T >= kRatioTmin;
      }
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
      return F1 < 1.0f; // once F saturates every later term is 0
    });
    if (D.ratio) { // the adjoint's replay starts from the sample's colour
      for (int k = 0; k < 3; ++k) V[k] = C[k];
    } else {
      float dF;
      Fcdf(L, dF);
      for (int k = 0; k < 3; ++k) V[k] += dF * prev[k];
    }
    const uint64_t at = ((uint64_t)py * A.W + px) * D.spp + s;
    if (D.path) D.path[at] = make_float4(V[0], V[1], V[2], 0.0f);
    D.colours[at] = make_float4(C[0], C[1], C[2], 0.0f);
  }
}

// The pixel: the mean of its samples, summed in sample order.
__global__ __launch_bounds__(256) void diff_fold(DiffArgs D) {
  const RenderArgs &A = D.A;
  const uint64_t p = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (p >= (uint64_t)A.W * A.H) return;
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t s = 0; s < D.spp; ++s) {
    const float4 c = D.colours[p * D.spp + s];
    acc[0] += c.x;
    acc[1] += c.y;
    acc[2] += c.z;
  }
  const float inv = 1.0f / (float)D.spp;
  D.out[p] = make_float4(acc[0] * inv, acc[1] * inv, acc[2] * inv, 1.0f);
}

__global__ __launch_bounds__(kDBlock) void diff_adjoint(DiffArgs D) {
  const RenderArgs &A = D.A;
  __shared__ float s_gdm[kTexMax];
#if !AARBF_DIFF_REGCACHE
  __shared__ uint32_t s_cache[kCache * kSlot * kDBlock];
#endif
  __shared__ uint32_t s_list[kList * kDBlock];
  AARBF_DSEG_LDS(s_list + threadIdx.x)
  load_levels(A);
  for (uint32_t j = threadIdx.x; j < kTexMax; j += kDBlock) s_gdm[j] = 0.0f;
  __syncthreads();
  const bool lds_dm = A.dm_n <= kTexMax;
  auto add_dm = [&](uint32_t j, float v) {
    if (lds_dm) atomicAdd(&s_gdm[j], v);
    else atomicAdd(&D.g_dm[j], v);
  };
  // A particle covers runs of consecutive march points, so its per-point
  // adjoints (position, radius, alpha^3, colour^2.2) are summed in a small
  // per-lane cache and added to global memory when another particle takes
  // the entry; the chain rule through the radiusmap / colormap lookups
  // follows per particle in diff_chain.  A few particles cover each point
  // and the next point's are mostly the same, so a four-entry associative
  // cache in registers keeps most of them (a one-entry cache flushed on 83 %
  // of the updates); the LDS cache is the alternative build.
#if AARBF_DIFF_REGCACHE
  // four entries in registers, fully associative (static indices only),
  // first in first out
  uint32_t ck[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
  float cv[4][8];
  auto flush_e = [&](const uint32_t p, const float (&v)[8]) {
    atomicAdd(&D.g_pos[3 * p], v[0]);
    atomicAdd(&D.g_pos[3 * p + 1], v[1]);
    atomicAdd(&D.g_pos[3 * p + 2], v[2]);
#pragma unroll
    for (uint32_t w = 0; w < 5; ++w) atomicAdd(&D.acc5[5 * (uint64_t)p + w], v[3 + w]);
  };
  auto flush_all = [&]() {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (ck[e] != 0xffffffffu) {
        flush_e(ck[e], cv[e]);
        ck[e] = 0xffffffffu;
      }
  };
  auto accumulate = [&](uint32_t p, const float (&v)[8]) {
    bool hit = false;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (ck[e] == p) {
        hit = true;
#pragma unroll
        for (int w = 0; w < 8; ++w) cv[e][w] += v[w];
      }
    if (!hit) { // the oldest entry leaves, the others move down one
      if (ck[3] != 0xffffffffu) flush_e(ck[3], cv[3]);
#pragma unroll
      for (int e = 3; e > 0; --e) {
        ck[e] = ck[e - 1];
#pragma unroll
        for (int w = 0; w < 8; ++w) cv[e][w] = cv[e - 1][w];
      }
      ck[0] = p;
#pragma unroll
      for (int w = 0; w < 8; ++w) cv[0][w] = v[w];
    }
  };
  auto reset_cache = [&]() {
#pragma unroll
    for (int e = 0; e < 4; ++e) ck[e] = 0xffffffffu;
  };
#else
  uint32_t *cache = s_cache + threadIdx.x; // [kCache][kSlot] words, stride kDBlock
  auto word = [&](uint32_t e, uint32_t w) -> float & { return *(float *)&cache[(e * kSlot + w) * kDBlock]; };
  auto flush = [&](uint32_t e) {
    const uint32_t p = cache[e * kSlot * kDBlock];
    if (p == 0xffffffffu) return;
    cache[e * kSlot * kDBlock] = 0xffffffffu;
    atomicAdd(&D.g_pos[3 * p], word(e, 1));
    atomicAdd(&D.g_pos[3 * p + 1], word(e, 2));
    atomicAdd(&D.g_pos[3 * p + 2], word(e, 3));
    for (uint32_t w = 0; w < 5; ++w) atomicAdd(&D.acc5[5 * (uint64_t)p + w], word(e, 4 + w));
  };
  auto flush_all = [&]() {
    for (uint32_t e = 0; e < kCache; ++e) flush(e);
  };
  // the cache entry of particle p (words strided by kDBlock)
  auto accumulate = [&](uint32_t p, const float (&v)[8]) {
    const uint32_t e = p & (kCache - 1u);
    if (cache[e * kSlot * kDBlock] != p) {
      flush(e);
      cache[e * kSlot * kDBlock] = p;
      for (uint32_t w = 1; w < kSlot; ++w) word(e, w) = 0.0f;
    }
    for (uint32_t w = 0; w < 8; ++w) word(e, 1 + w) += v[w];
  };
  auto reset_cache = [&]() {
    for (uint32_t e = 0; e < kCache; ++e) cache[e * kSlot * kDBlock] = 0xffffffffu;
  };
#endif
  int px, py;
  uint32_t s;
  if (lane_sample(D, px, py, s)) {
    const float4 gi = D.gimg[(uint64_t)px + (uint64_t)A.W * py];
    const float inv = 1.0f / (float)D.spp;
    const float g[3] = {gi.x * inv, gi.y * inv, gi.z * inv};
    do {
      reset_cache();
      // pass 1: the total of dC/dL_k over the path (kept by the forward when
      // it had a path buffer)
      float L = 0.0f, total = 0.0f, prev[3] = {0, 0, 0}, T1 = 1.0f;
      bool have = false;
      if (D.path) {
        const float4 V = D.path[((uint64_t)py * A.W + px) * D.spp + s];
        total = g[0] * V.x + g[1] * V.y + g[2] * V.z;
        have = true;
      } else march(D, W, px, py, s, [&](float, float, float, const DPay &pl, auto) {
        float sl, f, dF;
        uint32_t j0, j1;
        float Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, sl, j0, j1, f);
        float c[3];
        colour(pl, c);
        if (D.ratio) { // total = g . C
          total += (T1 * Dv) * (g[0] * c[0] + g[1] * c[1] + g[2] * c[2]);
          T1 = T1 * (1.0f - Dv);
          have = true;
          return T1 >= kRatioTmin;
        }
        if (have) {
          Fcdf(L, dF);
          total += dF * (g[0] * (prev[0] - c[0]) + g[1] * (prev[1] - c[1]) + g[2] * (prev[2] - c[2]));
        }
        L += Dv;
        prev[0] = c[0];
        prev[1] = c[1];
        prev[2] = c[2];
        have = true;
        return Fcdf(L, dF) < 1.0f;
      });
      if (!have) break;
      if (!D.path && !D.ratio) {
        float dF;
        Fcdf(L, dF);
        total += dF * (g[0] * prev[0] + g[1] * prev[1] + g[2] * prev[2]);
      }
      // pass 2: the chain rule at every point.  Marching: dC/dD_j = total -
      // prefix.  Ratio tracking: with q the g-weighted colour still to come
      // (q = g . C before the first point), q_j = (q_{j-1} - D_j g.c_j) /
      // (1 - D_j) and dC/dD_j = T_{j-1} (g.c_j - q_j); q_j = 0 at the point
      // that ends the sample.
      float Lp = 0.0f, prefix = 0.0f, T = 1.0f, q = total;
      have = false;
      march(D, W, px, py, s, [&](float x, float y, float z, const DPay &pl, auto cover) {
        float sD, fD, dF;
        uint32_t j0, j1;
        float Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, sD, j0, j1, fD);
        float c[3];
        colour(pl, c);
        float aD, pj;
        bool more;
        if (D.ratio) {
          const float gc = g[0] * c[0] + g[1] * c[1] + g[2] * c[2];
          const float Tn = T * (1.0f - Dv);
          more = Tn >= kRatioTmin;
          const float qn = more ? (q - Dv * gc) * __builtin_amdgcn_rcpf(1.0f - Dv) : 0.0f;
          aD = T * (gc - qn);
          pj = T * Dv;
          T = Tn;
          q = qn;
        } else {
          if (have) {
            Fcdf(Lp, dF);
            prefix += dF * (g[0] * (prev[0] - c[0]) + g[1] * (prev[1] - c[1]) + g[2] * (prev[2] - c[2]));
          }
          aD = total - prefix;
          float F0 = Fcdf(Lp, dF);
          Lp += Dv;
          float F1 = Fcdf(Lp, dF);
          pj = F1 - F0;
          more = F1 < 1.0f; // later points: dC/dD and the step weights are 0
        }
        add_dm(j0, aD * (1.0f - fD));
        add_dm(j1, aD * fD);
        float a_rho = aD * sD, a_S[3];
        const float inv_rho = __builtin_amdgcn_rcpf(pl.rho);
        for (int k = 0; k < 3; ++k) {
          // c = m^(1/2.2), m = S / rho: dc/dm = (1/2.2) c / m, so
          // dc/dS = (1/2.2) c / S and dc/drho = -(1/2.2) c / rho (both 0
          // where S = 0, and c = 0 there)
          const float gcc = (g[k] * pj) * ((1.0f / 2.2f) * c[k]);
          a_S[k] = pl.S[k] > 0.0f ? gcc * __builtin_amdgcn_rcpf(pl.S[k]) : 0.0f;
          a_rho -= gcc * inv_rho;
        }
        auto scatter = [&](uint32_t p, const float4 &q, float e) {
          const float4 k = A.pc[p];
          const float2 mi = D.pm[p];
          const float rbf = fexp(e), w = rbf * k.x;
          const float a_w = a_rho + a_S[0] * k.y + a_S[1] * k.z + a_S[2] * k.w;
          const float a_rbf = a_w * k.x;
          const float kk = a_rbf * ((2.0f * mi.x) * rbf); // d(rbf)/dP = 2 m rbf (P - x)
          const float v[8] = {kk * (q.x - x), kk * (q.y - y), kk * (q.z - z),
                              a_rbf * (rbf * (-2.0f * e) * mi.y), // a_r: d(rbf)/dr = -2 e / r rbf
                              a_w * rbf,                           // a_a3
                              a_S[0] * w, a_S[1] * w, a_S[2] * w}; // a_c22
          accumulate(p, v);
        };
        cover(scatter);
        prev[0] = c[0];
        prev[1] = c[1];
        prev[2] = c[2];
        have = true;
        return more;
      });
      flush_all();
    } while (false);
  }
  __syncthreads();
  if (lds_dm)
    for (uint32_t j = threadIdx.x; j < A.dm_n; j += kDBlock)
      if (s_gdm[j] != 0.0f) atomicAdd(&D.g_dm[j], s_gdm[j]);
}

// The chain rule of each particle's summed adjoints through its radiusmap
// and colormap lookups: r = radius * radiusmap(w), alpha^3 and rgb^2.2 from
// colormap(w).  Texel gradients are summed per block in LDS first.
constexpr int kCBlock = 256;
__global__ __launch_bounds__(kCBlock) void diff_chain(DiffArgs D) {
  const RenderArgs &A = D.A;
  __shared__ float s_gcm[4 * kTexMax], s_grm[kTexMax];
  for (uint32_t j = threadIdx.x; j < 4 * kTexMax; j += kCBlock) s_gcm[j] = 0.0f;
  for (uint32_t j = threadIdx.x; j < kTexMax; j += kCBlock) s_grm[j] = 0.0f;
  __syncthreads();
  const bool lds_cm = A.cm_n <= kTexMax, lds_rm = D.rm_n <= kTexMax;
  auto add_cm = [&](uint32_t j, float v) {
    if (lds_cm) atomicAdd(&s_gcm[j], v);
    else atomicAdd(&D.g_cm[j], v);
  };
  auto add_rm = [&](uint32_t j, float v) {
    if (lds_rm) atomicAdd(&s_grm[j], v);
    else atomicAdd(&D.g_rm[j], v);
  };
  const uint32_t p = blockIdx.x * kCBlock + threadIdx.x;
  if (p < A.numParticles) {
    const float *a = D.acc5 + 5 * (uint64_t)p;
    const float a_r = a[0], a_a3 = a[1];
    if (a_r != 0.0f || a_a3 != 0.0f || a[2] != 0.0f || a[3] != 0.0f || a[4] != 0.0f) {
      const float4 raw = D.particles[p];
      const float4 k = A.pc[p];
      float sl, f, a_attr = 0.0f;
      uint32_t k0, k1;
      // r = radius * radiusmap(attribute)
      const float rmv = tf_grad(D.rmap, D.rm_n, 0, raw.w, sl, k0, k1, f);
      D.g_rad[p] += a_r * rmv;
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
          a_c = cv > 0.0f ? a[2 + c2] * 2.2f * (c22 / cv) : 0.0f; // d(c^2.2)/dc = 2.2 c^2.2 / c
        }
        add_cm(4 * k0 + c2, a_c * (1.0f - f));
        add_cm(4 * k1 + c2, a_c * f);
        a_attr += a_c * sl;
      }
      D.g_attr[p] += a_attr;
    }
  }
  __syncthreads();
  if (lds_cm)
    for (uint32_t j = threadIdx.x; j < 4 * A.cm_n; j += kCBlock)
      if (s_gcm[j] != 0.0f) atomicAdd(&D.g_cm[j], s_gcm[j]);
  if (lds_rm)
    for (uint32_t j = threadIdx.x; j < D.rm_n; j += kCBlock)
      if (s_grm[j] != 0.0f) atomicAdd(&D.g_rm[j], s_grm[j]);
}

__global__ __launch_bounds__(256) void diff_zero(float *p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = 0.0f;
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
  // scratch: baked payload (32 B per particle), adjoint sums (20 B per
  // particle), RBF coefficients (8 B per particle), then the per-sample
  // colours of the forward (16 B per sample)
  const size_t part_bytes = ((32 + 20 + 8) * (size_t)(n ? n : 1) + 255) & ~(size_t)255;
  const size_t samples = (size_t)W * (size_t)H * spp;
  float4 *pr = (float4 *)scratch(ctx, part_bytes + 16 * samples);
  if (!pr) return set_error(ctx, AARBF_ERR_OUT_OF_MEMORY, "aarbf_diff: payload");
  D.acc5 = (float *)(pr + 2 * (size_t)(n ? n : 1));
  D.pm = (const float2 *)(D.acc5 + 5 * (size_t)(n ? n : 1));
  D.colours = (float4 *)((char *)pr + part_bytes);
  float4 *pcol = pr + (n ? n : 1);
  if (n)
    diff_bake<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const float4 *)pc->particles, (const float *)pc->radii, n,
                                                         rm->table, rm->n, cm->table, cm->n, pr, pcol,
                                                         (float2 *)D.pm);
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
  if (bvh->num_levels > kStackLevels) A.bvh.groups = nullptr; // the march then queries every point
  D.particles = (const float4 *)pc->particles;
  D.radii = (const float *)pc->radii;
  D.rmap = rm->table;
  D.rm_n = rm->n;
  D.spp = spp;
  D.ratio = ctx->diff_estimator == 1 ? 1u : 0u;
  return AARBF_OK;
}

// The march's walk stacks in LDS: levels 0 .. top-1 of the tree per lane
// (a tree that has no quantised groups is never walked: two words).  Sized
// to the tree, which keeps more blocks resident than the 16-level maximum.
static size_t diff_stack_bytes(const DiffArgs &D) {
  const uint32_t levels = D.A.bvh.groups ? std::max<uint32_t>(D.A.bvh.top, 2u) : 2u;
  return sizeof(uint32_t) * kDBlock * levels;
}

// Blocks of kDBlock lanes over the (pixel, sample) pairs of whole 8 x 8
// tiles (lane_sample drops the pixels past the image).
static uint32_t diff_blocks(int W, int H, uint32_t spp) {
  const uint64_t tiles = (uint64_t)((W + 7) / 8) * (uint64_t)((H + 7) / 8);
  return (uint32_t)((tiles * 64u * spp + kDBlock - 1) / kDBlock);
}

extern "C" {

int aarbf_set_diff_estimator(aarbf_context *ctx, int estimator) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, estimator == 0 || estimator == 1,
                "aarbf_set_diff_estimator: 0 (expected marching) or 1 (ratio tracking)");
  ctx->diff_estimator = estimator;
  return AARBF_OK;
}

int aarbf_diff_forward(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc, int W, int H,
                       uint32_t spp, uint64_t image, uint64_t path) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, image, "aarbf_diff_forward: null image");
  DiffArgs D{};
  int rc = diff_setup(ctx, rg, pc, W, H, spp, D);
  if (rc != AARBF_OK) return rc;
  D.out = (float4 *)image;
  D.path = (float4 *)path;
  diff_forward<<<diff_blocks(W, H, spp), kDBlock, diff_stack_bytes(D), ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_forward");
  const uint64_t npix = (uint64_t)W * (uint64_t)H;
  diff_fold<<<(uint32_t)((npix + 255) / 256), 256, 0, ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_fold");
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
  const uint64_t n5 = 5 * (uint64_t)D.A.numParticles;
  if (n5) diff_zero<<<(uint32_t)((n5 + 255) / 256), 256, 0, ctx->stream>>>(D.acc5, n5);
  diff_adjoint<<<diff_blocks(W, H, spp), kDBlock, diff_stack_bytes(D), ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_adjoint");
  if (D.A.numParticles)
    diff_chain<<<(D.A.numParticles + kCBlock - 1) / kCBlock, kCBlock, 0, ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_chain");
  return AARBF_OK;
}

} // extern "C"
