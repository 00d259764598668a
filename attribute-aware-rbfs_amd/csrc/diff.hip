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
      t = aa_flight(t, 1.0f - xi, A.unit);
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

// ---- one wave per (pixel, sample): expected marching --------------------------
// The lanes of a wave take 64 consecutive march points of one sample at a
// time.  A cooperative stretch query (a level-synchronous walk of the
// quantised tree, 64 child boxes per pass, the frontier in LDS) collects
// every particle whose widened sphere meets the ray stretch [ta, tb] into
// LDS, and each lane sums its point's payload over them.  The march's
// optical depth is a prefix sum over the lanes, so a sample that crosses a
// dense cluster costs its points / 64 passes, not one lane's serial march
// (the serial schedule waits for its slowest samples).  The stretch grows
// where it holds few particles and is halved where more than kWC meet it;
// a point covered by more than kWC particles is summed by its lane alone
// with the serial point query.  The adjoint's lanes then each own up to two
// of the stretch's particles and sum their gradients over the chunk's
// points in registers: one set of atomics per particle and stretch.
constexpr uint32_t kWC = 128; // particles of a stretch
constexpr uint32_t kWF = 512; // frontier of the stretch query

struct WaveMem {
  uint32_t cidx[kWC];
  float ctin[kWC], ctout[kWC];
  float4 cP[kWC];  // x, y, z, m = -4.5 / r^2
  float4 cK[kWC];  // alpha^3, rgb^2.2
  float cinv[kWC]; // 1 / r
  uint32_t front[2][kWF];
  float4 px[64]; // adjoint: a chunk's points (x, y, z, t; t = NaN where not covered)
  float4 pa[64]; // adjoint: their a_rho, a_S
  float tp[64];  // adjoint: the chunk's point parameters (ascending)
};

// k steps of the serial march's recurrence t <- t + unit (one binary32
// rounding per step).  Inside a binade every step adds the same number of
// units in the last place (unit / ulp rounded), unless that quotient is a
// tie, so a run of steps that stays in the binade is one integer multiply;
// binade crossings, ties, zero and subnormals are stepped one addition at
// a time.
__device__ inline float rec_advance(float t, float u, uint32_t k) {
  const uint32_t bu = __float_as_uint(u), eu = (bu >> 23) & 0xffu, mu = (bu & 0x7fffffu) | 0x800000u;
  while (k > 0u) {
    const float t1 = t + u;
    const uint32_t b0 = __float_as_uint(t), b1 = __float_as_uint(t1), e0 = (b0 >> 23) & 0xffu;
    bool single = k == 1u || !(t > 0.0f) || e0 == 0u || e0 == 0xffu || ((b1 >> 23) & 0xffu) != e0 || eu == 0u;
    if (!single && eu < e0) { // unit / ulp = mu / 2^(e0 - eu): a tie when its fraction is exactly 1/2
      const uint32_t sh = e0 - eu;
      single = sh <= 24u && (mu & ((1u << sh) - 1u)) == (1u << (sh - 1u));
    }
    if (single) {
      t = t1;
      --k;
      continue;
    }
    const uint32_t M0 = b0 & 0x7fffffu, R = (b1 & 0x7fffffu) - M0;
    uint32_t n = k;
    if (R > 0u) n = min(k, (0x7fffffu - M0) / R); // steps that stay in the binade
    if (n <= 1u) {
      t = t1;
      --k;
      continue;
    }
    t = __uint_as_float((b0 & 0xff800000u) | (M0 + n * R));
    k -= n;
  }
  return t;
}

__device__ inline float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_incl_scan(float v, uint32_t lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const float y = __shfl_up(v, o, 64);
    if (lane >= (uint32_t)o) v += y;
  }
  return v;
}

// Every particle whose sphere, widened by m, meets the ray stretch o + t d,
// t in [ta, tb], into W (entry / exit parameters, payload); nc = how many.
// False when the frontier or the list overflows.  Called by the whole wave.
__device__ inline bool wave_query(const DiffArgs &D, WaveMem &W, float3 o, float3 d, float ta, float tb, float m,
                                  uint32_t lane, uint32_t &nc) {
  const RenderArgs &A = D.A;
  const uint64_t lt = lane ? (~0ull >> (64u - lane)) : 0ull;
  SegQuery q;
  q.ax = o.x + ta * d.x;
  q.ay = o.y + ta * d.y;
  q.az = o.z + ta * d.z;
  q.dx = d.x;
  q.dy = d.y;
  q.dz = d.z;
  q.ix = safe_inv(d.x);
  q.iy = safe_inv(d.y);
  q.iz = safe_inv(d.z);
  q.m = m;
  q.len = (tb - ta) + 4.0f * m;
  const QGroup *__restrict__ Qg = A.bvh.groups;
  const uint32_t log2K = A.bvh.log2K, K = 1u << log2K, top = A.bvh.top;
  const float4 *root = A.bvh.nodes + g_lvl_ofs[top];
  uint32_t nf = seg_box(q, root[0], root[1]) ? 1u : 0u, cur = 0;
  if (lane == 0) W.front[0][0] = 0;
  __syncthreads();
  for (uint32_t lvl = top; lvl >= 1u && nf > 0u; --lvl) {
    const uint32_t total = nf << log2K;
    uint32_t nn = 0;
    for (uint32_t base = 0; base < total; base += 64u) {
      const uint32_t j = base + lane;
      bool hit = false;
      uint32_t child = 0;
      if (j < total) {
        const uint32_t node = W.front[cur][j >> log2K], k = j & (K - 1u);
        const QGroup *gp = Qg + g_lvl_qofs[lvl] + node;
        const uint32_t first = node << log2K, mcnt = min(K, g_lvl_cnt[lvl - 1u] - first);
        auto qb = [&](uint32_t field) { return (gp->q[2u * field + (k >> 2)] >> ((k & 3u) * 8u)) & 255u; };
        const uint32_t lx = qb(0), hx = qb(3);
        child = first + k;
        if (k < mcnt && lx <= hx) {
          const float4 ga = gp->a, gb = gp->b;
          float4 lo, hi;
          lo.x = ga.x + (float)lx * ga.w;
          lo.y = ga.y + (float)qb(1) * gb.x;
          lo.z = ga.z + (float)qb(2) * gb.y;
          hi.x = ga.x + (float)hx * ga.w;
          hi.y = ga.y + (float)qb(4) * gb.x;
          hi.z = ga.z + (float)qb(5) * gb.y;
          lo.w = hi.w = 0.0f;
          hit = seg_box(q, lo, hi);
        }
      }
      const uint64_t mk = __ballot(hit);
      if (hit) {
        const uint32_t at = nn + (uint32_t)__popcll(mk & lt);
        if (at < kWF) W.front[cur ^ 1u][at] = child;
      }
      nn += (uint32_t)__popcll(mk);
    }
    __syncthreads();
    if (nn > kWF) return false;
    cur ^= 1u;
    nf = nn;
  }
  // the leaves' particles
  nc = 0;
  const uint32_t ppl = A.ppl, total = nf * ppl;
  for (uint32_t base = 0; base < total; base += 64u) {
    const uint32_t j = base + lane;
    bool hit = false;
    uint32_t p = 0;
    float tin = 0.0f, tout = 0.0f;
    float4 P = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (j < total) {
      p = W.front[cur][j / ppl] * ppl + j % ppl;
      if (p < A.numParticles) {
        P = A.pr[p];
        if (P.w > 0.0f) {
          const float vx = P.x - o.x, vy = P.y - o.y, vz = P.z - o.z;
          const float tc = vx * d.x + vy * d.y + vz * d.z;
          const float hx = vx - tc * d.x, hy = vy - tc * d.y, hz = vz - tc * d.z;
          const float rr = P.w + m;
          const float w2 = rr * rr - (hx * hx + hy * hy + hz * hz);
          if (w2 >= 0.0f) {
            const float w = sqrtf(w2);
            tin = tc - w - m;
            tout = tc + w + m;
            hit = tout >= ta && tin <= tb;
          }
        }
      }
    }
    const uint64_t mk = __ballot(hit);
    if (hit) {
      const uint32_t at = nc + (uint32_t)__popcll(mk & lt);
      if (at < kWC) {
        const float2 mi = D.pm[p];
        W.cidx[at] = p;
        W.ctin[at] = tin;
        W.ctout[at] = tout;
        W.cP[at] = make_float4(P.x, P.y, P.z, mi.x);
        W.cK[at] = A.pc[p];
        W.cinv[at] = mi.y;
      }
    }
    nc += (uint32_t)__popcll(mk);
  }
  __syncthreads();
  return nc <= kWC;
}

// The sample of block b (pixel-major over 8 x 8 pixel tiles, the samples of
// a pixel on consecutive blocks).  False past the image.
__device__ inline bool block_sample(const DiffArgs &D, int &px, int &py, uint32_t &s) {
  const RenderArgs &A = D.A;
  const uint64_t b = blockIdx.x;
  s = (uint32_t)(b % D.spp);
  const uint64_t p = b / D.spp;
  const uint32_t tiles_x = ((uint32_t)A.W + 7u) / 8u;
  const uint64_t tile = p >> 6;
  const uint32_t in = (uint32_t)(p & 63u);
  px = (int)((uint32_t)(tile % tiles_x) * 8u + (in & 7u));
  py = (int)((uint32_t)(tile / tiles_x) * 8u + (in >> 3));
  return px < A.W && py < A.H;
}

// The wave march of sample s of pixel (px, py): fn(i0, i1, t, x, y, z, pl,
// dense) for each run of lanes [i0, i1) whose points share a stretch, in
// march order (lane i holds point t; pl its payload, count 0 where nothing
// covers it; dense: the run is one point summed by the serial query).  fn
// returns false to end the sample.  flush() runs before a new stretch
// replaces the candidates.
template <class Fn, class Flush>
__device__ inline void wave_march(const DiffArgs &D, WaveMem &W, int px, int py, uint32_t s, float3 &o_out,
                                  float3 &d_out, Fn fn, Flush flush) {
  const RenderArgs &A = D.A;
  const uint32_t lane = threadIdx.x;
  uint32_t fid = (A.frameID + s) & 0xffffu;
  uint32_t rng = aa_rng_seed(fid, (uint32_t)px, (uint32_t)py, (uint32_t)A.W);
  const float r1 = aa_lcg_next(&rng);
  const float sx = ((float)px + 0.5f) / (float)A.W, sy = ((float)py + 0.5f) / (float)A.H;
  float3 d;
  d.x = (A.d00.x + sx * A.du.x) + sy * A.dv.x;
  d.y = (A.d00.y + sx * A.du.y) + sy * A.dv.y;
  d.z = (A.d00.z + sx * A.du.z) + sy * A.dv.z;
  const float len = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
  d.x = d.x / len;
  d.y = d.y / len;
  d.z = d.z / len;
  float te, tx;
  aabb_intersection(A.pos, d, 0.0f, 10000.0f, A.lb, A.rt, te, tx);
  const float3 o = A.pos;
  o_out = o;
  d_out = d;
  if (!(te < tx)) return;
  const float m = 2e-5f * (1.0f + fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) + fmaxf(fabsf(te), fabsf(tx)));
  float t_last = te + A.unit * r1; // the first point is t_last + unit
  uint32_t k_done = 0, nc = 0;
  float sa = AA_INF_F, sb = -AA_INF_F, lenq = 64.0f * A.unit;
  bool have_stretch = false;
  for (;;) {
    const float t = rec_advance(t_last, A.unit, lane + 1u);
    const bool valid = t < tx && k_done + lane + 1u <= A.max_depth;
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid)); // a prefix: t grows with the lane
    if (nvalid == 0u) break;
    uint32_t i0 = 0;
    while (i0 < nvalid) {
      const float ti0 = __shfl(t, (int)i0, 64);
      bool dense = false;
      if (!(have_stretch && ti0 >= sa && ti0 < sb)) {
        if (have_stretch) flush();
        float ql = lenq;
        bool ok = false;
        for (;;) {
          const float tb = fminf(ti0 + ql, tx);
          ok = wave_query(D, W, o, d, ti0, tb, m, lane, nc);
          if (ok) {
            sa = ti0;
            sb = tb >= tx ? AA_INF_F : tb;
            lenq = nc == 0u ? 4.0f * ql : (nc < kWC / 4u ? 2.0f * ql : ql);
            break;
          }
          ql *= 0.5f;
          if (ql < 0.5f * A.unit) break;
        }
        have_stretch = ok;
        if (!ok) { // more than kWC particles at point ti0
          dense = true;
          lenq = A.unit;
        }
      }
      const uint32_t i1 = dense ? i0 + 1u : (uint32_t)__popcll(__ballot(valid && t < sb));
      const bool act = lane >= i0 && lane < i1;
      const float x = o.x + t * d.x, y = o.y + t * d.y, z = o.z + t * d.z;
      DPay pl;
      pl.count = 0;
      pl.rho = pl.S[0] = pl.S[1] = pl.S[2] = 0.0f;
      if (dense) {
        if (act)
          for_covering(D, x, y, z, [&](uint32_t p, const float4 &, float e) {
            const float4 k = A.pc[p];
            const float w = fexp(e) * k.x;
            pl.count += 1;
            pl.rho += w;
            pl.S[0] += k.y * w;
            pl.S[1] += k.z * w;
            pl.S[2] += k.w * w;
          });
      } else {
        const float tlo = __shfl(t, (int)i0, 64), thi = __shfl(t, (int)i1 - 1, 64);
        for (uint32_t c = 0; c < nc; ++c) {
          const float cin = W.ctin[c], cout = W.ctout[c];
          if (!(cin <= thi && cout >= tlo)) continue; // meets none of the run's points
          if (act && cin <= t && t <= cout) {
            const float4 P = W.cP[c];
            const float e = aa_dist2(P.x, P.y, P.z, x, y, z) * P.w;
            if (P.w < 0.0f && e > -4.5f) {
              const float4 k = W.cK[c];
              const float w = fexp(e) * k.x;
              pl.count += 1;
              pl.rho += w;
              pl.S[0] += k.y * w;
              pl.S[1] += k.z * w;
              pl.S[2] += k.w * w;
            }
          }
        }
      }
      if (!act) pl.count = 0;
      if (!fn(i0, i1, t, x, y, z, pl, dense, nc)) {
        if (have_stretch) flush();
        return;
      }
      i0 = i1;
      if (dense) have_stretch = false;
    }
    if (nvalid < 64u) break;
    t_last = __shfl(t, 63, 64);
    k_done += 64u;
  }
  if (have_stretch) flush();
}

// The march state the lanes share: the optical depth so far, and the last
// covered point's F'(L) and colour (its term of the path vector waits for
// the next covered point's colour).
struct WaveAcc {
  float L = 0.0f;
  bool have = false;
  float pdF = 0.0f, pc[3] = {0.0f, 0.0f, 0.0f};
};

// Per lane of a run: D and the colour where covered; the run's inclusive
// optical depth, F before and after, F'(L) after; the next covered lane's
// colour (cn, valid where nxt) and the run's first / last covered lanes.
struct WaveRun {
  bool cov, nxt;
  float Dv, sD, fD, c[3], cn[3], F0, F1, dF1, Lend;
  uint32_t j0, j1;
  uint64_t covm;
  uint32_t first, last;
};

__device__ inline void wave_run(const DiffArgs &D, const WaveAcc &acc, const DPay &pl, uint32_t lane, WaveRun &r) {
  const RenderArgs &A = D.A;
  r.cov = pl.count > 0 && pl.rho > 0.0f;
  r.Dv = 0.0f;
  r.sD = r.fD = 0.0f;
  r.j0 = r.j1 = 0;
  r.c[0] = r.c[1] = r.c[2] = 0.0f;
  if (r.cov) {
    r.Dv = tf_grad(A.dmap, A.dm_n, 0, pl.rho, r.sD, r.j0, r.j1, r.fD);
    colour(pl, r.c);
  }
  const float Lin = acc.L + wave_incl_scan(r.Dv, lane);
  r.Lend = __shfl(Lin, 63, 64);
  float dF0;
  r.F1 = Fcdf(Lin, r.dF1);
  r.F0 = Fcdf(Lin - r.Dv, dF0);
  r.covm = __ballot(r.cov);
  r.first = r.covm ? (uint32_t)__ffsll((unsigned long long)r.covm) - 1u : 0u;
  r.last = r.covm ? 63u - (uint32_t)__clzll((long long)r.covm) : 0u;
  const uint64_t above = lane == 63u ? 0ull : (r.covm & (~0ull << (lane + 1u)));
  r.nxt = r.cov && above != 0ull;
  const int nl = above ? (int)__ffsll((unsigned long long)above) - 1 : 0;
  for (int k = 0; k < 3; ++k) r.cn[k] = __shfl(r.c[k], nl, 64);
}

__global__ __launch_bounds__(64) void diff_forward_w(DiffArgs D) {
  const RenderArgs &A = D.A;
  __shared__ WaveMem W;
  load_levels(A);
  __syncthreads();
  int px, py;
  uint32_t s;
  if (!block_sample(D, px, py, s)) return; // the whole block
  const uint32_t lane = threadIdx.x;
  WaveAcc acc;
  float C[3] = {0.0f, 0.0f, 0.0f}, V[3] = {0.0f, 0.0f, 0.0f};
  float3 o, d;
  wave_march(
      D, W, px, py, s, o, d,
      [&](uint32_t, uint32_t, float, float, float, float, const DPay &pl, bool, uint32_t) {
        WaveRun r;
        wave_run(D, acc, pl, lane, r);
        for (int k = 0; k < 3; ++k) {
          C[k] += wave_sum(r.cov ? (r.F1 - r.F0) * r.c[k] : 0.0f);
          V[k] += wave_sum(r.nxt ? r.dF1 * (r.c[k] - r.cn[k]) : 0.0f);
        }
        if (!r.covm) return true;
        const float cf[3] = {__shfl(r.c[0], (int)r.first, 64), __shfl(r.c[1], (int)r.first, 64),
                             __shfl(r.c[2], (int)r.first, 64)};
        if (acc.have)
          for (int k = 0; k < 3; ++k) V[k] += acc.pdF * (acc.pc[k] - cf[k]);
        acc.have = true;
        acc.pdF = __shfl(r.dF1, (int)r.last, 64);
        for (int k = 0; k < 3; ++k) acc.pc[k] = __shfl(r.c[k], (int)r.last, 64);
        acc.L = r.Lend;
        // the serial march stops after the first point where F reaches 1
        return __shfl(r.F1, (int)r.last, 64) < 1.0f;
      },
      [&]() {});
  (void)o;
  (void)d;
  if (acc.have)
    for (int k = 0; k < 3; ++k) V[k] += acc.pdF * acc.pc[k];
  if (lane == 0) {
    const uint64_t at = ((uint64_t)py * A.W + px) * D.spp + s;
    if (D.path) D.path[at] = make_float4(V[0], V[1], V[2], 0.0f);
    D.colours[at] = make_float4(C[0], C[1], C[2], 0.0f);
  }
}

__global__ __launch_bounds__(64) void diff_adjoint_w(DiffArgs D) {
  const RenderArgs &A = D.A;
  __shared__ WaveMem W;
  __shared__ float s_gdm[kTexMax];
  load_levels(A);
  for (uint32_t j = threadIdx.x; j < kTexMax; j += 64u) s_gdm[j] = 0.0f;
  __syncthreads();
  const bool lds_dm = A.dm_n <= kTexMax;
  auto add_dm = [&](uint32_t j, float v) {
    if (lds_dm) atomicAdd(&s_gdm[j], v);
    else atomicAdd(&D.g_dm[j], v);
  };
  int px, py;
  uint32_t s;
  const uint32_t lane = threadIdx.x;
  if (block_sample(D, px, py, s)) {
    const float4 gi = D.gimg[(uint64_t)px + (uint64_t)A.W * py];
    const float inv = 1.0f / (float)D.spp;
    const float g[3] = {gi.x * inv, gi.y * inv, gi.z * inv};
    const float4 PV = D.path[((uint64_t)py * A.W + px) * D.spp + s];
    const float total = g[0] * PV.x + g[1] * PV.y + g[2] * PV.z;
    WaveAcc acc;
    float prefix = 0.0f;
    // each lane owns particles lane and lane + 64 of the current stretch
    float ac[2][8];
    for (int u = 0; u < 2; ++u)
      for (int w = 0; w < 8; ++w) ac[u][w] = 0.0f;
    uint32_t cur_nc = 0;
    auto flush_p = [&](uint32_t p, const float (&v)[8]) {
      atomicAdd(&D.g_pos[3 * p], v[0]);
      atomicAdd(&D.g_pos[3 * p + 1], v[1]);
      atomicAdd(&D.g_pos[3 * p + 2], v[2]);
      for (uint32_t w = 0; w < 5; ++w) atomicAdd(&D.acc5[5 * (uint64_t)p + w], v[3 + w]);
    };
    auto flush = [&]() {
      for (uint32_t u = 0; u < 2; ++u) {
        const uint32_t c = lane + 64u * u;
        bool any = false;
        for (int w = 0; w < 8; ++w) any |= ac[u][w] != 0.0f;
        if (c < cur_nc && any) flush_p(W.cidx[c], ac[u]);
        for (int w = 0; w < 8; ++w) ac[u][w] = 0.0f;
      }
      __syncthreads(); // the candidates may be replaced next
    };
    // the chain rule at one point: v = d(loss) / d(particle) for a covering
    // particle (P, m; k; 1/r) at x with RBF exponent e
    auto grad = [&](const float4 &P, const float4 &k, float inv_r, float x, float y, float z, float e,
                    const float4 &a, float (&v)[8]) {
      const float rbf = fexp(e), w = rbf * k.x;
      const float a_w = a.x + a.y * k.y + a.z * k.z + a.w * k.w;
      const float a_rbf = a_w * k.x;
      const float kk = a_rbf * ((2.0f * P.w) * rbf); // d(rbf)/dP = 2 m rbf (P - x)
      v[0] = kk * (P.x - x);
      v[1] = kk * (P.y - y);
      v[2] = kk * (P.z - z);
      v[3] = a_rbf * (rbf * (-2.0f * e) * inv_r); // d(rbf)/dr = -2 e / r rbf
      v[4] = a_w * rbf;
      v[5] = a.y * w;
      v[6] = a.z * w;
      v[7] = a.w * w;
    };
    float3 o, d;
    wave_march(
        D, W, px, py, s, o, d,
        [&](uint32_t i0, uint32_t i1, float t, float x, float y, float z, const DPay &pl, bool dense, uint32_t nc) {
          cur_nc = dense ? 0u : nc;
          WaveRun r;
          wave_run(D, acc, pl, lane, r);
          // dC/dD_j = total - (sum over the covered points k before j of
          // F'(L_k) g.(c_k - c_{k+1}))
          const float term = r.nxt ? r.dF1 * (g[0] * (r.c[0] - r.cn[0]) + g[1] * (r.c[1] - r.cn[1]) +
                                              g[2] * (r.c[2] - r.cn[2]))
                                   : 0.0f;
          const float ex = wave_incl_scan(term, lane) - term;
          float pterm = 0.0f;
          if (acc.have && r.covm) {
            const float cf[3] = {__shfl(r.c[0], (int)r.first, 64), __shfl(r.c[1], (int)r.first, 64),
                                 __shfl(r.c[2], (int)r.first, 64)};
            pterm = acc.pdF * (g[0] * (acc.pc[0] - cf[0]) + g[1] * (acc.pc[1] - cf[1]) + g[2] * (acc.pc[2] - cf[2]));
          }
          const float aD = total - ((prefix + (lane >= r.first ? pterm : 0.0f)) + ex);
          const float pj = r.F1 - r.F0;
          float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (r.cov) {
            add_dm(r.j0, aD * (1.0f - r.fD));
            add_dm(r.j1, aD * r.fD);
            float a_rho = aD * r.sD;
            float a_S[3];
            const float inv_rho = __builtin_amdgcn_rcpf(pl.rho);
            for (int k = 0; k < 3; ++k) {
              // c = (S / rho)^(1/2.2): dc/dS = (1/2.2) c / S, dc/drho = -(1/2.2) c / rho
              const float gcc = (g[k] * pj) * ((1.0f / 2.2f) * r.c[k]);
              a_S[k] = pl.S[k] > 0.0f ? gcc * __builtin_amdgcn_rcpf(pl.S[k]) : 0.0f;
              a_rho -= gcc * inv_rho;
            }
            a = make_float4(a_rho, a_S[0], a_S[1], a_S[2]);
          }
          prefix += pterm + wave_sum(term);
          if (r.covm) {
            acc.have = true;
            acc.pdF = __shfl(r.dF1, (int)r.last, 64);
            for (int k = 0; k < 3; ++k) acc.pc[k] = __shfl(r.c[k], (int)r.last, 64);
            acc.L = r.Lend;
          }
          if (dense) { // one point, summed by its lane: direct atomics per particle
            if (r.cov)
              for_covering(D, x, y, z, [&](uint32_t p, const float4 &, float e) {
                const float4 P4 = A.pr[p];
                const float2 mi = D.pm[p];
                float v[8];
                grad(make_float4(P4.x, P4.y, P4.z, mi.x), A.pc[p], mi.y, x, y, z, e, a, v);
                flush_p(p, v);
              });
          } else if (r.covm) {
            W.px[lane] = make_float4(x, y, z, r.cov ? t : AA_NAN_F);
            W.pa[lane] = a;
            W.tp[lane] = t;
            __syncthreads();
            for (uint32_t u = 0; u < 2; ++u) {
              const uint32_t c = lane + 64u * u;
              if (c >= nc) continue;
              const float cin = W.ctin[c], cout = W.ctout[c];
              const float4 P = W.cP[c], k = W.cK[c];
              const float inv_r = W.cinv[c];
              // the run's points inside [cin, cout]: from the first at or past cin
              uint32_t i = i0;
              for (uint32_t hi = i1; i < hi;) {
                const uint32_t mid = (i + hi) >> 1;
                if (W.tp[mid] < cin) i = mid + 1u;
                else hi = mid;
              }
              for (; i < i1 && W.tp[i] <= cout; ++i) {
                const float4 q = W.px[i];
                if (!(q.w == q.w)) continue; // not covered
                const float e = aa_dist2(P.x, P.y, P.z, q.x, q.y, q.z) * P.w;
                if (!(P.w < 0.0f && e > -4.5f)) continue;
                float v[8];
                grad(P, k, inv_r, q.x, q.y, q.z, e, W.pa[i], v);
                for (int w = 0; w < 8; ++w) ac[u][w] += v[w];
              }
            }
            __syncthreads();
          }
          return __shfl(r.F1, (int)r.last, 64) < 1.0f || !r.covm;
        },
        flush);
    (void)o;
    (void)d;
  }
  __syncthreads();
  if (lds_dm)
    for (uint32_t j = threadIdx.x; j < A.dm_n; j += 64u)
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
  A.inv_unit = 1.0f / pc->unit;
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
  A.dm_nf = (float)dm->n;
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

// The wave march: expected marching over a tree with quantised groups.
static bool wave_schedule(const aarbf_context *ctx, const DiffArgs &D) {
  return ctx->tune_diff_waves == 1 && D.ratio == 0 && D.A.bvh.groups != nullptr;
}

// One block (one wave) per (pixel, sample) of whole 8 x 8 tiles.
static uint32_t wave_blocks(int W, int H, uint32_t spp) {
  return (uint32_t)((uint64_t)((W + 7) / 8) * (uint64_t)((H + 7) / 8) * 64u * spp);
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
  if (wave_schedule(ctx, D)) {
    diff_forward_w<<<wave_blocks(W, H, spp), 64, 0, ctx->stream>>>(D);
    AARBF_CHECK_LAUNCH(ctx, "diff_forward_w");
  } else {
    diff_forward<<<diff_blocks(W, H, spp), kDBlock, diff_stack_bytes(D), ctx->stream>>>(D);
    AARBF_CHECK_LAUNCH(ctx, "diff_forward");
  }
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
  if (wave_schedule(ctx, D) && D.path) { // the wave march starts from the forward's path vector
    diff_adjoint_w<<<wave_blocks(W, H, spp), 64, 0, ctx->stream>>>(D);
    AARBF_CHECK_LAUNCH(ctx, "diff_adjoint_w");
  } else {
    diff_adjoint<<<diff_blocks(W, H, spp), kDBlock, diff_stack_bytes(D), ctx->stream>>>(D);
    AARBF_CHECK_LAUNCH(ctx, "diff_adjoint");
  }
  if (D.A.numParticles)
    diff_chain<<<(D.A.numParticles + kCBlock - 1) / kCBlock, kCBlock, 0, ctx->stream>>>(D);
  AARBF_CHECK_LAUNCH(ctx, "diff_chain");
  return AARBF_OK;
}

} // extern "C"
