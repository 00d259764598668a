// bvh.hip -- GenRBFBounds (deviceCodeBounds.hlsl:29-82) and the
// Hilbert-packed BVH that replaces the driver's BLAS/TLAS
// (hostCode.cpp:468-469,781-794).
//
// The leaves arrive in Hilbert order, so a bottom-up build that merges K
// consecutive nodes per parent gives spatially coherent nodes without any
// sorting (the packed R-tree construction).  Topology is implicit: node j
// of level l+1 has children [jK, jK+K) of level l, so no child pointers are
// stored and a depth-first walk can be done without a stack.
#include "aarbf_detmath.h"
#include "internal.h"

namespace aarbf {
namespace {

__device__ inline float tf_r(const float4 *__restrict__ t, uint32_t n, float u) {
  uint32_t j0, j1;
  float f;
  aa_tf_coords(u, n, &j0, &j1, &f);
  return aa_lerp_tex(t[j0].x, t[j1].x, f);
}

// One thread per leaf, as the reference's 1D compute launch.
__global__ __launch_bounds__(256) void gen_bounds_kernel(const float4 *__restrict__ particles,
                                                         const float *__restrict__ radii,
                                                         uint32_t numParticles, uint32_t ppl,
                                                         float rbfRadius, const float4 *__restrict__ rmap,
                                                         uint32_t rm_n, float *__restrict__ aabbs,
                                                         uint32_t numAABBs) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= numAABBs) return;
  float lo[3] = {AA_INF_F, AA_INF_F, AA_INF_F}, hi[3] = {-AA_INF_F, -AA_INF_F, -AA_INF_F};
  float *o = aabbs + 6 * (uint64_t)c;
  if ((uint64_t)c * ppl > numParticles) { // ':45' compares with '>'
    for (int d = 0; d < 6; ++d) o[d] = AA_NAN_F;
    return;
  }
  for (uint32_t i = 0; i < ppl; ++i) {
    uint64_t p = (uint64_t)c * ppl + i;
    if (p >= numParticles) break;
    float4 q = particles[p];
    float rp = tf_r(rmap, rm_n, q.w);
    if (rp == 0.0f) continue;
    float e = (radii ? radii[p] : rbfRadius) * rp;
    float a[3] = {q.x - e, q.y - e, q.z - e}, b[3] = {q.x + e, q.y + e, q.z + e};
    for (int d = 0; d < 3; ++d) {
      if (i == 0) {
        lo[d] = a[d];
        hi[d] = b[d];
      } else {
        lo[d] = fminf(lo[d], a[d]);
        hi[d] = fmaxf(hi[d], b[d]);
      }
    }
  }
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2])
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = AA_NAN_F;
  for (int d = 0; d < 3; ++d) {
    o[d] = lo[d];
    o[3 + d] = hi[d];
  }
}

// Level 0: each leaf box widened by 2^-20 of its largest magnitude per axis,
// so that the box test never rejects a point the exact |P-x| < r test
// accepts; NaN (inactive) leaves become empty boxes.
__global__ __launch_bounds__(256) void bvh_leaves(const float *__restrict__ aabbs, uint32_t n,
                                                  float4 *__restrict__ out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const float *b = aabbs + 6 * (uint64_t)j;
  float lo[3], hi[3];
  bool nan = false;
  for (int d = 0; d < 3; ++d) {
    lo[d] = b[d];
    hi[d] = b[3 + d];
    nan |= (lo[d] != lo[d]) || (hi[d] != hi[d]);
  }
  float4 L, H;
  float *l = &L.x, *h = &H.x;
  for (int d = 0; d < 3; ++d) {
    if (!nan && lo[d] <= hi[d]) {
      float m = fmaxf(fabsf(lo[d]), fabsf(hi[d]));
      l[d] = lo[d] - m * 9.5367431640625e-07f;
      h[d] = hi[d] + m * 9.5367431640625e-07f;
    } else {
      l[d] = AA_INF_F;
      h[d] = -AA_INF_F;
    }
  }
  L.w = 0.0f;
  H.w = 0.0f;
  out[2 * (uint64_t)j] = L;
  out[2 * (uint64_t)j + 1] = H;
}

// Level l+1 from level l: exact union of K consecutive children.
__global__ __launch_bounds__(256) void bvh_level(const float4 *__restrict__ child, uint32_t n_child,
                                                 uint32_t K, float4 *__restrict__ out, uint32_t n) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float4 L = make_float4(AA_INF_F, AA_INF_F, AA_INF_F, 0.0f), H = make_float4(-AA_INF_F, -AA_INF_F, -AA_INF_F, 0.0f);
  uint32_t end = min(j * K + K, n_child);
  for (uint32_t c = j * K; c < end; ++c) {
    float4 a = child[2 * (uint64_t)c], b = child[2 * (uint64_t)c + 1];
    L.x = fminf(L.x, a.x);
    L.y = fminf(L.y, a.y);
    L.z = fminf(L.z, a.z);
    H.x = fmaxf(H.x, b.x);
    H.y = fmaxf(H.y, b.y);
    H.z = fmaxf(H.z, b.z);
  }
  out[2 * (uint64_t)j] = L;
  out[2 * (uint64_t)j + 1] = H;
}

// Dequantised bound: lo + q * step (two roundings; the traversal evaluates
// exactly the same expression).
__device__ inline float deq(float lo, float step, uint32_t q) { return lo + (float)q * step; }

// One QGroup per node of level l >= 1 from the float boxes of levels l and
// l-1.  Child bounds are rounded outwards and then corrected until the
// dequantised box contains the float box.
__global__ __launch_bounds__(256) void bvh_qgroups(const float4 *__restrict__ parent, uint32_t n_parent,
                                                   const float4 *__restrict__ child, uint32_t n_child, uint32_t K,
                                                   QGroup *__restrict__ out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_parent) return;
  float4 P0 = parent[2 * (uint64_t)j], P1 = parent[2 * (uint64_t)j + 1];
  const bool pempty = !(P0.x <= P1.x);
  float plo[3] = {P0.x, P0.y, P0.z}, phi[3] = {P1.x, P1.y, P1.z}, step[3];
  for (int a = 0; a < 3; ++a) {
    if (pempty) {
      plo[a] = 0.0f;
      step[a] = 1.0f;
      continue;
    }
    float sa = fmaxf((phi[a] - plo[a]) / 255.0f, 1.17549435e-38f);
    while (deq(plo[a], sa, 255u) < phi[a]) sa = sa * 1.0000005f + 1e-37f;
    step[a] = sa;
  }
  uint32_t qb[6][8];
  for (uint32_t k = 0; k < 8; ++k) {
    uint64_t c = (uint64_t)j * K + k;
    bool empty = pempty || k >= K || c >= n_child;
    float4 C0 = make_float4(0, 0, 0, 0), C1 = C0;
    if (!empty) {
      C0 = child[2 * c];
      C1 = child[2 * c + 1];
      empty = !(C0.x <= C1.x);
    }
    if (empty) {
      for (int a = 0; a < 3; ++a) {
        qb[a][k] = 255u;
        qb[3 + a][k] = 0u;
      }
      continue;
    }
    float clo[3] = {C0.x, C0.y, C0.z}, chi[3] = {C1.x, C1.y, C1.z};
    for (int a = 0; a < 3; ++a) {
      float fl = floorf((clo[a] - plo[a]) / step[a]);
      uint32_t ql = fl <= 0.0f ? 0u : (fl >= 255.0f ? 255u : (uint32_t)fl);
      while (ql > 0u && deq(plo[a], step[a], ql) > clo[a]) --ql;
      float fh = ceilf((chi[a] - plo[a]) / step[a]);
      uint32_t qh = fh <= 0.0f ? 0u : (fh >= 255.0f ? 255u : (uint32_t)fh);
      while (qh < 255u && deq(plo[a], step[a], qh) < chi[a]) ++qh;
      qb[a][k] = ql;
      qb[3 + a][k] = qh;
    }
  }
  QGroup g;
  g.a = make_float4(plo[0], plo[1], plo[2], step[0]);
  g.b = make_float4(step[1], step[2], 0.0f, 0.0f);
  for (int f = 0; f < 6; ++f) {
    g.q[2 * f] = qb[f][0] | (qb[f][1] << 8) | (qb[f][2] << 16) | (qb[f][3] << 24);
    g.q[2 * f + 1] = qb[f][4] | (qb[f][5] << 8) | (qb[f][6] << 16) | (qb[f][7] << 24);
  }
  out[j] = g;
}

int build_levels(aarbf_context *ctx, BVH *b, const float *aabbs) {
  uint32_t n0 = b->counts[0];
  bvh_leaves<<<(n0 + 255) / 256, 256, 0, ctx->stream>>>(aabbs, n0, b->nodes + b->level_ofs[0]);
  for (uint32_t l = 1; l < b->num_levels; ++l) {
    uint32_t n = b->counts[l];
    bvh_level<<<(n + 255) / 256, 256, 0, ctx->stream>>>(b->nodes + b->level_ofs[l - 1], b->counts[l - 1], b->K,
                                                        b->nodes + b->level_ofs[l], n);
  }
  if (b->groups)
    for (uint32_t l = 1; l < b->num_levels; ++l) {
      uint32_t n = b->counts[l];
      bvh_qgroups<<<(n + 255) / 256, 256, 0, ctx->stream>>>(b->nodes + b->level_ofs[l], n,
                                                            b->nodes + b->level_ofs[l - 1], b->counts[l - 1], b->K,
                                                            b->groups + b->group_ofs[l]);
    }
  AARBF_CHECK_LAUNCH(ctx, "bvh build");
  return AARBF_OK;
}

} // namespace
} // namespace aarbf

using namespace aarbf;

extern "C" {

int aarbf_gen_bounds(aarbf_context *ctx, const aarbf_bounds_constants *bc) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, bc, "aarbf_gen_bounds: null constants");
  TF *rm = get_tf(ctx, bc->radiusmap);
  AARBF_REQUIRE(ctx, rm, "aarbf_gen_bounds: radiusmap is not a transfer-function handle");
  AARBF_REQUIRE(ctx, bc->particlesPerLeaf >= 1, "aarbf_gen_bounds: particlesPerLeaf must be >= 1");
  AARBF_REQUIRE(ctx, bc->numAABBs == 0 || bc->aabbs, "aarbf_gen_bounds: null aabbs");
  AARBF_REQUIRE(ctx, bc->numParticles == 0 || bc->particles, "aarbf_gen_bounds: null particles");
  if (bc->numAABBs == 0) return AARBF_OK;
  hipSetDevice(ctx->device);
  gen_bounds_kernel<<<(bc->numAABBs + 255) / 256, 256, 0, ctx->stream>>>(
      (const float4 *)bc->particles, (const float *)bc->radii, bc->numParticles, bc->particlesPerLeaf,
      bc->rbfRadius, rm->table, rm->n, (float *)bc->aabbs, bc->numAABBs);
  AARBF_CHECK_LAUNCH(ctx, "gen_bounds_kernel");
  return AARBF_OK;
}

int aarbf_bvh_build(aarbf_context *ctx, uint64_t aabbs, uint32_t numAABBs, uint32_t branching, uint64_t *handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, handle && aabbs && numAABBs >= 1, "aarbf_bvh_build: need at least one leaf box");
  if (branching == 0) branching = 8;
  AARBF_REQUIRE(ctx, branching >= 2 && branching <= 16 && (branching & (branching - 1)) == 0,
                "aarbf_bvh_build: branching must be 2, 4, 8 or 16");
  AARBF_REQUIRE(ctx, numAABBs <= (1u << 30), "aarbf_bvh_build: too many leaves");
  hipSetDevice(ctx->device);
  BVH *b = new (std::nothrow) BVH();
  if (!b) return AARBF_ERR_OUT_OF_MEMORY;
  b->K = branching;
  b->log2K = 0;
  while ((1u << b->log2K) < branching) ++b->log2K;
  uint32_t n = numAABBs;
  uint64_t total = 0;
  uint32_t L = 0;
  // at least one internal level above the leaves
  do {
    b->counts[L] = n;
    b->level_ofs[L] = 2 * total;
    total += n;
    ++L;
    n = (n + branching - 1) / branching;
  } while (L < 2 || b->counts[L - 1] > 1);
  b->num_levels = L;
  b->total_nodes = total;
  hipError_t e = hipMalloc((void **)&b->nodes, sizeof(float4) * 2 * total);
  if (e == hipSuccess && branching <= 8) {
    uint64_t g = 0;
    for (uint32_t l = 1; l < L; ++l) {
      b->group_ofs[l] = g;
      g += b->counts[l];
    }
    b->total_groups = g;
    e = hipMalloc((void **)&b->groups, sizeof(QGroup) * g);
  }
  if (e != hipSuccess) {
    if (b->nodes) hipFree(b->nodes);
    delete b;
    return hip_error(ctx, e, "aarbf_bvh_build");
  }
  int st = build_levels(ctx, b, (const float *)aabbs);
  if (st) {
    hipFree(b->nodes);
    if (b->groups) hipFree(b->groups);
    delete b;
    return st;
  }
  *handle = reinterpret_cast<uint64_t>(b);
  ctx->bvhs.insert(*handle);
  return AARBF_OK;
}

int aarbf_bvh_refit(aarbf_context *ctx, uint64_t handle, uint64_t aabbs) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  BVH *b = get_bvh(ctx, handle);
  AARBF_REQUIRE(ctx, b && aabbs, "aarbf_bvh_refit: unknown handle or null aabbs");
  hipSetDevice(ctx->device);
  return build_levels(ctx, b, (const float *)aabbs);
}

int aarbf_bvh_destroy(aarbf_context *ctx, uint64_t handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  BVH *b = get_bvh(ctx, handle);
  AARBF_REQUIRE(ctx, b, "aarbf_bvh_destroy: unknown handle");
  hipStreamSynchronize(ctx->stream);
  hipFree(b->nodes);
  if (b->groups) hipFree(b->groups);
  delete b;
  ctx->bvhs.erase(handle);
  return AARBF_OK;
}

int aarbf_bvh_levels(aarbf_context *ctx, uint64_t handle, uint32_t *num_levels, uint32_t *counts, uint32_t max_levels) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  BVH *b = get_bvh(ctx, handle);
  AARBF_REQUIRE(ctx, b && num_levels, "aarbf_bvh_levels: unknown handle");
  *num_levels = b->num_levels;
  for (uint32_t l = 0; counts && l < b->num_levels && l < max_levels; ++l) counts[l] = b->counts[l];
  return AARBF_OK;
}

int aarbf_bvh_read_groups(aarbf_context *ctx, uint64_t handle, uint32_t level, uint32_t *groups) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  BVH *b = get_bvh(ctx, handle);
  AARBF_REQUIRE(ctx, b && groups && level >= 1 && level < b->num_levels, "aarbf_bvh_read_groups: bad arguments");
  AARBF_REQUIRE(ctx, b->groups, "aarbf_bvh_read_groups: no quantised groups (branching > 8)");
  AARBF_HIP(ctx, hipMemcpyAsync(groups, b->groups + b->group_ofs[level], sizeof(QGroup) * b->counts[level],
                                hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AARBF_OK;
}

int aarbf_bvh_read_level(aarbf_context *ctx, uint64_t handle, uint32_t level, float *boxes) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  BVH *b = get_bvh(ctx, handle);
  AARBF_REQUIRE(ctx, b && boxes && level < b->num_levels, "aarbf_bvh_read_level: bad arguments");
  AARBF_HIP(ctx, hipMemcpyAsync(boxes, b->nodes + b->level_ofs[level], sizeof(float4) * 2 * b->counts[level],
                                hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AARBF_OK;
}

} // extern "C"
