// wavefront.hip -- the default RBF render schedule: wavefront rounds of a
// gather kernel and a step kernel over a pool of per-pixel ray states.
//
// Every pixel owns one ray slot; a slot walks through the pixel's spp
// samples (primary ray, optional NEE shadow ray) exactly as
// ParticleRBFRayGen does (deviceCodeRBF.hlsl:206-336), so images are bit for
// bit those of the one-lane-per-pixel kernel.  What changes is how a
// tracker step finds the particles covering its point:
//
//  * wf_gather (one lane per ray that needs it) walks the BVH once for the
//    rest of the ray and keeps the kCand particles whose sphere -- widened
//    by a rounding margin -- the ray enters first, with their conservative
//    entry/exit parameters [tin, tout].  The segment ends where the first
//    candidate that did not fit would enter, so all particles that can
//    cover any step before the segment end are in the list.  The list is
//    stored in ascending particle order plus a permutation by tin.
//  * wf_step (one lane per active ray) advances its tracker for up to a
//    step budget.  The active set {k : tin_k <= t <= tout_k} is kept as a
//    bitmask updated by entry/exit events, so a step in empty space costs
//    only the RNG draw, the logarithm and two comparisons; particles in the
//    active set get the exact |P - x| < r test in ascending order, which
//    reproduces the point query's covering set and summation order.
//    Where more than kCand particles crowd the start of a segment the
//    segment is marked dense and its steps use batched point queries.
//
// Rays that run past their segment, start a shadow ray or a new sample go
// back to the gather queue; the host alternates the two kernels until no
// ray is left.  Both kernels are dense (every lane has the same kind of
// work), which is what the one-kernel schedule could not achieve.
//
// Slot k renders the samples of pixel k (tile order) one after another,
// which keeps the rays of a wave on neighbouring pixels.  Samples are
// claimed through per-pixel counters, so a slot whose pixel has no samples
// left helps pixels that still have some (visited in order through a
// shared pointer); without this the pixels with the longest paths would
// leave the last rounds nearly empty.  Each sample writes its colour to a
// per-sample buffer and wf_fold forms the running mean of every pixel in
// sample order, exactly as the sequential loop does.
#include "render_common.h"

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace aarbf {
namespace {

// Candidates per ray segment.  12 keeps the step kernel's LDS at 10 KB per
// wave, so with <= 128 VGPRs four waves fit on each SIMD (16 gave longer
// segments but three waves, and was slower).
constexpr uint32_t kCand = 12;
constexpr int kGBlock = 128; // gather kernel block
#ifndef AARBF_GATHER_WAVES
#define AARBF_GATHER_WAVES 1 // occupancy target of the gather kernel (waves per SIMD)
#endif
constexpr int kSBlock = 64;  // step kernel block (one wave)
constexpr uint64_t kPassSamples = 64; // samples per colour-buffer pass
constexpr uint32_t kDmLds = 64;       // density-map texels cached in LDS
// Gather queues up to this length go to wf_gather_wave (one wave per ray):
// late in a call the few rays left each need one long tree walk, which one
// lane alone runs at memory latency.
#ifndef AARBF_WAVE_GATHER
#define AARBF_WAVE_GATHER 16384
#endif
constexpr uint32_t kWaveGather = AARBF_WAVE_GATHER;
#ifndef AARBF_WAVE_SPLIT
#define AARBF_WAVE_SPLIT 32
#endif
constexpr uint32_t kWaveSplit = AARBF_WAVE_SPLIT;   // subtrees the wave splits a walk into (at least)
constexpr uint32_t kWaveFront = 8u * kWaveSplit; // breadth-first frontier entries (K <= 8)
#ifndef AARBF_WAVE_SLICE
#define AARBF_WAVE_SLICE 32
#endif
constexpr uint32_t kWaveSlice = AARBF_WAVE_SLICE; // nodes a lane opens between shared-bound updates
// Empty-space steps a marching lane may take per iteration of the step
// loop: a wave iterates as long as its slowest lane, and most steps are
// empty-space steps, so taking several per iteration cuts the iterations
// (and the per-iteration overhead every lane pays) roughly by this factor.
// Delta tracking takes up to kFlightBlock per iteration, to its next
// flight anchor (10 VALU instructions per step: LCG, 1 - xi, product,
// compares, selects).
#ifndef AARBF_FAST_STEPS
#define AARBF_FAST_STEPS 24
#endif
constexpr int kFastSteps = AARBF_FAST_STEPS;
// delta tracking's flight anchors: every kFlightBlock steps of a ray
constexpr uint32_t kFlightBlock = AA_FLIGHT_BLOCK;
// Step queues at least this long let sparse waves leave early (RenderArgs
// min_live); shorter ones (the end of a pass) would only add rounds.
constexpr uint32_t kMinLiveQueue = 65536;

constexpr uint32_t F_SHADOW = 1u, F_PENDING = 2u, F_DENSE = 4u;
// F_HEAD: the segment is the pixel's head list (shared by the primary rays
// of all its samples, which follow the same line); F_HEADGATHER: the next
// gather builds that head list from the ray's first possible position.
constexpr uint32_t F_HEAD = 8u, F_HEADGATHER = 16u;
// F_MARCH: the ray's view uses stochastic marching (else delta tracking)
constexpr uint32_t F_MARCH = 32u;

// Ray-slot state, structure of arrays (one entry per slot).
struct WFState {
  float4 *od0;  // o.xyz, t
  float4 *od1;  // d.xyz, tmax
  uint32_t *pix; // pixel (tile order) of the slot's sample
  float4 *alb;  // primary albedo
  uint4 *u;     // trng, i, s | flags << 16, seg_n
  float4 *f;    // LHS, RHS, shw, r2
  float4 *seg;  // seg_t0, seg_end, rx, unused
  float2 *fl;   // delta tracking: flight anchor t_base, product P (aa_flight_pos)
  uint32_t *cidx; // [kCand][nslots] candidate particle index (ascending)
  float *ctin;    // [kCand][nslots]
  float *ctout;   // [kCand][nslots]
  uint8_t *ceo;   // [kCand][nslots] candidate positions ordered by tin
  uint32_t *hidx; // [kCand][npix] head list of the pixel's primary ray
  float *htin, *htout;
  uint8_t *heo;
  float4 *head;   // [npix] head t0, head segment end, n (bits), 1 = valid | 2 = dense (bits)
  float4 *col;    // [s1 - s0][npix] colour of each sample, w = its noise value
  uint32_t *next_s; // [npix] next unclaimed sample of each pixel
  uint32_t *visit;  // next pixel for slots that help
  uint32_t nslots; // == npix: slot k owns pixel k
  uint32_t npix;  // pixels in tile order (8x8 tiles of each view, padded)
  uint32_t s0, s1; // samples of this pass
};

struct Lane {
  float3 o, d;
  float t, tmax;
  uint32_t pix;
  float4 alb;
  uint32_t trng, i, s, flags, seg_n;
  float LHS, RHS, shw, r2;
  float seg_t0, seg_end, rx;
  float tb, P; // delta tracking: t = aa_flight_pos(tb, P, unit)
};

__device__ inline void load_lane(const WFState &S, uint32_t slot, Lane &L) {
  float4 a = S.od0[slot], b = S.od1[slot];
  L.o = make_float3(a.x, a.y, a.z);
  L.t = a.w;
  L.d = make_float3(b.x, b.y, b.z);
  L.tmax = b.w;
  L.pix = S.pix[slot];
  L.alb = S.alb[slot];
  uint4 u = S.u[slot];
  L.trng = u.x;
  L.i = u.y;
  L.s = u.z & 0xffffu;
  L.flags = u.z >> 16;
  L.seg_n = u.w;
  float4 f = S.f[slot];
  L.LHS = f.x;
  L.RHS = f.y;
  L.shw = f.z;
  L.r2 = f.w;
  float4 g = S.seg[slot];
  L.seg_t0 = g.x;
  L.seg_end = g.y;
  L.rx = g.z;
  const float2 h = S.fl[slot];
  L.tb = h.x;
  L.P = h.y;
}

__device__ inline void save_lane(const WFState &S, uint32_t slot, const Lane &L) {
  S.od0[slot] = make_float4(L.o.x, L.o.y, L.o.z, L.t);
  S.od1[slot] = make_float4(L.d.x, L.d.y, L.d.z, L.tmax);
  S.pix[slot] = L.pix;
  S.alb[slot] = L.alb;
  S.u[slot] = make_uint4(L.trng, L.i, (L.s & 0xffffu) | (L.flags << 16), L.seg_n);
  S.f[slot] = make_float4(L.LHS, L.RHS, L.shw, L.r2);
  S.seg[slot] = make_float4(L.seg_t0, L.seg_end, L.rx, 0.0f);
  S.fl[slot] = make_float2(L.tb, L.P);
}

// The view a pixel of the combined pixel space belongs to: with B (a
// batch) looked up in the device table, else the call's one view, read from
// the kernel arguments.
template <bool B>
__device__ inline const ReqArgs &view(const RenderArgs &A, uint32_t pix) {
  if (B) return A.req[A.tile_req[pix >> 6]];
  return A.v0;
}

// Stochastic marching (else delta tracking) for the ray's view.
template <bool B>
__device__ inline bool is_march(const RenderArgs &A, const Lane &L) {
  return B ? (L.flags & F_MARCH) != 0u : A.v0.blue != 0u;
}

// Pixel coordinates within its view.
__device__ inline void pix_xy(const ReqArgs &R, uint32_t pix, int &px, int &py) {
  uint32_t tile = (pix >> 6) - R.tile0, w = pix & 63u;
  px = (int)((tile % R.tiles_x) * 8u + (w & 7u));
  py = (int)((tile / R.tiles_x) * 8u + (w >> 3));
}

// One past the last sample of this pass that pixel pix has to render.
template <bool B>
__device__ inline uint32_t sample_end(const RenderArgs &A, const WFState &S, uint32_t pix) {
  return B ? min(view<B>(A, pix).spp, S.s1) : S.s1;
}

// Set-up of sample L.s (deviceCodeRBF.hlsl:207-280).  Returns true when the
// primary ray enters the clip box and must be tracked.
__device__ inline bool start_sample(const RenderArgs &A, const ReqArgs &R, int px, int py, Lane &L) {
  const bool marching = R.blue != 0;
  uint32_t fid = (R.frameID + L.s) & 0xffffu;
  uint32_t rng = aa_rng_seed(fid, (uint32_t)px, (uint32_t)py, (uint32_t)R.W);
  float sx = ((float)px + 0.5f) / (float)R.W;
  float sy = ((float)py + 0.5f) / (float)R.H;
  float3 dir;
  dir.x = (R.d00.x + sx * R.du.x) + sy * R.dv.x;
  dir.y = (R.d00.y + sx * R.du.y) + sy * R.dv.y;
  dir.z = (R.d00.z + sx * R.du.z) + sy * R.dv.z;
  float len = sqrtf((dir.x * dir.x + dir.y * dir.y) + dir.z * dir.z);
  dir.x = dir.x / len;
  dir.y = dir.y / len;
  dir.z = dir.z / len;
  float tenter, texit;
  aabb_intersection(R.pos, dir, 0.0f, 10000.0f, A.lb, A.rt, tenter, texit);
  float r0;
  if (marching) {
    int gx = (int)(fid % 8u), gy = (int)((fid / 8u) % 16u);
    int cx = gx * 256 + px % 256, cy = gy * 256 + py % 256;
    uint8_t v = 0;
    if (cx < R.stbn_w && cy < R.stbn_h) v = R.stbn[4 * ((uint64_t)cy * R.stbn_w + cx)];
    r0 = (float)v / 255.0f;
  } else {
    r0 = aa_lcg_next(&rng);
  }
  L.trng = marching ? 0u : rng;
  float r1 = aa_lcg_next(&rng);
  L.r2 = aa_lcg_next(&rng);
  L.rx = r0;
  L.i = 0;
  L.alb = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  L.shw = 0.0f;
  L.flags = marching ? F_MARCH : 0u;
  if (!(tenter < texit)) return false;
  L.o = R.pos;
  L.d = dir;
  L.tmax = texit;
  L.t = marching ? tenter + A.unit * r1 : 0.0f;
  L.tb = L.t; // the ray's first flight anchor
  L.P = 1.0f;
  L.seg_t0 = marching ? tenter : 0.0f; // where any sample's primary ray can start
  L.LHS = 0.0f;
  L.RHS = marching ? -aa_logf(1.0f - r0 * 0.99f) : 0.0f;
  return true;
}

// Shading and compositing over black (':316-324'); the colour goes to the
// sample's entry of the colour buffer, the running mean (':325-328') is
// formed by wf_fold.
template <bool STATS, bool B>
__device__ inline void finish_sample(const RenderArgs &A, const WFState &S, Lane &L, float vis, Counters &c) {
  const float amb = view<B>(A, L.pix).amb;
  float4 color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (L.alb.w == 1.0f) {
    color.x = (L.alb.x * vis) * (1.0f - amb) + L.alb.x * amb;
    color.y = (L.alb.y * vis) * (1.0f - amb) + L.alb.y * amb;
    color.z = (L.alb.z * vis) * (1.0f - amb) + L.alb.z * amb;
    color.w = 1.0f;
  }
  if (B && A.lat) {
    // finished samples per view, one atomic per view present among the
    // active lanes; the sample that completes its view stamps the clock
    const uint32_t r = A.tile_req[L.pix >> 6];
    bool todo = true;
    const int lane = threadIdx.x & 63;
    for (uint64_t m = __ballot(todo); m; m = __ballot(todo)) {
      const int leader = __ffsll((unsigned long long)m) - 1;
      const uint32_t rl = __shfl(r, leader, 64);
      const uint64_t mine = __ballot(todo && r == rl);
      if (lane == leader) {
        const unsigned long long k = (unsigned long long)__popcll(mine);
        if (atomicAdd(&A.lat[2 * rl], k) + k == A.req[rl].samples) A.lat[2 * rl + 1] = wall_clock64();
      }
      if (r == rl) todo = false;
    }
  }
  if (STATS) {
    c.samples++;
    if (L.i >= A.max_depth) c.exhausted++;
  }
  float ra = color.w + 1.0f * (1.0f - color.w);
  if (ra != 0.0f) {
    color.x = (color.x * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
    color.y = (color.y * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
    color.z = (color.z * color.w + (0.0f * 1.0f) * (1.0f - color.w)) / ra;
    color.w = ra;
  }
  // ra is 1 for any alpha in {0, 1}, so only rgb is kept; w holds the
  // sample's noise value for the showNoise image
  S.col[(uint64_t)(L.s - S.s0) * S.npix + L.pix] = make_float4(color.x, color.y, color.z, L.rx);
}

constexpr uint32_t kNoPix = 0xffffffffu;

// Claims the next sample for a slot: first those of the pixel it is on
// (its own, pixel == slot, or the one it helps), then, visiting pixels in
// order, those of the first pixel with samples left.  Every sample is
// claimed exactly once (atomic per-pixel counters).  Returns false when
// the slot finds no work (all pixels visited, or after kProbes pixels
// without samples left: the owners finish them).
template <bool B>
__device__ inline bool take_item(const RenderArgs &A, const WFState &S, Lane &L) {
  constexpr int kProbes = 32;
  uint32_t p = L.pix;
  for (int probe = 0; probe < kProbes; ++probe) {
    if (p != kNoPix) {
      uint32_t s = atomicAdd(&S.next_s[p], 1u);
      if (s < sample_end<B>(A, S, p)) {
        L.pix = p;
        L.s = s;
        return true;
      }
    }
    // next pixel to visit, one atomic per wave for the lanes here
    uint64_t mask = __ballot(1);
    int lane = threadIdx.x & 63;
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(S.visit, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t q = base + (uint32_t)__popcll(mask & lt);
    if (q >= S.npix) break;
    p = __hip_atomic_load(&S.next_s[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sample_end<B>(A, S, q) ? q : kNoPix;
  }
  L.pix = kNoPix;
  return false;
}

// Starts samples until one needs tracking.  Returns 2 when the new ray
// needs a gather, 1 when it can step at once (the pixel's head list is
// built), 0 when the slot has no work left.  The head list is built by the
// pixel's sample 0; a later sample that starts before it is ready gathers
// its own list.
template <bool STATS, bool B>
__device__ inline int next_ray(const RenderArgs &A, const WFState &S, Lane &L, Counters &c) {
  for (;;) {
    if (!take_item<B>(A, S, L)) return 0;
    const ReqArgs &R = view<B>(A, L.pix);
    int px, py;
    pix_xy(R, L.pix, px, py);
    if (start_sample(A, R, px, py, L)) {
      float4 h = S.head[L.pix];
      uint32_t hf = __float_as_uint(h.w);
      if (hf & 1u) {
        L.flags |= F_HEAD | ((hf & 2u) ? F_DENSE : 0u);
        L.seg_t0 = h.x;
        L.seg_end = h.y;
        L.seg_n = __float_as_uint(h.z);
        return 1;
      }
      if (L.s == 0) L.flags |= F_HEADGATHER;
      L.seg_end = -AA_INF_F;
      return 2;
    }
    finish_sample<STATS, B>(A, S, L, 1.0f, c);
  }
}

__device__ inline void push(uint32_t *list, uint32_t *count, uint32_t slot) {
  // one atomic per wave for all lanes pushing to this list
  uint64_t mask = __ballot(1);
  int lane = threadIdx.x & 63;
  int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  list[base + (uint32_t)__popcll(mask & lt)] = slot;
}

template <bool STATS>
__device__ inline void flush_stats(const RenderArgs &A, const Counters &c) {
  if (!STATS) return;
  uint32_t v[8] = {c.samples, c.steps, c.queries, c.nodes, c.leaves, c.particles, c.covering, c.exhausted};
  for (int k = 0; k < 8; ++k) {
    uint64_t sum = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0 && sum) atomicAdd((unsigned long long *)&A.stats[k], (unsigned long long)sum);
  }
}

// ---- initial rays ------------------------------------------------------------
// Head lists are invalid until a sample 0 builds them (start of a call).
__global__ __launch_bounds__(256) void wf_head_reset(WFState S) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p < S.npix) S.head[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Per pass: every pixel on its view's image has samples s0..min(spp, s1)-1
// to claim, the padding pixels of partial tiles none.
template <bool B>
__global__ __launch_bounds__(256) void wf_pass_reset(RenderArgs A, WFState S) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= S.npix) return;
  const ReqArgs &R = view<B>(A, p);
  int px, py;
  pix_xy(R, p, px, py);
  S.next_s[p] = (px < R.W && py < R.H && S.s0 < R.spp) ? S.s0 : S.s1;
}

// Slot k starts on its own pixel k.
template <bool STATS, bool B>
__global__ __launch_bounds__(256) void wf_init(RenderArgs A, WFState S, uint32_t *listA, uint32_t *listG,
                                               uint32_t *counts) {
  uint32_t slot = blockIdx.x * 256u + threadIdx.x;
  Counters c = {};
  bool push_a = false, push_g = false;
  if (slot < S.nslots) {
    Lane L;
    L.pix = slot;
    L.seg_n = 0;
    L.seg_t0 = 0.0f;
    int r = next_ray<STATS, B>(A, S, L, c);
    if (r) {
      save_lane(S, slot, L);
      push_a = true;
      push_g = r == 2;
    }
  }
  if (push_a) push(listA, &counts[0], slot);
  if (push_g) push(listG, &counts[1], slot);
  flush_stats<STATS>(A, c);
}

// Running mean of every pixel over the pass's samples, in sample order
// (deviceCodeRBF.hlsl:325-335).
template <bool B>
__global__ __launch_bounds__(256) void wf_fold(RenderArgs A, WFState S) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= S.npix) return;
  const ReqArgs &R = view<B>(A, p);
  int px, py;
  pix_xy(R, p, px, py);
  if (px >= R.W || py >= R.H || S.s0 >= R.spp) return;
  uint64_t ofs = (uint64_t)px + (uint64_t)R.W * py;
  float4 acc = R.accum[ofs];
  float rx = 0.0f;
  const uint32_t s1 = min(R.spp, S.s1);
  for (uint32_t s = S.s0; s < s1; ++s) {
    float4 col = S.col[(uint64_t)(s - S.s0) * S.npix + p];
    rx = col.w;
    uint32_t aid = (R.accumID + s) & 0xffffu;
    float w1 = 1.0f / (float)aid;
    float w0 = (float)((int)aid - 1) / (float)aid;
    float4 fin;
    fin.x = w1 * col.x + w0 * acc.x;
    fin.y = w1 * col.y + w0 * acc.y;
    fin.z = w1 * col.z + w0 * acc.z;
    fin.w = w1 * 1.0f + w0 * acc.w;
    acc = fin;
  }
  R.accum[ofs] = acc;
  R.image[ofs] = R.showNoise ? make_float4(rx, rx, rx, 1.0f) : acc;
}

// ---- gather --------------------------------------------------------------------
// One lane per ray, persistent: lanes take rays from the queue through a
// shared head counter (one atomic per wave for the lanes that take one).
// Walk lengths differ by an order of magnitude between rays, and a wave
// steps as long as its longest walk, so a wave in which fewer than
// kGatherRefill lanes are still walking stops stepping: the lanes whose
// walks ended write their lists and take new rays together, and the wave
// steps on.  Each ray's walk is the same as with one ray per lane for the
// whole launch, so the lists are too.  Measured on the 1080p workload
// (ms per frame): one ray per lane 360, refilling below 16 / 32 / 48 / 56 /
// 62 walking lanes 351 / 350 / 348 / 350 / 369 (walk budget 256; 512 and
// 1024 were 4-6 % slower with refills).
#ifndef AARBF_GATHER_REFILL
#define AARBF_GATHER_REFILL 48
#endif
constexpr uint32_t kGatherRefill = AARBF_GATHER_REFILL;

template <bool STATS, int Q>
__global__ __launch_bounds__(kGBlock, AARBF_GATHER_WAVES) void wf_gather(RenderArgs A, WFState S, const uint32_t *__restrict__ listG,
                                                     const uint32_t *__restrict__ countG,
                                                     uint32_t *__restrict__ listL, uint32_t *__restrict__ countL,
                                                     uint32_t budget, uint32_t *__restrict__ head) {
  __shared__ uint32_t s_idx[kCand * kGBlock];
  __shared__ float s_tin[kCand * kGBlock];
  // the walk's per-level visiting words for levels 1 .. top-1, kGBlock
  // words each (sized at launch to the tree, which keeps more blocks
  // resident)
  extern __shared__ uint32_t s_stack[];
  const uint32_t n_g = *countG;
  if (n_g <= kWaveGather) return;            // wf_gather_wave takes short queues
  if (blockIdx.x * kGBlock >= n_g) return; // the whole block: no barrier is skipped
  load_levels(A);
  Counters c = {};
  uint32_t *gi = s_idx + threadIdx.x;
  float *gtin = s_tin + threadIdx.x;
  uint32_t *stk = s_stack + threadIdx.x;
  const QGroup *__restrict__ Qg = A.bvh.groups;
  const uint32_t log2K = A.bvh.log2K, K = 1u << log2K, top = A.bvh.top;
  // the lane's ray and walk
  uint32_t slot = 0, pix = 0, flags = 0, n = 0, maxpos = 0, nodes0 = 0;
  uint32_t lvl = 0, base = 0, ord = 0, opened = 0;
  uint4 u = make_uint4(0u, 0u, 0u, 0u);
  bool head_list = false, handed = false;
  float3 o = make_float3(0.0f, 0.0f, 0.0f), d = o;
  float t0 = 0.0f, tmax = 0.0f, m = 0.0f, maxtin = -AA_INF_F, t_cut = AA_INF_F;
  SegQuery q;
  // conservative entry/exit of the particle's sphere widened by m; false
  // when the line misses it.  Exit parameters are not kept during the
  // walk (LDS for more resident blocks): the output recomputes them with
  // the same operations.
  auto chord = [&](const float4 &P, float &tin, float &tout) -> bool {
    float vx = P.x - o.x, vy = P.y - o.y, vz = P.z - o.z;
    float tc = vx * d.x + vy * d.y + vz * d.z;
    float hx = vx - tc * d.x, hy = vy - tc * d.y, hz = vz - tc * d.z;
    float rr = P.w + m;
    float w2 = rr * rr - (hx * hx + hy * hy + hz * hz);
    if (!(w2 >= 0.0f)) return false;
    float w = sqrtf(w2);
    tin = tc - w - m;
    tout = tc + w + m;
    return true;
  };
  auto take = [&](uint32_t p, const float4 &P) {
    if (!(P.w > 0.0f)) return;
    float tin, tout;
    if (!chord(P, tin, tout)) return;
    if (!(tout >= t0) || !(tin <= tmax) || !(tin < t_cut)) {
      if (tout >= t0 && tin <= tmax) t_cut = fminf(t_cut, tin);
      return;
    }
    if (n < kCand) {
      gi[n * kGBlock] = p;
      gtin[n * kGBlock] = tin;
      if (tin > maxtin) {
        maxtin = tin;
        maxpos = n;
      }
      ++n;
      return;
    }
    if (tin < maxtin) { // evict the latest-entering candidate
      t_cut = fminf(t_cut, maxtin);
      gi[maxpos * kGBlock] = p;
      gtin[maxpos * kGBlock] = tin;
      maxtin = -AA_INF_F;
      for (uint32_t k = 0; k < kCand; ++k) {
        float v = gtin[k * kGBlock];
        if (v > maxtin) {
          maxtin = v;
          maxpos = k;
        }
      }
    } else {
      t_cut = fminf(t_cut, tin);
    }
    q.len = fminf(tmax, t_cut + 4.0f * m) - t0;
  };
  // Front-to-back walk over the quantised tree: the hit children of a
  // node are visited in order of where the segment enters their box, so
  // the list fills with the nearest candidates first and t_cut prunes the
  // rest of the tree early.  The per-level visiting order is kept as
  // nibbles (child + 1) in an LDS word.
  auto ordered = [&](uint32_t lv, uint32_t node) -> uint32_t {
    if (STATS) c.nodes += min(K, g_lvl_cnt[lv - 1] - (node << log2K));
    return ftb_children(A, q, lv, node);
  };
  auto leaf_batch = [&](uint32_t lb, uint32_t od) {
    if (A.ppl == 1) {
      float4 P[8], co, cs;
      uint32_t pid[8];
      if (Q == 2 || (Q == 0 && A.q16)) ld_chunk(A, lb, co, cs); // the group's leaves share an FP16 chunk
#pragma unroll
      for (uint32_t r = 0; r < 8; ++r) {
        uint32_t nib = (od >> (4u * r)) & 15u;
        pid[r] = nib ? lb + nib - 1u : 0xffffffffu;
        if (nib) P[r] = ld_pr_in<Q>(A, pid[r], co, cs);
      }
#pragma unroll
      for (uint32_t r = 0; r < 8; ++r)
        if (pid[r] != 0xffffffffu) {
          if (STATS) c.particles++;
          take(pid[r], P[r]);
        }
      return;
    }
    for (; od; od >>= 4) {
      uint32_t leaf = lb + (od & 15u) - 1u;
      uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
      for (uint32_t p = p0; p < p1; ++p) {
        if (STATS) c.particles++;
        take(p, ld_pr<Q>(A, p));
      }
    }
  };
  // set-up of ray `gid` of the queue; false when its segment misses the
  // root box (the walk is over at once)
  auto start = [&](uint32_t gid) -> bool {
    slot = listG[gid];
    float4 a0 = S.od0[slot], a1 = S.od1[slot];
    o = make_float3(a0.x, a0.y, a0.z);
    d = make_float3(a1.x, a1.y, a1.z);
    u = S.u[slot];
    flags = u.z >> 16;
    head_list = (flags & F_HEADGATHER) != 0;
    pix = S.pix[slot];
    t0 = head_list ? S.seg[slot].x : a0.w;
    tmax = a1.w;
    q.ax = o.x + t0 * d.x;
    q.ay = o.y + t0 * d.y;
    q.az = o.z + t0 * d.z;
    q.dx = d.x;
    q.dy = d.y;
    q.dz = d.z;
    q.ix = safe_inv(d.x);
    q.iy = safe_inv(d.y);
    q.iz = safe_inv(d.z);
    m = 2e-5f * (1.0f + fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) + fmaxf(fabsf(t0), fabsf(tmax)));
    q.m = m;
    q.len = tmax - t0 + 4.0f * m;
    n = 0;
    maxpos = 0;
    maxtin = -AA_INF_F;
    t_cut = AA_INF_F;
    handed = false;
    opened = 0;
    if (STATS) c.queries++;
    nodes0 = c.nodes;
    const float4 *root = A.bvh.nodes + g_lvl_ofs[top];
    if (STATS) c.nodes++;
    if (!seg_box(q, root[0], root[1])) return false;
    lvl = top - 1;
    base = 0;
    ord = ordered(top, 0);
    return true;
  };
  // one pass of the walk; false when the walk is over (or handed over: a
  // walk that opens more than `budget` nodes hands its ray to the
  // long-walk list, which wf_gather_wave takes in the same round, so a
  // launch lasts as long as its typical walks, not its longest)
  auto advance = [&]() -> bool {
    if (lvl == 0) { // a whole leaf group, in entry order
      leaf_batch(base, ord);
      ord = 0;
    }
    uint32_t nib = ord & 15u;
    if (nib == 0) {
      if (lvl == top - 1) return false;
      ++lvl;
      ord = stk[(lvl - 1u) * kGBlock];
      base = ((base >> log2K) >> log2K) << log2K;
      return true;
    }
    if (listL && ++opened > budget) {
      handed = true;
      return false;
    }
    ord >>= 4;
    uint32_t node = base + nib - 1u;
    stk[(lvl - 1u) * kGBlock] = ord; // lvl >= 1 here
    base = node << log2K;
    ord = ordered(lvl, node);
    --lvl;
    return true;
  };
  // the ray's list and segment end into the slot state
  auto finish = [&]() {
    if (STATS) {
      uint32_t cost = c.nodes - nodes0;
      uint32_t b = cost ? 31u - (uint32_t)__clz(cost) : 0u;
      atomicAdd((unsigned long long *)&A.stats[8 + min(b, 17u)], 1ull);
    }
    float seg_end;
    if (t_cut <= t0 + 2.0f * A.unit) { // crowded: point queries for a short stretch
      flags |= F_DENSE;
      seg_end = t0 + 4.0f * A.unit;
      n = 0;
    } else {
      flags &= ~F_DENSE;
      uint32_t *oidx = head_list ? S.hidx : S.cidx;
      float *otin = head_list ? S.htin : S.ctin, *otout = head_list ? S.htout : S.ctout;
      uint8_t *oeo = head_list ? S.heo : S.ceo;
      const uint64_t row = head_list ? pix : slot, stride = head_list ? S.npix : S.nslots;
      seg_end = t_cut;
      // ascending particle order (insertion sort, n <= kCand)
      for (uint32_t a = 1; a < n; ++a) {
        uint32_t pi = gi[a * kGBlock];
        float ti = gtin[a * kGBlock];
        uint32_t b = a;
        while (b > 0 && gi[(b - 1) * kGBlock] > pi) {
          gi[b * kGBlock] = gi[(b - 1) * kGBlock];
          gtin[b * kGBlock] = gtin[(b - 1) * kGBlock];
          --b;
        }
        gi[b * kGBlock] = pi;
        gtin[b * kGBlock] = ti;
      }
      // entry order: position of rank r in tin order (ties by position)
      for (uint32_t k = 0; k < n; ++k) {
        float tk = gtin[k * kGBlock];
        uint32_t r = 0;
        for (uint32_t j = 0; j < n; ++j) {
          float tj = gtin[j * kGBlock];
          r += (tj < tk) || (tj == tk && j < k);
        }
        const uint32_t pk = gi[k * kGBlock];
        float tin, tout;
        chord(ld_pr<Q>(A, pk), tin, tout); // the exit parameter take() computed
        oeo[r * stride + row] = (uint8_t)k;
        oidx[k * stride + row] = pk;
        otin[k * stride + row] = tk;
        otout[k * stride + row] = tout;
      }
    }
    if (head_list) {
      flags = (flags & ~F_HEADGATHER) | F_HEAD;
      S.head[pix] = make_float4(t0, seg_end, __uint_as_float(n),
                                 __uint_as_float(1u | ((flags & F_DENSE) ? 2u : 0u)));
    } else {
      flags &= ~F_HEAD;
    }
    S.u[slot] = make_uint4(u.x, u.y, (u.z & 0xffffu) | (flags << 16), n);
    float4 g = S.seg[slot];
    S.seg[slot] = make_float4(t0, seg_end, g.z, g.w);
  };
  // rays from the queue for the lanes that ask, in queue order
  const int lane = threadIdx.x & 63;
  auto fetch = [&](bool want) -> uint32_t {
    const uint64_t mask = __ballot(want);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t b = 0;
    if (lane == leader) b = atomicAdd(head, (uint32_t)__popcll(mask));
    b = __shfl(b, leader, 64);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    return b + (uint32_t)__popcll(mask & lt);
  };
  // lanes: walking (a ray whose walk goes on), done (a ray whose walk
  // ended), or idle (the queue is empty)
  bool walking = false, done = false, idle = false;
  for (;;) {
    // done lanes write their lists; they and the lanes without a ray take
    // new rays until a walk does not end at its start
    if (done) {
      if (handed) push(listL, countL, slot);
      else finish();
      done = false;
    }
    bool want = !walking && !idle;
    while (__ballot(want)) {
      const uint32_t gid = fetch(want);
      if (want) {
        if (gid >= n_g) {
          idle = true;
        } else if (start(gid)) {
          walking = true;
        } else {
          finish(); // the segment misses the root box: an empty list
        }
      }
      want = !walking && !idle;
    }
    if (!__ballot(walking)) break;
    // step the walks while enough lanes walk (all of them once the queue
    // is empty)
    const bool drain = __ballot(idle) != 0;
    do {
      if (walking && !advance()) {
        walking = false;
        done = true;
      }
    } while ((uint32_t)__popcll(__ballot(walking)) >= (drain ? 1u : kGatherRefill));
  }
  flush_stats<STATS>(A, c);
}

// Tracker steps per ray in a round: fewer live rays -> more steps each, so
// the tail of a pass costs few rounds.  Only the split of a ray's steps
// over rounds depends on it, not the image.
__device__ inline uint32_t round_cap(uint32_t base_cap, uint32_t nslots, uint32_t n_active) {
  const uint64_t b = base_cap;
  return (uint32_t)min(65536ull, max(b, b * nslots / (4ull * max(n_active, 1u))));
}


// ---- gather, one wave per ray ---------------------------------------------------
// The same candidate list and segment end as wf_gather, for short queues.
// The wave first expands the tree breadth first from the root (64 child
// boxes per pass) until at least kWaveSplit subtrees meet the segment, then every
// lane walks its share of them front to back as wf_gather does, keeping its
// own kCand nearest entries.  The lanes share a bound on the result's
// segment end (the smallest of their own cuts, an LDS atomic): a lane's cut
// is the entry of a particle that does not fit among its own kCand, so at
// least kCand + 1 particles enter by then and no particle entering later
// can be in the result or end its segment; boxes entered after it (with the
// rounding margin) are skipped.  A merge then takes the kCand smallest
// entries over all lanes; the segment ends at the next entry or the
// smallest cut, whichever comes first -- the (kCand + 1)-th smallest entry
// overall, as in wf_gather.  Lists that differ only among equal entries
// give the same image: no particle entering at the segment end can cover a
// step before it.
__device__ inline int ord_key(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ inline float ord_val(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

template <bool STATS>
__global__ __launch_bounds__(64) void wf_gather_wave(RenderArgs A, WFState S, const uint32_t *__restrict__ listG,
                                                     const uint32_t *__restrict__ countG, uint32_t limit) {
  __shared__ uint32_t s_idx[kCand * 64];
  __shared__ float s_tin[kCand * 64];
  extern __shared__ uint32_t s_stack[]; // levels 0 .. top-1, 64 words each (sized at launch)
  __shared__ uint32_t s_front[2][kWaveFront];
  __shared__ uint32_t s_ridx[kCand], s_rk[kCand];
  __shared__ float s_rtin[kCand];
  __shared__ int s_bound;
  const uint32_t n_g = *countG;
  if (n_g > limit || blockIdx.x >= n_g) return; // the whole block
  load_levels(A);
  const uint32_t lane = threadIdx.x;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  Counters c = {};
  const QGroup *__restrict__ Qg = A.bvh.groups;
  const uint32_t log2K = A.bvh.log2K, K = 1u << log2K, top = A.bvh.top;
  uint32_t *gi = s_idx + lane, *stk = s_stack + lane;
  float *gtin = s_tin + lane;
  for (uint32_t item = blockIdx.x; item < n_g; item += gridDim.x) {
    const uint32_t slot = listG[item];
    const float4 a0 = S.od0[slot], a1 = S.od1[slot];
    const float3 o = make_float3(a0.x, a0.y, a0.z), d = make_float3(a1.x, a1.y, a1.z);
    const uint4 u = S.u[slot];
    uint32_t flags = u.z >> 16;
    const bool head = (flags & F_HEADGATHER) != 0;
    const uint32_t pix = S.pix[slot];
    const float t0 = head ? S.seg[slot].x : a0.w, tmax = a1.w;
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
    const float m = 2e-5f * (1.0f + fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) + fmaxf(fabsf(t0), fabsf(tmax)));
    q.m = m;
    q.len = tmax - t0 + 4.0f * m;
    uint32_t n = 0;
    float t_cut = AA_INF_F;
    if (STATS && lane == 0) c.queries++;
    if (lane == 0) s_bound = ord_key(AA_INF_F);
    // breadth first from the root
    const float4 *root = A.bvh.nodes + g_lvl_ofs[top];
    uint32_t nf = 0, lvl = top, cur = 0;
    if (STATS && lane == 0) c.nodes++;
    if (seg_box(q, root[0], root[1])) {
      if (lane == 0) s_front[0][0] = 0;
      nf = 1;
    }
    __syncthreads();
    while (nf > 0 && nf < kWaveSplit && lvl > 1) {
      const uint32_t total = nf << log2K; // < kWaveSplit * 8 = kWaveFront
      uint32_t nn = 0;
      for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t j = base + lane;
        bool hit = false;
        uint32_t child = 0;
        if (j < total) {
          const uint32_t node = s_front[cur][j >> log2K], k = j & (K - 1u);
          // this lane's child only: its bytes read from memory (k varies
          // across the lanes, so a copy of the group would be indexed
          // dynamically and land in scratch)
          const QGroup *gp = Qg + g_lvl_qofs[lvl] + node;
          const float4 ga = gp->a, gb = gp->b;
          auto qb = [&](uint32_t field) { return (gp->q[2u * field + (k >> 2)] >> ((k & 3u) * 8u)) & 255u; };
          const uint32_t first = node << log2K, mcnt = min(K, g_lvl_cnt[lvl - 1] - first);
          const uint32_t lx = qb(0), hx = qb(3);
          child = first + k;
          if (k < mcnt && lx <= hx) {
            float4 lo, hi;
            lo.x = ga.x + (float)lx * ga.w;
            lo.y = ga.y + (float)qb(1) * gb.x;
            lo.z = ga.z + (float)qb(2) * gb.y;
            hi.x = ga.x + (float)hx * ga.w;
            hi.y = ga.y + (float)qb(4) * gb.x;
            hi.z = ga.z + (float)qb(5) * gb.y;
            float tn;
            hit = seg_box_t(q, lo, hi, tn);
          }
          if (STATS && k < mcnt) c.nodes++;
        }
        const uint64_t mk = __ballot(hit);
        if (hit) s_front[cur ^ 1u][nn + (uint32_t)__popcll(mk & lt)] = child;
        nn += (uint32_t)__popcll(mk);
      }
      __syncthreads();
      cur ^= 1u;
      nf = nn;
      --lvl;
    }
    // each lane's subtrees, front to back
    auto cut_now = [&]() { return fminf(t_cut, ord_val(__hip_atomic_load(&s_bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))); };
    auto chord = [&](const float4 &P, float &tin, float &tout) -> bool {
      float vx = P.x - o.x, vy = P.y - o.y, vz = P.z - o.z;
      float tc = vx * d.x + vy * d.y + vz * d.z;
      float hx = vx - tc * d.x, hy = vy - tc * d.y, hz = vz - tc * d.z;
      float rr = P.w + m;
      float w2 = rr * rr - (hx * hx + hy * hy + hz * hz);
      if (!(w2 >= 0.0f)) return false;
      float w = sqrtf(w2);
      tin = tc - w - m;
      tout = tc + w + m;
      return true;
    };
    auto lower_cut = [&](float v) {
      if (v < t_cut) {
        t_cut = v;
        atomicMin(&s_bound, ord_key(v));
      }
    };
    auto take = [&](uint32_t p, const float4 &P) {
      if (!(P.w > 0.0f)) return;
      float tin, tout;
      if (!chord(P, tin, tout)) return;
      if (!(tout >= t0) || !(tin <= tmax) || !(tin < t_cut)) {
        if (tout >= t0 && tin <= tmax) lower_cut(tin);
        return;
      }
      // the lane's list stays sorted by entry; a full list drops its last
      uint32_t b = n;
      if (n == kCand) {
        const float last = gtin[(kCand - 1u) * 64];
        if (!(tin < last)) {
          lower_cut(tin);
          return;
        }
        lower_cut(last);
        b = kCand - 1u;
      } else {
        ++n;
      }
      for (; b > 0 && gtin[(b - 1u) * 64] > tin; --b) {
        gi[b * 64] = gi[(b - 1u) * 64];
        gtin[b * 64] = gtin[(b - 1u) * 64];
      }
      gi[b * 64] = p;
      gtin[b * 64] = tin;
    };
    auto leaf_batch = [&](uint32_t lb, uint32_t ord) {
      if (A.ppl == 1) {
        float4 P[8], co, cs;
        uint32_t pid[8];
        if (A.q16) ld_chunk(A, lb, co, cs); // the group's leaves share an FP16 chunk
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) {
          uint32_t nib = (ord >> (4u * r)) & 15u;
          pid[r] = nib ? lb + nib - 1u : 0xffffffffu;
          if (nib) P[r] = ld_pr_in(A, pid[r], co, cs);
        }
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r)
          if (pid[r] != 0xffffffffu) {
            if (STATS) c.particles++;
            take(pid[r], P[r]);
          }
        return;
      }
      for (; ord; ord >>= 4) {
        uint32_t leaf = lb + (ord & 15u) - 1u;
        uint32_t p0 = leaf * A.ppl, p1 = min(p0 + A.ppl, A.numParticles);
        for (uint32_t p = p0; p < p1; ++p) {
          if (STATS) c.particles++;
          take(p, ld_pr(A, p));
        }
      }
    };
    auto children = [&](uint32_t l, uint32_t node) {
      q.len = fminf(tmax, cut_now() + 4.0f * m) - t0;
      if (STATS) c.nodes += K;
      return ftb_children(A, q, l, node);
    };
    // The (kCand + 1)-th smallest entry over the lanes' sorted lists (+inf
    // while they hold fewer), by rounds that compare the lanes' heads; with
    // `out`, the kCand smallest are written to s_ridx / s_rtin as well.
    auto union_cut = [&](bool out, uint32_t &taken) -> float {
      uint32_t hk = 0;
      taken = 0;
      for (uint32_t r = 0; r <= kCand; ++r) {
        float v = hk < n ? gtin[hk * 64] : AA_INF_F;
        uint32_t who = lane;
        for (int o2 = 32; o2 > 0; o2 >>= 1) {
          const float v2 = __shfl_xor(v, o2, 64);
          const uint32_t w2 = __shfl_xor(who, o2, 64);
          if (v2 < v || (v2 == v && w2 < who)) {
            v = v2;
            who = w2;
          }
        }
        if (!(v < AA_INF_F)) return AA_INF_F; // fewer than kCand + 1 entries in all
        if (r == kCand) return v;
        if (lane == who) {
          if (out) {
            s_ridx[r] = gi[hk * 64];
            s_rtin[r] = gtin[hk * 64];
          }
          ++hk;
        }
        ++taken;
      }
      return AA_INF_F;
    };
    // Every lane walks its subtrees (frontier entries lane, lane + 64, ...)
    // front to back in slices of kWaveSlice opened nodes; after each slice
    // the (kCand + 1)-th smallest entry over all lanes becomes the shared
    // bound, so a lane whose own subtrees hold few particles still prunes.
    {
      uint32_t fi = lane, L = lvl - 1u, l = 0, base = 0, ord = 0;
      bool have = false, done = fi >= nf;
      for (;;) {
        uint32_t left = kWaveSlice;
        while (!done && left) {
          if (!have) { // the next subtree, its root at level lvl >= 1
            const uint32_t R = s_front[cur][fi];
            l = L;
            base = R << log2K;
            ord = children(lvl, R);
            have = true;
            --left;
            continue;
          }
          if (l == 0) { // a whole leaf group, in entry order
            leaf_batch(base, ord);
            ord = 0;
          }
          const uint32_t nib = ord & 15u;
          if (nib == 0) {
            if (l == L) {
              have = false;
              fi += 64u;
              done = fi >= nf;
              continue;
            }
            ++l;
            ord = stk[(l - 1u) * 64u];
            base = ((base >> log2K) >> log2K) << log2K;
            continue;
          }
          ord >>= 4;
          const uint32_t node = base + nib - 1u;
          stk[(l - 1u) * 64u] = ord; // l >= 1 here
          base = node << log2K;
          ord = children(l, node);
          --l;
          --left;
        }
        if (__all(done)) break;
        uint32_t tk;
        const float ub = union_cut(false, tk);
        if (lane == 0 && ub < ord_val(s_bound)) s_bound = ord_key(ub);
        __syncthreads();
      }
    }
    float cut_all = t_cut;
    for (int o2 = 32; o2 > 0; o2 >>= 1) cut_all = fminf(cut_all, __shfl_xor(cut_all, o2, 64));
    uint32_t nres = 0;
    const float t_next = union_cut(true, nres);
    t_cut = fminf(t_next, cut_all);
    __syncthreads();
    // the list out, one lane per entry: entry r (the r-th smallest entry)
    // goes to its position k in ascending particle order, and entry order
    // rank r' (ties by that position) maps to k
    const bool crowded = t_cut <= t0 + 2.0f * A.unit; // point queries for a short stretch
    const uint32_t nout = crowded ? 0u : nres;
    uint32_t *oidx = head ? S.hidx : S.cidx;
    float *otin = head ? S.htin : S.ctin, *otout = head ? S.htout : S.ctout;
    uint8_t *oeo = head ? S.heo : S.ceo;
    const uint64_t row = head ? pix : slot, stride = head ? S.npix : S.nslots;
    if (lane < nout) {
      const uint32_t pi = s_ridx[lane];
      uint32_t k = 0;
      for (uint32_t j = 0; j < nout; ++j) k += s_ridx[j] < pi;
      s_rk[lane] = k;
    }
    __syncthreads();
    if (lane < nout) {
      const uint32_t pi = s_ridx[lane], k = s_rk[lane];
      const float ti = s_rtin[lane];
      uint32_t r = 0;
      for (uint32_t j = 0; j < nout; ++j) {
        const float tj = s_rtin[j];
        r += (tj < ti) || (tj == ti && s_rk[j] < k);
      }
      float tin, tout;
      chord(ld_pr(A, pi), tin, tout); // the exit parameter take() computed
      oeo[r * stride + row] = (uint8_t)k;
      oidx[k * stride + row] = pi;
      otin[k * stride + row] = ti;
      otout[k * stride + row] = tout;
    }
    if (lane == 0) {
      float seg_end;
      if (crowded) {
        flags |= F_DENSE;
        seg_end = t0 + 4.0f * A.unit;
      } else {
        flags &= ~F_DENSE;
        seg_end = t_cut;
      }
      if (head) {
        flags = (flags & ~F_HEADGATHER) | F_HEAD;
        S.head[pix] = make_float4(t0, seg_end, __uint_as_float(nout),
                                  __uint_as_float(1u | ((flags & F_DENSE) ? 2u : 0u)));
      } else {
        flags &= ~F_HEAD;
      }
      S.u[slot] = make_uint4(u.x, u.y, (u.z & 0xffffu) | (flags << 16), nout);
      const float4 g = S.seg[slot];
      S.seg[slot] = make_float4(t0, seg_end, g.z, g.w);
    }
    __syncthreads(); // s_ridx, s_bound and the frontier are reused by the next ray
  }
  flush_stats<STATS>(A, c);
}

// ---- step ------------------------------------------------------------------------
// Rays whose segment has a candidate list; rays on a dense segment are
// handed to the dense queue (wf_dense_wave) untouched, which keeps the BVH
// walk's registers out of this kernel.
// Q: the particle payload, 0 = as A holds it, 1 = FP32, 2 = FP16 (the
// candidate tests compiled for one payload: 2 % faster with FP16).
template <bool STATS, int WPE, bool B, int Q>
__global__ __launch_bounds__(kSBlock, WPE) void wf_step(RenderArgs A, WFState S, const uint32_t *__restrict__ listA,
                                                   const uint32_t *__restrict__ countA, uint32_t *__restrict__ outA,
                                                   uint32_t *__restrict__ outG, uint32_t *__restrict__ counts,
                                                   uint32_t *__restrict__ outD, uint32_t *__restrict__ countD,
                                                   uint32_t step_cap) {
  __shared__ uint32_t s_idx[kCand * kSBlock];
  __shared__ float s_tin[kCand * kSBlock];
  __shared__ float s_tout[kCand * kSBlock];
  __shared__ uint8_t s_eo[kCand * kSBlock];
  __shared__ float s_dm[kDmLds];
  const uint32_t n_a = *countA;
  if (blockIdx.x * kSBlock >= n_a) return; // the whole block: no barrier is skipped
  step_cap = round_cap(step_cap, S.nslots, n_a);
  // density map red channel, read at every accepted-candidate step
  const bool dm_lds = A.dm_n <= kDmLds;
  if (dm_lds)
    for (uint32_t j = threadIdx.x; j < A.dm_n; j += kSBlock) s_dm[j] = A.dmap[j].x;
  __syncthreads();
  auto density = [&](float u) {
    if (!dm_lds) return tf_r(A.dmap, A.dm_n, u);
    uint32_t j0, j1;
    float f;
    aa_tf_coords_nf(u, A.dm_n, A.dm_nf, &j0, &j1, &f);
    return aa_lerp_tex(s_dm[j0], s_dm[j1], f);
  };
  Counters c = {};
  uint64_t simt[3] = {0, 0, 0};
  // early leaving only while the queue can fill the GPU again
  const bool may_leave = A.min_live != 0u && n_a >= kMinLiveQueue;
  // grid-stride, so a fixed grid (a captured graph) covers any count
  for (uint32_t base = blockIdx.x * kSBlock; base < n_a; base += gridDim.x * kSBlock) {
  const uint32_t gid = base + threadIdx.x;
  int outcome = 0; // 0 nothing, 1 keep active, 2 active + gather, 3 dense queue
  uint32_t slot = 0;
  if (gid < n_a) {
    slot = listA[gid];
    if ((S.u[slot].z >> 16) & F_DENSE) {
      outcome = 3;
    }
  }
  if (gid < n_a && outcome == 0) {
    Lane L;
    load_lane(S, slot, L);
    uint32_t *li = s_idx + threadIdx.x;
    float *ltin = s_tin + threadIdx.x, *ltout = s_tout + threadIdx.x;
    uint8_t *leo = s_eo + threadIdx.x;
    uint32_t n = 0;
    // active set at the current t
    uint32_t act = 0, eptr = 0;
    float t_enter = AA_INF_F, t_exit = AA_INF_F;
    auto enter = [&](float t) {
      while (eptr < n) {
        uint32_t k = leo[eptr * kSBlock];
        if (!(ltin[k * kSBlock] <= t)) break;
        float to = ltout[k * kSBlock];
        if (t <= to) {
          act |= 1u << k;
          t_exit = fminf(t_exit, to);
        }
        ++eptr;
      }
      t_enter = eptr < n ? ltin[leo[eptr * kSBlock] * kSBlock] : AA_INF_F;
    };
    // A draw that lands before lim with no active particle is a step in
    // empty space: the query at it covers nothing, so only t and i move.
    // With delta tracking a draw lands before lim while the product of
    // its 1 - xi since the anchor stays above T = aa_flight_bound(tb, lim).
    // lim = min(tmax, segment end, next entry)
    float lim = 0.0f, T = 0.0f;
    const float inv_unit = A.inv_unit;
    auto leave = [&](float t) {
      t_exit = AA_INF_F;
      uint32_t m = act;
      while (m) {
        uint32_t k = (uint32_t)__ffs(m) - 1u;
        m &= m - 1u;
        float to = ltout[k * kSBlock];
        if (to < t)
          act &= ~(1u << k);
        else
          t_exit = fminf(t_exit, to);
      }
    };
    // the candidate list of the lane's ray into LDS, active set at L.t; a
    // pixel's head list already there (the previous sample of the same
    // pixel ended on it) is not loaded again
    uint32_t lds_head = 0xffffffffu; // pixel whose head list the lane's LDS holds
    auto setup = [&]() {
      n = (L.flags & F_DENSE) ? 0u : L.seg_n;
      const bool hd = (L.flags & F_HEAD) != 0;
      if (!(hd && lds_head == L.pix)) {
        const uint32_t *iidx = hd ? S.hidx : S.cidx;
        const float *itin = hd ? S.htin : S.ctin, *itout = hd ? S.htout : S.ctout;
        const uint8_t *ieo = hd ? S.heo : S.ceo;
        const uint64_t row = hd ? L.pix : slot, stride = hd ? S.npix : S.nslots;
        for (uint32_t k = 0; k < n; ++k) {
          li[k * kSBlock] = iidx[k * stride + row];
          ltin[k * kSBlock] = itin[k * stride + row];
          ltout[k * kSBlock] = itout[k * stride + row];
          leo[k * kSBlock] = ieo[k * stride + row];
        }
        lds_head = hd ? L.pix : 0xffffffffu;
      }
      act = 0;
      eptr = 0;
      t_exit = AA_INF_F;
      enter(L.t);
      lim = fminf(fminf(L.tmax, L.seg_end), t_enter);
      T = aa_flight_bound(L.tb, lim, inv_unit);
    };
    setup();

    uint32_t steps = 0, it = 0;
    for (;;) {
      // a wave whose lanes have mostly left (for gathers) stops after at
      // least one iteration; its remaining rays go on in the next round,
      // packed into full waves, instead of stepping at low lane occupancy
      if (may_leave && ++it > 1u && (uint32_t)__popcll(__ballot(1)) < A.min_live) {
        outcome = 1;
        break;
      }
      if (STATS) { // SIMT accounting: iterations, lanes in them, iterations touching candidates
        uint64_t live = __ballot(1);
        if ((uint32_t)__ffsll((unsigned long long)live) - 1u == (threadIdx.x & 63u)) {
          simt[0] += 1;
          simt[1] += (uint32_t)__popcll(live);
        }
        uint64_t touch = __ballot(act != 0);
        if (touch && (uint32_t)__ffsll((unsigned long long)live) - 1u == (threadIdx.x & 63u)) simt[2] += 1;
      }
      const bool pending = (L.flags & F_PENDING) != 0;
      // The next free-flight draw, computed by every lane (nearly all of
      // them step in each iteration) and committed only where a step is
      // taken.  Marching steps a fixed distance and draws nothing; delta
      // tracking multiplies the draw's 1 - xi into the product P (its
      // position is needed only where the step is not in empty space).
      uint32_t ns = L.trng;
      const bool march = is_march<B>(A, L);
      const float tm = L.t + A.unit;
      const float pn = march ? 1.0f : L.P * flight_x(ns);
      const bool can = !pending && L.i < A.max_depth && steps < step_cap;
      if (can && act == 0u && (march ? tm < lim : pn > T)) {
        L.trng = ns;
        uint32_t taken = 1;
        // more draws in the same iteration, each committed only when it
        // too lands in empty space before the next event (the first that
        // does not is drawn again by the next iteration), within the depth
        // and step budgets left after this one
        const uint32_t room = min(A.max_depth - (L.i + 1u), step_cap - (steps + 1u));
        if (march) {
          L.t = tm;
          // the budgets are checked per step only where fewer than
          // kFastSteps - 1 steps are left (a rolled loop, rarely taken)
          auto draw = [&]() {
            const float tn2 = L.t + A.unit;
            if (!(tn2 < lim)) return false;
            L.t = tn2;
            ++taken;
            return true;
          };
          if (room >= (uint32_t)(kFastSteps - 1)) {
#pragma unroll
            for (int u = 1; u < kFastSteps; ++u)
              if (!draw()) break;
          } else {
#pragma nounroll
            for (uint32_t u = 1; u <= room; ++u)
              if (!draw()) break;
          }
        } else {
          L.P = pn;
          // delta tracking: up to the next flight anchor (L.i a multiple
          // of kFlightBlock), where the position is formed once
          const uint32_t left = min(room, kFlightBlock - 1u - L.i % kFlightBlock);
          // one region per draw, not nested (a lane that stops is masked
          // off the later ones, a wave skips a region no lane enters), so
          // no join copies and no saved masks per draw
          bool go = true;
#pragma unroll
          for (uint32_t u = 1; u < kFlightBlock; ++u) {
            if (go) {
              uint32_t ns2 = L.trng;
              const float p2 = L.P * flight_x(ns2);
              go = (p2 > T) & (u <= left);
              if (go) {
                L.trng = ns2;
                L.P = p2;
                ++taken;
              }
            }
          }
        }
        L.i += taken;
        steps += taken;
        if (STATS) c.steps += taken;
        if (!march && L.i % kFlightBlock == 0u) { // the anchor
          L.t = aa_flight_pos(L.tb, L.P, A.unit);
          L.tb = L.t;
          L.P = 1.0f;
          T = aa_flight_bound(L.tb, lim, inv_unit);
        }
        continue;
      }
      bool end = false;
      if (!pending) {
        if (!(L.i < A.max_depth)) {
          end = true;
        } else if (steps >= step_cap) {
          outcome = 1;
          break;
        } else {
          ++steps;
          if (STATS) c.steps++;
          L.trng = ns;
          if (march) {
            L.t = tm;
          } else {
            L.P = pn;
            L.t = aa_flight_pos(L.tb, pn, A.unit);
          }
          if (L.t >= L.tmax) end = true;
          else L.flags |= F_PENDING;
        }
      }
      if (!end) {
        // the point query at L.t
        if (!(L.t < L.seg_end)) {
          outcome = 2;
          break;
        }
        Payload pl;
        float x = L.o.x + L.t * L.d.x, y = L.o.y + L.t * L.d.y, z = L.o.z + L.t * L.d.z;
        bool moved = false; // lim or the flight anchor: the bound is due
        if (L.t >= t_enter) {
          enter(L.t);
          lim = fminf(fminf(L.tmax, L.seg_end), t_enter);
          moved = true;
        }
        if (L.t > t_exit) leave(L.t);
        pl.count = 0;
        pl.dens = 0.0f;
        pl.r = pl.g = pl.b = 0.0f;
        auto test = [&](const float4 &qq, const float4 &kk, float pm) {
          if (STATS) c.particles++;
          const float e = aa_dist2(qq.x, qq.y, qq.z, x, y, z) * pm;
          if (e > -4.5f) {
            float wa = aa_expf_mid(e) * kk.x;
            if (STATS) c.covering++;
            pl.count += 1;
            pl.dens += wa;
            pl.r += kk.y * wa;
            pl.g += kk.z * wa;
            pl.b += kk.w * wa;
            if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
          }
        };
        // two active candidates per pass (ascending list position, i.e.
        // ascending particle index), their loads in one round trip (three
        // or four per pass measured 1-2 % slower with FP32 and 2-6 % with
        // FP16: the absent ones are loaded and decoded too)
        uint32_t m = act;
        while (m) {
          const uint32_t k0 = (uint32_t)__ffs(m) - 1u;
          m &= m - 1u;
          const bool two = m != 0u;
          const uint32_t k1 = two ? (uint32_t)__ffs(m) - 1u : k0;
          m &= m - 1u;
          float4 q0, c0, q1, c1;
          float m0, m1;
          ld_cand<Q>(A, li[k0 * kSBlock], q0, c0, m0);
          ld_cand<Q>(A, li[k1 * kSBlock], q1, c1, m1);
          test(q0, c0, m0);
          if (two) test(q1, c1, m1);
        }
        L.flags &= ~F_PENDING;
        const bool shadow = (L.flags & F_SHADOW) != 0;
        if (is_march<B>(A, L)) {
          if (pl.count > 0 && pl.dens > 0.0f) {
            float D = density(pl.dens);
            L.LHS += D;
            if (L.LHS > L.RHS && !shadow) {
              float3 col = colour_of(A, pl);
              L.alb = make_float4(col.x, col.y, col.z, 1.0f);
              end = true;
            } else if (shadow) {
              L.shw += 1.0f - aa_expf(-D);
              if (L.shw >= 1.0f) {
                L.shw = 1.0f;
                end = true;
              }
            }
          }
        } else if (pl.count > 0) {
          float D = density(pl.dens);
          float xi2 = aa_lcg_next(&L.trng);
          if (xi2 < D) {
            if (shadow) {
              L.shw = 1.0f;
            } else {
              float3 col = colour_of(A, pl);
              L.alb = make_float4(col.x, col.y, col.z, 1.0f);
            }
            end = true;
          }
        }
        if (!end) {
          ++L.i;
          if (!march) {
            if (L.i % kFlightBlock == 0u) { // the anchor, at the query's position
              L.tb = L.t;
              L.P = 1.0f;
              moved = true;
            }
            if (moved) T = aa_flight_bound(L.tb, lim, inv_unit);
          }
          continue;
        }
      }
      // ---- the ray ended
      L.flags &= ~F_PENDING;
      const ReqArgs &R = view<B>(A, L.pix);
      if (!(L.flags & F_SHADOW) && L.alb.w > 0.0f && R.amb != 1.0f) {
        // NEE shadow ray (':288-311'); the step counter carries over
        float3 so = make_float3(R.pos.x + L.d.x * L.t, R.pos.y + L.d.y * L.t, R.pos.z + L.d.z * L.t);
        float ste, stx;
        aabb_intersection(so, R.light, 0.0f, 10000.0f, A.lb, A.rt, ste, stx);
        L.o = so;
        L.d = R.light;
        L.tmax = stx;
        L.flags = F_SHADOW | (L.flags & F_MARCH);
        L.shw = 0.0f;
        L.LHS = 0.0f;
        L.t = is_march<B>(A, L) ? -A.unit * L.r2 : 0.0f;
        L.tb = L.t; // the shadow ray's first flight anchor
        L.P = 1.0f;
        L.seg_end = -AA_INF_F;
        outcome = 2;
        break;
      }
      float vis = (L.flags & F_SHADOW) ? 1.0f - L.shw : 1.0f;
      finish_sample<STATS, B>(A, S, L, vis, c);
      outcome = next_ray<STATS, B>(A, S, L, c);
      if (outcome == 1 && !(L.flags & F_DENSE)) {
        // the pixel's next sample can step at once: go on with it here
        setup();
        outcome = 0;
        continue;
      }
      break;
    }
    // empty-space steps leave the position to the anchor and the product
    if (outcome && !is_march<B>(A, L)) L.t = aa_flight_pos(L.tb, L.P, A.unit);
    if (outcome) save_lane(S, slot, L);
  }
  // compaction: lanes with an outcome append to the next lists
  if (outcome == 1 || outcome == 2) push(outA, &counts[0], slot);
  if (outcome == 2) push(outG, &counts[1], slot);
  if (outcome == 3) push(outD, countD, slot);
  }
  flush_stats<STATS>(A, c);
  if (STATS)
    for (int k = 0; k < 3; ++k) {
      uint64_t sum = simt[k];
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if ((threadIdx.x & 63) == 0 && sum)
        atomicAdd((unsigned long long *)&A.stats[26 + k], (unsigned long long)sum);
    }
}

// ---- majorant tracking (render mode 2) -------------------------------------------
// Delta tracking with local majorants instead of the reference's global
// majorant of 1 per unit.  Between two entry/exit events of the candidate
// list the density is at most
//     mu = M(sum of alpha^3 over the active candidates),
// M(u) = max of the density map over [0, u] (each covering particle adds
// rbf * alpha^3 <= alpha^3), and mu = 0 where no candidate is active.  A
// free flight of optical depth -ln(1 - xi) is spent at rate mu / unit, so
// empty space costs nothing and faint regions few steps; a tentative
// collision is accepted with probability D / mu.  The estimator has the
// same expectation as the reference's (delta tracking is unbiased for any
// majorant that bounds the density) but consumes random numbers
// differently, so images agree with the reference in distribution only;
// they are checked against the FP64 quadrature.  Free flights do not
// survive a segment end or the step budget (the exponential distribution
// is memoryless, so a fresh draw there is exact).
template <bool STATS, bool B>
__global__ __launch_bounds__(kSBlock, 4) void wf_step_m(RenderArgs A, WFState S, const uint32_t *__restrict__ listA,
                                                   const uint32_t *__restrict__ countA, uint32_t *__restrict__ outA,
                                                   uint32_t *__restrict__ outG, uint32_t *__restrict__ counts,
                                                   uint32_t *__restrict__ outD, uint32_t *__restrict__ countD,
                                                   uint32_t step_cap) {
  __shared__ uint32_t s_idx[kCand * kSBlock];
  __shared__ float s_tin[kCand * kSBlock];
  __shared__ float s_tout[kCand * kSBlock];
  __shared__ float s_a3[kCand * kSBlock];
  __shared__ uint8_t s_eo[kCand * kSBlock];
  __shared__ float s_dm[kDmLds], s_dmax[kDmLds];
  const uint32_t n_a = *countA;
  if (blockIdx.x * kSBlock >= n_a) return; // the whole block: no barrier is skipped
  step_cap = round_cap(step_cap, S.nslots, n_a);
  const bool dm_lds = A.dm_n <= kDmLds;
  if (dm_lds)
    for (uint32_t j = threadIdx.x; j < A.dm_n; j += kSBlock) s_dm[j] = A.dmap[j].x;
  __syncthreads();
  if (dm_lds && threadIdx.x == 0) { // prefix maxima of the texels
    float m = 0.0f;
    for (uint32_t j = 0; j < A.dm_n; ++j) {
      m = fmaxf(m, s_dm[j]);
      s_dmax[j] = m;
    }
  }
  __syncthreads();
  auto density = [&](float u) {
    if (!dm_lds) return tf_r(A.dmap, A.dm_n, u);
    uint32_t j0, j1;
    float f;
    aa_tf_coords_nf(u, A.dm_n, A.dm_nf, &j0, &j1, &f);
    return aa_lerp_tex(s_dm[j0], s_dm[j1], f);
  };
  // max of the density map over [0, u]
  auto dmax = [&](float u) {
    if (!dm_lds) return 1.0f;
    uint32_t j0, j1;
    float f;
    aa_tf_coords(u, A.dm_n, &j0, &j1, &f);
    return fmaxf(aa_lerp_tex(s_dm[j0], s_dm[j1], f), s_dmax[j0]);
  };
  Counters c = {};
  for (uint32_t base = blockIdx.x * kSBlock; base < n_a; base += gridDim.x * kSBlock) {
  const uint32_t gid = base + threadIdx.x;
  int outcome = 0; // 0 nothing, 1 keep active, 2 active + gather, 3 dense queue
  uint32_t slot = 0;
  if (gid < n_a) {
    slot = listA[gid];
    if ((S.u[slot].z >> 16) & F_DENSE) outcome = 3;
  }
  if (gid < n_a && outcome == 0) {
    Lane L;
    load_lane(S, slot, L);
    uint32_t *li = s_idx + threadIdx.x;
    float *ltin = s_tin + threadIdx.x, *ltout = s_tout + threadIdx.x, *la3 = s_a3 + threadIdx.x;
    uint8_t *leo = s_eo + threadIdx.x;
    uint32_t n = 0, act = 0, eptr = 0;
    float t_enter = AA_INF_F, t_exit = AA_INF_F, mu = 0.0f;
    auto remajor = [&]() {
      float s3 = 0.0f;
      t_exit = AA_INF_F;
      for (uint32_t m = act; m; m &= m - 1u) {
        uint32_t k = (uint32_t)__ffs(m) - 1u;
        s3 += la3[k * kSBlock];
        t_exit = fminf(t_exit, ltout[k * kSBlock]);
      }
      mu = act ? fminf(1.0f, dmax(s3) * 1.00001f) : 0.0f;
    };
    auto enter = [&](float t) {
      while (eptr < n) {
        uint32_t k = leo[eptr * kSBlock];
        if (!(ltin[k * kSBlock] <= t)) break;
        if (t < ltout[k * kSBlock]) act |= 1u << k;
        ++eptr;
      }
      t_enter = eptr < n ? ltin[leo[eptr * kSBlock] * kSBlock] : AA_INF_F;
    };
    auto leave = [&](float t) {
      for (uint32_t m = act; m; m &= m - 1u) {
        uint32_t k = (uint32_t)__ffs(m) - 1u;
        if (ltout[k * kSBlock] <= t) act &= ~(1u << k);
      }
    };
    auto setup = [&]() {
      n = (L.flags & F_DENSE) ? 0u : L.seg_n;
      const bool hd = (L.flags & F_HEAD) != 0;
      const uint32_t *iidx = hd ? S.hidx : S.cidx;
      const float *itin = hd ? S.htin : S.ctin, *itout = hd ? S.htout : S.ctout;
      const uint8_t *ieo = hd ? S.heo : S.ceo;
      const uint64_t row = hd ? L.pix : slot, stride = hd ? S.npix : S.nslots;
      for (uint32_t k = 0; k < n; ++k) {
        uint32_t p = iidx[k * stride + row];
        li[k * kSBlock] = p;
        ltin[k * kSBlock] = itin[k * stride + row];
        ltout[k * kSBlock] = itout[k * stride + row];
        leo[k * kSBlock] = ieo[k * stride + row];
        la3[k * kSBlock] = ld_pc(A, p).x;
      }
      act = 0;
      eptr = 0;
      enter(L.t);
      leave(L.t);
      remajor();
    };
    setup();
    float tau = -1.0f; // optical depth left of the current free flight
    uint32_t steps = 0;
    // A ray the dense-segment kernel left with a drawn step (majorant 1)
    // first has that tentative collision tested, with acceptance D / 1.
    bool pend = (L.flags & F_PENDING) != 0;
    L.flags &= ~F_PENDING;
    for (;;) {
      bool end = false;
      if (pend) {
        tau = 0.0f;
      } else if (tau < 0.0f) {
        if (!(L.i < A.max_depth)) {
          end = true;
        } else if (steps >= step_cap) {
          outcome = 1;
          break;
        } else {
          ++steps;
          if (STATS) c.steps++;
          float xi = aa_lcg_next(&L.trng);
          tau = -aa_log_step(1.0f - xi);
        }
      }
      if (!end) {
        const float ev = fminf(fminf(t_enter, t_exit), fminf(L.seg_end, L.tmax));
        const float dt = pend ? 0.0f : (mu > 0.0f ? tau * (A.unit / mu) : AA_INF_F);
        const float mu_c = pend ? 1.0f : mu;
        pend = false;
        if (L.t + dt < ev || dt == 0.0f) {
          // tentative collision
          L.t = L.t + dt;
          tau = -1.0f;
          ++L.i;
          const float x = L.o.x + L.t * L.d.x, y = L.o.y + L.t * L.d.y, z = L.o.z + L.t * L.d.z;
          Payload pl;
          pl.count = 0;
          pl.dens = 0.0f;
          pl.r = pl.g = pl.b = 0.0f;
          for (uint32_t m = act; m; m &= m - 1u) {
            uint32_t k = (uint32_t)__ffs(m) - 1u;
            uint32_t p = li[k * kSBlock];
            float4 qq = ld_pr(A, p);
            float4 kk = ld_pc(A, p);
            if (STATS) c.particles++;
            const float e = aa_dist2(qq.x, qq.y, qq.z, x, y, z) * A.pm[p];
            if (e > -4.5f) {
              float wa = aa_expf_mid(e) * kk.x;
              if (STATS) c.covering++;
              pl.count += 1;
              pl.dens += wa;
              pl.r += kk.y * wa;
              pl.g += kk.z * wa;
              pl.b += kk.w * wa;
              if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
            }
          }
          if (pl.count > 0) {
            float D = density(pl.dens);
            float xi2 = aa_lcg_next(&L.trng);
            if (xi2 * mu_c < D) {
              if (L.flags & F_SHADOW) {
                L.shw = 1.0f;
              } else {
                float3 col = colour_of(A, pl);
                L.alb = make_float4(col.x, col.y, col.z, 1.0f);
              }
              end = true;
            }
          }
        } else {
          // reach the next event
          if (mu > 0.0f) tau = fmaxf(0.0f, tau - (ev - L.t) * (mu / A.unit));
          L.t = fmaxf(L.t, ev);
          if (!(L.t < L.tmax)) {
            end = true; // left the clip box
          } else if (!(L.t < L.seg_end)) {
            outcome = 2; // the next segment needs a gather
            break;
          } else {
            if (L.t >= t_enter) enter(L.t);
            leave(L.t);
            remajor();
          }
        }
      }
      if (!end) continue;
      // ---- the ray ended
      const ReqArgs &R = view<B>(A, L.pix);
      if (!(L.flags & F_SHADOW) && L.alb.w > 0.0f && R.amb != 1.0f) {
        float3 so = make_float3(R.pos.x + L.d.x * L.t, R.pos.y + L.d.y * L.t, R.pos.z + L.d.z * L.t);
        float ste, stx;
        aabb_intersection(so, R.light, 0.0f, 10000.0f, A.lb, A.rt, ste, stx);
        L.o = so;
        L.d = R.light;
        L.tmax = stx;
        L.flags = F_SHADOW;
        L.shw = 0.0f;
        L.LHS = 0.0f;
        L.t = 0.0f;
        L.seg_end = -AA_INF_F;
        outcome = 2;
        break;
      }
      float vis = (L.flags & F_SHADOW) ? 1.0f - L.shw : 1.0f;
      finish_sample<STATS, B>(A, S, L, vis, c);
      outcome = next_ray<STATS, B>(A, S, L, c);
      if (outcome == 1 && !(L.flags & F_DENSE)) {
        setup();
        tau = -1.0f;
        outcome = 0;
        continue;
      }
      break;
    }
    if (outcome) save_lane(S, slot, L);
  }
  if (outcome == 1 || outcome == 2) push(outA, &counts[0], slot);
  if (outcome == 2) push(outG, &counts[1], slot);
  if (outcome == 3) push(outD, countD, slot);
  }
  flush_stats<STATS>(A, c);
}

// ---- dense segments: one wave per ray ---------------------------------------------
// Rays in crowded clusters (more than kCand candidates where the segment
// starts) are few, but a point query there finds hundreds of particles, and
// one lane walking them alone leaves the GPU waiting on a handful of rays.
// Here a whole wave runs one ray: the tracker arithmetic is uniform across
// the wave, and each point query is a level-by-level walk whose frontier
// (nodes holding the point, in ascending order) lives in LDS, 64 child
// boxes tested per pass.  The particles of the leaves holding the point are
// tested 64 at a time and the covering ones folded in ascending index
// order, so the payload equals the serial query's bit for bit.
constexpr uint32_t kFront = 2048; // frontier entries per level

template <bool STATS, bool B>
__global__ __launch_bounds__(kSBlock) void wf_dense_wave(RenderArgs A, WFState S, const uint32_t *__restrict__ listD,
                                                        const uint32_t *__restrict__ countD,
                                                        uint32_t *__restrict__ outA, uint32_t *__restrict__ outG,
                                                        uint32_t *__restrict__ counts, const uint32_t *countA,
                                                        uint32_t step_cap) {
  __shared__ uint32_t s_front[2][kFront];
  __shared__ float s_dm[kDmLds];
  const uint32_t n_d = *countD;
  if (blockIdx.x >= n_d) return; // the whole block: no barrier is skipped
  step_cap = round_cap(step_cap, S.nslots, *countA);
  load_levels(A);
  const bool dm_lds = A.dm_n <= kDmLds;
  if (dm_lds)
    for (uint32_t j = threadIdx.x; j < A.dm_n; j += kSBlock) s_dm[j] = A.dmap[j].x;
  __syncthreads();
  auto density = [&](float u) {
    if (!dm_lds) return tf_r(A.dmap, A.dm_n, u);
    uint32_t j0, j1;
    float f;
    aa_tf_coords_nf(u, A.dm_n, A.dm_nf, &j0, &j1, &f);
    return aa_lerp_tex(s_dm[j0], s_dm[j1], f);
  };
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  Counters c = {};
  const float4 *__restrict__ N = A.bvh.nodes;
  const uint32_t log2K = A.bvh.log2K, top = A.bvh.top;

  // payload at (x, y, z); false when a frontier overflows
  auto query = [&](float x, float y, float z, Payload &pl) -> bool {
    pl.count = 0;
    pl.dens = 0.0f;
    pl.r = pl.g = pl.b = 0.0f;
    const float4 *root = N + g_lvl_ofs[top];
    if (!inside(root[0], root[1], x, y, z)) return true;
    uint32_t *cur = s_front[0], *nxt = s_front[1];
    if (lane == 0) cur[0] = 0;
    uint32_t nf = 1;
    __syncthreads();
    for (uint32_t lvl = top; lvl >= 1; --lvl) {
      const float4 *cl = N + g_lvl_ofs[lvl - 1];
      const uint32_t cnt = g_lvl_cnt[lvl - 1], total = nf << log2K;
      uint32_t nn = 0;
      for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t j = base + lane;
        bool ok = false;
        uint32_t child = 0;
        if (j < total) {
          child = (cur[j >> log2K] << log2K) + (j & ((1u << log2K) - 1u));
          ok = child < cnt && inside(cl[2 * child], cl[2 * child + 1], x, y, z);
        }
        const uint64_t m = __ballot(ok);
        const uint32_t pos = nn + (uint32_t)__popcll(m & lt);
        if (ok && pos < kFront) nxt[pos] = child;
        nn += (uint32_t)__popcll(m);
      }
      if (nn > kFront) return false;
      __syncthreads();
      uint32_t *t = cur;
      cur = nxt;
      nxt = t;
      nf = nn;
      if (nf == 0) return true;
    }
    // cur: the leaves whose box holds the point, ascending
    const uint32_t ppl = A.ppl, total = nf * ppl;
    for (uint32_t base = 0; base < total; base += 64u) {
      const uint32_t j = base + lane;
      bool cov = false;
      float wa = 0.0f, wr = 0.0f, wg = 0.0f, wb = 0.0f;
      if (j < total) {
        const uint32_t p = cur[j / ppl] * ppl + j % ppl;
        if (p < A.numParticles) {
          const float4 q = ld_pr(A, p);
          const float e = aa_dist2(q.x, q.y, q.z, x, y, z) * A.pm[p];
          if (e > -4.5f) {
            const float4 k = ld_pc(A, p);
            wa = aa_expf_mid(e) * k.x;
            wr = k.y * wa;
            wg = k.z * wa;
            wb = k.w * wa;
            cov = true;
          }
        }
      }
      // fold the covering particles in ascending order
      for (uint64_t m = __ballot(cov); m; m &= m - 1ull) {
        const int l = __ffsll((unsigned long long)m) - 1;
        pl.count += 1;
        pl.dens += __shfl(wa, l, 64);
        pl.r += __shfl(wr, l, 64);
        pl.g += __shfl(wg, l, 64);
        pl.b += __shfl(wb, l, 64);
        if (!A.visAttr) pl.dens = fminf(pl.dens, 1.0f);
      }
      if (STATS && lane == 0) c.particles += min(64u, total - base);
    }
    if (STATS && lane == 0) c.covering += pl.count;
    return true;
  };

  // one ray per wave at a time; a fixed grid (a captured graph) loops
  for (uint32_t item = blockIdx.x; item < n_d; item += gridDim.x) {
  const uint32_t slot = listD[item];
  Lane L;
  load_lane(S, slot, L);
  int outcome = 0;
  uint32_t steps = 0;
  for (;;) {
    const bool pending = (L.flags & F_PENDING) != 0;
    bool end = false;
    if (!pending) {
      if (!(L.i < A.max_depth)) {
        end = true;
      } else if (steps >= step_cap) {
        outcome = 1;
        break;
      } else {
        ++steps;
        if (STATS && lane == 0) c.steps++;
        if (is_march<B>(A, L)) {
          L.t = L.t + A.unit;
        } else {
          float xi = aa_lcg_next(&L.trng);
          L.P = L.P * (1.0f - xi);
          L.t = aa_flight_pos(L.tb, L.P, A.unit);
        }
        if (L.t >= L.tmax) end = true;
        else L.flags |= F_PENDING;
      }
    }
    if (!end) {
      if (!(L.t < L.seg_end)) {
        outcome = 2;
        break;
      }
      Payload pl;
      const float x = L.o.x + L.t * L.d.x, y = L.o.y + L.t * L.d.y, z = L.o.z + L.t * L.d.z;
      if (STATS && lane == 0) c.queries++;
      if (!query(x, y, z, pl)) point_query<false>(A, x, y, z, pl, c); // frontier overflow: serial walk
      L.flags &= ~F_PENDING;
      const bool shadow = (L.flags & F_SHADOW) != 0;
      if (is_march<B>(A, L)) {
        if (pl.count > 0 && pl.dens > 0.0f) {
          float D = density(pl.dens);
          L.LHS += D;
          if (L.LHS > L.RHS && !shadow) {
            float3 col = colour_of(A, pl);
            L.alb = make_float4(col.x, col.y, col.z, 1.0f);
            end = true;
          } else if (shadow) {
            L.shw += 1.0f - aa_expf(-D);
            if (L.shw >= 1.0f) {
              L.shw = 1.0f;
              end = true;
            }
          }
        }
      } else if (pl.count > 0) {
        float D = density(pl.dens);
        float xi2 = aa_lcg_next(&L.trng);
        if (xi2 < D) {
          if (shadow) {
            L.shw = 1.0f;
          } else {
            float3 col = colour_of(A, pl);
            L.alb = make_float4(col.x, col.y, col.z, 1.0f);
          }
          end = true;
        }
      }
      if (!end) {
        ++L.i;
        if (!is_march<B>(A, L) && L.i % kFlightBlock == 0u) { // the flight anchor
          L.tb = L.t;
          L.P = 1.0f;
        }
        continue;
      }
    }
    // ---- the ray ended
    L.flags &= ~F_PENDING;
    const ReqArgs &R = view<B>(A, L.pix);
    if (!(L.flags & F_SHADOW) && L.alb.w > 0.0f && R.amb != 1.0f) {
      float3 so = make_float3(R.pos.x + L.d.x * L.t, R.pos.y + L.d.y * L.t, R.pos.z + L.d.z * L.t);
      float ste, stx;
      aabb_intersection(so, R.light, 0.0f, 10000.0f, A.lb, A.rt, ste, stx);
      L.o = so;
      L.d = R.light;
      L.tmax = stx;
      L.flags = F_SHADOW | (L.flags & F_MARCH);
      L.shw = 0.0f;
      L.LHS = 0.0f;
      L.t = is_march<B>(A, L) ? -A.unit * L.r2 : 0.0f;
      L.tb = L.t; // the shadow ray's first flight anchor
      L.P = 1.0f;
      L.seg_end = -AA_INF_F;
      outcome = 2;
      break;
    }
    if (lane == 0) {
      float vis = (L.flags & F_SHADOW) ? 1.0f - L.shw : 1.0f;
      finish_sample<STATS, B>(A, S, L, vis, c);
      outcome = next_ray<STATS, B>(A, S, L, c);
    }
    break;
  }
  if (lane == 0) {
    if (outcome) save_lane(S, slot, L);
    if (outcome == 1 || outcome == 2) push(outA, &counts[0], slot);
    if (outcome == 2) push(outG, &counts[1], slot);
  }
  }
  flush_stats<STATS>(A, c);
}

__global__ void wf_zero(uint32_t *c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

// Clears the list counts and the visiting pointer.
__global__ void wf_reset_counts(uint32_t *c) {
  if (threadIdx.x < 16) c[threadIdx.x] = 0u;
}

// ---- the view table -------------------------------------------------------------
constexpr uint32_t kViewPack = 8;
struct ViewPack {
  ReqArgs v[kViewPack];
  uint32_t first, n;
};

// Copies up to kViewPack views (passed by value) into the device table.
__global__ void wf_put_views(ReqArgs *dst, ViewPack P) {
  if (threadIdx.x < P.n) dst[P.first + threadIdx.x] = P.v[threadIdx.x];
}

// The view of every tile: the last view whose first tile is <= the tile.
__global__ __launch_bounds__(256) void wf_tile_views(const ReqArgs *__restrict__ V, uint32_t nv,
                                                     uint32_t *__restrict__ tile_req, uint32_t ntiles) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t lo = 0, hi = nv - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (V[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  tile_req[t] = lo;
}

// Latency record: per view (samples finished, clock at the last) and the
// clock at the start.
__global__ void wf_lat_start(unsigned long long *lat, uint32_t nv) {
  for (uint32_t i = threadIdx.x; i < 2 * nv; i += blockDim.x) lat[i] = 0ull;
  if (threadIdx.x == 0) lat[2 * nv] = (unsigned long long)wall_clock64();
}

} // namespace

// Launches one gather / step round on list pair `cur`.  Grids: gather and
// step from the host's counts (gb, sb), the dense kernel from its count read
// back after the step (db == 0), or all three fixed (graph capture; the
// kernels read their counts on the device and loop over them).
// Step budget per ray per round for a pass over nslots slots: the tuned
// base for queues up to 8x the lanes the GPU keeps resident (4 waves per
// SIMD), scaled by (queue / that)^(2/3) above it and rounded to the nearest
// power of two, at most 65536.  A round's step launch ends with a tail in
// which its last waves run alone; a long queue has many waves per resident
// slot, so its tail is a small part of the launch and longer rounds (fewer
// exits at the budget, fewer launches) pay off.  Measured with waves
// leaving at 16 live lanes: 1920x1080 (2.1M slots, 8x the resident lanes
// of 256 CUs) best at 8192, eight such frames in one batch (16.6M, 63x) at
// 32768 (16384: 2.5 % slower, 65536: 1 %).
static uint32_t step_budget(const aarbf_context *ctx, uint64_t nslots) {
  const double resident = 16.0 * 64.0 * (double)std::max(ctx->cus, 1); // lanes: 4 waves x 4 SIMDs per CU
  const double f = std::max(1.0, std::cbrt(std::pow((double)nslots / (8.0 * resident), 2.0)));
  const double cap = std::exp2(std::round(std::log2((double)ctx->tune_step_cap * f)));
  return (uint32_t)std::min(65536.0, std::max((double)ctx->tune_step_cap, cap));
}

template <bool B>
static int launch_round(aarbf_context *ctx, const RenderArgs &A, const WFState &S, uint32_t *const lists[4],
                        uint32_t *listD, uint32_t *listL, uint32_t *cnt, int cur, uint32_t nG, uint32_t gb,
                        uint32_t sb, uint32_t db, uint32_t budget, uint32_t *nD_out) {
  const bool st = ctx->stats;
  uint32_t *lA = lists[cur], *lG = lists[cur + 1], *nA2 = lists[2 - cur], *nG2 = lists[3 - cur];
  uint32_t *cA = cnt + (cur ? 2 : 0), *cN = cnt + (cur ? 0 : 2);
  const bool tm = ctx->timing && db == 0;
  if (tm) AARBF_HIP(ctx, hipEventRecord(ctx->tev[0], ctx->stream));
  // gathers: one lane per ray for long queues, one wave per ray for short
  // ones (each kernel returns at once on the other's queues).  Host-driven
  // rounds pass the queue length (nG) and launch the one that applies;
  // captured rounds (nG == ~0u) launch both with fixed grids.
  if (gb > 0) {
    const bool fixed = nG == ~0u;
    const uint32_t wfix = (uint32_t)std::min<uint64_t>(kWaveGather, 16ull * (uint64_t)std::max(ctx->cus, 1));
    const size_t wstack = sizeof(uint32_t) * 64 * std::max<uint32_t>(A.bvh.top, 2u);
    if (fixed || nG > kWaveGather) {
      // long queue: one lane per ray, then one wave per ray for the walks
      // that ran over the budget (a fixed grid looping over their count)
      const size_t stack_bytes = sizeof(uint32_t) * kGBlock * std::max<uint32_t>(A.bvh.top, 2u) - sizeof(uint32_t) * kGBlock;
      uint32_t *cL = cnt + 6;
      wf_zero<<<1, 64, 0, ctx->stream>>>(cL, 2); // long-walk count, queue head (cnt[7])
      // persistent lanes: enough blocks to fill the GPU, each taking rays
      // from the queue
      const uint32_t gp = std::min<uint32_t>(gb, 8u * (uint32_t)std::max(ctx->cus, 1));
      if (st)
        wf_gather<true, 0><<<gp, kGBlock, stack_bytes, ctx->stream>>>(A, S, lG, cA + 1, listL, cL, budget, cnt + 7);
      else if (A.q16)
        wf_gather<false, 2><<<gp, kGBlock, stack_bytes, ctx->stream>>>(A, S, lG, cA + 1, listL, cL, budget, cnt + 7);
      else
        wf_gather<false, 1><<<gp, kGBlock, stack_bytes, ctx->stream>>>(A, S, lG, cA + 1, listL, cL, budget, cnt + 7);
      AARBF_CHECK_LAUNCH(ctx, "wf_gather");
      if (st)
        wf_gather_wave<true><<<wfix, 64, wstack, ctx->stream>>>(A, S, listL, cL, ~0u);
      else
        wf_gather_wave<false><<<wfix, 64, wstack, ctx->stream>>>(A, S, listL, cL, ~0u);
      AARBF_CHECK_LAUNCH(ctx, "wf_gather_wave");
    }
    if (fixed || nG <= kWaveGather) {
      const uint32_t wb = fixed ? wfix : nG;
      if (st)
        wf_gather_wave<true><<<wb, 64, wstack, ctx->stream>>>(A, S, lG, cA + 1, kWaveGather);
      else
        wf_gather_wave<false><<<wb, 64, wstack, ctx->stream>>>(A, S, lG, cA + 1, kWaveGather);
      AARBF_CHECK_LAUNCH(ctx, "wf_gather_wave");
    }
  }
  wf_zero<<<1, 64, 0, ctx->stream>>>(cN, 2);
  wf_zero<<<1, 64, 0, ctx->stream>>>(cnt + 4, 1);
  if (tm) AARBF_HIP(ctx, hipEventRecord(ctx->tev[1], ctx->stream));
  const uint32_t cap = step_budget(ctx, S.nslots);
  if (ctx->render_mode == 2 && !A.blue) {
    if (st)
      wf_step_m<true, B><<<sb, kSBlock, 0, ctx->stream>>>(A, S, lA, cA, nA2, nG2, cN, listD, cnt + 4, cap);
    else
      wf_step_m<false, B><<<sb, kSBlock, 0, ctx->stream>>>(A, S, lA, cA, nA2, nG2, cN, listD, cnt + 4, cap);
  } else if (st) {
    wf_step<true, 1, B, 0><<<sb, kSBlock, 0, ctx->stream>>>(A, S, lA, cA, nA2, nG2, cN, listD, cnt + 4, cap);
  } else {
    auto step = [&](auto payload) {
      constexpr int Q = decltype(payload)::value;
      switch (ctx->tune_step_waves) {
      case 5:
        wf_step<false, 5, B, Q><<<sb, kSBlock, 0, ctx->stream>>>(A, S, lA, cA, nA2, nG2, cN, listD, cnt + 4, cap);
        break;
      case 6:
        wf_step<false, 6, B, Q><<<sb, kSBlock, 0, ctx->stream>>>(A, S, lA, cA, nA2, nG2, cN, listD, cnt + 4, cap);
        break;
      default: // 4 waves per SIMD
        wf_step<false, 4, B, Q><<<sb, kSBlock, 0, ctx->stream>>>(A, S, lA, cA, nA2, nG2, cN, listD, cnt + 4, cap);
      }
    };
    if (A.q16)
      step(std::integral_constant<int, 2>{});
    else
      step(std::integral_constant<int, 1>{});
  }
  AARBF_CHECK_LAUNCH(ctx, "wf_step");
  if (tm) AARBF_HIP(ctx, hipEventRecord(ctx->tev[2], ctx->stream));
  uint32_t nD = db;
  if (db == 0) {
    // the dense queue (filled by the step kernel) sizes its own grid
    AARBF_HIP(ctx, hipMemcpyAsync(&nD, cnt + 4, 4, hipMemcpyDeviceToHost, ctx->stream));
    AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (nD > 0) {
    // one wave per dense ray
    if (st)
      wf_dense_wave<true, B><<<nD, kSBlock, 0, ctx->stream>>>(A, S, listD, cnt + 4, nA2, nG2, cN, cA, cap);
    else
      wf_dense_wave<false, B><<<nD, kSBlock, 0, ctx->stream>>>(A, S, listD, cnt + 4, nA2, nG2, cN, cA, cap);
    AARBF_CHECK_LAUNCH(ctx, "wf_dense_wave");
  }
  if (tm) AARBF_HIP(ctx, hipEventRecord(ctx->tev[3], ctx->stream));
  if (nD_out) *nD_out = nD;
  return AARBF_OK;
}

// Fixed grids of the captured rounds: enough blocks for every CU at the
// kernels' occupancy, each looping over its share of the queue.
static void graph_grids(const aarbf_context *ctx, uint64_t nslots, uint32_t &gb, uint32_t &sb, uint32_t &db) {
  const uint64_t cus = (uint64_t)std::max(ctx->cus, 1);
  gb = (uint32_t)std::min<uint64_t>((nslots + kGBlock - 1) / kGBlock, 12 * cus);
  sb = (uint32_t)std::min<uint64_t>((nslots + kSBlock - 1) / kSBlock, 32 * cus);
  db = (uint32_t)std::min<uint64_t>(nslots, 8 * cus);
}

// The gather / step rounds of one pass, until no ray is left.  Host-driven:
// the two queue lengths are read back after each round and size the next
// round's grids.  With graph rounds R > 0: R rounds with fixed grids are
// captured once into a hipGraph (cached in the context, keyed by every
// argument the kernels see) and replayed, with one read-back per replay.
template <bool B>
static int run_rounds(aarbf_context *ctx, const RenderArgs &A, const WFState &S, uint32_t *const lists[4],
                      uint32_t *listD, uint32_t *listL, uint32_t *cnt) {
  uint32_t h[2];
  AARBF_HIP(ctx, hipMemcpyAsync(h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint32_t nA = h[0], nG = h[1];
  const int R = ctx->tune_graph_rounds;
  if (R > 0 && !ctx->timing) {
    uint32_t gb, sb, db;
    graph_grids(ctx, S.nslots, gb, sb, db);
    struct Key {
      RenderArgs A;
      WFState S;
      uint32_t *lists[4], *listD, *listL, *cnt;
      uint32_t gb, sb, db;
      int R, batch, stats, mode, waves, cap, budget;
    } key;
    memset(&key, 0, sizeof(key));
    key.A = A;
    key.S = S;
    for (int k = 0; k < 4; ++k) key.lists[k] = lists[k];
    key.listD = listD;
    key.cnt = cnt;
    key.gb = gb;
    key.sb = sb;
    key.db = db;
    key.R = R;
    key.batch = B;
    key.stats = ctx->stats;
    key.mode = ctx->render_mode;
    key.waves = ctx->tune_step_waves;
    key.cap = ctx->tune_step_cap;
    key.budget = ctx->tune_walk_budget;
    key.listL = listL;
    const std::string kb((const char *)&key, sizeof(key));
    hipGraphExec_t exec = nullptr;
    for (auto &g : ctx->graphs)
      if (g.first == kb) exec = g.second;
    if (!exec) {
      AARBF_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rc = AARBF_OK;
      for (int r = 0; r < R && rc == AARBF_OK; ++r)
        rc = launch_round<B>(ctx, A, S, lists, listD, listL, cnt, (r & 1) ? 2 : 0, ~0u, gb, sb, db,
                             (uint32_t)ctx->tune_walk_budget, nullptr);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
      if (rc != AARBF_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      AARBF_HIP(ctx, e);
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      AARBF_HIP(ctx, ei);
      if (ctx->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(ctx->graphs.front().second);
        ctx->graphs.erase(ctx->graphs.begin());
      }
      ctx->graphs.emplace_back(kb, exec);
    }
    // R is even: after a replay the next lists are pair 0 again
    while (nA > 0) {
      AARBF_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
      AARBF_HIP(ctx, hipMemcpyAsync(h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
      AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
      nA = h[0];
    }
    return AARBF_OK;
  }
  int cur = 0; // lists[cur] = A, lists[cur+1] = G; the other pair receives
  // The walk budget starts at the tuned value and grows after a round in
  // which more than 1/32 of a long gather queue ran over it (x2; x8 above
  // 1/8): in scenes whose walks are long (dense clusters) the one-wave
  // gather would otherwise take too many rays.
  uint32_t budget = (uint32_t)ctx->tune_walk_budget;
  while (nA > 0) {
    uint32_t nD = 0;
    const uint32_t gb = (nG + kGBlock - 1) / kGBlock, sb = (nA + kSBlock - 1) / kSBlock;
    int rc = launch_round<B>(ctx, A, S, lists, listD, listL, cnt, cur, nG, gb, sb, 0, budget, &nD);
    if (rc != AARBF_OK) return rc;
    uint32_t c8[8];
    AARBF_HIP(ctx, hipMemcpyAsync(c8, cnt, sizeof(c8), hipMemcpyDeviceToHost, ctx->stream));
    AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    h[0] = c8[cur ? 0 : 2];
    h[1] = c8[(cur ? 0 : 2) + 1];
    if (nG > kWaveGather && 32ull * c8[6] > (uint64_t)nG)
      budget = std::min<uint32_t>((8ull * c8[6] > (uint64_t)nG ? 8u : 2u) * budget, 65536u);
    if (ctx->timing) {
      float e[3];
      for (int k = 0; k < 3; ++k) AARBF_HIP(ctx, hipEventElapsedTime(&e[k], ctx->tev[k], ctx->tev[k + 1]));
      if (nG > 0) {
        ctx->tms[0] += e[0];
        ctx->tlaunch[0] += 1;
      }
      ctx->tms[1] += e[1];
      ctx->tms[2] += e[2];
      ctx->tlaunch[1] += 1;
      ctx->tlaunch[2] += nD > 0;
    }
    nA = h[0];
    nG = h[1];
    cur = 2 - cur;
  }
  return AARBF_OK;
}


// Host driver of the wavefront rounds over one or more views.  The views'
// 8x8 tiles are laid out one after another in a combined pixel space with a
// slot per pixel, so views of any size share every round.  Grows the slot
// state in the context's scratch pool; reads the two queue lengths back
// after each round.  Samples are rendered in passes of at most
// kPassSamples (the colour buffer holds one pass), each pass folded into
// the views' accum buffers before the next.
int render_wavefront(aarbf_context *ctx, const RenderArgs &A0, const ReqArgs *views, uint32_t nv, bool batch) {
  // captured rounds read the view from the table, so one graph serves
  // every frame of a view
  if (ctx->tune_graph_rounds > 0) batch = true;
  std::vector<ReqArgs> V(views, views + nv);
  uint64_t ntiles = 0;
  uint32_t max_spp = 0;
  for (ReqArgs &R : V) {
    R.tile0 = (uint32_t)std::min<uint64_t>(ntiles, 0xffffffffu);
    ntiles += (uint64_t)R.tiles_x * (uint64_t)((R.H + 7) / 8);
    max_spp = std::max(max_spp, R.spp);
  }
  const uint64_t npix = ntiles * 64u;
  const uint64_t nslots = npix;
  AARBF_REQUIRE(ctx, npix < (1ull << 28), "aarbf_render_rbf: image too large (2^28 pixels per call)");
  const uint64_t chunk = std::min<uint64_t>(max_spp, kPassSamples);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t o_od0 = take(16 * nslots), o_od1 = take(16 * nslots), o_pix = take(4 * nslots), o_alb = take(16 * nslots),
         o_u = take(16 * nslots), o_f = take(16 * nslots), o_seg = take(16 * nslots), o_fl = take(8 * nslots),
         o_cidx = take(4 * kCand * nslots), o_ctin = take(4 * kCand * nslots), o_ctout = take(4 * kCand * nslots),
         o_ceo = take(kCand * nslots), o_hidx = take(4 * kCand * npix), o_htin = take(4 * kCand * npix),
         o_htout = take(4 * kCand * npix), o_heo = take(kCand * npix), o_head = take(16 * npix),
         o_col = take(16 * npix * chunk), o_next = take(4 * npix), o_l0 = take(4 * nslots), o_l1 = take(4 * nslots), o_l2 = take(4 * nslots),
         o_l3 = take(4 * nslots), o_l4 = take(4 * nslots), o_l5 = take(4 * nslots), o_cnt = take(64), o_view = take(sizeof(ReqArgs) * nv),
         o_tview = take(batch ? 4 * ntiles : 4);
  char *base = (char *)scratch(ctx, off);
  if (!base) return set_error(ctx, AARBF_ERR_OUT_OF_MEMORY, "aarbf_render_rbf: wavefront state");
  WFState S;
  S.od0 = (float4 *)(base + o_od0);
  S.od1 = (float4 *)(base + o_od1);
  S.pix = (uint32_t *)(base + o_pix);
  S.alb = (float4 *)(base + o_alb);
  S.u = (uint4 *)(base + o_u);
  S.f = (float4 *)(base + o_f);
  S.seg = (float4 *)(base + o_seg);
  S.fl = (float2 *)(base + o_fl);
  S.cidx = (uint32_t *)(base + o_cidx);
  S.ctin = (float *)(base + o_ctin);
  S.ctout = (float *)(base + o_ctout);
  S.ceo = (uint8_t *)(base + o_ceo);
  S.hidx = (uint32_t *)(base + o_hidx);
  S.htin = (float *)(base + o_htin);
  S.htout = (float *)(base + o_htout);
  S.heo = (uint8_t *)(base + o_heo);
  S.head = (float4 *)(base + o_head);
  S.col = (float4 *)(base + o_col);
  S.next_s = (uint32_t *)(base + o_next);
  S.nslots = (uint32_t)nslots;
  S.npix = (uint32_t)npix;
  uint32_t *lists[4] = {(uint32_t *)(base + o_l0), (uint32_t *)(base + o_l1), (uint32_t *)(base + o_l2),
                        (uint32_t *)(base + o_l3)};
  uint32_t *listD = (uint32_t *)(base + o_l4);
  uint32_t *listL = (uint32_t *)(base + o_l5); // walks over the budget, for wf_gather_wave
  // [0..1] and [2..3]: (A, G) counts of the two list pairs; [4]: dense; [6]: long walks; [8]: visiting pointer
  uint32_t *cnt = (uint32_t *)(base + o_cnt);
  S.visit = cnt + 8;
  const bool st = ctx->stats;
  RenderArgs A = A0;
  if (batch) {
    // per-view fields live in the table; cleared here so that the kernel
    // arguments (and a captured graph's key) do not change from frame to
    // frame.  blue stays: the host picks mode 2 only when no view marches.
    A.frameID = A.accumID = A.spp = 0;
    A.showNoise = 0;
    A.W = A.H = 0;
    A.amb = 0.0f;
    A.light = A.pos = A.d00 = A.du = A.dv = make_float3(0.0f, 0.0f, 0.0f);
    A.stbn = nullptr;
    A.stbn_w = A.stbn_h = 0;
    A.image = A.accum = nullptr;
    A.v0 = ReqArgs{};
  } else {
    A.v0 = V[0];
  }
  A.req = nullptr;
  A.tile_req = nullptr;
  A.nreq = nv;
  A.lat = nullptr;
  A.min_live = (uint32_t)ctx->tune_step_min_live;
  AARBF_REQUIRE(ctx, batch || nv == 1, "render_wavefront: several views need the batch schedule");
  if (batch) {
    // the view table and the view of every tile, then the latency record
    A.req = (const ReqArgs *)(base + o_view);
    A.tile_req = (const uint32_t *)(base + o_tview);
    for (uint32_t i = 0; i < nv; i += kViewPack) {
      ViewPack P{};
      P.first = i;
      P.n = std::min(kViewPack, nv - i);
      for (uint32_t k = 0; k < P.n; ++k) P.v[k] = V[i + k];
      wf_put_views<<<1, 64, 0, ctx->stream>>>((ReqArgs *)A.req, P);
    }
    wf_tile_views<<<(uint32_t)((ntiles + 255) / 256), 256, 0, ctx->stream>>>(A.req, nv, (uint32_t *)A.tile_req,
                                                                              (uint32_t)ntiles);
    if (ctx->lat_cap < 2 * nv + 1) {
      if (ctx->lat_buf) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->lat_buf);
        ctx->lat_buf = nullptr;
        ctx->lat_cap = 0;
      }
      AARBF_HIP(ctx, hipMalloc(&ctx->lat_buf, sizeof(unsigned long long) * (2 * nv + 1)));
      ctx->lat_cap = 2 * nv + 1;
    }
    A.lat = ctx->lat_buf;
    wf_lat_start<<<1, 256, 0, ctx->stream>>>(A.lat, nv);
  }
  ctx->lat_n = batch ? nv : 0;
  AARBF_CHECK_LAUNCH(ctx, "view table");
  const uint32_t pix_blocks = (uint32_t)((npix + 255) / 256), init_blocks = (uint32_t)((nslots + 255) / 256);
  wf_head_reset<<<pix_blocks, 256, 0, ctx->stream>>>(S);
  for (uint32_t s0 = 0; s0 < max_spp; s0 += (uint32_t)chunk) {
    S.s0 = s0;
    S.s1 = (uint32_t)std::min<uint64_t>(max_spp, s0 + chunk);
    if (ctx->timing) AARBF_HIP(ctx, hipEventRecord(ctx->tev[0], ctx->stream));
    wf_reset_counts<<<1, 64, 0, ctx->stream>>>(cnt);
    if (batch) {
      wf_pass_reset<true><<<pix_blocks, 256, 0, ctx->stream>>>(A, S);
      if (st)
        wf_init<true, true><<<init_blocks, 256, 0, ctx->stream>>>(A, S, lists[0], lists[1], cnt);
      else
        wf_init<false, true><<<init_blocks, 256, 0, ctx->stream>>>(A, S, lists[0], lists[1], cnt);
    } else {
      wf_pass_reset<false><<<pix_blocks, 256, 0, ctx->stream>>>(A, S);
      if (st)
        wf_init<true, false><<<init_blocks, 256, 0, ctx->stream>>>(A, S, lists[0], lists[1], cnt);
      else
        wf_init<false, false><<<init_blocks, 256, 0, ctx->stream>>>(A, S, lists[0], lists[1], cnt);
    }
    AARBF_CHECK_LAUNCH(ctx, "wf_init");
    if (ctx->timing) {
      AARBF_HIP(ctx, hipEventRecord(ctx->tev[1], ctx->stream));
      AARBF_HIP(ctx, hipEventSynchronize(ctx->tev[1]));
      float e;
      AARBF_HIP(ctx, hipEventElapsedTime(&e, ctx->tev[0], ctx->tev[1]));
      ctx->tms[3] += e;
      ctx->tlaunch[3] += 1;
    }
    int rc = batch ? run_rounds<true>(ctx, A, S, lists, listD, listL, cnt)
                   : run_rounds<false>(ctx, A, S, lists, listD, listL, cnt);
    if (rc != AARBF_OK) return rc;
    if (ctx->timing) AARBF_HIP(ctx, hipEventRecord(ctx->tev[0], ctx->stream));
    if (batch)
      wf_fold<true><<<pix_blocks, 256, 0, ctx->stream>>>(A, S);
    else
      wf_fold<false><<<pix_blocks, 256, 0, ctx->stream>>>(A, S);
    AARBF_CHECK_LAUNCH(ctx, "wf_fold");
    if (ctx->timing) {
      AARBF_HIP(ctx, hipEventRecord(ctx->tev[1], ctx->stream));
      AARBF_HIP(ctx, hipEventSynchronize(ctx->tev[1]));
      float e;
      AARBF_HIP(ctx, hipEventElapsedTime(&e, ctx->tev[0], ctx->tev[1]));
      ctx->tms[3] += e;
    }
  }
  return AARBF_OK;
}

} // namespace aarbf
