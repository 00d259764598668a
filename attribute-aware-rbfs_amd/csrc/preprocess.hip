// preprocess.hip -- the reference's host pre-process (hostCode.cpp:183-250)
// as device passes: space-time AABB, 48-bit Hilbert codes, a stable LSD
// radix sort on the codes with the (x, y, z, w) tie order, and attribute
// normalisation.
#include <cstring>
#include <vector>

#include "aarbf_detmath.h"
#include "internal.h"

namespace aarbf {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems; // elements per radix block

__device__ inline float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-wide min/max of `comps` values per thread.  Result in lane 0 of
// wave 0's registers is written by the caller via the returned arrays.
template <int C>
__device__ void block_minmax(float (&mn)[C], float (&mx)[C]) {
  __shared__ float smn[C][kBlock / 64], smx[C][kBlock / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c = 0; c < C; ++c) {
    mn[c] = wave_min(mn[c]);
    mx[c] = wave_max(mx[c]);
    if (lane == 0) {
      smn[c][w] = mn[c];
      smx[c][w] = mx[c];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < C; ++c)
      for (int k = 1; k < kBlock / 64; ++k) {
        mn[c] = fminf(mn[c], smn[c][k]);
        mx[c] = fmaxf(mx[c], smx[c][k]);
      }
}

// Partial min/max of the xyz (C=3, offset 0) or w (C=1, offset 3) columns;
// partials[block] = {min..., max...}; nan_flag set if any value is NaN.
template <int C, int OFS>
__global__ __launch_bounds__(kBlock) void minmax_partial(const float4 *__restrict__ p, uint64_t n,
                                                         float *__restrict__ partials,
                                                         uint32_t *__restrict__ nan_flag) {
  float mn[C], mx[C];
  for (int c = 0; c < C; ++c) {
    mn[c] = AA_INF_F;
    mx[c] = -AA_INF_F;
  }
  bool nan = false;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
    float4 q = p[i];
    float v[4] = {q.x, q.y, q.z, q.w};
    for (int c = 0; c < C; ++c) {
      float x = v[OFS + c];
      nan |= (x != x);
      mn[c] = fminf(mn[c], x);
      mx[c] = fmaxf(mx[c], x);
    }
  }
  if (__any(nan) && (threadIdx.x & 63) == 0) atomicOr(nan_flag, 1u);
  block_minmax<C>(mn, mx);
  if (threadIdx.x == 0)
    for (int c = 0; c < C; ++c) {
      partials[2 * C * blockIdx.x + c] = mn[c];
      partials[2 * C * blockIdx.x + C + c] = mx[c];
    }
}

template <int C>
__global__ __launch_bounds__(kBlock) void minmax_final(const float *__restrict__ partials, int nparts,
                                                       float *__restrict__ out) {
  float mn[C], mx[C];
  for (int c = 0; c < C; ++c) {
    mn[c] = AA_INF_F;
    mx[c] = -AA_INF_F;
  }
  for (int b = threadIdx.x; b < nparts; b += kBlock)
    for (int c = 0; c < C; ++c) {
      mn[c] = fminf(mn[c], partials[2 * C * b + c]);
      mx[c] = fmaxf(mx[c], partials[2 * C * b + C + c]);
    }
  block_minmax<C>(mn, mx);
  if (threadIdx.x == 0)
    for (int c = 0; c < C; ++c) {
      out[c] = mn[c];
      out[C + c] = mx[c];
    }
}

// hostCode.cpp:200-207: ((p - min) / (max - min)) * 65536, truncated to an
// integer.  NaN or negative quotients (zero-extent axis) give 0.
__device__ inline uint32_t quantise16(float p, float lo, float hi) {
  float t = (p - lo) / (hi - lo);
  t = t * 65536.0f;
  if (!(t >= 0.0f)) return 0u;
  return (uint32_t)t; // t <= 65536
}

// Hilbert index of hilbert_c2i(3, 16, coord) (hilbert.h:181-225), computed
// on the interleaved word: bit 3b+d of `w` is bit b of axis d.  The three
// coordinates are first packed by shift-and-OR as the reference does, which
// is what makes a 17-bit value 65536 spill into the next axis.
__device__ inline uint64_t hilbert_index(uint32_t c0, uint32_t c1, uint32_t c2) {
  uint64_t packed = ((((uint64_t)c2 << 16) | c1) << 16) | c0;
  uint32_t a0 = (uint32_t)(packed & 0xffff), a1 = (uint32_t)((packed >> 16) & 0xffff),
           a2 = (uint32_t)((packed >> 32) & 0xffff);
  // interleave: spread 16 bits to every third bit
  auto spread = [](uint64_t v) {
    v &= 0xffffull;
    v = (v | (v << 16)) & 0x0000ff0000ffull;
    v = (v | (v << 8)) & 0x00f00f00f00full;
    v = (v | (v << 4)) & 0x0c30c30c30c3ull;
    v = (v | (v << 2)) & 0x249249249249ull;
    return v;
  };
  uint64_t w = spread(a0) | (spread(a1) << 1) | (spread(a2) << 2);
  w ^= w >> 3;
  uint64_t index = 0;
  uint32_t rot = 0, flip = 0;
#pragma unroll
  for (int b = 15; b >= 0; --b) {
    uint32_t g = (uint32_t)(w >> (3 * b)) & 7u;
    uint32_t x = flip ^ g;
    uint32_t h = ((x >> rot) | (x << (3 - rot))) & 7u;
    index = (index << 3) | h;
    flip = 1u << rot;
    // advance by 1 + position of the lowest set bit of h's two low bits
    uint32_t step = (h & 1u) ? 2u : ((h & 2u) ? 3u : 1u);
    rot += step;
    rot = rot >= 3u ? rot - 3u : rot;
    rot = rot >= 3u ? rot - 3u : rot;
  }
  index ^= 0x124924924924ull;
  index ^= index >> 1;
  index ^= index >> 2;
  index ^= index >> 4;
  index ^= index >> 8;
  index ^= index >> 16;
  index ^= index >> 32;
  return index;
}

__global__ __launch_bounds__(kBlock) void hilbert_codes_kernel(const float4 *__restrict__ p, uint64_t n,
                                                               float lx, float ly, float lz, float hx,
                                                               float hy, float hz,
                                                               uint64_t *__restrict__ codes) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 q = p[i];
  codes[i] = hilbert_index(quantise16(q.x, lx, hx), quantise16(q.y, ly, hy), quantise16(q.z, lz, hz));
}

// ---- stable LSD radix sort of (uint64 key, uint32 value) ------------------

__global__ __launch_bounds__(kBlock) void radix_hist(const uint64_t *__restrict__ keys, uint64_t n,
                                                     uint32_t shift, uint32_t nb,
                                                     uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * kTile;
  for (int r = 0; r < kItems; ++r) {
    uint64_t i = base + (uint64_t)r * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void radix_scatter(const uint64_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ vals, uint64_t n,
                                                        uint32_t shift, uint32_t nb,
                                                        const uint32_t *__restrict__ offsets,
                                                        uint64_t *__restrict__ okeys,
                                                        uint32_t *__restrict__ ovals) {
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_cnt[kBlock / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  s_base[tid] = offsets[(uint64_t)tid * nb + blockIdx.x];
  for (int k = 0; k < kBlock / 64; ++k) s_cnt[k][tid] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t base = (uint64_t)blockIdx.x * kTile;
  for (int r = 0; r < kItems; ++r) {
    uint64_t i = base + (uint64_t)r * kBlock + tid;
    bool valid = i < n;
    uint64_t key = valid ? keys[i] : 0;
    uint32_t val = valid ? vals[i] : 0;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t m = __ballot(valid && bit);
      peers &= bit ? m : ~m;
    }
    uint32_t rank = __popcll(peers & lt);
    if (valid && rank == 0) s_cnt[wv][d] = __popcll(peers);
    __syncthreads();
    uint32_t total = 0;
    {
      uint32_t acc = 0;
      for (int k = 0; k < kBlock / 64; ++k) {
        uint32_t c = s_cnt[k][tid];
        s_cnt[k][tid] = acc;
        acc += c;
      }
      total = acc;
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[d] + s_cnt[wv][d] + rank;
      okeys[pos] = key;
      ovals[pos] = val;
    }
    __syncthreads();
    s_base[tid] += total;
    for (int k = 0; k < kBlock / 64; ++k) s_cnt[k][tid] = 0;
    __syncthreads();
  }
}

// Exclusive scan of uint32 (three phases: tile sums, scan of sums, tiles).
constexpr int kScanTile = 2048;

__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t ws[kBlock / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int k = 0; k < kBlock / 64; ++k) {
    if (k < w) off += ws[k];
    tot += ws[k];
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

__global__ __launch_bounds__(kBlock) void scan_tile_sums(const uint32_t *__restrict__ in, uint64_t n,
                                                         uint32_t *__restrict__ sums) {
  uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (int k = 0; k < kScanTile / kBlock; ++k) {
    uint64_t i = base + (uint64_t)threadIdx.x * (kScanTile / kBlock) + k;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void scan_sums_single(uint32_t *__restrict__ sums, uint32_t m) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < m; base += kBlock) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < m ? sums[i] : 0;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, &tot);
    if (i < m) sums[i] = carry + ex;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void scan_tiles(const uint32_t *__restrict__ in, uint64_t n,
                                                     const uint32_t *__restrict__ sums,
                                                     uint32_t *__restrict__ out) {
  constexpr int per = kScanTile / kBlock;
  uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * per;
  uint32_t v[per];
  uint32_t s = 0;
  for (int k = 0; k < per; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_exclusive_scan(s, &tot) + sums[blockIdx.x];
  for (int k = 0; k < per; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

int exclusive_scan(aarbf_context *ctx, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *sums) {
  uint32_t tiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
  scan_tile_sums<<<tiles, kBlock, 0, ctx->stream>>>(in, n, sums);
  scan_sums_single<<<1, kBlock, 0, ctx->stream>>>(sums, tiles);
  scan_tiles<<<tiles, kBlock, 0, ctx->stream>>>(in, n, sums, out);
  AARBF_CHECK_LAUNCH(ctx, "exclusive_scan");
  return AARBF_OK;
}

__global__ void iota_kernel(uint32_t *v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

// Monotone integer key of a float ('<' order, -0 before +0).
__device__ inline uint32_t float_key(float f) {
  uint32_t u = aa_f2u(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void gather_particles(const float4 *__restrict__ src, const uint32_t *__restrict__ perm,
                                 uint64_t n, float4 *__restrict__ dst) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ perm, uint64_t n,
                           uint32_t *__restrict__ dst) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void component_keys(const float4 *__restrict__ p, uint64_t n, int comp, uint64_t *__restrict__ k) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 q = p[i];
  float v = comp == 0 ? q.x : (comp == 1 ? q.y : (comp == 2 ? q.z : q.w));
  k[i] = float_key(v);
}

__global__ void gather_u64(const uint64_t *__restrict__ src, const uint32_t *__restrict__ perm, uint64_t n,
                           uint64_t *__restrict__ dst) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

constexpr uint32_t kShortRun = 64;

__device__ inline bool item_less(const float4 &a, uint32_t pa, const float4 &b, uint32_t pb) {
  uint32_t ka[4] = {float_key(a.x), float_key(a.y), float_key(a.z), float_key(a.w)};
  uint32_t kb[4] = {float_key(b.x), float_key(b.y), float_key(b.z), float_key(b.w)};
  for (int j = 0; j < 4; ++j)
    if (ka[j] != kb[j]) return ka[j] < kb[j];
  return pa < pb;
}

// Orders each run of equal codes by (x, y, z, w, original index) with an
// insertion sort; runs longer than kShortRun raise *long_run instead.
__global__ void fix_runs(const uint64_t *__restrict__ codes, uint64_t n, float4 *__restrict__ p,
                         uint32_t *__restrict__ perm, uint32_t *__restrict__ long_run) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  if (i > 0 && codes[i - 1] == codes[i]) return;
  if (codes[i + 1] != codes[i]) return;
  uint64_t end = i + 1;
  while (end < n && codes[end] == codes[i]) {
    ++end;
    if (end - i > kShortRun) {
      atomicOr(long_run, 1u);
      return;
    }
  }
  for (uint64_t a = i + 1; a < end; ++a) {
    float4 v = p[a];
    uint32_t pv = perm[a];
    uint64_t b = a;
    while (b > i && item_less(v, pv, p[b - 1], perm[b - 1])) {
      p[b] = p[b - 1];
      perm[b] = perm[b - 1];
      --b;
    }
    p[b] = v;
    perm[b] = pv;
  }
}

__global__ void normalize_kernel(float4 *__restrict__ p, uint64_t n, float lo, float den) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i].w = (p[i].w - lo) / den;
}

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Sorts (keys, vals) by the bits [0, bits) of keys with stable 8-bit passes.
// keys/vals are overwritten; tmp buffers must hold n entries.
int radix_sort(aarbf_context *ctx, uint64_t *keys, uint32_t *vals, uint64_t n, int bits, uint64_t *tkeys,
               uint32_t *tvals, uint32_t *hist, uint32_t *offs, uint32_t *sums) {
  uint32_t nb = blocks_for(n, kTile);
  for (int shift = 0; shift < bits; shift += 8) {
    radix_hist<<<nb, kBlock, 0, ctx->stream>>>(keys, n, shift, nb, hist);
    int st = exclusive_scan(ctx, hist, offs, (uint64_t)nb * 256, sums);
    if (st) return st;
    radix_scatter<<<nb, kBlock, 0, ctx->stream>>>(keys, vals, n, shift, nb, offs, tkeys, tvals);
    AARBF_CHECK_LAUNCH(ctx, "radix_scatter");
    std::swap(keys, tkeys);
    std::swap(vals, tvals);
  }
  if (((bits + 7) / 8) & 1) { // odd number of passes: result is in the temporaries
    AARBF_HIP(ctx, hipMemcpyAsync(tkeys, keys, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    AARBF_HIP(ctx, hipMemcpyAsync(tvals, vals, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return AARBF_OK;
}

} // namespace
} // namespace aarbf

using namespace aarbf;

extern "C" {

static int minmax_host(aarbf_context *ctx, const float4 *p, uint64_t n, int which, float *out, bool *nan) {
  uint32_t parts = (uint32_t)std::min<uint64_t>(2048, (n + kBlock - 1) / kBlock);
  if (parts == 0) parts = 1;
  char *s = (char *)scratch(ctx, parts * 6 * sizeof(float) + 64);
  if (!s) return set_error(ctx, AARBF_ERR_OUT_OF_MEMORY, "scratch");
  float *partials = (float *)(s + 64);
  uint32_t *flag = (uint32_t *)s;
  float *res = (float *)(s + 8);
  AARBF_HIP(ctx, hipMemsetAsync(flag, 0, 4, ctx->stream));
  if (which == 0) {
    minmax_partial<3, 0><<<parts, kBlock, 0, ctx->stream>>>(p, n, partials, flag);
    minmax_final<3><<<1, kBlock, 0, ctx->stream>>>(partials, parts, res);
  } else {
    minmax_partial<1, 3><<<parts, kBlock, 0, ctx->stream>>>(p, n, partials, flag);
    minmax_final<1><<<1, kBlock, 0, ctx->stream>>>(partials, parts, res);
  }
  AARBF_CHECK_LAUNCH(ctx, "minmax");
  float host[8];
  AARBF_HIP(ctx, hipMemcpyAsync(host, s, 32, hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint32_t f;
  std::memcpy(&f, host, 4);
  *nan = f != 0;
  int C = which == 0 ? 3 : 1;
  for (int c = 0; c < 2 * C; ++c) out[c] = host[2 + c];
  return AARBF_OK;
}

// hostCode.cpp:183-193.
int aarbf_spacetime_aabb(aarbf_context *ctx, uint64_t particles, uint64_t n, float aabb[6]) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, aabb && (n == 0 || particles), "aarbf_spacetime_aabb: null argument");
  if (n == 0) return AARBF_OK;
  hipSetDevice(ctx->device);
  float r[6];
  bool nan = false;
  int st = minmax_host(ctx, (const float4 *)particles, n, 0, r, &nan);
  if (st) return st;
  AARBF_REQUIRE(ctx, !nan, "aarbf_spacetime_aabb: NaN particle coordinate");
  for (int d = 0; d < 3; ++d) {
    aabb[d] = (aabb[d] < r[d]) ? aabb[d] : r[d];             // linalg::min
    aabb[3 + d] = (aabb[3 + d] < r[3 + d]) ? r[3 + d] : aabb[3 + d]; // linalg::max
  }
  return AARBF_OK;
}

// hostCode.cpp:196-209.
int aarbf_hilbert_codes(aarbf_context *ctx, uint64_t particles, uint64_t n, const float aabb[6], uint64_t codes) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, aabb && (n == 0 || (particles && codes)), "aarbf_hilbert_codes: null argument");
  if (n == 0) return AARBF_OK;
  hipSetDevice(ctx->device);
  hilbert_codes_kernel<<<blocks_for(n, kBlock), kBlock, 0, ctx->stream>>>(
      (const float4 *)particles, n, aabb[0], aabb[1], aabb[2], aabb[3], aabb[4], aabb[5], (uint64_t *)codes);
  AARBF_CHECK_LAUNCH(ctx, "hilbert_codes_kernel");
  return AARBF_OK;
}

// hostCode.cpp:212-221: sort by pair<code, float4>.
int aarbf_sort_particles(aarbf_context *ctx, uint64_t particles, uint64_t codes, uint64_t n, uint64_t perm_out) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, n == 0 || (particles && codes), "aarbf_sort_particles: null argument");
  AARBF_REQUIRE(ctx, n < (1ull << 32), "aarbf_sort_particles: more than 2^32-1 particles");
  if (n == 0) return AARBF_OK;
  hipSetDevice(ctx->device);
  uint32_t nb = blocks_for(n, kTile);
  uint64_t hist_n = (uint64_t)nb * 256;
  uint32_t scan_tiles_n = blocks_for(hist_n, kScanTile);
  // scratch layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t o_keys = take(n * 8), o_tkeys = take(n * 8), o_vals = take(n * 4), o_tvals = take(n * 4),
         o_hist = take(hist_n * 4), o_offs = take(hist_n * 4), o_sums = take(scan_tiles_n * 4 + 4),
         o_p = take(n * 16), o_flag = take(4), o_perm2 = take(n * 4);
  char *s = (char *)scratch(ctx, off);
  if (!s) return set_error(ctx, AARBF_ERR_OUT_OF_MEMORY, "aarbf_sort_particles: scratch");
  uint64_t *keys = (uint64_t *)(s + o_keys), *tkeys = (uint64_t *)(s + o_tkeys);
  uint32_t *vals = (uint32_t *)(s + o_vals), *tvals = (uint32_t *)(s + o_tvals);
  uint32_t *hist = (uint32_t *)(s + o_hist), *offs = (uint32_t *)(s + o_offs), *sums = (uint32_t *)(s + o_sums);
  float4 *ptmp = (float4 *)(s + o_p);
  uint32_t *flag = (uint32_t *)(s + o_flag), *perm2 = (uint32_t *)(s + o_perm2);
  float4 *P = (float4 *)particles;
  uint64_t *C = (uint64_t *)codes;

  AARBF_HIP(ctx, hipMemcpyAsync(keys, C, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  iota_kernel<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(vals, n);
  int st = radix_sort(ctx, keys, vals, n, 48, tkeys, tvals, hist, offs, sums);
  if (st) return st;
  // keys/vals now sorted by code; vals = original index
  gather_particles<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(P, vals, n, ptmp);
  AARBF_HIP(ctx, hipMemsetAsync(flag, 0, 4, ctx->stream));
  fix_runs<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(keys, n, ptmp, vals, flag);
  AARBF_CHECK_LAUNCH(ctx, "fix_runs");
  uint32_t long_run = 0;
  AARBF_HIP(ctx, hipMemcpyAsync(&long_run, flag, 4, hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (long_run) {
    // Full stable LSD sort by (code, x, y, z, w) starting from the original
    // order: w, z, y, x keys (32 bits each), then the code (48 bits).
    iota_kernel<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(vals, n);
    for (int comp = 3; comp >= 0; --comp) {
      // keys for the current order: component of P[vals[i]]
      gather_particles<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(P, vals, n, ptmp);
      component_keys<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(ptmp, n, comp, keys);
      st = radix_sort(ctx, keys, vals, n, 32, tkeys, tvals, hist, offs, sums);
      if (st) return st;
    }
    gather_u64<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(C, vals, n, keys);
    st = radix_sort(ctx, keys, vals, n, 48, tkeys, tvals, hist, offs, sums);
    if (st) return st;
    gather_particles<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(P, vals, n, ptmp);
  }
  AARBF_HIP(ctx, hipMemcpyAsync(P, ptmp, n * 16, hipMemcpyDeviceToDevice, ctx->stream));
  AARBF_HIP(ctx, hipMemcpyAsync(C, keys, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  if (perm_out) AARBF_HIP(ctx, hipMemcpyAsync((void *)perm_out, vals, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  (void)perm2;
  AARBF_CHECK_LAUNCH(ctx, "aarbf_sort_particles");
  return AARBF_OK;
}

// hostCode.cpp:227-240.
int aarbf_attribute_range(aarbf_context *ctx, uint64_t particles, uint64_t n, float range[2]) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, range && (n == 0 || particles), "aarbf_attribute_range: null argument");
  if (n == 0) return AARBF_OK;
  hipSetDevice(ctx->device);
  float r[2];
  bool nan = false;
  int st = minmax_host(ctx, (const float4 *)particles, n, 1, r, &nan);
  if (st) return st;
  AARBF_REQUIRE(ctx, !nan, "aarbf_attribute_range: NaN attribute");
  range[0] = (range[0] < r[0]) ? range[0] : r[0]; // std::min(w, min)
  range[1] = (range[1] < r[1]) ? r[1] : range[1]; // std::max(w, max)
  return AARBF_OK;
}

// hostCode.cpp:243-250.
int aarbf_normalize_attributes(aarbf_context *ctx, uint64_t particles, uint64_t n, const float range[2]) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, range && (n == 0 || particles), "aarbf_normalize_attributes: null argument");
  if (n == 0 || !(range[1] > range[0])) return AARBF_OK;
  hipSetDevice(ctx->device);
  normalize_kernel<<<blocks_for(n, 256), 256, 0, ctx->stream>>>((float4 *)particles, n, range[0],
                                                                 range[1] - range[0]);
  AARBF_CHECK_LAUNCH(ctx, "normalize_kernel");
  return AARBF_OK;
}

} // extern "C"
