// taa.hip -- CompositeGui's temporal anti-aliasing resolve
// (deviceCodeCommon.hlsl:27-43,67-115) plus the history copy the host does
// after presenting (hostCode.cpp:829-830).
//
// One thread per pixel on 16x16 tiles.  The 3x3 neighbourhood is read at
// exact texel centres, so the reference's linear sampler returns texels
// unchanged; borders clamp to the edge.  Each pixel reads only its own
// history texel, so the history copy is done in the same pass.
#include "aarbf_detmath.h"
#include "internal.h"

namespace aarbf {
namespace {

__device__ inline float3 rgb2ycocg(float3 c) {
  float co = c.x - c.z;
  float t = c.z + co / 2.0f;
  float cg = c.y - t;
  float y = t + cg / 2.0f;
  return make_float3(y, co, cg);
}

__device__ inline float3 ycocg2rgb(float3 c) {
  float t = c.x - c.z / 2.0f;
  float g = c.z + t;
  float b = t - c.y / 2.0f;
  float r = c.y + b;
  return make_float3(r, g, b);
}

__device__ inline float3 fmin3(float3 a, float3 b) { return make_float3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ inline float3 fmax3(float3 a, float3 b) { return make_float3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }

// sRGB transfer function of the GUI layer's R8G8B8A8_SRGB texels, with the
// renderer's deterministic power function (host restatements get the same
// bits).
__device__ inline float srgb_decode(uint32_t v) {
  float c = (float)v / 255.0f;
  return c <= 0.04045f ? c / 12.92f : aa_powf((c + 0.055f) / 1.055f, 2.4f);
}

__device__ inline uint32_t to8(float v) {
  v = fminf(fmaxf(v, 0.0f), 1.0f); // NaN -> 0
  return (uint32_t)(v * 255.0f + 0.5f);
}

__global__ __launch_bounds__(256) void taa_kernel(int W, int H, int disable, const float4 *__restrict__ image,
                                                  const uint32_t *__restrict__ gui, uint32_t *__restrict__ frame,
                                                  float4 *__restrict__ taa, float4 *__restrict__ prev) {
  int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= W || py >= H) return;
  uint64_t ofs = (uint64_t)px + (uint64_t)W * py;
  float4 center = image[ofs];
  float4 out;
  if (!disable) {
    float3 mn = rgb2ycocg(make_float3(center.x, center.y, center.z)), mx = mn;
    for (int iy = -1; iy <= 1; ++iy)
      for (int ix = -1; ix <= 1; ++ix) {
        if (ix == 0 && iy == 0) continue;
        int qx = min(max(px + ix, 0), W - 1), qy = min(max(py + iy, 0), H - 1);
        float4 c = image[(uint64_t)qx + (uint64_t)W * qy];
        float3 y = rgb2ycocg(make_float3(c.x, c.y, c.z));
        mn = fmin3(mn, y);
        mx = fmax3(mx, y);
      }
    float4 old = prev[ofs];
    float3 o = ycocg2rgb(fmax3(rgb2ycocg(make_float3(old.x, old.y, old.z)), mn));
    o = ycocg2rgb(fmin3(rgb2ycocg(o), mx));
    old.x = o.x;
    old.y = o.y;
    old.z = o.z;
    out.x = old.x + 0.2f * (center.x - old.x);
    out.y = old.y + 0.2f * (center.y - old.y);
    out.z = old.z + 0.2f * (center.z - old.z);
    out.w = old.w + 0.2f * (center.w - old.w);
    taa[ofs] = out;
    prev[ofs] = out;
  } else {
    out = center;
    prev[ofs] = taa[ofs];
  }
  if (frame) {
    float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (gui) {
      uint32_t v = gui[ofs];
      g = make_float4(srgb_decode(v & 255u), srgb_decode((v >> 8) & 255u), srgb_decode((v >> 16) & 255u),
                      (float)(v >> 24) / 255.0f);
    }
    // over(gui, pixel) (sharedCode.h:142-149)
    float ra = g.w + out.w * (1.0f - g.w);
    float4 c = g;
    if (ra != 0.0f) {
      c.x = (g.x * g.w + (out.x * out.w) * (1.0f - g.w)) / ra;
      c.y = (g.y * g.w + (out.y * out.w) * (1.0f - g.w)) / ra;
      c.z = (g.z * g.w + (out.z * out.w) * (1.0f - g.w)) / ra;
      c.w = ra;
    }
    frame[ofs] = to8(c.z) | (to8(c.y) << 8) | (to8(c.x) << 16) | (to8(c.w) << 24);
  }
}

} // namespace
} // namespace aarbf

using namespace aarbf;

extern "C" int aarbf_taa_resolve(aarbf_context *ctx, const aarbf_taa_constants *tc) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, tc, "aarbf_taa_resolve: null constants");
  int W = tc->fbSize.x, H = tc->fbSize.y;
  AARBF_REQUIRE(ctx, W > 0 && H > 0, "aarbf_taa_resolve: empty framebuffer");
  AARBF_REQUIRE(ctx, tc->imageTexture && tc->taaBuffer && tc->taaPrevBuffer, "aarbf_taa_resolve: null buffer");
  hipSetDevice(ctx->device);
  dim3 grid((W + 15) / 16, (H + 15) / 16);
  taa_kernel<<<grid, 256, 0, ctx->stream>>>(W, H, tc->disableTAA ? 1 : 0, (const float4 *)tc->imageTexture,
                                           (const uint32_t *)tc->guiTexture, (uint32_t *)tc->frameBuffer,
                                           (float4 *)tc->taaBuffer, (float4 *)tc->taaPrevBuffer);
  AARBF_CHECK_LAUNCH(ctx, "taa_kernel");
  return AARBF_OK;
}
